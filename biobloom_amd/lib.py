"""ctypes binding of libbbt_hip (include/bbt_hip.h).

The library is required: importing a classifier without the built library
raises, there is no host fallback.  Device buffers are passed as integer
device pointers (e.g. torch tensors' data_ptr()) and streams as integer
hipStream_t handles.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libbbt_hip.so")

SIMPLE, LENGTH, HARMONIC, BINOMIAL = 0, 1, 2, 3
STD, ORDERED, BESTHIT, SCORES = 0, 1, 2, 3
PATH_PLAIN, PATH_BATCHED = 0, 1
READS_ASCII, READS_2BIT = 0, 1
FLAG_TIME_WALK = 1
FLAG_GRAPH = 2


class BbtError(RuntimeError):
    pass


class FilterInfo(C.Structure):
    _fields_ = [("m", C.c_uint64), ("h", C.c_uint32), ("k", C.c_uint32),
                ("fpr", C.c_double), ("popcount", C.c_uint64),
                ("desired_fpr", C.c_double)]


class EvalParams(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("scorer", C.c_uint32), ("threshold", C.c_double),
                ("streak", C.c_uint32), ("paired", C.c_uint32),
                ("inclusive", C.c_uint32), ("path", C.c_uint32), ("flags", C.c_uint32),
                ("wave_min_len", C.c_uint32), ("dust", C.c_uint32), ("dust_t", C.c_uint32),
                ("dust_w", C.c_uint32), ("reserved", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("wave_min_len", C.c_uint32), ("refill_min", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("probe_ahead", C.c_uint32)]


class Reads(C.Structure):
    _fields_ = [("seq", C.c_void_p), ("seq_bytes", C.c_uint64), ("off", C.c_void_p),
                ("nreads", C.c_uint64), ("max_len", C.c_uint32),
                ("format", C.c_uint32), ("order", C.c_void_p)]


class EvalOut(C.Structure):
    _fields_ = [("hits", C.c_void_p), ("scores", C.c_void_p), ("cls", C.c_void_p),
                ("summary", C.c_void_p), ("counters", C.c_void_p)]


class MibfInfo(C.Structure):
    _fields_ = [("m", C.c_uint64), ("pop", C.c_uint64), ("pop_saturated", C.c_uint64),
                ("nids", C.c_uint32), ("nseeds", C.c_uint32), ("span", C.c_uint32),
                ("kmer_mode", C.c_uint32), ("direct", C.c_uint32), ("reserved", C.c_uint32)]


class MibfParams(C.Structure):
    _fields_ = [("score", C.c_double), ("multi_thresh", C.c_double),
                ("allowed_miss", C.c_uint32), ("min_nonsat", C.c_uint32),
                ("max_hits", C.c_uint32), ("flags", C.c_uint32)]


class MibfOut(C.Structure):
    _fields_ = [("nhits", C.c_void_p), ("hits", C.c_void_p), ("sat", C.c_void_p),
                ("counters", C.c_void_p)]


SERVE_BATCHED, SERVE_PLAIN = 0, 1


class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_uint32), ("mode", C.c_uint32), ("device_bytes", C.c_uint64),
                ("max_wave_reads", C.c_uint64), ("max_wave_bases", C.c_uint64),
                ("host_bytes", C.c_uint64)]


class RequestStats(C.Structure):
    _fields_ = [("nreads", C.c_uint64), ("wave", C.c_uint64), ("submit_us", C.c_double),
                ("admit_us", C.c_double), ("done_us", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("waves", C.c_uint64), ("requests_done", C.c_uint64),
                ("reads_done", C.c_uint64), ("max_waves_in_flight", C.c_uint64),
                ("arena_bytes", C.c_uint64), ("arena_high_water", C.c_uint64),
                ("mean_wave_reads", C.c_double), ("mean_wave_requests", C.c_double),
                ("submit_ms", C.c_double), ("pack_ms", C.c_double), ("device_ms", C.c_double),
                ("probes", C.c_uint64), ("lookups", C.c_uint64), ("lookups_issued", C.c_uint64)]


# exported symbols and their signatures: (restype, argtypes)
_SIGS = {
    "bbt_last_error": (C.c_char_p, []),
    "bbt_filters_load": (C.c_int32, [C.POINTER(C.c_char_p), C.c_uint32,
                                     C.POINTER(C.c_void_p)]),
    "bbt_filters_create": (C.c_int32, [C.c_uint32, C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_void_p)]),
    "bbt_filters_free": (C.c_int32, [C.c_void_p]),
    "bbt_filterset_size": (C.c_uint32, [C.c_void_p]),
    "bbt_filter_info_get": (C.c_int32, [C.c_void_p, C.c_uint32, C.POINTER(FilterInfo)]),
    "bbt_filter_id": (C.c_char_p, [C.c_void_p, C.c_uint32]),
    "bbt_filter_set_id": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_char_p]),
    "bbt_filter_device_bits": (C.c_void_p, [C.c_void_p, C.c_uint32]),
    "bbt_filter_upload": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64]),
    "bbt_filter_download": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64]),
    "bbt_filter_insert": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64,
                                      C.POINTER(C.c_uint64), C.c_void_p]),
    "bbt_filter_insert_ordered": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64,
                                              C.POINTER(C.c_uint64), C.c_void_p]),
    "bbt_filter_insert_subtract": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                               C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64),
                                               C.c_void_p]),
    "bbt_eval_subtract": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                      C.POINTER(Reads), C.c_uint32, C.c_double, C.c_uint32,
                                      C.c_void_p, C.c_void_p]),
    "bbt_filter_refresh": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "bbt_filter_store": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_char_p, C.c_double,
                                     C.c_uint64, C.c_uint64, C.c_uint64, C.c_char_p]),
    "bbt_calc_optimal_size": (C.c_uint64, [C.c_uint64, C.c_double, C.c_uint32]),
    "bbt_calc_optimal_hashnum": (C.c_uint32, [C.c_double]),
    "bbt_eval_batch": (C.c_int32, [C.c_void_p, C.POINTER(Reads), C.POINTER(EvalParams),
                                   C.POINTER(EvalOut), C.c_void_p]),
    "bbt_last_walk_ms": (C.c_double, [C.c_void_p]),
    "bbt_last_walk_kernel": (C.c_char_p, [C.c_void_p]),
    "bbt_measure_gather": (C.c_int32, [C.c_uint64, C.c_uint32, C.c_void_p,
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bbt_eval_batch_noalloc": (C.c_int32, [C.c_void_p, C.POINTER(Reads),
                                           C.POINTER(EvalParams), C.POINTER(EvalOut),
                                           C.c_void_p]),
    "bbt_workspace_create": (C.c_int32, [C.POINTER(C.c_void_p)]),
    "bbt_workspace_free": (C.c_int32, [C.c_void_p]),
    "bbt_eval_batch_ws": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(Reads),
                                      C.POINTER(EvalParams), C.POINTER(EvalOut), C.c_void_p]),
    "bbt_server_create": (C.c_int32, [C.c_void_p, C.POINTER(EvalParams),
                                      C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "bbt_server_submit": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                      C.POINTER(C.c_uint64)]),
    "bbt_server_wait": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_double]),
    "bbt_server_scores": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "bbt_server_drain": (C.c_int32, [C.c_void_p, C.c_double]),
    "bbt_server_result": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                      C.POINTER(RequestStats)]),
    "bbt_server_release": (C.c_int32, [C.c_void_p, C.c_uint64]),
    "bbt_server_get_stats": (C.c_int32, [C.c_void_p, C.POINTER(ServerStats)]),
    "bbt_server_destroy": (C.c_int32, [C.c_void_p]),
    "bbt_set_tuning": (C.c_int32, [C.c_void_p, C.POINTER(Tuning)]),
    "bbt_get_tuning": (C.c_int32, [C.c_void_p, C.POINTER(Tuning)]),
    "bbt_autotune": (C.c_int32, [C.c_void_p, C.POINTER(Reads), C.POINTER(EvalParams),
                                 C.POINTER(EvalOut), C.c_void_p, C.POINTER(Tuning)]),
    "bbt_tuning_save": (C.c_int32, [C.c_char_p]),
    "bbt_tuning_load": (C.c_int32, [C.c_char_p]),
    "bbt_tuning_lookup": (C.c_int32, [C.c_void_p, C.POINTER(Reads), C.POINTER(EvalParams),
                                      C.POINTER(Tuning)]),
    "bbt_summary_text": (C.c_int32, [C.c_void_p, C.POINTER(C.c_uint64), C.c_uint64,
                                     C.c_char_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "bbt_calc_min_count": (C.c_uint64, [C.c_uint64, C.c_double, C.c_double]),
    "bbt_binomial_score": (C.c_double, [C.c_uint64, C.c_uint64, C.c_double]),
    "bbt_mibf_build": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.POINTER(C.c_uint64), C.c_uint64,
                                   C.POINTER(C.c_char_p), C.c_uint32,
                                   C.POINTER(C.c_char_p), C.c_uint32, C.c_uint32,
                                   C.c_double, C.c_uint64, C.c_uint64,
                                   C.POINTER(C.c_void_p), C.c_void_p]),
    "bbt_mibf_store": (C.c_int32, [C.c_void_p, C.c_char_p]),
    "bbt_mibf_load": (C.c_int32, [C.c_char_p, C.POINTER(C.c_void_p), C.c_void_p]),
    "bbt_mibf_free": (C.c_int32, [C.c_void_p]),
    "bbt_mibf_info": (C.c_int32, [C.c_void_p, C.POINTER(MibfInfo)]),
    "bbt_mibf_name": (C.c_char_p, [C.c_void_p, C.c_uint32]),
    "bbt_mibf_seed": (C.c_char_p, [C.c_void_p, C.c_uint32]),
    "bbt_mibf_download": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "bbt_mibf_calc_size": (C.c_uint64, [C.c_uint64, C.c_uint32, C.c_double]),
    "bbt_mibf_summary_text": (C.c_int32, [C.c_void_p, C.POINTER(C.c_uint64), C.c_uint64,
                                          C.c_char_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "bbt_mibf_frame_probs": (C.c_int32, [C.c_void_p, C.c_uint32, C.POINTER(C.c_double)]),
    "bbt_mibf_min_count": (C.c_uint32, [C.c_uint64, C.c_double, C.c_double, C.c_uint32]),
    "bbt_mibf_query_batch": (C.c_int32, [C.c_void_p, C.POINTER(Reads), C.POINTER(MibfParams),
                                         C.POINTER(MibfOut), C.c_void_p]),
}

_lib = None


def load():
    """Loads libbbt_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BbtError(f"{LIB_PATH} is missing: build it with "
                           "`python -m biobloom_amd.build` (no host fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != 0:
        msg = load().bbt_last_error()
        raise BbtError(f"bbt status {status}: {msg.decode() if msg else ''}")


class FilterSet:
    """A set of Bloom filters resident in device memory (bbt_filterset)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def load(cls, bf_paths):
        L = load()
        arr = (C.c_char_p * len(bf_paths))(*[p.encode() for p in bf_paths])
        h = C.c_void_p()
        check(L.bbt_filters_load(arr, len(bf_paths), C.byref(h)))
        return cls(h.value)

    @classmethod
    def create(cls, m, h, k):
        L = load()
        n = len(m)
        h_ = C.c_void_p()
        check(L.bbt_filters_create(n, (C.c_uint64 * n)(*m), (C.c_uint32 * n)(*h),
                                   (C.c_uint32 * n)(*k), C.byref(h_)))
        return cls(h_.value)

    def close(self):
        if self._h:
            load().bbt_filters_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def __len__(self):
        return load().bbt_filterset_size(self._h)

    def info(self, i):
        out = FilterInfo()
        check(load().bbt_filter_info_get(self._h, i, C.byref(out)))
        return out

    def filter_id(self, i):
        return load().bbt_filter_id(self._h, i).decode()

    def set_id(self, i, name):
        check(load().bbt_filter_set_id(self._h, i, name.encode()))

    def device_bits(self, i):
        return load().bbt_filter_device_bits(self._h, i)

    def upload(self, i, host_bits):
        check(load().bbt_filter_upload(self._h, i, host_bits.ctypes.data, host_bits.nbytes))

    def download(self, i, out):
        check(load().bbt_filter_download(self._h, i, out.ctypes.data, out.nbytes))
        return out

    def insert(self, i, dseq_ptr, nbytes, stream=0):
        counts = (C.c_uint64 * 2)()
        check(load().bbt_filter_insert(self._h, i, C.c_void_p(dseq_ptr), nbytes, counts,
                                       C.c_void_p(stream)))
        return int(counts[0]), int(counts[1])

    def insert_subtract(self, i, sub, sub_i, dseq_ptr, nbytes, stream=0):
        """Insert every k-mer that filter sub_i of FilterSet sub does not
        contain -> (new, already present, removed)."""
        counts = (C.c_uint64 * 3)()
        check(load().bbt_filter_insert_subtract(self._h, i, sub._h, sub_i, C.c_void_p(dseq_ptr),
                                                nbytes, counts, C.c_void_p(stream)))
        return int(counts[0]), int(counts[1]), int(counts[2])

    def eval_subtract(self, i, sub, sub_i, reads, scorer, threshold, streak, pass_ptr,
                      stream=0):
        """evalRead of every read against filter i, k-mers of filter sub_i of
        FilterSet sub (or None) scoring nothing; one byte per read."""
        check(load().bbt_eval_subtract(self._h, i, sub._h if sub is not None else None, sub_i,
                                       C.byref(reads), scorer, threshold, streak,
                                       C.c_void_p(pass_ptr), C.c_void_p(stream)))

    def refresh(self, i, stream=0):
        check(load().bbt_filter_refresh(self._h, i, C.c_void_p(stream)))

    def store(self, i, prefix, desired_fpr, expected, num_entries, redundant, sources):
        check(load().bbt_filter_store(self._h, i, prefix.encode(), desired_fpr, expected,
                                      num_entries, redundant, sources.encode()))

    def eval_batch(self, reads, params, out, stream=0):
        check(load().bbt_eval_batch(self._h, C.byref(reads), C.byref(params),
                                    C.byref(out), C.c_void_p(stream)))

    def set_tuning(self, wave_min_len=0, refill_min=0, blocks_per_cu=0, probe_ahead=0):
        check(load().bbt_set_tuning(self._h, C.byref(Tuning(wave_min_len, refill_min,
                                                             blocks_per_cu, probe_ahead))))

    def get_tuning(self):
        t = Tuning()
        check(load().bbt_get_tuning(self._h, C.byref(t)))
        return tuning_dict(t)

    def autotune(self, reads, params, out, stream=0):
        t = Tuning()
        check(load().bbt_autotune(self._h, C.byref(reads), C.byref(params), C.byref(out),
                                  C.c_void_p(stream), C.byref(t)))
        d = tuning_dict(t)
        d["walk_ms"] = self.last_walk_ms()
        return d

    def tuning_lookup(self, reads, params):
        """The tuning table's knobs for this batch's geometry, or None."""
        t = Tuning()
        st = load().bbt_tuning_lookup(self._h, C.byref(reads), C.byref(params), C.byref(t))
        return tuning_dict(t) if st == 0 else None

    def last_walk_ms(self):
        return load().bbt_last_walk_ms(self._h)

    def last_walk_kernel(self):
        return (load().bbt_last_walk_kernel(self._h) or b"").decode()

    def summary_text(self, summary_counts, nrec):
        L = load()
        arr = (C.c_uint64 * len(summary_counts))(*[int(x) for x in summary_counts])
        need = C.c_uint64()
        check(L.bbt_summary_text(self._h, arr, nrec, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value + 1)
        check(L.bbt_summary_text(self._h, arr, nrec, buf, need.value + 1, None))
        return buf.value.decode()


def tuning_dict(t):
    return {"wave_min_len": t.wave_min_len, "refill_min": t.refill_min,
            "blocks_per_cu": t.blocks_per_cu, "probe_ahead": t.probe_ahead}


def tuning_save(path):
    check(load().bbt_tuning_save(str(path).encode()))


def tuning_load(path):
    check(load().bbt_tuning_load(str(path).encode()))


def measure_gather(nbytes, rounds=64, stream=0):
    """Random-line load rate (lines/s) over a device buffer of nbytes."""
    r, ms = C.c_double(), C.c_double()
    check(load().bbt_measure_gather(nbytes, rounds, C.c_void_p(stream), C.byref(r),
                                    C.byref(ms)))
    return r.value, ms.value


def params(mode=STD, scorer=SIMPLE, threshold=0.15, streak=3, paired=False,
           inclusive=False, path=PATH_BATCHED, flags=0, wave_min_len=0, dust=False,
           dust_t=20, dust_w=64):
    return EvalParams(mode, scorer, threshold, streak, int(paired), int(inclusive), path,
                      flags, wave_min_len, int(dust), dust_t, dust_w, 0)


class Mibf:
    """A multi-index Bloom filter resident in device memory (bbt_mibf)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def build(cls, dseq_ptr, drec_off_ptr, drec_id_ptr, rec_off, names, seeds,
              occupancy=0.5, entries=0, m=0, stream=0, kmer=None):
        """rec_off: host sequence of nrec + 1 offsets; every record ends with
        one non-ACGT separator byte.  seeds: list of '0'/'1' strings, or None
        with kmer = (k, number of hash values) for k-mer mode."""
        L = load()
        nrec = len(rec_off) - 1
        off = (C.c_uint64 * len(rec_off))(*[int(x) for x in rec_off])
        nm = (C.c_char_p * len(names))(*[n.encode() for n in names])
        if seeds is None:
            k, nh = kmer
            sd = None
        else:
            k, nh = 0, len(seeds)
            sd = (C.c_char_p * len(seeds))(*[x.encode() for x in seeds])
        h = C.c_void_p()
        check(L.bbt_mibf_build(C.c_void_p(dseq_ptr), C.c_void_p(drec_off_ptr),
                               C.c_void_p(drec_id_ptr), off, nrec, nm, len(names), sd,
                               nh, k, occupancy, entries, m, C.byref(h),
                               C.c_void_p(stream)))
        return cls(h.value)

    @classmethod
    def load(cls, path, stream=0):
        h = C.c_void_p()
        check(load().bbt_mibf_load(path.encode(), C.byref(h), C.c_void_p(stream)))
        return cls(h.value)

    def close(self):
        if self._h:
            load().bbt_mibf_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        out = MibfInfo()
        check(load().bbt_mibf_info(self._h, C.byref(out)))
        return out

    def name(self, i):
        r = load().bbt_mibf_name(self._h, i)
        return None if r is None else r.decode()

    def seeds(self):
        """spaced seeds, or None in k-mer mode"""
        info = self.info()
        if info.kmer_mode:
            return None
        return [load().bbt_mibf_seed(self._h, i).decode() for i in range(info.nseeds)]

    def download(self, bv, data):
        check(load().bbt_mibf_download(self._h, bv.ctypes.data, data.ctypes.data))

    def store(self, prefix):
        check(load().bbt_mibf_store(self._h, prefix.encode()))

    def summary_text(self, summary_counts, nrec):
        L = load()
        arr = (C.c_uint64 * len(summary_counts))(*[int(x) for x in summary_counts])
        need = C.c_uint64()
        check(L.bbt_mibf_summary_text(self._h, arr, nrec, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value + 1)
        check(L.bbt_mibf_summary_text(self._h, arr, nrec, buf, need.value + 1, None))
        return buf.value.decode()

    def frame_probs(self, allowed_miss):
        n = self.info().nids + 1
        arr = (C.c_double * n)()
        check(load().bbt_mibf_frame_probs(self._h, allowed_miss, arr))
        return list(arr)

    def query_batch(self, reads, params, out, stream=0):
        check(load().bbt_mibf_query_batch(self._h, C.byref(reads), C.byref(params),
                                          C.byref(out), C.c_void_p(stream)))


MIBF_PAIRED = 1
MIBF_BEST_AGREE = 2
MIBF_LINES = 4
MIBF_HIT_WORDS = 7


def mibf_params(score=0.1, multi_thresh=1.0, allowed_miss=0, min_nonsat=1, max_hits=8,
                paired=False, best_agree=False, lines=False):
    """lines: look up through the lines and ID array even when the filter has
    its direct-mapped table."""
    flags = ((MIBF_PAIRED if paired else 0) | (MIBF_BEST_AGREE if best_agree else 0)
             | (MIBF_LINES if lines else 0))
    return MibfParams(score, multi_thresh, allowed_miss, min_nonsat, max_hits, flags)


class Workspace:
    """Work queues for batches on one stream (bbt_workspace)."""

    def __init__(self):
        h = C.c_void_p()
        check(load().bbt_workspace_create(C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            load().bbt_workspace_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def eval_batch(self, fs, reads, params, out, stream=0):
        check(load().bbt_eval_batch_ws(fs.handle, self._h, C.byref(reads), C.byref(params),
                                       C.byref(out), C.c_void_p(stream)))


class Server:
    """Request-serving loop over a resident filter set (bbt_server)."""

    def __init__(self, fs, params, slots=4, mode=SERVE_BATCHED, device_bytes=1 << 30,
                 max_wave_reads=1 << 18, max_wave_bases=64 << 20, host_bytes=0):
        self.fs = fs  # keeps the filter set alive
        self.nf = len(fs)
        self.paired = bool(params.paired)
        cfg = ServerConfig(slots, mode, device_bytes, max_wave_reads, max_wave_bases,
                           host_bytes)
        h = C.c_void_p()
        check(load().bbt_server_create(fs.handle, C.byref(params), C.byref(cfg), C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            load().bbt_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, buf, off):
        """buf: uint8 numpy array, off: uint64 numpy offsets (nreads + 1)."""
        rid = C.c_uint64()
        check(load().bbt_server_submit(self._h, buf.ctypes.data, off.ctypes.data,
                                       len(off) - 1, C.byref(rid)))
        return rid.value

    def wait(self, rid, timeout=-1.0):
        check(load().bbt_server_wait(self._h, rid, timeout))

    def drain(self, timeout=-1.0):
        check(load().bbt_server_drain(self._h, timeout))

    def result(self, rid, nreads):
        import numpy as np
        nrec = nreads // 2 if self.paired else nreads
        hits = np.zeros(max(1, nreads * self.nf), np.uint8)
        cls = np.zeros(max(1, nrec), np.uint32)
        st = RequestStats()
        check(load().bbt_server_result(self._h, rid, hits.ctypes.data, cls.ctypes.data,
                                       C.byref(st)))
        return (hits[: nreads * self.nf].reshape(nreads, self.nf), cls[:nrec],
                {"wave": st.wave, "submit_us": st.submit_us, "admit_us": st.admit_us,
                 "done_us": st.done_us})

    def scores(self, rid, nreads):
        """evalScore values (nreads x nf) of a request to a BESTHIT / SCORES server."""
        import numpy as np
        out = np.zeros(max(1, nreads * self.nf), np.float64)
        check(load().bbt_server_scores(self._h, rid, out.ctypes.data))
        return out[: nreads * self.nf].reshape(nreads, self.nf)

    def release(self, rid):
        check(load().bbt_server_release(self._h, rid))

    def stats(self):
        st = ServerStats()
        check(load().bbt_server_get_stats(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}
