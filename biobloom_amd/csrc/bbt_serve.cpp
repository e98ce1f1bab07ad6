/*
 * bbt_serve.cpp -- request-serving loop over a resident filter set.
 *
 * The reference categorizer reads one file stream and classifies it in an
 * OpenMP loop (BioBloomClassifier::filter, BioBloomCategorizer/
 * BioBloomClassifier.cpp:33-90); it has no request API.  This runtime keeps
 * the filters resident in device memory and serves independent requests
 * (a batch of reads each) with continuous batching:
 *
 *   submit      copies the request's reads into a pinned host arena (waiting
 *               while the arena is full) and queues the request (FIFO); any
 *               thread may submit.
 *   dispatcher  forms a launch ("wave") from the head of the queue whenever a
 *               slot is free: requests are merged while the wave stays within
 *               the read / base caps and its device block (reads + outputs)
 *               fits in the device arena -- admission by device memory.  Each
 *               request's bases are copied to the device straight from the
 *               pinned arena (the offsets are rebased through the slot's
 *               pinned buffer), classified by the batched persistent path
 *               (bbt_eval_batch_ws on the slot's stream and workspace) and its
 *               per-read hits and per-record classes copied back.  Up to
 *               `slots` waves are in flight on separate streams, so copies of
 *               one overlap the kernels of another, and requests that arrive
 *               while waves run join the next wave.
 *   completion  polls the waves' events, scatters each wave's results into
 *               its requests, stamps their completion time and releases the
 *               slot and arena block.
 *
 * Plain mode (the per-request baseline): one request per launch, one slot,
 * the plain one-thread-per-walk kernel, in arrival order.
 *
 * A request's results do not depend on what it was batched with: every read
 * is classified independently (the reference's per-read loop), and the
 * kernels' counters and queues belong to the wave.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/bbt_hip.h"
#include "bbt_host.h"

using namespace bbt;

namespace {

using Clock = std::chrono::steady_clock;

bbt_status sfail(bbt_status c, const std::string &m) { return set_error(c, m); }

#define SHIP(expr)                                                                    \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return sfail(BBT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr uint64_t kAlign = 256;
uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

/* First-fit allocator over one device allocation. */
class Arena {
public:
    void reset(uint64_t size) {
        free_.clear();
        free_[0] = size;
        size_ = size;
        used_ = high_ = 0;
    }
    bool alloc(uint64_t n, uint64_t &off) {
        n = align_up(std::max<uint64_t>(n, 1), kAlign);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n)
                continue;
            off = it->first;
            const uint64_t rest = it->second - n;
            free_.erase(it);
            if (rest)
                free_[off + n] = rest;
            used_ += n;
            high_ = std::max(high_, used_);
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        n = align_up(std::max<uint64_t>(n, 1), kAlign);
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t size() const { return size_; }
    uint64_t high_water() const { return high_; }

private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t size_ = 0, used_ = 0, high_ = 0;
};

enum ReqState { PENDING = 0, ADMITTED = 1, DONE = 2, FAILED = 3 };

struct Request {
    uint64_t id = 0;
    uint64_t host_off = 0, bases = 0; /* the bases, in the pinned host arena */
    bool host_held = false;
    std::vector<uint64_t> off; /* nreads + 1, from 0 */
    uint64_t nreads = 0, nrec = 0;
    uint32_t max_len = 0;
    std::vector<uint8_t> hits;
    std::vector<uint32_t> cls;
    std::vector<double> scores; /* BESTHIT / SCORES: nreads * nf evalScore values */
    Clock::time_point t_submit, t_admit, t_done;
    uint64_t wave = 0;
    int state = PENDING;
    std::string error;
};

/* byte layout of one wave's device block */
struct WaveLayout {
    uint64_t seq = 0, off = 0, hits = 0, cls = 0, scores = 0, total = 0;
    uint64_t seq_bytes = 0;
    static WaveLayout of(uint64_t nreads, uint64_t bases, uint64_t nrec, uint32_t nf,
                         bool with_scores) {
        WaveLayout L;
        L.seq_bytes = align_up(bases, 8) + 8;
        uint64_t o = 0;
        L.seq = o;
        o = align_up(o + L.seq_bytes, kAlign);
        L.off = o;
        o = align_up(o + (nreads + 1) * 8, kAlign);
        L.hits = o;
        o = align_up(o + std::max<uint64_t>(1, nreads * nf), kAlign);
        L.cls = o;
        o = align_up(o + std::max<uint64_t>(1, nrec) * 4, kAlign);
        L.scores = o;
        if (with_scores)
            o = align_up(o + std::max<uint64_t>(1, nreads * nf) * 8, kAlign);
        L.total = o;
        return L;
    }
};

struct Slot {
    hipStream_t st = nullptr;
    bbt_workspace *ws = nullptr;
    hipEvent_t start = nullptr, done = nullptr; /* around the launch's device work */
    uint64_t *h_off = nullptr;
    uint8_t *h_hits = nullptr;
    uint32_t *h_cls = nullptr;
    double *h_scores = nullptr; /* BESTHIT / SCORES */
    bool busy = false;
    std::vector<std::shared_ptr<Request>> reqs;
    uint64_t block_off = 0, block_len = 0;
    uint64_t wave = 0;
    uint64_t nreads = 0, nrec = 0;
};

} // namespace

/* Helper threads that split a large copy, so a submit moves a request into
 * pinned memory at more than one core's bandwidth.  Copies from several
 * submitting threads take turns. */
class CopyPool {
  public:
    explicit CopyPool(unsigned helpers) {
        for (unsigned i = 0; i < helpers; ++i)
            m_th.emplace_back([this, i] { loop(i + 1); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> g(m_mu);
            m_stop = true;
        }
        m_cv.notify_all();
        for (std::thread &t : m_th)
            t.join();
    }
    void copy(char *dst, const char *src, size_t n) {
        if (m_th.empty() || n < (size_t(4) << 20)) {
            std::memcpy(dst, src, n);
            return;
        }
        std::lock_guard<std::mutex> turn(m_call);
        const unsigned parts = unsigned(m_th.size()) + 1;
        {
            std::lock_guard<std::mutex> g(m_mu);
            m_dst = dst;
            m_src = src;
            m_n = n;
            m_parts = parts;
            m_left = unsigned(m_th.size());
            ++m_gen;
        }
        m_cv.notify_all();
        part(0);
        std::unique_lock<std::mutex> g(m_mu);
        m_done.wait(g, [&] { return m_left == 0; });
    }

  private:
    void part(unsigned i) {
        const size_t a = m_n * i / m_parts, b = m_n * (i + 1) / m_parts;
        std::memcpy(m_dst + a, m_src + a, b - a);
    }
    void loop(unsigned i) {
        uint64_t seen = 0;
        while (true) {
            {
                std::unique_lock<std::mutex> g(m_mu);
                m_cv.wait(g, [&] { return m_stop || m_gen != seen; });
                if (m_stop)
                    return;
                seen = m_gen;
            }
            part(i);
            {
                std::lock_guard<std::mutex> g(m_mu);
                --m_left;
            }
            m_done.notify_all();
        }
    }
    std::vector<std::thread> m_th;
    std::mutex m_mu, m_call;
    std::condition_variable m_cv, m_done;
    char *m_dst = nullptr;
    const char *m_src = nullptr;
    size_t m_n = 0;
    unsigned m_parts = 1, m_left = 0;
    uint64_t m_gen = 0;
    bool m_stop = false;
};

struct bbt_server {
    int device = 0; /* the server's threads run on the creating thread's device */
    bbt_filterset *fs = nullptr;
    bbt_eval_params params{};
    bbt_server_config cfg{};
    uint32_t nf = 0;
    bool scores = false; /* BESTHIT / SCORES: evalScore values are returned too */
    char *darena = nullptr;
    Arena arena;
    char *harena = nullptr; /* pinned: bases of queued and in-flight requests */
    Arena hostmem;
    std::condition_variable cv_host;
    std::vector<Slot> slots;
    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> reqs;
    std::vector<int> inflight; /* slot indices in launch order */
    std::mutex mu;
    std::condition_variable cv_dispatch, cv_complete, cv_done;
    std::thread dispatcher, completer;
    std::unique_ptr<CopyPool> copier;
    uint64_t *dctr = nullptr; /* device: contains() calls, lookups, lookups issued, flag */
    bool stop = false;
    uint64_t next_id = 1, waves = 0, done_reqs = 0, done_reads = 0, max_inflight = 0;
    uint64_t wave_reads_sum = 0, wave_reqs_sum = 0;
    double submit_ms = 0, pack_ms = 0, device_ms = 0; /* totals */
    std::string error;
    Clock::time_point t0;

    void dispatch_loop();
    void complete_loop();
    bbt_status launch(Slot &s);
    /* with mu held: the wave's pinned host ranges go back to the arena */
    void release_host(Slot &s) {
        for (auto &r : s.reqs)
            if (r->host_held) {
                hostmem.release(r->host_off, r->bases);
                r->host_held = false;
            }
        cv_host.notify_all();
    }
};

namespace {

double us_since(Clock::time_point t0, Clock::time_point t) {
    return std::chrono::duration<double, std::micro>(t - t0).count();
}

void fail_wave(bbt_server *S, Slot &s, const std::string &msg) {
    for (auto &r : s.reqs) {
        r->state = FAILED;
        r->error = msg;
        r->t_done = Clock::now();
    }
}

} // namespace

bbt_status bbt_server::launch(Slot &s) {
    const bool paired = params.paired != 0;
    const Clock::time_point tp = Clock::now();
    /* pack the wave's reads into the pinned staging buffers */
    uint64_t nreads = 0, bases = 0;
    uint32_t max_len = 0;
    s.h_off[0] = 0;
    for (auto &r : s.reqs) {
        for (uint64_t i = 0; i < r->nreads; ++i)
            s.h_off[nreads + i + 1] = bases + r->off[i + 1];
        nreads += r->nreads;
        bases += r->bases;
        max_len = std::max(max_len, r->max_len);
    }
    const uint64_t nrec = paired ? nreads / 2 : nreads;
    const WaveLayout L = WaveLayout::of(nreads, bases, nrec, nf, scores);
    char *blk = darena + s.block_off;
    const double packed = std::chrono::duration<double, std::milli>(Clock::now() - tp).count();
    {
        std::lock_guard<std::mutex> g(mu);
        pack_ms += packed;
    }
    SHIP(hipEventRecord(s.start, s.st));
    /* the bases of each request, straight from the pinned arena (the padding
     * past the last base is never read) */
    uint64_t at = 0;
    for (auto &r : s.reqs) {
        if (r->bases)
            SHIP(hipMemcpyAsync(blk + L.seq + at, harena + r->host_off, r->bases,
                                hipMemcpyHostToDevice, s.st));
        at += r->bases;
    }
    SHIP(hipMemcpyAsync(blk + L.off, s.h_off, (nreads + 1) * 8, hipMemcpyHostToDevice, s.st));
    bbt_reads in;
    std::memset(&in, 0, sizeof(in));
    in.seq = blk + L.seq;
    in.seq_bytes = L.seq_bytes;
    in.off = reinterpret_cast<const uint64_t *>(blk + L.off);
    in.nreads = nreads;
    in.max_len = max_len;
    bbt_eval_out out;
    std::memset(&out, 0, sizeof(out));
    out.hits = reinterpret_cast<uint8_t *>(blk + L.hits);
    out.cls = reinterpret_cast<uint32_t *>(blk + L.cls);
    if (scores)
        out.scores = reinterpret_cast<double *>(blk + L.scores);
    out.counters = dctr; /* every launch adds to the server's totals */
    bbt_eval_params p = params;
    p.path = cfg.mode == BBT_SERVE_PLAIN ? BBT_PATH_PLAIN : BBT_PATH_BATCHED;
    p.flags = 0;
    bbt_status st = bbt_eval_batch_ws(fs, s.ws, &in, &p, &out, s.st);
    if (st != BBT_OK)
        return st;
    SHIP(hipMemcpyAsync(s.h_hits, out.hits, std::max<uint64_t>(1, nreads * nf),
                        hipMemcpyDeviceToHost, s.st));
    SHIP(hipMemcpyAsync(s.h_cls, out.cls, std::max<uint64_t>(1, nrec) * 4,
                        hipMemcpyDeviceToHost, s.st));
    if (scores)
        SHIP(hipMemcpyAsync(s.h_scores, out.scores, std::max<uint64_t>(1, nreads * nf) * 8,
                            hipMemcpyDeviceToHost, s.st));
    SHIP(hipEventRecord(s.done, s.st));
    s.nreads = nreads;
    s.nrec = nrec;
    return BBT_OK;
}

void bbt_server::dispatch_loop() {
    (void)hipSetDevice(device);
    const bool paired = params.paired != 0;
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        cv_dispatch.wait(lk, [&] {
            if (stop)
                return true;
            if (pending.empty())
                return false;
            for (const Slot &s : slots)
                if (!s.busy)
                    return true;
            return false;
        });
        if (stop)
            return;
        int si = -1;
        for (size_t i = 0; i < slots.size(); ++i)
            if (!slots[i].busy) {
                si = int(i);
                break;
            }
        Slot &s = slots[size_t(si)];
        /* the wave: a FIFO prefix of the queue within the caps and the arena */
        uint64_t nreads = 0, bases = 0, take = 0, block = 0;
        const size_t maxreq = cfg.mode == BBT_SERVE_PLAIN ? 1 : pending.size();
        for (size_t i = 0; i < pending.size() && i < maxreq; ++i) {
            const Request &r = *pending[i];
            if (nreads + r.nreads > cfg.max_wave_reads || bases + r.bases > cfg.max_wave_bases)
                break;
            const WaveLayout L = WaveLayout::of(nreads + r.nreads, bases + r.bases,
                                                paired ? (nreads + r.nreads) / 2 : nreads + r.nreads,
                                                nf, scores);
            uint64_t off = 0;
            if (!arena.alloc(L.total, off))
                break;
            arena.release(off, L.total);
            nreads += r.nreads;
            bases += r.bases;
            take = i + 1;
            block = L.total;
        }
        if (take == 0) {
            /* the head request waits for device memory held by waves in flight */
            cv_dispatch.wait(lk, [&] { return stop || !inflight.empty(); });
            if (stop)
                return;
            cv_complete.notify_all();
            cv_dispatch.wait_for(lk, std::chrono::microseconds(200));
            continue;
        }
        if (!arena.alloc(block, s.block_off)) /* cannot fail: checked above under the lock */
            continue;
        s.block_len = block;
        s.reqs.assign(pending.begin(), pending.begin() + long(take));
        pending.erase(pending.begin(), pending.begin() + long(take));
        s.busy = true;
        s.wave = ++waves;
        const Clock::time_point now = Clock::now();
        for (auto &r : s.reqs) {
            r->state = ADMITTED;
            r->t_admit = now;
            r->wave = s.wave;
        }
        wave_reads_sum += nreads;
        wave_reqs_sum += take;
        lk.unlock();
        bbt_status st = launch(s);
        lk.lock();
        if (st != BBT_OK) {
            error = last_error();
            fail_wave(this, s, error);
            release_host(s);
            arena.release(s.block_off, s.block_len);
            s.reqs.clear();
            s.busy = false;
            cv_done.notify_all();
            continue;
        }
        inflight.push_back(si);
        max_inflight = std::max<uint64_t>(max_inflight, inflight.size());
        cv_complete.notify_all();
    }
}

void bbt_server::complete_loop() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        cv_complete.wait(lk, [&] { return stop || !inflight.empty(); });
        if (stop && inflight.empty())
            return;
        /* poll every wave in flight; waves on different streams may finish
         * out of launch order */
        bool any = false;
        for (size_t k = 0; k < inflight.size();) {
            Slot &s = slots[size_t(inflight[k])];
            const hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) {
                ++k;
                continue;
            }
            any = true;
            const Clock::time_point now = Clock::now();
            if (q != hipSuccess) {
                error = std::string("wave failed: ") + hipGetErrorString(q);
                fail_wave(this, s, error);
            } else {
                float ms = 0;
                if (hipEventElapsedTime(&ms, s.start, s.done) == hipSuccess)
                    device_ms += ms;
                uint64_t rd = 0, rc = 0;
                for (auto &r : s.reqs) {
                    r->hits.assign(s.h_hits + rd * nf, s.h_hits + (rd + r->nreads) * nf);
                    r->cls.assign(s.h_cls + rc, s.h_cls + rc + r->nrec);
                    if (scores)
                        r->scores.assign(s.h_scores + rd * nf,
                                         s.h_scores + (rd + r->nreads) * nf);
                    rd += r->nreads;
                    rc += r->nrec;
                    r->t_done = now;
                    r->state = DONE;
                    ++done_reqs;
                    done_reads += r->nreads;
                }
            }
            release_host(s);
            arena.release(s.block_off, s.block_len);
            s.reqs.clear();
            s.busy = false;
            inflight.erase(inflight.begin() + long(k));
        }
        if (any) {
            cv_done.notify_all();
            cv_dispatch.notify_all();
            continue;
        }
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
    }
}

extern "C" {

bbt_status bbt_server_create(bbt_filterset *fs, const bbt_eval_params *p,
                             const bbt_server_config *cfg, bbt_server **out) {
    if (!fs || !p || !cfg || !out)
        return sfail(BBT_ERR_ARG, "bbt_server_create: null argument");
    if (cfg->slots == 0 || cfg->slots > 8 || cfg->max_wave_reads == 0 ||
        cfg->max_wave_bases == 0 || cfg->device_bytes == 0 || cfg->mode > BBT_SERVE_PLAIN)
        return sfail(BBT_ERR_ARG, "bbt_server_create: bad configuration");
    *out = nullptr;
    std::unique_ptr<bbt_server> S(new bbt_server());
    S->fs = fs;
    S->params = *p;
    S->cfg = *cfg;
    if (S->cfg.mode == BBT_SERVE_PLAIN)
        S->cfg.slots = 1;
    S->nf = bbt_filterset_size(fs);
    S->scores = p->mode == BBT_BESTHIT || p->mode == BBT_SCORES;
    S->t0 = Clock::now();
    SHIP(hipGetDevice(&S->device));
    const uint64_t rd = S->cfg.max_wave_reads, bs = S->cfg.max_wave_bases;
    const WaveLayout Lmax = WaveLayout::of(rd, bs, rd, S->nf, S->scores);
    if (Lmax.total > cfg->device_bytes)
        return sfail(BBT_ERR_ARG, "bbt_server_create: device_bytes below one full wave (" +
                                      std::to_string(Lmax.total) + " bytes)");
    SHIP(hipMalloc(&S->darena, cfg->device_bytes));
    SHIP(hipMalloc(&S->dctr, 4 * sizeof(uint64_t)));
    SHIP(hipMemset(S->dctr, 0, 4 * sizeof(uint64_t)));
    S->arena.reset(cfg->device_bytes);
    const uint64_t hb = align_up(cfg->host_bytes ? cfg->host_bytes
                                                 : bs * (uint64_t(S->cfg.slots) + 2),
                                 kAlign);
    if (hb < align_up(bs, kAlign))
        return sfail(BBT_ERR_ARG, "bbt_server_create: host_bytes below one full wave");
    SHIP(hipHostMalloc(reinterpret_cast<void **>(&S->harena), hb));
    S->hostmem.reset(hb);
    S->slots.resize(S->cfg.slots);
    for (Slot &s : S->slots) {
        SHIP(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
        SHIP(hipEventCreate(&s.start));
        SHIP(hipEventCreate(&s.done));
        bbt_status st = bbt_workspace_create(&s.ws);
        if (st != BBT_OK)
            return st;
        SHIP(hipHostMalloc(reinterpret_cast<void **>(&s.h_off), (rd + 1) * 8));
        SHIP(hipHostMalloc(reinterpret_cast<void **>(&s.h_hits), std::max<uint64_t>(1, rd * S->nf)));
        SHIP(hipHostMalloc(reinterpret_cast<void **>(&s.h_cls), std::max<uint64_t>(1, rd) * 4));
        if (S->scores)
            SHIP(hipHostMalloc(reinterpret_cast<void **>(&s.h_scores),
                               std::max<uint64_t>(1, rd * S->nf) * 8));
    }
    const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
    S->copier.reset(new CopyPool(std::min(3u, hc - 1)));
    bbt_server *raw = S.release();
    raw->dispatcher = std::thread([raw] { raw->dispatch_loop(); });
    raw->completer = std::thread([raw] { raw->complete_loop(); });
    *out = raw;
    return BBT_OK;
}

bbt_status bbt_server_submit(bbt_server *S, const char *seq, const uint64_t *off, uint64_t nreads,
                             uint64_t *req_id) {
    if (!S || !off || (!seq && nreads) || !req_id)
        return sfail(BBT_ERR_ARG, "bbt_server_submit: null argument");
    if (nreads == 0)
        return sfail(BBT_ERR_ARG, "bbt_server_submit: empty request");
    if (S->params.paired && nreads % 2)
        return sfail(BBT_ERR_ARG, "bbt_server_submit: paired request with an odd read count");
    const Clock::time_point t_in = Clock::now();
    auto r = std::make_shared<Request>();
    const uint64_t b0 = off[0], bases = off[nreads] - off[0];
    if (nreads > S->cfg.max_wave_reads || bases > S->cfg.max_wave_bases)
        return sfail(BBT_ERR_ARG, "bbt_server_submit: request larger than one wave");
    r->off.resize(nreads + 1);
    for (uint64_t i = 0; i <= nreads; ++i) {
        if (i && off[i] < off[i - 1])
            return sfail(BBT_ERR_ARG, "bbt_server_submit: offsets must not decrease");
        r->off[i] = off[i] - b0;
        if (i)
            r->max_len = std::max<uint32_t>(r->max_len, uint32_t(off[i] - off[i - 1]));
    }
    r->nreads = nreads;
    r->nrec = S->params.paired ? nreads / 2 : nreads;
    r->bases = bases;
    {
        /* room in the pinned arena: requests wait here while queued and
         * in-flight requests hold it (host-side admission) */
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_host.wait(lk, [&] { return S->stop || S->hostmem.alloc(bases, r->host_off); });
        if (S->stop)
            return sfail(BBT_ERR_ARG, "bbt_server_submit: server is stopping");
        r->host_held = true;
    }
    const Clock::time_point t_copy = Clock::now();
    S->copier->copy(S->harena + r->host_off, seq + b0, bases);
    {
        std::lock_guard<std::mutex> g(S->mu);
        r->id = S->next_id++;
        r->t_submit = t_in;
        S->submit_ms += std::chrono::duration<double, std::milli>(Clock::now() - t_copy).count();
        S->reqs[r->id] = r;
        S->pending.push_back(r);
    }
    S->cv_dispatch.notify_all();
    *req_id = r->id;
    return BBT_OK;
}

bbt_status bbt_server_wait(bbt_server *S, uint64_t req_id, double timeout_s) {
    if (!S)
        return sfail(BBT_ERR_ARG, "bbt_server_wait: null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(req_id);
    if (it == S->reqs.end())
        return sfail(BBT_ERR_ARG, "bbt_server_wait: unknown request");
    std::shared_ptr<Request> r = it->second;
    auto ready = [&] { return r->state == DONE || r->state == FAILED; };
    if (timeout_s < 0)
        S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
        return sfail(BBT_ERR_HIP, "bbt_server_wait: timed out");
    if (r->state == FAILED)
        return sfail(BBT_ERR_HIP, "request failed: " + r->error);
    return BBT_OK;
}

bbt_status bbt_server_drain(bbt_server *S, double timeout_s) {
    if (!S)
        return sfail(BBT_ERR_ARG, "bbt_server_drain: null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto idle = [&] {
        if (!S->pending.empty() || !S->inflight.empty())
            return false;
        for (const Slot &s : S->slots)
            if (s.busy)
                return false;
        return true;
    };
    if (timeout_s < 0)
        S->cv_done.wait(lk, idle);
    else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), idle))
        return sfail(BBT_ERR_HIP, "bbt_server_drain: timed out");
    if (!S->error.empty())
        return sfail(BBT_ERR_HIP, "server error: " + S->error);
    return BBT_OK;
}

bbt_status bbt_server_scores(bbt_server *S, uint64_t req_id, double *scores) {
    if (!S || !scores)
        return sfail(BBT_ERR_ARG, "bbt_server_scores: null argument");
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->reqs.find(req_id);
    if (it == S->reqs.end())
        return sfail(BBT_ERR_ARG, "bbt_server_scores: unknown request");
    const Request &r = *it->second;
    if (r.state != DONE)
        return sfail(r.state == FAILED ? BBT_ERR_HIP : BBT_ERR_ARG,
                     r.state == FAILED ? "request failed: " + r.error
                                       : std::string("bbt_server_scores: request not done"));
    if (!S->scores)
        return sfail(BBT_ERR_ARG, "bbt_server_scores: the server's mode has no scores");
    std::memcpy(scores, r.scores.data(), r.scores.size() * 8);
    return BBT_OK;
}

bbt_status bbt_server_result(bbt_server *S, uint64_t req_id, uint8_t *hits, uint32_t *cls,
                             bbt_request_stats *stats) {
    if (!S)
        return sfail(BBT_ERR_ARG, "bbt_server_result: null server");
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->reqs.find(req_id);
    if (it == S->reqs.end())
        return sfail(BBT_ERR_ARG, "bbt_server_result: unknown request");
    const Request &r = *it->second;
    if (r.state != DONE)
        return sfail(r.state == FAILED ? BBT_ERR_HIP : BBT_ERR_ARG,
                     r.state == FAILED ? "request failed: " + r.error
                                       : std::string("bbt_server_result: request not done"));
    if (hits)
        std::memcpy(hits, r.hits.data(), r.hits.size());
    if (cls)
        std::memcpy(cls, r.cls.data(), r.cls.size() * 4);
    if (stats) {
        stats->nreads = r.nreads;
        stats->wave = r.wave;
        stats->submit_us = us_since(S->t0, r.t_submit);
        stats->admit_us = us_since(S->t0, r.t_admit);
        stats->done_us = us_since(S->t0, r.t_done);
    }
    return BBT_OK;
}

bbt_status bbt_server_release(bbt_server *S, uint64_t req_id) {
    if (!S)
        return sfail(BBT_ERR_ARG, "bbt_server_release: null server");
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->reqs.find(req_id);
    if (it == S->reqs.end())
        return sfail(BBT_ERR_ARG, "bbt_server_release: unknown request");
    if (it->second->state == PENDING || it->second->state == ADMITTED)
        return sfail(BBT_ERR_ARG, "bbt_server_release: request still in progress");
    S->reqs.erase(it);
    return BBT_OK;
}

bbt_status bbt_server_get_stats(bbt_server *S, bbt_server_stats *out) {
    if (!S || !out)
        return sfail(BBT_ERR_ARG, "bbt_server_get_stats: null argument");
    std::lock_guard<std::mutex> g(S->mu);
    out->waves = S->waves;
    out->requests_done = S->done_reqs;
    out->reads_done = S->done_reads;
    out->max_waves_in_flight = S->max_inflight;
    out->arena_bytes = S->arena.size();
    out->arena_high_water = S->arena.high_water();
    out->mean_wave_reads = S->waves ? double(S->wave_reads_sum) / double(S->waves) : 0.0;
    out->mean_wave_requests = S->waves ? double(S->wave_reqs_sum) / double(S->waves) : 0.0;
    out->submit_ms = S->submit_ms;
    out->pack_ms = S->pack_ms;
    out->device_ms = S->device_ms;
    /* the counters as the finished launches left them (read while none runs
     * for exact totals, e.g. after bbt_server_drain) */
    uint64_t c[4] = {0, 0, 0, 0};
    SHIP(hipMemcpy(c, S->dctr, sizeof(c), hipMemcpyDeviceToHost));
    out->probes = c[0];
    out->lookups = c[1];
    out->lookups_issued = c[2];
    return BBT_OK;
}

bbt_status bbt_server_destroy(bbt_server *S) {
    if (!S)
        return BBT_OK;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->cv_host.notify_all();
    S->cv_dispatch.notify_all();
    S->cv_complete.notify_all();
    if (S->dispatcher.joinable())
        S->dispatcher.join();
    /* waves already launched finish before their buffers go */
    for (Slot &s : S->slots)
        if (s.st)
            (void)hipStreamSynchronize(s.st);
    if (S->completer.joinable()) {
        {
            std::lock_guard<std::mutex> g(S->mu);
            S->inflight.clear();
        }
        S->cv_complete.notify_all();
        S->completer.join();
    }
    for (Slot &s : S->slots) {
        if (s.ws)
            bbt_workspace_free(s.ws);
        if (s.done)
            (void)hipEventDestroy(s.done);
        if (s.start)
            (void)hipEventDestroy(s.start);
        if (s.st)
            (void)hipStreamDestroy(s.st);
        (void)hipHostFree(s.h_off);
        (void)hipHostFree(s.h_hits);
        (void)hipHostFree(s.h_cls);
        if (s.h_scores)
            (void)hipHostFree(s.h_scores);
    }
    (void)hipFree(S->darena);
    (void)hipFree(S->dctr);
    (void)hipHostFree(S->harena);
    delete S;
    return BBT_OK;
}

} // extern "C"
