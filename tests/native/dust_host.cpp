/*
 * dust_host.cpp -- compiles the device sdust pass (bbt_dust_core.h) for the
 * CPU so the tests can check it against the oracle without a GPU, with the
 * interval-list capacity of either device kernel.  Test infrastructure only.
 */
#include <cstdint>
#include <vector>

#include "bbt_dust_core.h"

using namespace bbt;

namespace {

/* SEP: the new entries of a perfect() call kept apart and merged (the
 * device-memory kernel's view), or shifted into place one by one (the LDS
 * kernel's view). */
template <bool SEP>
struct HostView {
    static constexpr bool kSepNew = SEP;
    uint16_t cv_[64], cw_[64], c_[64];
    std::vector<uint16_t> ring_;
    std::vector<DustPerf> P_, N_;
    uint32_t rm, pc;
    HostView(int W, uint32_t pcap) {
        uint32_t cap = 64;
        while (int(cap) < W - 2)
            cap <<= 1;
        ring_.assign(cap, 0);
        rm = cap - 1;
        pc = pcap;
        P_.assign(pcap + 1, DustPerf{0, 0, 0, 0});
        N_.assign(W, DustPerf{0, 0, 0, 0});
    }
    DustPerf N(uint32_t k) const { return N_[k]; }
    void setN(uint32_t k, const DustPerf &x) { N_[k] = x; }
    uint16_t &cv(unsigned t) { return cv_[t]; }
    uint16_t &cw(unsigned t) { return cw_[t]; }
    uint16_t &c(unsigned t) { return c_[t]; }
    uint16_t &ring(uint32_t i) { return ring_[i]; }
    uint32_t rmask() const { return rm; }
    uint32_t pcap() const { return pc; }
    DustPerf P(uint32_t j) const { return P_[j]; }
    void setP(uint32_t j, const DustPerf &x) { P_[j] = x; }
};

struct HostSrc {
    const uint8_t *s;
    unsigned at(uint32_t i) const { return s[i]; }
};

struct HostEmit {
    uint8_t *mask;
    void operator()(int a, int f) const {
        for (int p = a; p < f; ++p)
            mask[p] = 1;
    }
};

} // namespace

extern "C" {

/* Marks mask[p] = 1 for every low-complexity window start p of seq[0, n).
 * pcap = 0: the list capacity and the view of the device-memory kernel (never
 * reached); else that many entries, in the LDS kernel's view.  Returns 1, or 0 when the
 * list would outgrow pcap (the device then runs the read again on the
 * device-memory kernel). */
int dust_host(const uint8_t *seq, uint32_t n, int T, int W, uint32_t pcap, uint8_t *mask) {
    HostSrc src{seq};
    HostEmit em{mask};
    if (pcap == 0) {
        HostView<true> v(W, uint32_t(dust_pcap_bound(W)));
        return dust_core(src, n, T, W, v, em) ? 1 : 0;
    }
    HostView<false> v(W, pcap);
    return dust_core(src, n, T, W, v, em) ? 1 : 0;
}

/* The same with the separate new-entry list for any capacity (1: taken). */
int dust_host_sep(const uint8_t *seq, uint32_t n, int T, int W, uint32_t pcap, uint8_t *mask) {
    HostSrc src{seq};
    HostEmit em{mask};
    HostView<true> v(W, pcap);
    return dust_core(src, n, T, W, v, em) ? 1 : 0;
}

uint32_t dust_host_fast_pcap() { return kDustFastP; }
int dust_host_fast_max_w() { return kDustFastW; }

}
