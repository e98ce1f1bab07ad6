/*
 * bbt_dust.cpp -- CPU restatement of the symmetric DUST low-complexity
 * masker the reference vendors (Common/sdust.c:1-170, the minimap2 `sdust`
 * of H. Li; algorithm of Morgulis et al., J. Comput. Biol. 13(5), 2006) and of
 * its C++ wrapper's query (Common/SDust.hpp:17-55).  TEST INFRASTRUCTURE ONLY:
 * the checker of the dust parity tests.
 *
 * Definitions (triplet words of 3 bases; counts c_t of each of the 64 words):
 *   score of a run of l words = sum_t c_t (c_t - 1) / 2 / (l - 1); a run is
 *   "above threshold" when 10 * sum > T * (l - 1).
 *   A perfect interval is above threshold and scores at least as high as every
 *   perfect interval inside it; the masked bases are the union of the perfect
 *   intervals that fit in a window of W bases, reported as merged
 *   [start, finish) base intervals (adjacent intervals merge).
 * The procedure below follows sdust.c step for step (window shift with the
 * suffix length L, the perfect-interval list sorted by descending start, the
 * flush of intervals that leave the window), including that an N resets only
 * the word and run length (sdust.c:156-160), not the triplet window.
 * Pinned by a brute-force model of the definition (tests/dust_model.py).
 */
#include <algorithm>
#include <cstring>
#include <deque>
#include <vector>

#include "bbt_oracle.h"

namespace {

struct Perf {
    int start, finish, r, l;
};

inline int nt4(unsigned char c) {
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

class Sdust {
  public:
    Sdust(int T, int W) : T_(T), W_(W) {
        std::memset(cv_, 0, sizeof(cv_));
        std::memset(cw_, 0, sizeof(cw_));
    }

    void run(const unsigned char *s, int n) {
        int l = 0;
        unsigned t = 0;
        for (int i = 0; i <= n; ++i) {
            const int b = i < n ? nt4(s[i]) : 4;
            if (b < 4) {
                ++l;
                t = ((t << 2) | unsigned(b)) & 63u;
                if (l >= 3) {
                    const int start = std::max(l - W_, 0) + (i + 1 - l);
                    flush(start);
                    shift(int(t));
                    if (rw_ * 10 > L_ * T_)
                        perfect(start);
                }
            } else {
                int start = std::max(l - W_ + 1, 0) + (i + 1 - l);
                while (!P_.empty())
                    flush(start++);
                l = 0;
                t = 0;
            }
        }
    }

    const std::vector<std::pair<int, int>> &intervals() const { return res_; }

  private:
    /* sdust.c:75-95 */
    void shift(int t) {
        if (int(w_.size()) >= W_ - 3 + 1) {
            const int s = w_.front();
            w_.pop_front();
            rw_ -= --cw_[s];
            if (L_ > int(w_.size())) {
                --L_;
                rv_ -= --cv_[s];
            }
        }
        w_.push_back(t);
        ++L_;
        rw_ += cw_[t]++;
        rv_ += cv_[t]++;
        if (cv_[t] * 10 > T_ * 2) {
            int s;
            do {
                s = w_[w_.size() - size_t(L_)];
                rv_ -= --cv_[s];
                --L_;
            } while (s != t);
        }
    }

    /* sdust.c:97-111: the perfect interval with the smallest start leaves the
     * window; merge it into the last reported interval or report it */
    void flush(int start) {
        if (P_.empty() || P_.back().start >= start)
            return;
        const Perf &p = P_.back();
        if (!res_.empty() && p.start <= res_.back().second)
            res_.back().second = std::max(res_.back().second, p.finish);
        else
            res_.push_back({p.start, p.finish});
        while (!P_.empty() && P_.back().start < start)
            P_.pop_back();
    }

    /* sdust.c:113-137: candidate suffixes of the window ending at the last word */
    void perfect(int start) {
        int c[64];
        std::memcpy(c, cv_, sizeof(c));
        int r = rv_, max_r = 0, max_l = 0;
        const int wn = int(w_.size());
        for (int i = wn - L_ - 1; i >= 0; --i) {
            const int t = w_[size_t(i)];
            r += c[t]++;
            const int new_r = r, new_l = wn - i - 1;
            if (new_r * 10 > T_ * new_l) {
                size_t j = 0;
                for (; j < P_.size() && P_[j].start >= i + start; ++j) {
                    const Perf &p = P_[j];
                    if (max_r == 0 || p.r * max_l > max_r * p.l) {
                        max_r = p.r;
                        max_l = p.l;
                    }
                }
                if (max_r == 0 || new_r * max_l >= max_r * new_l) {
                    max_r = new_r;
                    max_l = new_l;
                    P_.insert(P_.begin() + long(j), Perf{i + start, wn + 2 + start, new_r, new_l});
                }
            }
        }
    }

    int T_, W_;
    std::deque<int> w_;
    int cv_[64], cw_[64];
    int rv_ = 0, rw_ = 0, L_ = 0;
    std::vector<Perf> P_;
    std::vector<std::pair<int, int>> res_;
};

} // namespace

extern "C" {

int64_t orc_sdust(const char *seq, uint64_t len, int T, int W, uint32_t *starts,
                  uint32_t *finishes, uint64_t cap) {
    Sdust d(T, W);
    d.run(reinterpret_cast<const unsigned char *>(seq), int(len));
    const auto &iv = d.intervals();
    if (iv.size() > cap)
        return -1;
    for (size_t i = 0; i < iv.size(); ++i) {
        starts[i] = uint32_t(iv[i].first);
        finishes[i] = uint32_t(iv[i].second);
    }
    return int64_t(iv.size());
}

} // extern "C"
