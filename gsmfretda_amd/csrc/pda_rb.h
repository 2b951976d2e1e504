// pda_rb.h -- the Rao-Blackwellised (RB) form of the Monte-Carlo PDA walk.
//
// The plain walk (mc_kernel, src/mc.cu:83-197) draws, per dwell window, an
// acceptor count X_w ~ Binomial(m_w, E'_w).  Without distance broadening
// (v = 0) E'_w is the efficiency of the window's state, so the acceptor
// total of a (burst, resample) unit given its walk is a sum of one binomial
// per state, Binomial(M_s, p_s) with M_s the photon counts of the windows
// spent in state s.  The RB form keeps the walk (the same keyed draws of
// states, dwell times and background) and replaces the acceptor draw by the
// exact distribution of the total: P(a) = conv_s Binomial(M_s, p_s)(a),
// spread into the E histogram at E = a / F.  The expected histogram is smooth
// in E; its derivative in the rates is estimated by the likelihood ratio of
// the walk's draws.
//
// Host-and-device code: the kernels and the host build used by the tests run
// the same operations in the same order.
#pragma once
#include <cstdint>

#include "pda_math.h"
#include "pda_types.h"

namespace pda {

constexpr int kRbMaxStates = 4;
constexpr int kRbMaxPairs = kRbMaxStates * (kRbMaxStates - 1);
constexpr int kRbMaxCount = 4096;  // photons per burst the RB form accepts

// Per parameter set, built on the host from the Model.
struct RbModel {
    int32_t n, K;              // states, rates (n (n - 1))
    int32_t pad0[2];
    float p[kRbMaxStates];     // acceptor probability per state (Model::ep0)
    double lp[kRbMaxStates];   // log p, log (1 - p)
    double lq[kRbMaxStates];
    double dscale[kRbMaxStates];  // 1 / (p (1 - p)): d pmf / dp = pmf (k - m p) dscale
    // score of the walk's draws per window (s -> j) (likelihood ratio):
    // pair index of rate s -> j in the args' off-diagonal list, 2 / K(j, s),
    // and 1 / sum_{j' != s} K(j', s)
    int32_t pair[kRbMaxStates][kRbMaxStates];
    float two_over_k[kRbMaxStates][kRbMaxStates];
    float inv_sum[kRbMaxStates];
    // acceptor totals of a unit are at most rb_total_bound(P) for a burst of
    // P photons: c = max(gamma - crosstalk, 1 - dexdir, 0)
    float coef;
    int32_t pad1;
};

// Largest acceptor total of a unit in a burst of P photons: the windows
// count disjoint photon runs and window w contributes at most
// floor(c n_w + 1/2), which is <= n_w when c <= 1.
PDA_HD int rb_total_bound(uint32_t P, float coef) {
    if (coef <= 1.0f) return int(P);
    return int(ceil((double(coef) + 0.5) * double(P)));
}

// Args index (after the n efficiencies) of the rate s -> j, K(j, s): row j,
// column s of genMatK's row-major off-diagonal list (src/eigenhelper.cpp:4-21).
PDA_HD int rate_pair(int n, int s, int j) { return j * (n - 1) + (s < j ? s : s - 1); }

// log k! for k = 0..count, summed in order (lf[k] = lf[k-1] + log k).
PDA_HD void log_factorials(double* lf, int count) {
    lf[0] = 0.0;
    for (int k = 1; k <= count; ++k) lf[k] = lf[k - 1] + dlog(double(k));
}

// Binomial(m, p)(k) in double from log factorials; exact 0/1 at p = 0, 1.
PDA_HD double rb_binom_pmf(int m, int k, const double* lf, float p, double lp, double lq) {
    if (p <= 0.0f) return k == 0 ? 1.0 : 0.0;
    if (p >= 1.0f) return k == m ? 1.0 : 0.0;
    double l = lf[m] - lf[k];
    l = l - lf[m - k];
    l = l + double(k) * lp;
    l = l + double(m - k) * lq;
    return dexp(l);
}

// One output of Z = Y (*) X: sum_{k} Y[k] X[a - k], k ascending over the
// overlap, float accumulation from 0.
template <class T>
PDA_HD T rb_conv_at(const T* Y, int ylen, const T* X, int xlen, int a) {
    int k0 = a - (xlen - 1);
    if (k0 < 0) k0 = 0;
    const int k1 = a < ylen - 1 ? a : ylen - 1;
    T z = T(0);
    for (int k = k0; k <= k1; ++k) z = z + Y[k] * X[a - k];
    return z;
}

// Histogram bin of a acceptors out of F corrected photons, as the plain
// walk bins acc / F (-1: dropped).
PDA_HD int rb_bin(int a, float F, int B) {
    const float E = F > 0.0f ? float(a) / F : 0.0f;
    if (!(E >= 0.0f && E < 1.0f)) return -1;
    const int b = int(floor(double(E) * double(B)));
    return b < B ? b : -1;
}

// Binomial pmfs tabulated per parameter set for counts below kRbPmfTable:
// entry (s, m, k) at s * kRbPmfStride + m (m + 1) / 2 + k, the float value of
// rb_binom_pmf(m, k, ...).
constexpr int kRbPmfTable = 256;
constexpr int kRbPmfStride = kRbPmfTable * (kRbPmfTable + 1) / 2;
PDA_HD int rb_pmf_index(int s, int m, int k) { return s * kRbPmfStride + m * (m + 1) / 2 + k; }

// Pmf memo: a unit's pmf depends only on its photon counts per state (and
// the set's efficiencies), so the pmfs of every count vector with acceptor
// total Mt < kRbMemoCap are computed once per forward and looked up by the
// units; units with larger totals compute their own.
constexpr int kRbMemoCap = 48;
// Units that compute their own pmf are listed in two lists: totals below
// kRbBigSplit (pmfs of at most kRbBigSplit entries: small LDS buffers, many
// workgroups per CU) and the rest (buffers for the largest total any burst
// of the run allows).
constexpr int kRbBigSplit = 256;

// C(x, y) for x >= 0 (0 when x < y); y <= kRbMaxStates in closed form with
// constant divisors (each product of y consecutive integers is divisible
// by y!, and stays below 2^32 for the x < 64 used here).
PDA_HD uint32_t rb_choose(int x, int y) {
    const uint32_t v = uint32_t(x);
    switch (y) {
        case 0: return 1u;
        case 1: return v;
        case 2: return v * (v - 1u) / 2u;
        case 3: return v * (v - 1u) * (v - 2u) / 6u;
        case 4: return v * (v - 1u) * (v - 2u) * (v - 3u) / 24u;
        default: {
            uint32_t c = 1;
            for (int i = 0; i < y; ++i) c = c * uint32_t(x - i) / uint32_t(i + 1);
            return x < y ? 0u : c;
        }
    }
}

// Lexicographic rank of the count vector (m_0 .. m_{n-1}) among all
// n-vectors of counts >= 0 with sum <= L: for each position i, the vectors
// that agree before i and have a smaller count at i, counted by the hockey
// stick identity sum_{w <= W} C(w + k, k) = C(W + k + 1, k + 1).
template <int N>
PDA_HD uint32_t rb_rank_n(const int* m, int L) {
    uint32_t r = 0;
    int left = L;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const int k = N - 1 - i;
        r += rb_choose(left + k + 1, k + 1) - rb_choose(left - m[i] + k + 1, k + 1);
        left -= m[i];
    }
    return r;
}
PDA_HD uint32_t rb_rank(const int* m, int n, int L) {
    switch (n) {
        case 1: return rb_rank_n<1>(m, L);
        case 2: return rb_rank_n<2>(m, L);
        case 3: return rb_rank_n<3>(m, L);
        default: return rb_rank_n<4>(m, L);
    }
}

// Fixed-point weight of a probability (2^-32 units), rounded half up; the
// 32-bit conversion whenever the value fits (P < 1), the same result.
PDA_HD uint64_t rb_fixed(float P) {
    const double x = double(P) * 4294967296.0 + 0.5;
    return x < 4294967296.0 ? uint64_t(uint32_t(x)) : uint64_t(x);
}

// Summary of one unit's walk: per-state photon counts (16 bits each), F.
PDA_HD uint2 rb_pack(const uint32_t* cnt) {
    uint2 v;
    v.x = (cnt[0] & 0xFFFFu) | (cnt[1] << 16);
    v.y = (cnt[2] & 0xFFFFu) | (cnt[3] << 16);
    return v;
}
PDA_HD uint32_t rb_count(uint2 v, int s) {
    const uint32_t w = s < 2 ? v.x : v.y;
    return (s & 1) ? (w >> 16) : (w & 0xFFFFu);
}

}  // namespace pda
