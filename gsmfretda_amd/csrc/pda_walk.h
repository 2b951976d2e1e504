// pda_walk.h -- one (burst, resample) unit of the Monte-Carlo PDA walk:
// the body of gSMFRETda's mc_kernel (src/mc.cu:92-196) over the resident
// photon layout of pda_types.h.  Host-and-device code: the gfx950 walk
// kernels run it per lane; a host build of the same source checks the
// algorithm against the oracle without a GPU.
//
// Two forms of the same computation:
//   * general: every branch of the reference (distance broadening, BTPE,
//     inversion restarts, Poisson chunking for large background means);
//   * fast: the common case only -- no broadening (v == 0, so the broadened
//     efficiency and its binomial inversion constants are per-state tables),
//     inversion without restarts, background means <= 64.  A unit that would
//     leave the common case is abandoned (replay = true) and recomputed from
//     scratch by the general form; the keyed random streams make the replay
//     produce exactly the draws the fast form would have consumed.
#pragma once
#include <type_traits>

#include "pda_math.h"
#include "pda_types.h"

namespace pda {

constexpr uint32_t kMaxWindows = 1u << 22;

// drawDisIdx (src/gen_rand.cuh:35-50) on precomputed running sums: the first
// i with cum[i] > u, else n - 1.  n is the same for every lane (the model's),
// so the loop runs n passes with no early exit; unrolled to kMaxStates it
// would keep one compare mask per bound live in scalar registers.
PDA_HD int first_above(const float* cum, int n, float u) {
    int idx = n - 1;
    bool found = false;
#pragma unroll 1
    for (int i = 0; i < n; ++i) {
        const bool hit = !found && cum[i] > u;
        idx = hit ? i : idx;
        found = found || hit;
    }
    return idx;
}

struct WalkOut {
    float acc;    // sum of acceptor draws (float accumulation, as mcE[t] += ai)
    float F;      // sum of corrected photon counts
    bool replay;  // fast form only: the unit needs the general form
};

// Binomial(n, ep0[s]) by inversion from the per-state tables; -1 when the
// general sampler is needed (BTPE regime, n beyond the table, a restart).
// Same operations as binomial()/binomial_inversion() in pda_math.h.
PDA_HD int64_t binomial_table(int n, int s, const Model& M, const BinomTable& tab, uint32_t hi,
                              uint32_t lo) {
    if (n <= 0) return 0;
    const double pp = M.pp[s];
    if (n >= kBinomTable || !(n < kInversionMaxN || pp * double(n) <= 30.0)) return -1;
    const double q = M.qq[s];
    double px = tab.qn[s][n];
    const int bound = tab.bound[s][n];
    double U = uniform53(hi, lo);
    int X = 0;  // < kBinomTable: 32-bit counters convert to double exactly
    while (U > px) {
        ++X;
        if (X > bound) return -1;
        U -= px;
        px = (double(n - X + 1) * pp * px) / (double(X) * q);
    }
    return M.flip[s] ? int64_t(n - X) : int64_t(X);
}

static_assert(kInversionMaxN >= kBinomTable, "every tabulated count draws by inversion");

constexpr int64_t kU53Max = (int64_t(1) << 53) - 1;  // largest 53-bit grid index

// Rows of the guided inversion: binomial_table's X as a function of the
// uniform's grid index u (U = u 2^-53 exactly, uniform53).
//
// The loop keeps U_0 = U, U_{i+1} = fl(U_i - px_i) and goes past i while
// U_i > px_i, where px_i does not depend on U.  Rounding is monotone, so
// every U_i is a non-decreasing function of U, and so is X(U): X >= k + 1
// exactly for U at or above some grid point.  Locating it: every subtraction
// rounds a value in (0, 1), so U_i is within i 2^-54 of U - S_{i-1}, with
// S_k = px_0 + ... + px_k exact; hence U > S_k + k 2^-54 gives X >= k + 1
// and U <= S_k - k 2^-54 gives X <= k.
//
// The rows hold C_k, an approximation of S_k computed in parallel (below),
// with |C_k - S_k| + k 2^-54 < 2^-42 = kGuideBand grid steps, and
//   up_k = floor(C_k 2^53) + kGuideBand + 1:  u >= up_k  =>  X >= k + 1
//   lo_k = ceil(C_k 2^53) - kGuideBand - 1:   u <= lo_k  =>  X <= k
// The error bound, with eps = 2^-53 and k <= 127: the loop's px_k is
// qn rho_0 ... rho_{k-1} (rho_i = (m - i) p / ((i + 1) q), qn = q^m as
// tabulated) times 1 + t with |t| <= gamma(4k) (four roundings a step); here
// each ratio is rounded three times and the prefix product of k of them and
// qn k times more, also gamma(4k); both are within 2 gamma(4k + 1) < 1019 eps
// relative of the same exact product, whose sum over k is qn / q^m <= 1.0001.
// Any order of summing k + 1 such terms adds at most gamma(k) < 128 eps.
// With k 2^-54 <= 64 eps that is < 1211 eps < 2048 eps.
//
// A row ends at the bound (then a sentinel makes X = bound + 1 a replay, as
// the loop's "X > bound") or at the first k whose lo_k covers the whole grid
// (X <= k for every uniform).  guide[j] is the number of leading entries
// with up <= j 2^47, the least X of any u in bucket j.
constexpr int64_t kGuideBand = 2048;

// The device builds a row with one wave (guide_kernel): chunks of 64
// entries, lane l of chunk c holding k = 64 c + l; ratio r_k (r_0 = 1),
// Hillis-Steele inclusive scans (offsets 1, 2, ..., 32) of the ratios'
// product and of px_k = qn P_k, each chunk continuing from the last lane of
// the one before.  This host function performs the same operations in the
// same order, so its rows equal the device's bit for bit.
inline int build_guide_row_host(const Model& M, const BinomTable& tab, int s, int m, ThrPair* thr,
                                uint8_t* guide) {
    const double pp = M.pp[s];
    const double q = M.qq[s];
    const int bound = tab.bound[s][m];
    const double qn = tab.qn[s][m];
    double carryP = 1.0, carryC = 0.0;
    int end = -1;
    bool covered = false;
    for (int c = 0; c * kWave < kBinomTable && end < 0; ++c) {
        double P[kWave], C[kWave], t[kWave];
        for (int l = 0; l < kWave; ++l) {
            const int k = c * kWave + l;
            P[l] = k == 0 ? 1.0 : (double(m - k + 1) * pp) / (double(k) * q);
        }
        for (int off = 1; off < kWave; off <<= 1) {
            for (int l = 0; l < kWave; ++l) t[l] = l >= off ? P[l] * P[l - off] : P[l];
            for (int l = 0; l < kWave; ++l) P[l] = t[l];
        }
        for (int l = 0; l < kWave; ++l) C[l] = qn * (P[l] * carryP);
        for (int off = 1; off < kWave; off <<= 1) {
            for (int l = 0; l < kWave; ++l) t[l] = l >= off ? C[l] + C[l - off] : C[l];
            for (int l = 0; l < kWave; ++l) C[l] = t[l];
        }
        for (int l = 0; l < kWave && end < 0; ++l) {
            const int k = c * kWave + l;
            const double cs = (C[l] + carryC) * 0x1p53;
            const int64_t lo = int64_t(ceil(cs)) - kGuideBand - 1;
            thr[k] = ThrPair{int64_t(floor(cs)) + kGuideBand + 1, lo};
            if (lo >= kU53Max) covered = true;
            if (covered || k >= bound) end = k;
        }
        carryP = P[kWave - 1] * carryP;
        carryC = C[kWave - 1] + carryC;
    }
    thr[end + 1] = ThrPair{INT64_MAX, covered ? INT64_MAX : int64_t(-1)};
    for (int j = 0; j < kGuideBuckets; ++j) {
        const int64_t u0 = int64_t(j) << kGuideShift;
        int x = 0;
        while (x <= end && thr[x].up <= u0) ++x;
        guide[j] = uint8_t(x);
    }
    return end + 1;
}

// binomial_table by the thresholds: the same X for every uniform, or -1
// where binomial_table returns -1 and where u falls between lo_X and up_X
// (2 kGuideBand + 2 grid points around each threshold, ~5e-13 of the grid:
// the general form then runs the loop).
// Two dependent loads instead of X + 1 loop passes whose count differs from
// lane to lane.
PDA_HD int64_t binomial_guided(int n, int s, const Model& M, const BinomGuide& G, uint32_t hi,
                               uint32_t lo) {
    if (n <= 0) return 0;
    if (n >= kBinomTable) return -1;
    const int64_t u = (int64_t(hi >> 5) << 26) | int64_t(lo >> 6);
    const size_t row = size_t(s) * kBinomTable + size_t(n);
    int X = G.guide[row * kGuideBuckets + size_t(u >> kGuideShift)];
    const ThrPair* t = G.thr + guide_row_offset(s, n);
    ThrPair e = t[X];
    while (u >= e.up) e = t[++X];
    if (u > e.lo) return -1;
    return M.flip[s] ? int64_t(n - X) : int64_t(X);
}

PDA_HD int64_t binomial_fast(int n, int s, const Model& M, const BinomTable& tab, uint32_t hi,
                             uint32_t lo) {
    return binomial_table(n, s, M, tab, hi, lo);
}

PDA_HD int64_t binomial_fast(int n, int s, const Model& M, const BinomGuide& g, uint32_t hi,
                             uint32_t lo) {
    return binomial_guided(n, s, M, g, hi, lo);
}

// Division of a unit index by the resample count without a divide
// instruction sequence: for d >= 1, l = ceil(log2 d) and m = ceil(2^(31+l) / d)
// (< 2^32), floor(u m / 2^(31+l)) = floor(u / d) for every u < 2^31 (m d =
// 2^(31+l) + e with 0 <= e < d, so u m / 2^(31+l) exceeds u / d by less than
// u / 2^(31+l) * e / d < 2^-l <= 1 / d, which cannot reach the next integer).
// Unit indices stay below 2^31 (every run checks it).  Built once per wave.
struct UnitDiv {
    uint32_t m, sh;
};
PDA_HD UnitDiv make_unit_div(uint32_t d) {
    uint32_t l = 0;
    while (l < 32u && (uint64_t(1) << l) < uint64_t(d)) ++l;
    const uint64_t m = ((uint64_t(1) << (31u + l)) + uint64_t(d) - 1u) / uint64_t(d);
    return UnitDiv{uint32_t(m), 31u + l};
}
PDA_HD uint32_t unit_div(uint32_t u, const UnitDiv& q) {
    return uint32_t((uint64_t(u) * uint64_t(q.m)) >> q.sh);
}

template <class T>
struct NoDeduce {
    using type = T;
};

// background_count() for means <= 64 (one inversion chunk); -1 otherwise
// (also for caps of 2^30 photons or more, which a 32-bit counter could pass).
// The inversion's first two steps run for every lane without a branch: their
// divisions by 1 and 2 are exact (the identity and a power-of-two scaling),
// so pk * lam and (pk * lam) * 0.5 round to the loop's values; the loop then
// continues from k = 3 only for the lanes whose uniform lies beyond the
// cumulative probability of 2 photons, so a wave no longer runs the loop as
// often as its longest draw needs.
PDA_HD float background_single(float lam, float cap, float u_first, float u_cap) {
    if (lam > 64.0f || !(cap < 0x1p30f)) return -1.0f;
    int32_t k = 0;
    bool over = false;
    if (lam > 0.0f) {
        const float pk0 = fexp(-lam);
        const float pk1 = pk0 * lam;
        const float cdf1 = pk0 + pk1;
        const float pk2 = (pk1 * lam) * 0.5f;
        const float cdf2 = cdf1 + pk2;
        float pk = pk0, cdf = pk0;
        bool go = u_first > pk0;
        if (go) {  // k = 1
            k = 1;
            over = 1.0f > cap;
            pk = pk1;
            cdf = cdf1;
            go = !over && u_first > cdf1;
        }
        if (go) {  // k = 2
            k = 2;
            over = 2.0f > cap;
            pk = pk2;
            cdf = cdf2;
            go = !over && u_first > cdf2;
        }
        while (go) {
            ++k;
            if (float(k) > cap) {
                over = true;
                break;
            }
            pk = (pk * lam) / float(k);
            cdf = cdf + pk;
            go = u_first > cdf;
        }
    }
    if (!over && float(k) <= cap) return float(k);
    return floorf(cap * u_cap);
}

// Walk state of one unit between windows.  The fast form keeps the window
// ends in 32 bits (ticks < 2^31; a larger end is replayed), the general
// form in 64 bits as the reference.
template <bool kFast>
struct WalkState {
    using Tick = typename std::conditional<kFast, uint32_t, int64_t>::type;
    int s;        // current state
    float tsum;   // mcSpendTime
    Tick w0;      // start tick of the current window (bin0clk_t)
    uint32_t c;   // photon cursor (binIdxStart)
    uint32_t w;   // window index
    float acc, F;
};

enum StepResult : int { kStepMore = 0, kStepDone = 1, kStepReplay = 2 };

// What one window drew, for the counts form (kCounts): the state it dwelt
// in, the next state, the dwell time and the photon count the binomial
// acceptor draw would have taken (int(m), <= 0 draws nothing).
struct WindowInfo {
    int s, j;
    float tau;
    int32_t m;
};

// Start of a unit (src/mc.cu:92-116): initial state from genMatP's inflow.
// Returns kStepDone when the unit has no window that counts anything
// (zero or one photon, or T <= 0: E = 0).
template <bool kFast>
PDA_HD int begin_unit(const Model& M, const UnitKey& key, uint32_t P, float T,
                      WalkState<kFast>& st) {
    st.s = 0;
    st.tsum = 0.0f;
    st.w0 = 0;
    st.c = 0;
    st.w = 0;
    st.acc = 0.0f;
    st.F = 0.0f;
    if (P <= 1u || !(T > 0.0f)) return kStepDone;
    st.s = M.init_state >= 0
               ? M.init_state
               : first_above(M.p0cum, M.n,
                             uniform24(draw_block(key, kInitWindow, kSubMain).w[0]));
    return kStepMore;
}

// One dwell window of the walk (one pass of the while loop, src/mc.cu:117-191).
// tick/link point at the burst's photons in LDS or in global memory.
// kCounts (general form only): no acceptor draw; the window is described in
// *wi instead, for the Rao-Blackwellised form that sums the binomial pmf.
// Tab: BinomTable (inversion loop) or BinomGuide (thresholds) for the fast
// form's binomial draw; unused by the general form.
template <bool kFast, bool kCounts = false, class Tab = BinomTable>
PDA_HD int step_window(const Model& M, const typename NoDeduce<Tab>::type* tab, const UnitKey& key,
                       uint32_t P,
                       uint32_t stop_rel, float T, const uint32_t* tick, const uint2* link,
                       const RunArgs& a, WalkState<kFast>& st, WindowInfo* wi = nullptr) {
    using Tick = typename WalkState<kFast>::Tick;
    const int n = M.n;
    const int s = st.s;
    const uint32_t w = st.w;
    const uint32_t last = P - 1u;
    const Block A = draw_block(key, w, kSubMain);
    // next state (drawJ_Si2Sj, src/gen_rand.cuh:78-89); forced for n == 2
    int j;
    if (n == 2) {
        j = 1 - s;
    } else {
        const int jj = first_above(M.nscum[s], n - 1, uniform24(A.w[0]));
        j = jj >= s ? jj + 1 : jj;
    }
    // dwell at the chosen transition's rate (drawTau, src/mc.cu:122)
    const float tau = flog(uniform24(A.w[1])) / (-M.krate[s][j]);
    const float tsum = st.tsum + tau;
    Tick w1;
    uint32_t w1c;
    if (!(tsum < T)) {
        w1 = Tick(stop_rel);
        w1c = stop_rel;
    } else {
        const float wf = float(st.w0) + tsum / a.clk_p;  // src/mc.cu:132
        if (kFast) {
            if (wf >= 0x1p31f) return kStepReplay;
            w1 = Tick(uint32_t(wf));
            w1c = uint32_t(w1);
        } else {
            const int64_t w64 = wf >= 0x1p62f ? (int64_t(1) << 62) : int64_t(wf);
            w1 = Tick(w64);
            w1c = w64 <= 0 ? 0u : (w64 >= int64_t(kTickSentinel) ? kTickSentinel : uint32_t(w64));
        }
    }
    // window close: first channel photon at or after w1, else the last photon
    uint32_t lo = st.c, hi = P;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tick[mid] < w1c) lo = mid + 1u;
        else hi = mid;
    }
    uint32_t close = last;
    if (lo < P) {
        const uint32_t f1 = link[lo].x & kIndexMask;
        if (f1 < close) close = f1;
    }
    // sticky-flag count over [c, close) (src/mc.cu:171-180)
    const uint2 L = link[st.c];
    const uint32_t f = L.x & kIndexMask;
    uint32_t cnt = 0;
    if (f < close) cnt = (L.y < close ? L.y : close) - f;
    const uint32_t colour = L.x >> 30;
    const float nd = colour == 1u ? float(cnt) : 0.0f;
    const float na = colour == 2u ? float(cnt) : 0.0f;
    const float fv = M.gmx * nd + M.omdd * na;
    float m;
    if (!a.bg_on) {
        m = floorf(fv + 0.5f);
    } else {  // background photons (src/mc.cu:153-166)
        const Block Bk = draw_block(key, w, kSubBackground);
        float dt;
        if (kFast) dt = float(int32_t(uint32_t(w1) - uint32_t(st.w0))) * a.clk_p;
        else dt = float(int64_t(w1) - int64_t(st.w0)) * a.clk_p;
        float bga, bgd;
        if (kFast) {
            bga = background_single(dt * a.bg_ad, M.omdd * na, uniform24(Bk.w[0]),
                                    uniform24(Bk.w[2]));
            bgd = background_single(dt * a.bg_dd, M.gmx * nd, uniform24(Bk.w[1]),
                                    uniform24(Bk.w[3]));
            if (bga < 0.0f || bgd < 0.0f) return kStepReplay;
        } else {
            bga = background_count(dt * a.bg_ad, M.omdd * na, uniform24(Bk.w[0]),
                                   uniform24(Bk.w[2]), key, w, kSubPoissonA);
            bgd = background_count(dt * a.bg_dd, M.gmx * nd, uniform24(Bk.w[1]),
                                   uniform24(Bk.w[3]), key, w, kSubPoissonD);
        }
        m = floorf(((fv - bgd) - bga) + 0.5f);
    }
    int64_t x;
    if (kCounts) {
        x = 0;
        wi->s = s;
        wi->j = j;
        wi->tau = tau;
        wi->m = int32_t(m);
    } else if (kFast) {
        x = binomial_fast(int(m), s, M, *tab, A.w[2], A.w[3]);
        if (x < 0) return kStepReplay;
    } else {
        // broadened efficiency (drawE, src/gen_rand.cuh:63-68)
        float de;
        if (M.v[s] == 0.0f) {
            de = M.ep0[s];
        } else {
            const Block C = draw_block(key, w, kSubNormal);
            de = distance_efficiency(gauss(C.w[0], C.w[1]) * M.v[s] + M.base[s], M.r0);
        }
        x = binomial(int(m), double(de), A.w[2], A.w[3], key, w);
    }
    st.F += fv;
    st.acc += float(x);
    st.tsum = tsum;
    st.w0 = w1;
    st.s = j;
    st.c = close;
    st.w = w + 1u;
    // loop condition of the next window; every window after the cursor
    // reached the last photon is empty
    if (close == last || !(T > tsum) || st.w >= kMaxWindows) return kStepDone;
    return kStepMore;
}

// One (burst, resample) unit: the body of mc_kernel (src/mc.cu:92-196).
template <bool kFast, class Tab = BinomTable>
PDA_HD WalkOut walk_unit(const Model& M, const typename NoDeduce<Tab>::type* tab, const UnitKey& key,
                         uint32_t P, uint32_t stop_rel, float T, const uint32_t* tick,
                         const uint2* link, const RunArgs& a) {
    WalkState<kFast> st;
    int r = begin_unit<kFast>(M, key, P, T, st);
    while (r == kStepMore)
        r = step_window<kFast, false, Tab>(M, tab, key, P, stop_rel, T, tick, link, a, st);
    return WalkOut{st.acc, st.F, r == kStepReplay};
}

}  // namespace pda
