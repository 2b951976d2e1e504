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

// background_count() for means <= 64 (one inversion chunk); -1 otherwise
// (also for caps of 2^30 photons or more, which a 32-bit counter could pass).
PDA_HD float background_single(float lam, float cap, float u_first, float u_cap) {
    if (lam > 64.0f || !(cap < 0x1p30f)) return -1.0f;
    int32_t k = 0;
    bool over = false;
    if (lam > 0.0f) {
        float pk = fexp(-lam);
        float cdf = pk;
        while (u_first > cdf) {
            ++k;
            if (float(k) > cap) {
                over = true;
                break;
            }
            pk = (pk * lam) / float(k);
            cdf = cdf + pk;
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
template <bool kFast, bool kCounts = false>
PDA_HD int step_window(const Model& M, const BinomTable* tab, const UnitKey& key, uint32_t P,
                       uint32_t stop_rel, float T, const uint32_t* tick, const uint2* link,
                       const RunArgs& a, WalkState<kFast>& st, WindowInfo* wi = nullptr) {
    static_assert(!(kFast && kCounts), "the counts form walks in the general form");
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
        x = binomial_table(int(m), s, M, *tab, A.w[2], A.w[3]);
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
template <bool kFast>
PDA_HD WalkOut walk_unit(const Model& M, const BinomTable* tab, const UnitKey& key, uint32_t P,
                         uint32_t stop_rel, float T, const uint32_t* tick, const uint2* link,
                         const RunArgs& a) {
    WalkState<kFast> st;
    int r = begin_unit<kFast>(M, key, P, T, st);
    while (r == kStepMore) r = step_window<kFast>(M, tab, key, P, stop_rel, T, tick, link, a, st);
    return WalkOut{st.acc, st.F, r == kStepReplay};
}

}  // namespace pda
