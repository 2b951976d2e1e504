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
#include "pda_math.h"
#include "pda_types.h"

namespace pda {

constexpr uint32_t kMaxWindows = 1u << 22;

// drawDisIdx (src/gen_rand.cuh:35-50) on precomputed running sums: the first
// i with cum[i] > u, else n - 1.
PDA_HD int first_above(const float* cum, int n, float u) {
    int idx = n - 1;
    bool found = false;
#pragma unroll
    for (int i = 0; i < kMaxStates; ++i) {
        if (i < n && !found && cum[i] > u) {
            idx = i;
            found = true;
        }
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
    if (!(pp * double(n) <= 30.0) || n >= kBinomTable) return -1;
    const double q = M.qq[s];
    double px = tab.qn[s][n];
    const int64_t bound = tab.bound[s][n];
    double U = uniform53(hi, lo);
    int64_t X = 0;
    while (U > px) {
        ++X;
        if (X > bound) return -1;
        U -= px;
        px = (double(n - X + 1) * pp * px) / (double(X) * q);
    }
    return M.flip[s] ? int64_t(n) - X : X;
}

// background_count() for means <= 64 (one inversion chunk); -1 otherwise.
PDA_HD float background_single(float lam, float cap, float u_first, float u_cap) {
    if (lam > 64.0f) return -1.0f;
    int64_t k = 0;
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

// One (burst, resample) unit: the body of mc_kernel (src/mc.cu:92-196).
// tick/link point at the burst's photons in LDS or in global memory.
template <bool kFast>
PDA_HD WalkOut walk_unit(const Model& M, const BinomTable* tab, const UnitKey& key, uint32_t P,
                         uint32_t stop_rel, float T, const uint32_t* tick, const uint2* link,
                         const RunArgs& a) {
    WalkOut out{0.0f, 0.0f, false};
    // Zero or one photon: every window counts nothing, E = 0.
    if (P <= 1u || !(T > 0.0f)) return out;
    const int n = M.n;
    const uint32_t last = P - 1u;
    int s = first_above(M.p0cum, n, uniform24(draw_block(key, kInitWindow, kSubMain).w[0]));
    float tsum = 0.0f;
    int64_t w0 = 0;
    uint32_t c = 0;
    for (uint32_t w = 0; T > tsum && w < kMaxWindows; ++w) {
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
        tsum = tsum + flog(uniform24(A.w[1])) / (-M.krate[s][j]);
        int64_t w1;
        if (!(tsum < T)) {
            w1 = int64_t(stop_rel);
        } else {
            const float wf = float(w0) + tsum / a.clk_p;  // src/mc.cu:132
            w1 = wf >= 0x1p62f ? (int64_t(1) << 62) : int64_t(wf);
        }
        const uint32_t w1c = w1 <= 0 ? 0u : (w1 >= int64_t(kTickSentinel) ? kTickSentinel
                                                                          : uint32_t(w1));
        // window close: first channel photon at or after w1, else the last photon
        uint32_t lo = c, hi = P;
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
        const uint2 L = link[c];
        const uint32_t f = L.x & kIndexMask;
        uint32_t cnt = 0;
        if (f < close) cnt = (L.y < close ? L.y : close) - f;
        const uint32_t colour = L.x >> 30;
        const float nd = colour == 1u ? float(cnt) : 0.0f;
        const float na = colour == 2u ? float(cnt) : 0.0f;
        const float fv = M.gmx * nd + M.omdd * na;
        out.F += fv;
        float m;
        if (!a.bg_on) {
            m = floorf(fv + 0.5f);
        } else {  // background photons (src/mc.cu:153-166)
            const Block Bk = draw_block(key, w, kSubBackground);
            const float dt = float(w1 - w0) * a.clk_p;
            float bga, bgd;
            if (kFast) {
                bga = background_single(dt * a.bg_ad, M.omdd * na, uniform24(Bk.w[0]),
                                        uniform24(Bk.w[2]));
                bgd = background_single(dt * a.bg_dd, M.gmx * nd, uniform24(Bk.w[1]),
                                        uniform24(Bk.w[3]));
                if (bga < 0.0f || bgd < 0.0f) {
                    out.replay = true;
                    return out;
                }
            } else {
                bga = background_count(dt * a.bg_ad, M.omdd * na, uniform24(Bk.w[0]),
                                       uniform24(Bk.w[2]), key, w, kSubPoissonA);
                bgd = background_count(dt * a.bg_dd, M.gmx * nd, uniform24(Bk.w[1]),
                                       uniform24(Bk.w[3]), key, w, kSubPoissonD);
            }
            m = floorf(((fv - bgd) - bga) + 0.5f);
        }
        int64_t x;
        if (kFast) {
            x = binomial_table(int(m), s, M, *tab, A.w[2], A.w[3]);
            if (x < 0) {
                out.replay = true;
                return out;
            }
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
        out.acc += float(x);
        w0 = w1;
        s = j;
        c = close;
        if (c == last) break;  // every later window is empty
    }
    return out;
}

}  // namespace pda
