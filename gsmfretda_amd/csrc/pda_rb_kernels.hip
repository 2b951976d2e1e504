// pda_rb_kernels.hip -- gfx950 kernels of the Rao-Blackwellised (RB) PDA form.
//
//   rb_walk_kernel   the walk of every (burst, resample) unit with the keyed
//                    draws of the plain kernel (src/mc.cu:92-191) but no
//                    acceptor draw; per unit it stores the photon counts per
//                    state, F, and (for rate gradients) the score of the
//                    walk's draws
//   rb_pmf_kernel    one wave per burst: for each of its units, the pmf of
//                    the acceptor total (a convolution of one binomial per
//                    state, in LDS, lanes over outcomes) spread into the E
//                    histogram as 2^-32 fixed point -- or, for the adjoint,
//                    the cotangent-weighted pmf and its derivatives
//   rb_finish_kernel reduces the histogram rows, computes chi2 on the
//                    expected counts, resets the counters
#include <hip/hip_runtime.h>

#include "pda_math.h"
#include "pda_rb.h"
#include "pda_types.h"
#include "pda_walk.h"

namespace pda {

constexpr uint32_t kRbPmfCounter = 8;  // work[8]: bursts taken by rb_pmf_kernel

template <bool kScores>
__global__ __launch_bounds__(kWave) void rb_walk_kernel(
    const BurstDesc* __restrict__ bursts, const uint32_t* __restrict__ ticks,
    const uint2* __restrict__ links, const ModelBlob* __restrict__ blob,
    const RbModel* __restrict__ rbm, const RunArgs a, uint2* __restrict__ cnt_out,
    float* __restrict__ F_out, float* __restrict__ S_out, uint32_t* __restrict__ work) {
    __shared__ Model sM;
    __shared__ RbModel sR;
    __shared__ uint32_t s_cnt[kRbMaxStates][kWave];
    __shared__ float s_S[kRbMaxPairs][kWave];
    const uint32_t lane = threadIdx.x;
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(&blob->m);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&sM);
        for (uint32_t i = lane; i < sizeof(Model) / 4; i += kWave) dst[i] = src[i];
        const uint32_t* src2 = reinterpret_cast<const uint32_t*>(rbm);
        uint32_t* dst2 = reinterpret_cast<uint32_t*>(&sR);
        for (uint32_t i = lane; i < sizeof(RbModel) / 4; i += kWave) dst2[i] = src2[i];
    }
    __syncthreads();
    const int n = sR.n, K = sR.K;

    // the work-stealing loop of walk_kernel (general form, no replays)
    const uint32_t queue = blockIdx.x % a.queues;
    const uint32_t q_begin = uint32_t(uint64_t(a.nunits) * queue / a.queues);
    const uint32_t q_end = uint32_t(uint64_t(a.nunits) * (queue + 1u) / a.queues);
    uint32_t pool_next = 0, pool_end = 0;
    bool busy = false;
    uint32_t u = 0;
    WalkState<false> st{};
    UnitKey key{a.seed, a.set, 0u, 0u};
    BurstDesc d{0u, 0u, 0u, 0.0f};
    bool drained = false;
    auto store = [&](uint32_t unit) {
        uint2 packed;  // rb_pack of the four counts (states >= n hold 0)
        packed.x = (s_cnt[0][lane] & 0xFFFFu) | (s_cnt[1][lane] << 16);
        packed.y = (s_cnt[2][lane] & 0xFFFFu) | (s_cnt[3][lane] << 16);
        cnt_out[unit] = packed;
        F_out[unit] = st.F;
        if (kScores)
            for (int i = 0; i < K; ++i) S_out[size_t(unit) * size_t(K) + i] = s_S[i][lane];
    };
    while (true) {
        bool need = !busy;
        while (!drained) {
            const uint64_t want = __ballot(need);
            if (want == 0) break;
            if (pool_next >= pool_end) {
                uint32_t off = 0;
                if (lane == 0) off = atomicAdd(&work[queue], a.batch);
                off = __shfl(off, 0);
                if (off >= q_end - q_begin) {
                    drained = true;
                    break;
                }
                pool_next = q_begin + off;
                pool_end = pool_next + a.batch < q_end ? pool_next + a.batch : q_end;
            }
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                uint32_t(want >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(want), 0u));
            const uint32_t avail = pool_end - pool_next;
            if (need && rank < avail) {
                u = pool_next + rank;
                const uint32_t bl = u / a.R;
                const uint32_t b = a.begin + bl;
                d = bursts[b];
                key.r = u - bl * a.R;
                key.b = b;
                for (int s = 0; s < kRbMaxStates; ++s) s_cnt[s][lane] = 0u;
                if (kScores)
                    for (int i = 0; i < K; ++i) s_S[i][lane] = 0.0f;
                if (begin_unit<false>(sM, key, d.P, d.T, st) == kStepMore) {
                    busy = true;
                    need = false;
                } else {
                    store(u);
                }
            }
            const uint32_t taken = uint32_t(__popcll(want));
            pool_next += taken < avail ? taken : avail;
        }
        if (!__any(busy)) break;
        if (busy) {
            WindowInfo wi;
            const int r = step_window<false, true>(sM, nullptr, key, d.P, d.stop_rel, d.T,
                                                   ticks + d.off, links + d.off, a, st, &wi);
            if (wi.m > 0) s_cnt[wi.s][lane] += uint32_t(wi.m);
            if (kScores) {
                // d/dK(j', s) log[P(j | s) K(j, s) exp(-K(j, s) tau)]
                //   = [j' = j] (2 / K(j, s) - tau) - 1 / sum_{j''} K(j'', s)
                s_S[sR.pair[wi.s][wi.j]][lane] += sR.two_over_k[wi.s][wi.j] - wi.tau;
                for (int j2 = 0; j2 < n; ++j2)
                    if (j2 != wi.s) s_S[sR.pair[wi.s][j2]][lane] -= sR.inv_sum[wi.s];
            }
            if (r != kStepMore) {
                store(u);
                busy = false;
            }
        }
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Z[0 .. xl + ylen - 1) = Y (*) X, lanes over outputs.
__device__ __forceinline__ void conv_into(float* Z, const float* Y, int ylen, const float* X,
                                          int xl, uint32_t lane) {
    const int nl = xl + ylen - 1;
    for (int q = int(lane); q < nl; q += kWave) Z[q] = rb_conv_at(Y, ylen, X, xl, q);
}

// One wave per burst (taken from work[8]); its R units in order.  Dynamic
// LDS: kGrad ? 8 : 3 buffers of `cap` floats (X, free, Y [, dY, D0..D3]).
template <bool kGrad>
__global__ __launch_bounds__(kWave) void rb_pmf_kernel(
    const RbModel* __restrict__ rbm, const double* __restrict__ lf, const RunArgs a,
    const uint2* __restrict__ cnt_in, const float* __restrict__ F_in,
    const float* __restrict__ S_in, const double* __restrict__ g,
    unsigned long long* __restrict__ hist_acc, double* __restrict__ grad,
    uint32_t* __restrict__ work, int cap) {
    extern __shared__ float s_buf[];
    __shared__ unsigned long long s_h[kMaxBins];
    __shared__ double s_g[kMaxBins];
    __shared__ RbModel sR;
    const uint32_t lane = threadIdx.x;
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(rbm);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&sR);
        for (uint32_t i = lane; i < sizeof(RbModel) / 4; i += kWave) dst[i] = src[i];
        for (uint32_t i = lane; i < uint32_t(a.B); i += kWave) {
            if (kGrad) s_g[i] = g[i];
            else s_h[i] = 0ull;
        }
    }
    __syncthreads();
    const int n = sR.n, K = sR.K, B = a.B;
    const uint32_t nb = a.nunits / a.R;
    double gp_acc = 0.0, gk_acc = 0.0;  // lane s: d/dp_s; lane i: d/dk_i
    while (true) {
        uint32_t bl = 0;
        if (lane == 0) bl = atomicAdd(&work[kRbPmfCounter], 1u);
        bl = __shfl(bl, 0);
        if (bl >= nb) break;
        double H = 0.0, Sg = 0.0, Ac = 0.0;
        for (uint32_t r = 0; r < a.R; ++r) {
            const uint32_t u = bl * a.R + r;
            const uint2 c = cnt_in[u];
            const float F = F_in[u];
            // buffer roles (wave-uniform indices into s_buf)
            int bX = 0, bFree = 1, bY = 2, bdY = 3;
            int bD[kRbMaxStates] = {4, 5, 6, 7};
            uint32_t dmask = 0;
            int xl = 1;
            if (lane == 0) s_buf[bX * cap] = 1.0f;
#pragma unroll
            for (int s = 0; s < kRbMaxStates; ++s) {
                const int m = s < n ? int(rb_count(c, s)) : 0;
                if (m == 0) continue;
                float* Y = s_buf + bY * cap;
                float* dY = s_buf + bdY * cap;
                for (int k = int(lane); k <= m; k += kWave) {
                    const double b = rb_binom_pmf(m, k, lf, sR.p[s], sR.lp[s], sR.lq[s]);
                    Y[k] = float(b);
                    if (kGrad)
                        dY[k] = float(b * ((double(k) - double(m) * double(sR.p[s])) * sR.dscale[s]));
                }
                __syncthreads();
                if (xl == 1) {  // X = [1]: the product is Y itself
                    const int t = bX;
                    bX = bY;
                    bY = t;
                    if (kGrad) {
                        const int t2 = bD[s];
                        bD[s] = bdY;
                        bdY = t2;
                    }
                    xl = m + 1;
                    dmask |= 1u << s;
                    continue;
                }
                const float* X = s_buf + bX * cap;
                if (kGrad) {
                    conv_into(s_buf + bD[s] * cap, dY, m + 1, X, xl, lane);  // D_s = X (*) dB_s
                    for (int t = 0; t < kRbMaxStates; ++t) {
                        if (!(dmask & (1u << t))) continue;
                        conv_into(s_buf + bFree * cap, Y, m + 1, s_buf + bD[t] * cap, xl, lane);
                        __syncthreads();
                        const int t2 = bD[t];
                        bD[t] = bFree;
                        bFree = t2;
                    }
                }
                conv_into(s_buf + bFree * cap, Y, m + 1, X, xl, lane);
                __syncthreads();
                const int t = bX;
                bX = bFree;
                bFree = t;
                xl += m;
                dmask |= 1u << s;
            }
            const float* P = s_buf + bX * cap;
            if (!kGrad) {
                for (int q = int(lane); q < xl; q += kWave) {
                    const int bin = rb_bin(q, F, B);
                    if (bin >= 0) atomicAdd(&s_h[bin], (unsigned long long)rb_fixed(P[q]));
                }
            } else {
                double hp = 0.0, dp[kRbMaxStates] = {0.0, 0.0, 0.0, 0.0};
                for (int q = int(lane); q < xl; q += kWave) {
                    const int bin = rb_bin(q, F, B);
                    const double G = bin >= 0 ? s_g[bin] : 0.0;
                    hp += G * double(P[q]);
#pragma unroll
                    for (int t = 0; t < kRbMaxStates; ++t)
                        if (dmask & (1u << t)) dp[t] += G * double(s_buf[bD[t] * cap + q]);
                }
                const double h = wave_sum(hp);
#pragma unroll
                for (int t = 0; t < kRbMaxStates; ++t) {
                    if (!(dmask & (1u << t))) continue;
                    const double v = wave_sum(dp[t]);
                    if (int(lane) == t) gp_acc += v;
                }
                H += h;
                if (int(lane) < K) {
                    const double si = double(S_in[size_t(u) * size_t(K) + lane]);
                    Sg += si;
                    Ac += h * si;
                }
            }
            __syncthreads();
        }
        // likelihood-ratio estimate with the leave-one-out baseline of the
        // burst's other resamples: R / (R - 1) sum_r (h_r - mean h) S_r
        if (kGrad && int(lane) < K && a.R > 1)
            gk_acc += (double(a.R) / double(a.R - 1)) * (Ac - H * Sg / double(a.R));
    }
    if (!kGrad) {
        __syncthreads();
        unsigned long long* row = hist_acc + size_t(blockIdx.x & 7u) * kMaxBins;
        for (uint32_t i = lane; i < uint32_t(B); i += kWave)
            if (s_h[i]) atomicAdd(&row[i], s_h[i]);
    } else {
        if (int(lane) < n) atomicAdd(&grad[lane], gp_acc);
        if (int(lane) < K) atomicAdd(&grad[n + int(lane)], gk_acc);
    }
}

// Device result of an RB run, followed by uint64 M[B] (2^-32 units).
struct RbResultHeader {
    double chi2;
    int32_t effN;
    int32_t B;
};

// One workgroup: sums the 8 rows (leaving them zero), chi2 of the expected
// counts M_b / R against the observed histogram with the reference's rules
// (src/streamWorker.cpp:197-206), counter reset.
__global__ __launch_bounds__(kWave) void rb_finish_kernel(
    const RunArgs a, const float* __restrict__ obs, unsigned long long* __restrict__ hist_acc,
    uint32_t* __restrict__ work, RbResultHeader* __restrict__ res) {
    __shared__ double s_t[kMaxBins];
    __shared__ unsigned long long s_M[kMaxBins];
    const uint32_t lane = threadIdx.x;
    unsigned long long* M = reinterpret_cast<unsigned long long*>(res + 1);
    for (uint32_t b = lane; b < uint32_t(a.B); b += kWave) {
        unsigned long long v = 0;
        for (uint32_t x = 0; x < 8u; ++x) v += atomicExch(&hist_acc[size_t(x) * kMaxBins + b], 0ull);
        s_M[b] = v;
        M[b] = v;
        double t = 0.0;
        if (v > 0) {
            const double m = double(v) * 0x1p-32 / double(a.R);
            const double dd = double(obs[b]) - m;
            t = dd * dd / m;
        }
        s_t[b] = t;
    }
    __syncthreads();
    if (lane == 0) {
        int effN = a.B - a.left_pad;
        double c = 0.0;
        for (int b = a.left_pad; b < a.B; ++b) {
            if (s_M[b] > 0) c += s_t[b];
            else --effN;
        }
        res->chi2 = c / double(effN - a.nstates * (a.nstates + 1));
        res->effN = effN;
        res->B = a.B;
    }
    if (lane <= kRbPmfCounter) atomicExch(&work[lane], 0u);
}

hipError_t rb_occupancy(int which, int dyn_lds, int* per_cu) {
    switch (which) {
        case 0: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_walk_kernel<false>, kWave, 0);
        case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_walk_kernel<true>, kWave, 0);
        case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_pmf_kernel<false>, kWave, size_t(dyn_lds));
        default: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_pmf_kernel<true>, kWave, size_t(dyn_lds));
    }
}

hipError_t launch_rb_walk(bool scores, uint32_t grid, const BurstDesc* bursts, const uint32_t* ticks,
                          const uint2* links, const ModelBlob* blob, const RbModel* rbm,
                          const RunArgs& a, uint2* cnt, float* F, float* S, uint32_t* work,
                          hipStream_t stream) {
    if (scores)
        hipLaunchKernelGGL(rb_walk_kernel<true>, dim3(grid), dim3(kWave), 0, stream, bursts, ticks,
                           links, blob, rbm, a, cnt, F, S, work);
    else
        hipLaunchKernelGGL(rb_walk_kernel<false>, dim3(grid), dim3(kWave), 0, stream, bursts,
                           ticks, links, blob, rbm, a, cnt, F, S, work);
    return hipGetLastError();
}

hipError_t launch_rb_pmf(bool grad_mode, uint32_t grid, int cap, const RbModel* rbm,
                         const double* lf, const RunArgs& a, const uint2* cnt, const float* F,
                         const float* S, const double* g, unsigned long long* hist_acc,
                         double* grad, uint32_t* work, hipStream_t stream) {
    const size_t lds = sizeof(float) * size_t(cap) * (grad_mode ? 8u : 3u);
    if (grad_mode)
        hipLaunchKernelGGL(rb_pmf_kernel<true>, dim3(grid), dim3(kWave), lds, stream, rbm, lf, a,
                           cnt, F, S, g, hist_acc, grad, work, cap);
    else
        hipLaunchKernelGGL(rb_pmf_kernel<false>, dim3(grid), dim3(kWave), lds, stream, rbm, lf, a,
                           cnt, F, S, g, hist_acc, grad, work, cap);
    return hipGetLastError();
}

hipError_t launch_rb_finish(const RunArgs& a, const float* obs, unsigned long long* hist_acc,
                            uint32_t* work, void* res, hipStream_t stream) {
    hipLaunchKernelGGL(rb_finish_kernel, dim3(1), dim3(kWave), 0, stream, a, obs, hist_acc, work,
                       static_cast<RbResultHeader*>(res));
    return hipGetLastError();
}

}  // namespace pda
