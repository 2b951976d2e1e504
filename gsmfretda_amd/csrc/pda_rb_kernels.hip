// pda_rb_kernels.hip -- gfx950 kernels of the Rao-Blackwellised (RB) PDA form.
//
//   rb_walk_kernel   the walk of every (burst, resample) unit with the keyed
//                    draws of the plain kernel (src/mc.cu:92-191) but no
//                    acceptor draw; per unit it stores the photon counts per
//                    state, F, and (for rate gradients) the score of the
//                    walk's draws
//   rb_pmf_kernel    one wave per burst, 4 units at a time on groups of 16
//                    lanes (the whole wave per unit for large bursts): the
//                    pmf of each unit's acceptor total (a convolution of one
//                    binomial per state, in LDS, lanes over outcomes) spread
//                    into the E histogram as 2^-32 fixed point -- or, for the
//                    adjoint, the cotangent-weighted pmf and its derivatives,
//                    or, for the Jacobian, the derivatives of every bin
//   rb_finish_kernel reduces the histogram rows, computes chi2 on the
//                    expected counts, resets the counters
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pda_math.h"
#include "pda_rb.h"
#include "pda_types.h"
#include "pda_walk.h"

namespace pda {

// rb_pmf_kernel takes bursts in batches from up to kAuxQueues counters
// after the walk's (workgroup g from counter g % kAuxQueues, one per XCD,
// each over a contiguous eighth of the bursts): one counter and one grab per
// burst made the pass wait on a million serialised atomics at 10^6 bursts.
// In the derivative passes batches shrink to 1 as a queue empties and a
// workgroup whose queue is empty takes single bursts from the others, so the
// pass ends with every wave within about one burst of the last (the forward,
// whose bursts are cheap, keeps fixed batches: fewer grabs measured faster).
constexpr uint32_t kRbPmfCounter = kMaxQueues * kCounterStride;
constexpr uint32_t kRbMaxBurstBatch = 8;

// Jacobian scratch in LDS (doubles, after the float buffers): per bin the
// efficiency and rate derivatives, the current unit's histogram per lane
// group, the burst's histogram, the burst's score sums, the current units'
// scores per lane group.
struct JacLds {
    double* JE;  // [B][kRbMaxStates]
    double* JK;  // [B][kRbMaxPairs]
    double* hu;  // [kWave / kRbSmallLanes][B]
    double* Hb;  // [B]
    double* Sk;  // [kRbMaxPairs]
    double* Su;  // [kWave / kRbSmallLanes][kRbMaxPairs]
};
__host__ __device__ inline size_t rb_jac_doubles(int B) {
    return size_t(B) * (kRbMaxStates + kRbMaxPairs + kWave / 16 + 1) +
           size_t(kRbMaxPairs) * (1 + kWave / 16);
}

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
                if (lane == 0) off = atomicAdd(&work[queue * kCounterStride], a.batch);
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

// Sum over the L lanes of a lane group (every lane gets the total).
template <int L>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, L);
    return v;
}

constexpr int kRbSmallLanes = 16;  // lanes per unit for small bursts (4 units per wave)
constexpr int kRbSmallCap = 256;   // acceptor totals below this use the small layout

// All R units of one burst, kWave / L at a time (one unit per group of L
// lanes, lanes over acceptor totals).  Per group, LDS buffers X (the pmf so
// far), Y (the next state's binomial pmf) and Z (a free buffer) of `cap`
// floats each; the group's buffer roles are swapped instead of copied.  The
// in-place forward (kRbFwdInPlace) writes each 16-lane product back over X
// and has no Z.
//
// Forward (kGrad = false): the unit's pmf P(a), a = 0..Mt, goes into the LDS
// histogram at E = a / F in 2^-32 fixed point.
//
// Adjoint (kMode = kRbGrad): with G(a) the cotangent of a's bin, the unit
// contributes h = sum_a G(a) P(a) and, per state s with m_s photons,
//   d h / d p_s = m_s sum_c P_s(c) (G(c + 1) - G(c))
// where P_s is the pmf with one trial of state s removed,
// P(a) = q_s P_s(a) + p_s P_s(a - 1).  P_s follows from P by a first-order
// recurrence run as a lane-parallel scan of affine maps, upwards when
// p_s <= 1/2 and downwards otherwise (the direction in which the recurrence
// damps rounding errors).  Per burst, the rate gradient is the likelihood
// ratio R / (R - 1) sum_r (h_r - mean h) S_r.
//
// Jacobian (kMode = kRbJac): the adjoint for every indicator cotangent
// G_b(a) = [bin(a) = b] at once.  G_b(c + 1) - G_b(c) is nonzero only where
// c and c + 1 fall in different bins, so d M_b / d p_s gathers m_s P_s(c) at
// the unit's bin boundaries; the rate columns use the unit's histogram h_rb
// in place of h_r (sum_r h_rb S_r here, the mean term per burst).
// kRbFwdInPlace: the forward with the 16-lane products written back over X
// from registers (two buffers per group instead of three, more registers):
// chosen when that lets more waves share a CU.
constexpr int kRbFwd = 0, kRbGrad = 1, kRbJac = 2, kRbFwdInPlace = 3;

template <int L, int kMode>
__device__ __forceinline__ void rb_burst(const RbModel& sR, const double* __restrict__ lf,
                                         const float* __restrict__ pmf_tab,
                                         const RunArgs& a, uint32_t bl,
                                         const uint2* __restrict__ cnt_in,
                                         const float* __restrict__ F_in,
                                         const float* __restrict__ S_in, const double* s_g,
                                         unsigned long long* s_h, float* s_buf, int nbuf, int cap,
                                         uint32_t lane, double& gp_acc, double& gk_acc,
                                         const JacLds& jl, float* __restrict__ pmf_io,
                                         const uint64_t* __restrict__ poff) {
    constexpr bool kGrad = kMode == kRbGrad || kMode == kRbJac;
    constexpr int G = kWave / L;
    const int g = int(lane) / L, lg = int(lane) % L;
    const int n = sR.n, K = sR.K, B = a.B;
    float* base = s_buf + g * nbuf * cap;
    // pmf_io: the forward stores each unit's pmf at poff[bl] + r * stride,
    // the derivative passes read it back instead of rebuilding it
    const uint64_t slot0 = pmf_io ? poff[bl] : 0u;
    const uint64_t stride = pmf_io ? (poff[bl + 1] - slot0) / a.R : 0u;
    double H = 0.0, Sg = 0.0, Ac = 0.0;  // lane lg of each group: rate pair lg
    for (uint32_t r0 = 0; r0 < a.R; r0 += G) {
        const uint32_t r = r0 + uint32_t(g);
        const bool act = r < a.R;
        const uint32_t u = bl * a.R + (act ? r : 0u);
        const uint2 c = act ? cnt_in[u] : make_uint2(0u, 0u);
        const float F = act ? F_in[u] : 0.0f;
        // the forward in the 16-lane layout convolves in place (X and Y
        // only); every other case writes products to Z
        constexpr bool kInPlace = kMode == kRbFwdInPlace && L == kRbSmallLanes;
        int bX = 0, bZ = 1, bY = kInPlace ? 1 : 2;
        int xl = 1;
        if (lg == 0) base[0] = 1.0f;
        if (kGrad && pmf_io) {
            // the pmf of the forward run: length 1 + sum of the counts
            for (int s = 0; s < kRbMaxStates; ++s) xl += (act && s < n) ? int(rb_count(c, s)) : 0;
            const float* src = pmf_io + slot0 + uint64_t(r) * stride;
            for (int q = lg; act && q < xl; q += L) base[q] = src[q];
            __syncthreads();
        } else {
#pragma unroll
            for (int s = 0; s < kRbMaxStates; ++s) {
                const int m = (act && s < n) ? int(rb_count(c, s)) : 0;
                float* Y = base + bY * cap;
                for (int k = lg; m > 0 && k <= m; k += L)
                    Y[k] = m < kRbPmfTable
                               ? pmf_tab[rb_pmf_index(s, m, k)]
                               : float(rb_binom_pmf(m, k, lf, sR.p[s], sR.lp[s], sR.lq[s]));
                __syncthreads();
                if (m > 0) {
                    if (xl == 1) {  // X = [1]: the product is Y itself, bit for bit
                        const int t = bX;
                        bX = bY;
                        bY = t;
                        xl = m + 1;
                    } else if (kInPlace) {
                        // at most kRbSmallCap / L outputs per lane, held in
                        // registers until every read of X is done: the
                        // group's lanes are lanes of one wave, so all the
                        // loads below are issued before any store, and each
                        // store waits on the loads its value came from
                        constexpr int kOut = kRbSmallCap / kRbSmallLanes;
                        float* X = base + bX * cap;
                        float zr[kOut];
#pragma unroll
                        for (int i = 0; i < kOut; ++i) {
                            const int q = lg + i * L;
                            zr[i] = q < xl + m ? rb_conv_at(Y, m + 1, X, xl, q) : 0.0f;
                        }
#pragma unroll
                        for (int i = 0; i < kOut; ++i) {
                            const int q = lg + i * L;
                            if (q < xl + m) X[q] = zr[i];
                        }
                        xl += m;
                    } else {
                        const float* X = base + bX * cap;
                        float* Z = base + bZ * cap;
                        for (int q = lg; q < xl + m; q += L) Z[q] = rb_conv_at(Y, m + 1, X, xl, q);
                        const int t = bX;
                        bX = bZ;
                        bZ = t;
                        xl += m;
                    }
                }
                __syncthreads();
            }
        }
        const float* P = base + bX * cap;
        if (!kGrad) {
            float* dst = pmf_io ? pmf_io + slot0 + uint64_t(r) * stride : nullptr;
            for (int q = lg; act && q < xl; q += L) {
                const int bin = rb_bin(q, F, B);
                if (bin >= 0) atomicAdd(&s_h[bin], (unsigned long long)rb_fixed(P[q]));
                if (dst) dst[q] = P[q];
            }
        } else {
            float* dG = base + bZ * cap;  // G(c + 1) - G(c)
            double hp = 0.0;
            double* hu = jl.hu + g * B;   // Jacobian: this unit's histogram
            for (int q = lg; q < xl; q += L) {
                const int b0 = rb_bin(q, F, B);
                if (kMode == kRbJac) {
                    if (act && b0 >= 0) atomicAdd(&hu[b0], double(P[q]));
                    continue;
                }
                const double G0 = b0 >= 0 ? s_g[b0] : 0.0;
                hp += G0 * double(P[q]);
                if (q + 1 < xl) {
                    const int b1 = rb_bin(q + 1, F, B);
                    dG[q] = float((b1 >= 0 ? s_g[b1] : 0.0) - G0);
                }
            }
            __syncthreads();
            const double h = group_sum<L>(hp);
            const int Mt = xl - 1;
#pragma unroll
            for (int s = 0; s < kRbMaxStates; ++s) {
                const int m = (act && s < n) ? int(rb_count(c, s)) : 0;
                double part = 0.0;
                if (m > 0) {  // uniform within the group
                    const double p = double(sR.p[s]), q = 1.0 - p;
                    const bool up = p <= 0.5;
                    const double alpha = up ? -p / q : -q / p;
                    const double inv = up ? 1.0 / q : 1.0 / p;
                    double carry = 0.0;
                    for (int j0 = 0; j0 < Mt; j0 += L) {
                        const int j = j0 + lg;
                        const bool v = j < Mt;
                        double A = v ? alpha : 1.0;
                        double Bv = v ? double(P[up ? j : Mt - j]) * inv : 0.0;
#pragma unroll
                        for (int d = 1; d < L; d <<= 1) {
                            const double A2 = __shfl_up(A, d, L);
                            const double B2 = __shfl_up(Bv, d, L);
                            if (lg >= d) {
                                Bv = A * B2 + Bv;
                                A = A * A2;
                            }
                        }
                        const double x = A * carry + Bv;  // P_s(up ? j : Mt - 1 - j)
                        carry = __shfl(x, L - 1, L);
                        const int ci = up ? j : Mt - 1 - j;
                        if (kMode == kRbJac) {
                            const int b0 = rb_bin(ci, F, B), b1 = rb_bin(ci + 1, F, B);
                            if (v && b0 != b1) {
                                const double w = double(m) * x;
                                if (b1 >= 0) atomicAdd(&jl.JE[b1 * kRbMaxStates + s], w);
                                if (b0 >= 0) atomicAdd(&jl.JE[b0 * kRbMaxStates + s], -w);
                            }
                        } else if (v) {
                            part += x * double(dG[ci]);
                        }
                    }
                    if (kMode == kRbGrad) part = group_sum<L>(part) * double(m);
                }
                if (kMode == kRbGrad && lg == s) gp_acc += part;
            }
            if (kMode == kRbJac) {
                // rates: JK[b][k] += sum_g hu_g[b] S_g[k] over the wave's G
                // units, and the burst's histogram Hb[b] += sum_g hu_g[b];
                // every entry owned by one lane (lanes over k = 0..K, the
                // column K being Hb), in a fixed order: no atomics
                if (lg < K) {
                    const double si = act ? double(S_in[size_t(u) * size_t(K) + size_t(lg)]) : 0.0;
                    Sg += si;
                    jl.Su[g * kRbMaxPairs + lg] = si;
                }
                __syncthreads();
                const int w = K + 1, per = kWave / w;
                const int k = int(lane) % w, bb = int(lane) / w;
                if (bb < per) {
                    double sk[G];
#pragma unroll
                    for (int q = 0; q < G; ++q) sk[q] = k < K ? jl.Su[q * kRbMaxPairs + k] : 1.0;
                    for (int b = bb; b < B; b += per) {
                        double acc = 0.0;
#pragma unroll
                        for (int q = 0; q < G; ++q) acc += jl.hu[q * B + b] * sk[q];
                        if (k < K) jl.JK[b * kRbMaxPairs + k] += acc;
                        else jl.Hb[b] += acc;
                    }
                }
                __syncthreads();
                for (int b = lg; b < B; b += L) hu[b] = 0.0;
            } else if (act) {
                H += h;
                if (lg < K) {
                    const double si = double(S_in[size_t(u) * size_t(K) + size_t(lg)]);
                    Sg += si;
                    Ac += h * si;
                }
            }
        }
        __syncthreads();
    }
    if (kMode == kRbGrad) {
#pragma unroll
        for (int o = L; o < kWave; o <<= 1) {
            H += __shfl_xor(H, o);
            Sg += __shfl_xor(Sg, o);
            Ac += __shfl_xor(Ac, o);
        }
        if (int(lane) < K && a.R > 1)
            gk_acc += (double(a.R) / double(a.R - 1)) * (Ac - H * Sg / double(a.R));
    } else if (kMode == kRbJac) {
        // the burst's mean term: JK[b][k] -= H_b S_k / R (the factor
        // R / (R - 1) is applied once, when the kernel flushes JK)
#pragma unroll
        for (int o = L; o < kWave; o <<= 1) Sg += __shfl_xor(Sg, o);
        if (int(lane) < K) jl.Sk[lane] = Sg;
        __syncthreads();
        const double inv_r = 1.0 / double(a.R);
        for (int i = int(lane); i < B * K; i += kWave) {
            const int b = i / K, k = i - b * K;
            jl.JK[b * kRbMaxPairs + k] -= jl.Hb[b] * jl.Sk[k] * inv_r;
        }
        __syncthreads();
        for (int b = int(lane); b < B; b += kWave) jl.Hb[b] = 0.0;
        __syncthreads();
    }
}

// pmf buffers per lane group: the forward convolves in X and Y (in place)
// in the 16-lane layout and in X, Y and Z in the whole-wave one; a
// derivative pass without stored pmfs rebuilds them in X, Y and Z; with
// stored pmfs the adjoint needs X and Z (the cotangent differences), the
// Jacobian X alone.  Fewer buffers leave LDS for more resident waves.
__host__ __device__ inline int rb_pmf_buffers(int mode, bool stored, bool small) {
    if (stored && mode == kRbJac) return 1;
    if (stored && mode == kRbGrad) return 2;
    return mode == kRbFwdInPlace && small ? 2 : 3;
}

// Float entries of the pmf buffers: 4 groups x nbuf x 256 (16-lane layout)
// or nbuf x cap_big (whole-wave layout), whichever is larger.
__host__ __device__ inline size_t rb_pmf_floats(int cap_big, int mode, bool stored) {
    const size_t small = size_t(kWave / kRbSmallLanes) *
                         size_t(rb_pmf_buffers(mode, stored, true)) * kRbSmallCap;
    const size_t big = size_t(rb_pmf_buffers(mode, stored, false)) * size_t(cap_big);
    return ((small > big ? small : big) + 1u) & ~size_t(1);  // even: doubles follow
}

// One wave per burst at a time (bursts from the counters after the walk's).
// Dynamic LDS: B 8-byte entries (the forward's fixed-point histogram or the
// adjoint's cotangent; none in the Jacobian) followed by the float buffers,
// then (Jacobian) the JacLds arrays.  The in-place forward, whose products
// sit in registers, is held to 128 VGPRs (4 waves per SIMD) so that its
// smaller LDS can let more waves in; the
// adjoint is held to 128 VGPRs (4 waves per SIMD instead of 3, a few values
// spilled to scratch): its deconvolution scans wait on shuffles and LDS, and
// the extra wave measured 8-14 % faster.  The Jacobian's LDS allows 3 waves
// per SIMD whatever its registers.
template <int kMode>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(
    kMode == kRbGrad || kMode == kRbFwdInPlace ? 4 : 1))) void rb_pmf_kernel(
    const RbModel* __restrict__ rbm, const double* __restrict__ lf,
    const float* __restrict__ pmf_tab, const RunArgs a, const BurstDesc* __restrict__ bursts,
    const uint2* __restrict__ cnt_in, const float* __restrict__ F_in,
    const float* __restrict__ S_in, const double* __restrict__ g,
    unsigned long long* __restrict__ hist_acc, double* __restrict__ grad,
    uint32_t* __restrict__ work, int cap_big, float* __restrict__ pmf_io,
    const uint64_t* __restrict__ poff) {
    constexpr bool kGrad = kMode == kRbGrad || kMode == kRbJac;
    extern __shared__ double s_dyn[];
    __shared__ RbModel sR;
    // B 8-byte entries first (the forward's histogram, the adjoint's
    // cotangent; none for the Jacobian)
    const int head = kMode == kRbJac ? 0 : a.B;
    unsigned long long* s_h = reinterpret_cast<unsigned long long*>(s_dyn);
    double* s_g = s_dyn;
    float* s_buf = reinterpret_cast<float*>(s_dyn + head);
    const int B = a.B;
    JacLds jl;
    const bool stored = pmf_io != nullptr;
    jl.JE = s_dyn + head + rb_pmf_floats(cap_big, kMode, stored) / 2;
    jl.JK = jl.JE + B * kRbMaxStates;
    jl.hu = jl.JK + B * kRbMaxPairs;
    jl.Hb = jl.hu + B * (kWave / kRbSmallLanes);
    jl.Sk = jl.Hb + B;
    jl.Su = jl.Sk + kRbMaxPairs;
    const uint32_t lane = threadIdx.x;
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(rbm);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&sR);
        for (uint32_t i = lane; i < sizeof(RbModel) / 4; i += kWave) dst[i] = src[i];
        for (uint32_t i = lane; i < uint32_t(head); i += kWave) {
            if (kMode == kRbGrad) s_g[i] = g[i];
            else s_h[i] = 0ull;
        }
        if (kMode == kRbJac)
            for (size_t i = lane; i < rb_jac_doubles(B); i += kWave) jl.JE[i] = 0.0;
    }
    __syncthreads();
    const uint32_t nb = a.nunits / a.R;
    // every counter must have a workgroup: fewer queues on a small grid
    const uint32_t nq = gridDim.x < kAuxQueues ? gridDim.x : kAuxQueues;
    const uint32_t waves_q = gridDim.x / nq;  // >= 1 workgroups per queue
    uint32_t queue = blockIdx.x % nq, visited = 1;
    uint32_t qb = uint32_t(uint64_t(nb) * queue / nq);
    uint32_t qn = uint32_t(uint64_t(nb) * (queue + 1u) / nq) - qb;
    uint32_t* counter = &work[kRbPmfCounter + queue * kCounterStride];
    // The forward takes fixed batches (about 16 grabs per workgroup).  The
    // derivative passes, several times dearer per burst, take guided batches
    // (a quarter of the wave's share of what its queue had left at its last
    // grab) and then single bursts from the other queues.
    constexpr bool kGuided = kGrad;
    uint32_t batch = kGuided ? qn / (waves_q * 4u) : nb / (gridDim.x * 16u);
    batch = batch < 1u ? 1u : (batch > kRbMaxBurstBatch ? kRbMaxBurstBatch : batch);
    double gp_acc = 0.0, gk_acc = 0.0;  // lane s: d/dp_s; lane i: d/dk_i
    uint32_t next = 0, end = 0;  // the wave's current batch of bursts
    while (true) {
        if (next >= end) {
            uint32_t off = 0;
            if (lane == 0) off = atomicAdd(counter, batch);
            off = __shfl(off, 0);
            if (off >= qn) {
                // queue drained: help the next one until every queue has
                // been seen drained
                if (!kGuided || visited == nq) break;
                ++visited;
                queue = queue + 1u == nq ? 0u : queue + 1u;
                qb = uint32_t(uint64_t(nb) * queue / nq);
                qn = uint32_t(uint64_t(nb) * (queue + 1u) / nq) - qb;
                counter = &work[kRbPmfCounter + queue * kCounterStride];
                batch = 1u;
                continue;
            }
            next = qb + off;
            end = off + batch < qn ? next + batch : qb + qn;
            if (kGuided) {
                const uint32_t left = off + batch < qn ? qn - off - batch : 0u;
                batch = left / (waves_q * 4u);
                batch = batch < 1u ? 1u : (batch > kRbMaxBurstBatch ? kRbMaxBurstBatch : batch);
            }
        }
        const uint32_t bl = next++;
        const int bound = rb_total_bound(bursts[a.begin + bl].P, sR.coef);
        if (bound < kRbSmallCap)
            rb_burst<kRbSmallLanes, kMode>(sR, lf, pmf_tab, a, bl, cnt_in, F_in, S_in, s_g, s_h,
                                           s_buf, rb_pmf_buffers(kMode, stored, true),
                                           kRbSmallCap, lane, gp_acc, gk_acc, jl,
                                           pmf_io, poff);
        else
            rb_burst<kWave, kMode>(sR, lf, pmf_tab, a, bl, cnt_in, F_in, S_in, s_g, s_h, s_buf,
                                   rb_pmf_buffers(kMode, stored, false), cap_big, lane,
                                   gp_acc, gk_acc, jl, pmf_io, poff);
    }
    if (kMode == kRbJac) {
        // grad holds J[b][i], i < n + K (efficiencies, then rates in args
        // order), B rows of kRbMaxStates + kRbMaxPairs
        __syncthreads();
        const int n = sR.n, K = sR.K, w = n + K;
        const double f = a.R > 1 ? double(a.R) / double(a.R - 1) : 0.0;
        for (int i = int(lane); i < B * w; i += kWave) {
            const int b = i / w, c = i - b * w;
            const double v = c < n ? jl.JE[b * kRbMaxStates + c] : f * jl.JK[b * kRbMaxPairs + (c - n)];
            if (v != 0.0) atomicAdd(&grad[size_t(b) * (kRbMaxStates + kRbMaxPairs) + size_t(c)], v);
        }
    } else if (!kGrad) {
        __syncthreads();
        unsigned long long* row = hist_acc + size_t(blockIdx.x & 7u) * kMaxBins;
        for (uint32_t i = lane; i < uint32_t(a.B); i += kWave)
            if (s_h[i]) atomicAdd(&row[i], s_h[i]);
    } else {
        // every group's lane s holds a share of d/dp_s
#pragma unroll
        for (int o = kRbSmallLanes; o < kWave; o <<= 1) gp_acc += __shfl_xor(gp_acc, o);
        if (int(lane) < sR.n) atomicAdd(&grad[lane], gp_acc);
        if (int(lane) < sR.K) atomicAdd(&grad[sR.n + int(lane)], gk_acc);
    }
}

// The binomial pmfs of every state for counts m < kRbPmfTable (one thread
// per entry), read by rb_pmf_kernel instead of recomputing them per unit.
__global__ void rb_table_kernel(const RbModel* __restrict__ rbm, const double* __restrict__ lf,
                                float* __restrict__ pmf_tab) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    const int n = rbm->n;
    if (i >= n * kRbPmfStride) return;
    const int s = i / kRbPmfStride, r = i - s * kRbPmfStride;
    // r = m (m + 1) / 2 + k with 0 <= k <= m
    int m = int((sqrt(8.0 * double(r) + 1.0) - 1.0) * 0.5);
    while (m * (m + 1) / 2 > r) --m;
    while ((m + 1) * (m + 2) / 2 <= r) ++m;
    const int k = r - m * (m + 1) / 2;
    pmf_tab[i] = float(rb_binom_pmf(m, k, lf, rbm->p[s], rbm->lp[s], rbm->lq[s]));
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
    for (uint32_t q = lane; q < kMaxQueues + kAuxQueues; q += kWave)
        atomicExch(&work[q * kCounterStride], 0u);
}

hipError_t rb_occupancy(int which, int dyn_lds, int* per_cu) {
    switch (which) {
        case 0: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_walk_kernel<false>, kWave, 0);
        case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_walk_kernel<true>, kWave, 0);
        case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_pmf_kernel<kRbFwd>, kWave, size_t(dyn_lds));
        case 5: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_pmf_kernel<kRbFwdInPlace>, kWave, size_t(dyn_lds));
        case 3: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_pmf_kernel<kRbGrad>, kWave, size_t(dyn_lds));
        default: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_pmf_kernel<kRbJac>, kWave, size_t(dyn_lds));
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

size_t rb_pmf_lds(int B, int cap_big, int mode, bool stored) {
    return sizeof(double) * size_t(mode == kRbJac ? 0 : B) +
           sizeof(float) * rb_pmf_floats(cap_big, mode, stored) +
           (mode == 2 ? sizeof(double) * rb_jac_doubles(B) : 0u);
}

hipError_t launch_rb_table(const RbModel* rbm, const double* lf, float* pmf_tab, int n,
                           hipStream_t stream) {
    const int total = n * kRbPmfStride;
    hipLaunchKernelGGL(rb_table_kernel, dim3(uint32_t((total + 255) / 256)), dim3(256), 0, stream,
                       rbm, lf, pmf_tab);
    return hipGetLastError();
}

hipError_t launch_rb_pmf(int mode, uint32_t grid, int cap_big, const RbModel* rbm,
                         const double* lf, const float* pmf_tab, const RunArgs& a,
                         const BurstDesc* bursts, const uint2* cnt, const float* F,
                         const float* S, const double* g, unsigned long long* hist_acc,
                         double* grad, uint32_t* work, float* pmf_io, const uint64_t* poff,
                         hipStream_t stream) {
    const size_t lds = rb_pmf_lds(a.B, cap_big, mode, pmf_io != nullptr);
    if (mode == kRbGrad)
        hipLaunchKernelGGL(rb_pmf_kernel<kRbGrad>, dim3(grid), dim3(kWave), lds, stream, rbm, lf,
                           pmf_tab, a, bursts, cnt, F, S, g, hist_acc, grad, work, cap_big,
                           pmf_io, poff);
    else if (mode == kRbJac)
        hipLaunchKernelGGL(rb_pmf_kernel<kRbJac>, dim3(grid), dim3(kWave), lds, stream, rbm, lf,
                           pmf_tab, a, bursts, cnt, F, S, g, hist_acc, grad, work, cap_big,
                           pmf_io, poff);
    else if (mode == kRbFwdInPlace)
        hipLaunchKernelGGL(rb_pmf_kernel<kRbFwdInPlace>, dim3(grid), dim3(kWave), lds, stream, rbm, lf,
                           pmf_tab, a, bursts, cnt, F, S, g, hist_acc, grad, work, cap_big,
                           pmf_io, poff);
    else
        hipLaunchKernelGGL(rb_pmf_kernel<kRbFwd>, dim3(grid), dim3(kWave), lds, stream, rbm, lf,
                           pmf_tab, a, bursts, cnt, F, S, g, hist_acc, grad, work, cap_big,
                           pmf_io, poff);
    return hipGetLastError();
}

hipError_t launch_rb_finish(const RunArgs& a, const float* obs, unsigned long long* hist_acc,
                            uint32_t* work, void* res, hipStream_t stream) {
    hipLaunchKernelGGL(rb_finish_kernel, dim3(1), dim3(kWave), 0, stream, a, obs, hist_acc, work,
                       static_cast<RbResultHeader*>(res));
    return hipGetLastError();
}

}  // namespace pda
