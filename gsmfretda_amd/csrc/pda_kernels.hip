// pda_kernels.hip -- gfx950 kernels of the Monte-Carlo PDA hot path.
//
// walk_kernel replaces gSMFRETda's mc_kernel (src/mc.cu:83-197) together
// with the host histogram and chi2 of src/streamWorker.cpp:184-206:
//   * one 64-lane workgroup (one wavefront) per tile of 64 consecutive
//     (burst, resample) units, burst-major, so the lanes of a wave are
//     resamples of the same one or two bursts;
//   * the tile's photons (burst-relative uint32 ticks + channel-run links)
//     are staged in LDS with coalesced loads, then each lane runs its own
//     dwell walk, locating every window end by a binary search in LDS
//     instead of rescanning photons from global memory;
//   * random numbers are stateless Philox4x32-10 blocks keyed by
//     (seed, set, window, sub-stream, resample, burst) -- no RNG state in
//     memory, nothing to initialise (replaces setup_kernel,
//     src/gen_rand.cuh:7-25);
//   * each workgroup keeps an LDS int32 histogram, adds it to the run's
//     accumulator with device-scope atomics, and the last workgroup to
//     finish computes chi2 and resets the accumulator and ticket, so one
//     launch is the whole evaluation and needs no memset.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pda_math.h"
#include "pda_types.h"
#include "pda_walk.h"

namespace pda {

// Which units a launch walks, and how.
enum Pass : int {
    kPassFast = 0,     // all units, fast form; abandoned units go to the replay list
    kPassReplay = 1,   // the replay list, general form; then chi2
    kPassGeneral = 2,  // all units, general form (broadening); then chi2
};

// Histogram flush and, in the last workgroup of the evaluation's final pass,
// the reduced chi2 of src/streamWorker.cpp:194-206 (reference float
// accumulation and FP64) plus the reset of every per-run counter.
template <int kPass>
__device__ void finish(int32_t* s_hist, const RunArgs& a, const float* obs, int32_t* hist_acc,
                       uint32_t* ticket, uint32_t* replay_count, uint32_t* work,
                       ResultHeader* result, uint32_t* s_last) {
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < uint32_t(a.B); i += kWave) {
        const int32_t v = s_hist[i];
        if (v) atomicAdd(&hist_acc[i], v);
    }
    if (kPass == kPassFast) return;
    __threadfence();
    if (lane == 0) *s_last = atomicAdd(ticket, 1u) == gridDim.x - 1u ? 1u : 0u;
    __syncthreads();
    if (!*s_last) return;
    __threadfence();
    int32_t* rhist = reinterpret_cast<int32_t*>(result + 1);
    for (uint32_t i = lane; i < uint32_t(a.B); i += kWave) {
        const int32_t v = atomicExch(&hist_acc[i], 0);
        s_hist[i] = v;
        rhist[i] = v;
    }
    __syncthreads();
    // per-bin terms in parallel (FP64, and the reference's double term of a
    // float difference), then lane 0 folds them in bin order
    double* s_t64 = reinterpret_cast<double*>(reinterpret_cast<uint32_t*>(s_hist) +
                                              ((uint32_t(a.B) + 1u) & ~1u) +
                                              ((a.cap + 1u) & ~1u));  // the link area
    double* s_t32 = s_t64 + a.B;
    const float Rf = float(a.R);
    for (int b = int(lane); b < a.B; b += kWave) {
        const int32_t Mb = s_hist[b];
        double t64 = 0.0, t32 = 0.0;
        if (Mb > 0) {
            const float Mf = float(Mb);
            const float dlt = obs[b] - Mf / Rf;
            t32 = (double(dlt) * double(dlt)) / double(Mf / Rf);
            const double md = double(Mb) / double(a.R);
            const double dd = double(obs[b]) - md;
            t64 = dd * dd / md;
        }
        s_t64[b] = t64;
        s_t32[b] = t32;
    }
    __syncthreads();
    if (lane == 0) {
        int effN = a.B - a.left_pad;
        float c32 = 0.0f;
        double c64 = 0.0;
        for (int b = a.left_pad; b < a.B; ++b) {
            if (s_hist[b] > 0) {
                c32 = float(double(c32) + s_t32[b]);
                c64 += s_t64[b];
            } else {
                --effN;
            }
        }
        const int dof = effN - a.nstates * (a.nstates + 1);
        result->chi2_ref = c32 / float(dof);
        result->chi2_64 = c64 / double(dof);
        result->effN = effN;
        result->units = int64_t(a.nunits);
        result->B = a.B;
        result->pad = 0;
        atomicExch(replay_count, 0u);
        for (uint32_t q = 0; q < kMaxQueues; ++q) atomicExch(&work[q], 0u);
        atomicExch(ticket, 0u);
    }
}

__device__ __forceinline__ void tally(int32_t* s_hist, const RunArgs& a, const WalkOut& o,
                                      uint32_t u, float* outE, int32_t* outAcc, float* outF) {
    const float E = o.F > 0.0f ? o.acc / o.F : 0.0f;
    if (E >= 0.0f && E < 1.0f) {
        const int bin = int(floor(double(E) * double(a.B)));
        if (bin < a.B) atomicAdd(&s_hist[bin], 1);
    }
    if (outE) outE[u] = E;
    if (outAcc) outAcc[u] = int32_t(o.acc);
    if (outF) outF[u] = o.F;
}

template <int kPass>
__global__ __launch_bounds__(kWave) void walk_kernel(
    const BurstDesc* __restrict__ bursts, const uint32_t* __restrict__ ticks,
    const uint2* __restrict__ links, const ModelBlob* __restrict__ blob, const RunArgs a,
    const float* __restrict__ obs, int32_t* __restrict__ hist_acc, uint32_t* __restrict__ ticket,
    uint32_t* __restrict__ replay, uint32_t* __restrict__ replay_count, uint32_t* __restrict__ work,
    ResultHeader* __restrict__ result, float* __restrict__ outE, int32_t* __restrict__ outAcc,
    float* __restrict__ outF) {
    extern __shared__ uint32_t smem[];
    __shared__ Model sM;
    __shared__ uint32_t s_last;
    const uint32_t lane = threadIdx.x;
    int32_t* s_hist = reinterpret_cast<int32_t*>(smem);
    const uint32_t hist_words = (uint32_t(a.B) + 1u) & ~1u;
    uint32_t* s_tick = smem + hist_words;
    uint2* s_link = reinterpret_cast<uint2*>(s_tick + ((a.cap + 1u) & ~1u));
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(&blob->m);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&sM);
        for (uint32_t i = lane; i < sizeof(Model) / 4; i += kWave) dst[i] = src[i];
        for (uint32_t i = lane; i < uint32_t(a.B); i += kWave) s_hist[i] = 0;
    }
    __syncthreads();

    if (kPass == kPassReplay) {
        // units abandoned by the fast pass, general form: one workgroup per
        // unit, its burst staged in LDS when it fits, lane 0 walks it
        const uint32_t count = *replay_count < a.replay_cap ? *replay_count : a.replay_cap;
        for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
            const uint32_t u = replay[i];
            const uint32_t bl = u / a.R;
            const uint32_t b = a.begin + bl;
            const BurstDesc d = bursts[b];
            const bool staged = d.P <= a.cap;
            if (staged) {
                for (uint32_t k = lane; k < d.P; k += kWave) {
                    s_tick[k] = ticks[d.off + k];
                    s_link[k] = links[d.off + k];
                }
            }
            __syncthreads();
            if (lane == 0) {
                const UnitKey key{a.seed, a.set, u - bl * a.R, b};
                const WalkOut o = walk_unit<false>(sM, nullptr, key, d.P, d.stop_rel, d.T,
                                                   staged ? s_tick : ticks + d.off,
                                                   staged ? s_link : links + d.off, a);
                tally(s_hist, a, o, u, outE, outAcc, outF);
            }
            __syncthreads();
        }
        finish<kPass>(s_hist, a, obs, hist_acc, ticket, replay_count, work, result, &s_last);
        return;
    }

    // Work stealing at two levels.  A wave takes batches of consecutive units
    // from the run's work counter; its lanes take units from the batch as
    // they need them.  One loop pass advances every busy lane by one window,
    // and a lane whose unit ends starts another in the same pass, so lanes
    // do not wait for the wave's longest unit and waves do not wait for the
    // grid's slowest chunk.  Photons are read through the L1/L2 caches.
    // The units are split into a.queues contiguous ranges with a counter
    // each; workgroup g draws from queue g % a.queues, i.e. from its own XCD's
    // counter when a.queues == 8 (workgroups are dispatched to the XCDs
    // round-robin), so the counters live at 8 addresses instead of one.
    constexpr bool kFastForm = kPass == kPassFast;
    const BinomTable* tab = &blob->tab;
    const uint32_t queue = blockIdx.x % a.queues;
    const uint32_t q_begin = uint32_t(uint64_t(a.nunits) * queue / a.queues);
    const uint32_t q_end = uint32_t(uint64_t(a.nunits) * (queue + 1u) / a.queues);
    uint32_t pool_next = 0, pool_end = 0;  // the wave's current batch
    bool busy = false;
    uint32_t u = 0;
    WalkState<kFastForm> st{};
    UnitKey key{a.seed, a.set, 0u, 0u};
    BurstDesc d{0u, 0u, 0u, 0.0f};
    bool drained = false;
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
                if (begin_unit<kFastForm>(sM, key, d.P, d.T, st) == kStepMore) {
                    busy = true;
                    need = false;
                } else {
                    tally(s_hist, a, WalkOut{0.0f, 0.0f, false}, u, outE, outAcc, outF);
                }
            }
            const uint32_t taken = uint32_t(__popcll(want));
            pool_next += taken < avail ? taken : avail;
        }
        if (!__any(busy)) break;
        if (busy) {
            const int r = step_window<kFastForm>(sM, tab, key, d.P, d.stop_rel, d.T,
                                                 ticks + d.off, links + d.off, a, st);
            if (r != kStepMore) {
                if (r == kStepReplay) {
                    const uint32_t slot = atomicAdd(replay_count, 1u);
                    if (slot < a.replay_cap) replay[slot] = u;
                } else {
                    tally(s_hist, a, WalkOut{st.acc, st.F, false}, u, outE, outAcc, outF);
                }
                busy = false;
            }
        }
    }
    finish<kPass>(s_hist, a, obs, hist_acc, ticket, replay_count, work, result, &s_last);
}

// Evaluates shared-library primitives for the bit-exactness tests.
struct BinomItem {
    int32_t n;
    uint32_t pad0;
    double p;
    uint32_t seed, set, r, b, w, hi, lo, pad1;
};
struct BgItem {
    float lam, cap, u1, u2;
    uint32_t seed, set, r, b, w, sub;
};

__global__ void primitive_kernel(int op, const void* __restrict__ in, void* __restrict__ out,
                                 int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    switch (op) {
        case 0: {
            const uint32_t* q = static_cast<const uint32_t*>(in) + 6 * i;
            const Block x = philox10(q[0], q[1], q[2], q[3], q[4], q[5]);
            uint32_t* o = static_cast<uint32_t*>(out) + 4 * i;
            o[0] = x.w[0]; o[1] = x.w[1]; o[2] = x.w[2]; o[3] = x.w[3];
            break;
        }
        case 1: static_cast<float*>(out)[i] = flog(static_cast<const float*>(in)[i]); break;
        case 2: static_cast<float*>(out)[i] = fexp(static_cast<const float*>(in)[i]); break;
        case 3: static_cast<double*>(out)[i] = dlog(static_cast<const double*>(in)[i]); break;
        case 4: static_cast<double*>(out)[i] = dexp(static_cast<const double*>(in)[i]); break;
        case 5: static_cast<float*>(out)[i] = cos_turns(static_cast<const float*>(in)[i]); break;
        case 6: {
            const uint32_t* q = static_cast<const uint32_t*>(in) + 2 * i;
            static_cast<float*>(out)[i] = gauss(q[0], q[1]);
            break;
        }
        case 7: {
            const BinomItem it = static_cast<const BinomItem*>(in)[i];
            const UnitKey key{it.seed, it.set, it.r, it.b};
            static_cast<int64_t*>(out)[i] = binomial(it.n, it.p, it.hi, it.lo, key, it.w);
            break;
        }
        case 8: {
            const BgItem it = static_cast<const BgItem*>(in)[i];
            const UnitKey key{it.seed, it.set, it.r, it.b};
            static_cast<float*>(out)[i] =
                background_count(it.lam, it.cap, it.u1, it.u2, key, it.w, it.sub);
            break;
        }
        case 9: static_cast<uint32_t*>(out)[i] = static_cast<const uint32_t*>(in)[i]; break;
        default: break;
    }
}

// ---- launchers (called by the C-ABI layer) ---------------------------------

// LDS of a walk workgroup: histogram | photon ticks | photon links, the
// link area doubling as the chi2 term buffer (2 x B doubles).
size_t walk_smem_bytes(int B, uint32_t cap) {
    const size_t link_words = std::max(2u * size_t(cap), 4u * size_t(B));
    return (size_t((B + 1) & ~1) + size_t((cap + 1u) & ~1u) + link_words) * 4u;
}

hipError_t walk_occupancy(int pass, int B, uint32_t cap, int* blocks_per_cu) {
    const size_t smem = walk_smem_bytes(B, cap);
    switch (pass) {
        case kPassFast:
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, walk_kernel<kPassFast>, kWave, smem);
        case kPassReplay:
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, walk_kernel<kPassReplay>, kWave, smem);
        default:
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, walk_kernel<kPassGeneral>, kWave, smem);
    }
}

hipError_t launch_walk(int pass, uint32_t grid, const BurstDesc* bursts, const uint32_t* ticks,
                       const uint2* links, const ModelBlob* blob, const RunArgs& a,
                       const float* obs, int32_t* hist_acc, uint32_t* ticket, uint32_t* replay,
                       uint32_t* replay_count, uint32_t* work, ResultHeader* result, float* outE,
                       int32_t* outAcc, float* outF, hipStream_t stream) {
    const size_t smem = walk_smem_bytes(a.B, a.cap);
#define PDA_LAUNCH(P)                                                                       \
    hipLaunchKernelGGL(walk_kernel<P>, dim3(grid), dim3(kWave), smem, stream, bursts, ticks, \
                       links, blob, a, obs, hist_acc, ticket, replay, replay_count, work,    \
                       result, outE, outAcc, outF)
    switch (pass) {
        case kPassFast: PDA_LAUNCH(kPassFast); break;
        case kPassReplay: PDA_LAUNCH(kPassReplay); break;
        default: PDA_LAUNCH(kPassGeneral); break;
    }
#undef PDA_LAUNCH
    return hipGetLastError();
}

hipError_t launch_primitive(int op, const void* in, void* out, int64_t n, hipStream_t stream) {
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(primitive_kernel, dim3(uint32_t(blocks)), dim3(256), 0, stream, op, in, out,
                       n);
    return hipGetLastError();
}

}  // namespace pda
