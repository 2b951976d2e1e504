// pda_kernels.hip -- gfx950 kernels of the Monte-Carlo PDA hot path.
//
// walk_kernel + finish_kernel replace gSMFRETda's mc_kernel (src/mc.cu:83-197)
// together with the host histogram and chi2 of src/streamWorker.cpp:184-206:
//   * one-wave workgroups; every lane streams through (burst, resample)
//     units taken from per-XCD work counters, advancing one dwell window
//     per loop pass, so lanes and waves never wait for the longest unit;
//   * each window end is located by a binary search over burst-relative
//     uint32 ticks and the window's sticky-flag count read from per-photon
//     channel-run links, instead of rescanning photons;
//   * random numbers are stateless Philox4x32-10 blocks keyed by
//     (seed, set, window, sub-stream, resample, burst) -- no RNG state in
//     memory, nothing to initialise (replaces setup_kernel,
//     src/gen_rand.cuh:7-25);
//   * the common case runs in a lean form (at most 72 VGPRs); the rare units
//     that need BTPE draws, inversion restarts or large background means are
//     listed and replayed in full by the same kernel in its general form
//     over the whole grid; finish_kernel then reduces the LDS histograms of
//     all workgroups and computes chi2 on the device.
#include <hip/hip_runtime.h>

#include "pda_math.h"
#include "pda_rb.h"
#include "pda_types.h"
#include "pda_walk.h"

namespace pda {

// Which form a walk launch runs.
enum Pass : int {
    kPassFast = 0,     // fast form; abandoned units go to the replay list
    kPassReplay = 1,   // general form over the units of the replay list
    kPassGeneral = 2,  // general form for every unit (models with broadening)
};

// Work counters of a replay pass: one per XCD, on the counter lines after the
// kMaxQueues lines of the walk (kAuxQueues of them).
constexpr uint32_t kReplayQueues = 8;
static_assert(kReplayQueues <= kAuxQueues, "replay counters live on the auxiliary lines");

// The per-unit outputs outE / outAcc / outF are requested together (all
// null or all set), so one test covers the three stores.
__device__ __forceinline__ void tally(int32_t* s_hist, const RunArgs& a, const WalkOut& o,
                                      uint32_t u, float* outE, int32_t* outAcc, float* outF) {
    const float E = o.F > 0.0f ? o.acc / o.F : 0.0f;
    if (E >= 0.0f && E < 1.0f) {
        const int bin = int(floor(double(E) * double(a.B)));
        if (bin < a.B) atomicAdd(&s_hist[bin], 1);
    }
    if (outE) {
        outE[u] = E;
        outAcc[u] = int32_t(o.acc);
        outF[u] = o.F;
    }
}

// The walk over every unit of the run (kPassFast, kPassGeneral), or over the
// units the fast pass abandoned (kPassReplay, launched right after it on the
// same stream, so the replay list and its count are final).  Each one-wave
// workgroup keeps an LDS histogram and adds it to its XCD's accumulator row
// at the end (workgroups are dispatched to the 8 XCDs round-robin, so the
// atomics of a launch spread over 8 x B addresses).  No completion ticket
// and no fence: the finish kernel, launched behind on the same stream, reads
// the rows.
// The fast forms are held to a register count that keeps many waves
// resident: their window loop waits on dependent loads, and each more
// resident wave per SIMD hides more of that latency than the registers it
// gives up cost.  The single-set walk fits 64 VGPRs (8 waves per SIMD, no
// spill to scratch; -2 % walk time against 7 waves on the 3-state
// workload, the LDS of 32 workgroups exactly 160 KB); the population walk
// keeps 72 (7 waves), where 64 would spill.
constexpr int kFastWavesPerEU = 8;
constexpr int kPopFastWavesPerEU = 7;

template <int kPass>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(
    kPass == kPassFast ? kFastWavesPerEU : 1))) void walk_kernel(
    const BurstDesc* __restrict__ bursts, const uint32_t* __restrict__ ticks,
    const uint2* __restrict__ links, const ModelBlob* __restrict__ blob, const BinomGuide guide,
    const RunArgs a, int32_t* __restrict__ hist_acc, uint32_t* __restrict__ replay,
    uint32_t* __restrict__ replay_count, uint32_t* __restrict__ work, float* __restrict__ outE,
    int32_t* __restrict__ outAcc, float* __restrict__ outF) {
    constexpr bool kFastForm = kPass == kPassFast;
    constexpr bool kReplay = kPass == kPassReplay;
    // units of this pass, units per grab, queues and their counters
    uint32_t nunits = a.nunits, batch = a.batch, queues = a.queues;
    uint32_t* counters = work;
    if (kReplay) {
        const uint32_t c = *replay_count;
        nunits = c < a.replay_cap ? c : a.replay_cap;
        // as many workgroups as the list needs a wave for, each queue served
        // by at least one of them
        const uint32_t need = (nunits + kWave - 1u) / kWave;
        const uint32_t alive = need < gridDim.x ? need : gridDim.x;
        if (blockIdx.x >= alive) return;  // uniform over the workgroup
        batch = kWave;
        queues = alive < kReplayQueues ? alive : kReplayQueues;
        counters = work + size_t(kMaxQueues) * kCounterStride;
    }
    __shared__ Model sM;
    __shared__ int32_t s_hist[kMaxBins];
    const uint32_t lane = threadIdx.x;
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(&blob->m);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&sM);
        for (uint32_t i = lane; i < sizeof(Model) / 4; i += kWave) dst[i] = src[i];
        for (uint32_t i = lane; i < uint32_t(a.B); i += kWave) s_hist[i] = 0;
    }
    __syncthreads();

    // Work stealing at two levels.  A wave takes batches of consecutive units
    // from a work counter; its lanes take units from the batch as they need
    // them.  One loop pass advances every busy lane by one window, and a lane
    // whose unit ends starts another in the same pass, so lanes do not wait
    // for the wave's longest unit and waves do not wait for the grid's
    // slowest batch.  The units are split into a.queues contiguous ranges
    // with a counter each, every counter on its own 256-byte line (device-
    // scope atomics to one line serialise); workgroup g draws from queue
    // g % a.queues, so with a multiple of 8 queues each queue is served by
    // one XCD and its bursts stay in that XCD's L2.  Batches are small
    // (about six grabs per wave): the wave that grabs last still walks a
    // whole batch, so the batch sets the grid's tail.  Photons are read
    // through L1/L2.  A replay pass takes its units the same way from the
    // replay list.
    const uint32_t queue = blockIdx.x % queues;
    const uint32_t q_begin = uint32_t(uint64_t(nunits) * queue / queues);
    const uint32_t q_end = uint32_t(uint64_t(nunits) * (queue + 1u) / queues);
    uint32_t pool_next = 0, pool_end = 0;  // the wave's current batch
    bool busy = false;
    uint32_t u = 0;
    WalkState<kFastForm> st{};
    UnitKey key{a.seed, a.set, 0u, 0u};
    BurstDesc d{0u, 0u, 0u, 0.0f};
    const UnitDiv divR = make_unit_div(a.R);
    bool drained = false;
    while (true) {
        bool need = !busy;
        while (!drained) {
            const uint64_t want = __ballot(need);
            if (want == 0) break;
            if (pool_next >= pool_end) {
                uint32_t off = 0;
                if (lane == 0) off = atomicAdd(&counters[queue * kCounterStride], batch);
                off = __shfl(off, 0);
                if (off >= q_end - q_begin) {
                    drained = true;
                    break;
                }
                pool_next = q_begin + off;
                pool_end = pool_next + batch < q_end ? pool_next + batch : q_end;
            }
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                uint32_t(want >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(want), 0u));
            const uint32_t avail = pool_end - pool_next;
            if (need && rank < avail) {
                u = kReplay ? replay[pool_next + rank] : pool_next + rank;
                const uint32_t bl = unit_div(u, divR);
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
            const int r = step_window<kFastForm, false, BinomGuide>(
                sM, &guide, key, d.P, d.stop_rel, d.T, ticks + d.off, links + d.off, a, st);
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
    __syncthreads();
    int32_t* row = hist_acc + size_t(blockIdx.x & 7u) * kMaxBins;
    for (uint32_t i = lane; i < uint32_t(a.B); i += kWave) {
        const int32_t v = s_hist[i];
        if (v) atomicAdd(&row[i], v);
    }
}

// Threshold rows of the fast walk's binomial draws, one wave per (state,
// photon count) row of a parameter set (the operations and their order are
// those of build_guide_row_host): blockIdx.x = state * kBinomTable + count,
// blockIdx.y = set; set i's state s takes table row blobs[i].m.guide_base + s.
// Runs on the stream after the model upload, so the walks that follow read
// finished rows.
__global__ __launch_bounds__(kWave) void guide_kernel(const ModelBlob* __restrict__ blobs,
                                                      ThrPair* __restrict__ thr,
                                                      uint8_t* __restrict__ guide) {
    __shared__ int64_t s_up[kBinomTable + 1];
    const ModelBlob* blob = blobs + blockIdx.y;
    const int s = int(blockIdx.x) / kBinomTable;
    const int m = int(blockIdx.x) % kBinomTable;
    if (s >= blob->m.n) return;  // uniform over the workgroup
    const int srow = blob->m.guide_base + s;
    const int lane = int(threadIdx.x);
    const double pp = blob->m.pp[s];
    const double q = blob->m.qq[s];
    const int bound = blob->tab.bound[s][m];
    const double qn = blob->tab.qn[s][m];
    ThrPair* t = thr + guide_row_offset(srow, m);
    double carryP = 1.0, carryC = 0.0;
    int end = -1;
    bool covered = false;
    for (int c = 0; c * kWave < kBinomTable && end < 0; ++c) {
        const int k = c * kWave + lane;
        double P = k == 0 ? 1.0 : (double(m - k + 1) * pp) / (double(k) * q);
        for (int off = 1; off < kWave; off <<= 1) {
            const double o = __shfl_up(P, off);
            if (lane >= off) P = P * o;
        }
        const double Pf = P * carryP;
        double C = qn * Pf;
        for (int off = 1; off < kWave; off <<= 1) {
            const double o = __shfl_up(C, off);
            if (lane >= off) C = C + o;
        }
        const double Cf = C + carryC;
        const double cs = Cf * 0x1p53;
        const int64_t up = int64_t(floor(cs)) + kGuideBand + 1;
        const int64_t lo = int64_t(ceil(cs)) - kGuideBand - 1;
        const bool full = lo >= kU53Max;
        const uint64_t stop = __ballot(full || k >= bound);
        const int first = stop ? int(__builtin_ctzll(stop)) : kWave;
        if (lane <= first && lane < kWave) {
            t[k] = ThrPair{up, lo};
            s_up[k] = up;
        }
        if (stop) {
            end = c * kWave + first;
            covered = __shfl(full ? 1 : 0, first) != 0;
        }
        carryP = __shfl(Pf, kWave - 1);
        carryC = __shfl(Cf, kWave - 1);
    }
    if (lane == 0) t[end + 1] = ThrPair{INT64_MAX, covered ? INT64_MAX : int64_t(-1)};
    __syncthreads();
    // guide: entries with up <= j 2^47, by binary search over the row
    const int64_t u0 = int64_t(lane) << kGuideShift;
    int lo_i = 0, hi_i = end + 1;  // count in [lo_i, hi_i]
    while (lo_i < hi_i) {
        const int mid = (lo_i + hi_i) >> 1;
        if (s_up[mid] <= u0) lo_i = mid + 1;
        else hi_i = mid;
    }
    guide[(size_t(srow) * kBinomTable + size_t(m)) * kGuideBuckets + size_t(lane)] = uint8_t(lo_i);
}

constexpr uint32_t kFinishWaves = 16;

// Finishes a run in one workgroup of 16 waves: reduces the 8 accumulator
// rows (leaving them zero for the next run), computes the reduced chi2 of
// src/streamWorker.cpp:194-206 in the reference's float accumulation and in
// FP64, and resets the run's counters (walk and replay queues, replay count).
__global__ __launch_bounds__(kWave * kFinishWaves) void finish_kernel(
    const RunArgs a, const float* __restrict__ obs, int32_t* __restrict__ hist_acc,
    uint32_t* __restrict__ replay_count, uint32_t* __restrict__ work,
    ResultHeader* __restrict__ result) {
    __shared__ int32_t s_hist[kMaxBins];
    __shared__ double s_t64[kMaxBins];
    __shared__ double s_t32[kMaxBins];
    const uint32_t tid = threadIdx.x;
    int32_t* rhist = reinterpret_cast<int32_t*>(result + 1);
    const float Rf = float(a.R);
    for (uint32_t b = tid; b < uint32_t(a.B); b += blockDim.x) {
        int32_t Mb = 0;
        for (uint32_t x = 0; x < 8u; ++x) Mb += atomicExch(&hist_acc[size_t(x) * kMaxBins + b], 0);
        s_hist[b] = Mb;
        rhist[b] = Mb;
        // per-bin terms: FP64, and the reference's double term of a float
        // difference (pow(float, 2) / float, src/streamWorker.cpp:200-201)
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
    if (tid == 0) {
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
        result->replays = int32_t(atomicExch(replay_count, 0u));
    }
    if (tid < kMaxQueues + kReplayQueues) atomicExch(&work[tid * kCounterStride], 0u);
}

// ---- population runs ---------------------------------------------------------
//
// A population of parameter sets (an optimiser generation) over the same
// resident bursts in one launch sequence: pop_walk_kernel walks every
// (set, burst, resample) unit, pop_replay_kernel recomputes the units the
// fast form abandoned, pop_finish_kernel computes each set's histogram and
// chi2.  The reference evaluates one set per stream and per mc::run_kernel
// call (src/gpuWorker.cpp:71-80); the units and their random streams are the
// same as a single-set run with set id set_base + s * set_stride.

__device__ __forceinline__ int e_bin(float acc, float F, int B) {
    const float E = F > 0.0f ? acc / F : 0.0f;
    if (!(E >= 0.0f && E < 1.0f)) return -1;
    const int bin = int(floor(double(E) * double(B)));
    return bin < B ? bin : -1;
}

// Adds an LDS histogram to its set's accumulator row of this XCD and clears it.
__device__ __forceinline__ void flush_hist(int32_t* h, int32_t* __restrict__ hist_acc,
                                           uint32_t set, uint32_t xrow, int B, uint32_t lane) {
    int32_t* row = hist_acc + (size_t(set) * 8u + xrow) * size_t(B);
    for (uint32_t i = lane; i < uint32_t(B); i += kWave) {
        const int32_t v = h[i];
        if (v) {
            atomicAdd(&row[i], v);
            h[i] = 0;
        }
    }
}

// Work is a flat list of batches: set s owns batches [s * bps, (s + 1) * bps),
// each covering up to a.batch consecutive units of that set; the list is
// split into a.queues contiguous ranges with a counter each, as in
// walk_kernel.  A wave's lanes may be walking units of two sets at once: the
// workgroup keeps two LDS model slots with an LDS histogram each, and a wave
// moves to a third set only after the older slot's lanes have finished, so it
// never drains to change sets.
// kGuided (fast form): binomial draws from the sets' threshold rows (guide),
// else by the inversion loop from each set's BinomTable.
template <int kPass, bool kGuided>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(
    kPass == kPassFast ? kPopFastWavesPerEU : 1))) void pop_walk_kernel(
    const BurstDesc* __restrict__ bursts, const uint32_t* __restrict__ ticks,
    const uint2* __restrict__ links, const PopControl* __restrict__ ctl,
    const ModelBlob* __restrict__ blobs, const BinomGuide guide, const RunArgs a,
    int32_t* __restrict__ hist_acc,
    uint64_t* __restrict__ replay, uint32_t* __restrict__ replay_count,
    uint32_t* __restrict__ work) {
    __shared__ Model sM[2];
    extern __shared__ int32_t s_hist[];  // [2][a.B]
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < 2u * uint32_t(a.B); i += kWave) s_hist[i] = 0;
    __syncthreads();

    constexpr bool kFastForm = kPass == kPassFast;
    const uint32_t seed = ctl->seed, set_base = ctl->set_base, set_stride = ctl->set_stride;
    const uint32_t queue = blockIdx.x % a.queues;
    const uint32_t nb = a.nsets * a.bps;
    const uint32_t q_begin = uint32_t(uint64_t(nb) * queue / a.queues);
    const uint32_t q_end = uint32_t(uint64_t(nb) * (queue + 1u) / a.queues);
    const uint32_t xrow = blockIdx.x & 7u;
    int set0 = -1, set1 = -1;    // wave-uniform: the set each LDS slot holds
    uint32_t cur = 0;            // slot that receives new units
    uint32_t pool_set = 0, pool_next = 0, pool_end = 0;
    bool busy = false;
    uint32_t my_slot = 0, my_set = 0, u = 0;
    WalkState<kFastForm> st{};
    UnitKey key{seed, 0u, 0u, 0u};
    BurstDesc d{0u, 0u, 0u, 0.0f};
    const UnitDiv divR = make_unit_div(a.R);
    bool drained = false;
    while (true) {
        bool need = !busy;
        while (!drained) {
            const uint64_t want = __ballot(need);
            if (want == 0) break;
            if (pool_next >= pool_end) {
                uint32_t k = 0;
                if (lane == 0) k = atomicAdd(&work[queue * kCounterStride], 1u);
                k = __shfl(k, 0);
                if (k >= q_end - q_begin) {
                    drained = true;
                    break;
                }
                k += q_begin;
                pool_set = k / a.bps;
                pool_next = (k - pool_set * a.bps) * a.batch;
                pool_end = pool_next + a.batch < a.nunits ? pool_next + a.batch : a.nunits;
            }
            if (int(pool_set) != (cur ? set1 : set0)) {
                const uint32_t other = cur ^ 1u;
                if (__any(busy && my_slot == other)) break;  // the older set is still walking
                const int old = other ? set1 : set0;
                if (old >= 0)
                    flush_hist(s_hist + other * uint32_t(a.B), hist_acc, uint32_t(old), xrow,
                               a.B, lane);
                const uint32_t* src = reinterpret_cast<const uint32_t*>(&blobs[pool_set].m);
                uint32_t* dst = reinterpret_cast<uint32_t*>(&sM[other]);
                for (uint32_t i = lane; i < sizeof(Model) / 4; i += kWave) dst[i] = src[i];
                __syncthreads();
                if (other) set1 = int(pool_set);
                else set0 = int(pool_set);
                cur = other;
            }
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                uint32_t(want >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(want), 0u));
            const uint32_t avail = pool_end - pool_next;
            if (need && rank < avail) {
                u = pool_next + rank;
                my_slot = cur;
                my_set = pool_set;
                const uint32_t bl = unit_div(u, divR);
                const uint32_t b = a.begin + bl;
                d = bursts[b];
                key.set = set_base + my_set * set_stride;
                key.r = u - bl * a.R;
                key.b = b;
                if (begin_unit<kFastForm>(sM[my_slot], key, d.P, d.T, st) == kStepMore) {
                    busy = true;
                    need = false;
                } else {
                    const int bin = e_bin(0.0f, 0.0f, a.B);
                    if (bin >= 0) atomicAdd(&s_hist[my_slot * uint32_t(a.B) + uint32_t(bin)], 1);
                }
            }
            const uint32_t taken = uint32_t(__popcll(want));
            pool_next += taken < avail ? taken : avail;
        }
        if (!__any(busy)) break;
        if (busy) {
            int r;
            if (kGuided) {
                const size_t base = size_t(sM[my_slot].guide_base);
                const BinomGuide g{guide.thr + base * kGuideStatePairs,
                                   guide.guide + base * kBinomTable * kGuideBuckets};
                r = step_window<kFastForm, false, BinomGuide>(sM[my_slot], &g, key, d.P,
                                                              d.stop_rel, d.T, ticks + d.off,
                                                              links + d.off, a, st);
            } else {
                r = step_window<kFastForm>(sM[my_slot], &blobs[my_set].tab, key, d.P, d.stop_rel,
                                           d.T, ticks + d.off, links + d.off, a, st);
            }
            if (r != kStepMore) {
                if (r == kStepReplay) {
                    const uint32_t slot = atomicAdd(replay_count, 1u);
                    if (slot < a.replay_cap) replay[slot] = (uint64_t(my_set) << 32) | u;
                } else {
                    const int bin = e_bin(st.acc, st.F, a.B);
                    if (bin >= 0) atomicAdd(&s_hist[my_slot * uint32_t(a.B) + uint32_t(bin)], 1);
                }
                busy = false;
            }
        }
    }
    __syncthreads();
    if (set0 >= 0) flush_hist(s_hist, hist_acc, uint32_t(set0), xrow, a.B, lane);
    if (set1 >= 0) flush_hist(s_hist + uint32_t(a.B), hist_acc, uint32_t(set1), xrow, a.B, lane);
}

// Recomputes abandoned (set, unit) pairs in the general form, one per lane,
// and adds them to row 0 of the set's accumulator.
__global__ __launch_bounds__(kWave) void pop_replay_kernel(
    const BurstDesc* __restrict__ bursts, const uint32_t* __restrict__ ticks,
    const uint2* __restrict__ links, const PopControl* __restrict__ ctl,
    const ModelBlob* __restrict__ blobs, const RunArgs a, int32_t* __restrict__ hist_acc,
    const uint64_t* __restrict__ replay, const uint32_t* __restrict__ replay_count) {
    const uint32_t count = *replay_count < a.replay_cap ? *replay_count : a.replay_cap;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const uint64_t e = replay[i];
        const uint32_t set = uint32_t(e >> 32), u = uint32_t(e);
        const uint32_t bl = u / a.R;
        const uint32_t b = a.begin + bl;
        const BurstDesc d = bursts[b];
        const UnitKey key{ctl->seed, ctl->set_base + set * ctl->set_stride, u - bl * a.R, b};
        const WalkOut o = walk_unit<false>(blobs[set].m, nullptr, key, d.P, d.stop_rel, d.T,
                                           ticks + d.off, links + d.off, a);
        const int bin = e_bin(o.acc, o.F, a.B);
        if (bin >= 0) atomicAdd(&hist_acc[size_t(set) * 8u * size_t(a.B) + uint32_t(bin)], 1);
    }
}

// One workgroup per set: reduces the set's 8 accumulator rows (leaving them
// zero), computes its chi2 as finish_kernel does, with the set's own degrees
// of freedom; workgroup 0 reports and resets the replay count and the work
// counters.
__global__ __launch_bounds__(kWave) void pop_finish_kernel(
    const ModelBlob* __restrict__ blobs, const RunArgs a, const float* __restrict__ obs,
    int32_t* __restrict__ hist_acc, uint32_t* __restrict__ replay_count,
    uint32_t* __restrict__ work, PopStatus* __restrict__ status, SetResult* __restrict__ results,
    int32_t* __restrict__ hists) {
    __shared__ int32_t s_M[kMaxBins];
    __shared__ double s_t64[kMaxBins];
    __shared__ double s_t32[kMaxBins];
    const uint32_t set = blockIdx.x, lane = threadIdx.x;
    const uint32_t B = uint32_t(a.B);
    int32_t* rows = hist_acc + size_t(set) * 8u * B;
    const float Rf = float(a.R);
    for (uint32_t b = lane; b < B; b += kWave) {
        int32_t Mb = 0;
        for (uint32_t x = 0; x < 8u; ++x) Mb += atomicExch(&rows[x * B + b], 0);
        s_M[b] = Mb;
        hists[size_t(set) * B + b] = Mb;
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
            if (s_M[b] > 0) {
                c32 = float(double(c32) + s_t32[b]);
                c64 += s_t64[b];
            } else {
                --effN;
            }
        }
        const int n = blobs[set].m.n;
        const int dof = effN - n * (n + 1);
        SetResult r;
        r.chi2_ref = c32 / float(dof);
        r.chi2_64 = c64 / double(dof);
        r.effN = effN;
        r.n = n;
        r.pad = 0;
        results[set] = r;
    }
    if (set == 0) {
        if (lane == 0) {
            const uint32_t c = atomicExch(replay_count, 0u);
            PopStatus st;
            st.replays = c;
            st.overflow = c > a.replay_cap ? 1u : 0u;
            st.pad[0] = st.pad[1] = 0;
            *status = st;
        }
        for (uint32_t q = lane; q < kMaxQueues; q += kWave) atomicExch(&work[q * kCounterStride], 0u);
    }
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

// ---- microbenchmarks of the algorithm's events ---------------------------------
//
// The weights of the algorithmic cost model: each lane repeats one event of
// the walk -- a Philox block, the scalar work of a dwell window, a
// binary-search step, a background draw, a capped-Poisson iteration, a
// binomial draw, a binomial inversion iteration, the per-unit start and
// tally -- with a loop-carried dependence, counts the events it performed
// and writes a checksum.  Launched with 256-thread workgroups at full
// occupancy, the loops are bound by VALU issue, so the time per event times
// the device's 32-bit lane-operation rate is the event's weight in
// lane-issue-cycles.  Inputs vary with a 32-bit LCG step per event (part of
// the measured cost, standing in for reading a Philox word).
enum MicroOp : int {
    kMbPhilox = 0,   // one Philox4x32-10 block
    kMbWindow = 1,   // a dwell window's scalar work (next state, dwell, window end,
                     // corrected count, its rounding and sums)
    kMbSearch = 2,   // one step of the window-end binary search (LDS-resident ticks,
                     // whole 7-step searches)
    kMbPoisson = 3,  // one capped-Poisson draw (and its inversion iterations)
    kMbBinom = 4,    // one binomial inversion draw from the q^m table (and its
                     // iterations)
    kMbUnit = 5,     // a unit's initial state draw and its tally into the histogram
    kMbOutcome = 6,  // one outcome of the Rao-Blackwellised spread: the bin of E = a / F
                     // and the 2^-32 fixed-point weight of its probability
};
constexpr int kMicroOps = 7;
constexpr uint32_t kMbThreads = 256;

__device__ __forceinline__ uint32_t lcg(uint32_t x) { return x * 1664525u + 1013904223u; }

template <int kOp>
__global__ __launch_bounds__(kMbThreads) void microbench_kernel(int iters, float param,
                                                                const double* __restrict__ qn,
                                                                const int32_t* __restrict__ bound,
                                                                uint64_t* __restrict__ out) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t h = gid * 0x9E3779B9u + 0x7F4A7C15u;
    uint64_t events = 0, inner = 0;
    __shared__ uint32_t s_tick[128];
    __shared__ int32_t s_hist[kMbThreads];
    if (kOp == kMbSearch || kOp == kMbUnit) {
        for (uint32_t i = threadIdx.x; i < 128u; i += blockDim.x) s_tick[i] = i * 97u;
        for (uint32_t i = threadIdx.x; i < kMbThreads; i += blockDim.x) s_hist[i] = 0;
        __syncthreads();
    }
    if (kOp == kMbPhilox) {
        // every counter and key word changes per block (no part of the
        // rounds can be hoisted out of the loop)
        Block x{{h, gid, 0u, 0u}};
        for (int i = 0; i < iters; ++i)
            x = philox10(uint32_t(i) ^ x.w[3], x.w[0], gid ^ x.w[2], x.w[1], h ^ uint32_t(i),
                         x.w[2] ^ x.w[3]);
        h = x.w[0] ^ x.w[1] ^ x.w[2] ^ x.w[3];
        events = uint64_t(iters);
    } else if (kOp == kMbWindow) {
        const float cum[2] = {0.4f, 1.0f};
        const float krate[3] = {900.0f, 3000.0f, 5500.0f};
        const float T = 1e-3f, clk = 12.5e-9f, gmx = 0.95f, omdd = 1.0f;
        float tsum = 0.0f, F = 0.0f, acc = 0.0f;
        uint32_t w0 = 0;
        int st = 0;
        for (int i = 0; i < iters; ++i) {
            h = lcg(h);
            const int jj = first_above(cum, 2, uniform24(h));
            const int j = jj >= st ? jj + 1 : jj;
            h = lcg(h);
            const float kk = j == 0 ? krate[0] : (j == 1 ? krate[1] : krate[2]);
            const float tau = flog(uniform24(h)) / (-kk);
            tsum = tsum + tau;
            const bool end = !(tsum < T);
            const float wf = float(w0) + tsum / clk;
            const uint32_t w1 = end ? 80000u : uint32_t(wf);
            const uint32_t cnt = (w1 - w0) & 31u;
            const float nd = (h & 1u) ? float(cnt) : 0.0f;
            const float na = (h & 1u) ? 0.0f : float(cnt);
            const float fv = gmx * nd + omdd * na;
            const float m = floorf(fv + 0.5f);
            F += fv;
            acc += m;
            w0 = end ? 0u : w1;
            tsum = end ? 0.0f : tsum;
            st = j;
        }
        h ^= fbits(F) ^ fbits(acc);
        events = uint64_t(iters);
    } else if (kOp == kMbSearch) {
        // whole lower_bound searches over 127 ticks (7 steps each), a new
        // key per search
        uint32_t acc = 0;
        for (int i = 0; i < iters; ++i) {
            h = lcg(h);
            const uint32_t key = h & 0x3FFFu;
            uint32_t lo = 0, hi = 127;
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_tick[mid] < key) lo = mid + 1u;
                else hi = mid;
            }
            acc += lo;
        }
        h ^= acc;
        events = uint64_t(iters) * 7u;
    } else if (kOp == kMbPoisson) {
        uint64_t it = 0;
        for (int i = 0; i < iters; ++i) {
            h = lcg(h);
            const float u1 = uniform24(h);
            h = lcg(h);
            const float k = background_single(param, 1e6f, u1, uniform24(h));
            it += uint64_t(k);
            h ^= fbits(k);
        }
        events = uint64_t(iters);
        inner = it;
    } else if (kOp == kMbBinom) {
        const int n = int(param);
        const double pp = 0.3, q = 0.7;
        uint64_t it = 0;
        for (int i = 0; i < iters; ++i) {
            h = lcg(h);
            const uint32_t lo32 = h;
            h = lcg(h);
            // the inversion of binomial_table (pda_walk.h) on a q^m table
            double px = qn[n];
            const int bnd = bound[n];
            double U = uniform53(h, lo32);
            int X = 0;
            while (U > px) {
                ++X;
                if (X > bnd) break;
                U -= px;
                px = (double(n - X + 1) * pp * px) / (double(X) * q);
            }
            it += uint64_t(X);
            h ^= uint32_t(X);
        }
        events = uint64_t(iters);
        inner = it;
    } else if (kOp == kMbUnit) {
        const float p0cum[3] = {0.2f, 0.7f, 1.0f};
        float acc = 3.0f, F = 7.0f;
        for (int i = 0; i < iters; ++i) {
            h = lcg(h);
            const int s0 = first_above(p0cum, 3, uniform24(h));
            acc += float(s0);
            F += float(h & 7u);
            const float E = F > 0.0f ? acc / F : 0.0f;
            if (E >= 0.0f && E < 1.0f) {
                // the histogram update; one address per lane (how often lanes
                // collide on a bin depends on the data, not on the algorithm)
                const int bin = int(floor(double(E) * 50.0));
                if (bin < 50) atomicAdd(&s_hist[threadIdx.x], bin + 1);
            }
            acc = E * 3.0f;
        }
        h ^= fbits(acc);
        events = uint64_t(iters);
    } else if (kOp == kMbOutcome) {
        // rb_bin and rb_fixed of pda_rb.h on fresh (a, F, P) per event; the
        // sums carry the results out
        uint64_t fx = 0;
        int32_t bins = 0;
        for (int i = 0; i < iters; ++i) {
            h = lcg(h);
            const int a = int(h & 31u);
            const float F = 8.0f + float(h >> 27) + param;
            const float P = uniform24(h ^ 0x5bd1e995u);
            bins += rb_bin(a, F, 50);
            fx += rb_fixed(P);
        }
        h ^= uint32_t(fx) ^ uint32_t(fx >> 32) ^ uint32_t(bins);
        events = uint64_t(iters);
    }
    out[3u * gid] = h;
    out[3u * gid + 1u] = events;
    out[3u * gid + 2u] = inner;
}

// ---- launchers (called by the C-ABI layer) ---------------------------------

hipError_t walk_occupancy(int pass, int* blocks_per_cu) {
    if (pass == kPassFast)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, walk_kernel<kPassFast>,
                                                            kWave, 0);
    if (pass == kPassReplay)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu,
                                                            walk_kernel<kPassReplay>, kWave, 0);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, walk_kernel<kPassGeneral>,
                                                        kWave, 0);
}

hipError_t launch_walk(int pass, uint32_t grid, const BurstDesc* bursts, const uint32_t* ticks,
                       const uint2* links, const ModelBlob* blob, const BinomGuide& guide,
                       const RunArgs& a, int32_t* hist_acc, uint32_t* replay,
                       uint32_t* replay_count, uint32_t* work, float* outE, int32_t* outAcc,
                       float* outF, hipStream_t stream) {
    if (pass == kPassFast)
        hipLaunchKernelGGL(walk_kernel<kPassFast>, dim3(grid), dim3(kWave), 0, stream, bursts,
                           ticks, links, blob, guide, a, hist_acc, replay, replay_count, work, outE,
                           outAcc, outF);
    else if (pass == kPassReplay)
        hipLaunchKernelGGL(walk_kernel<kPassReplay>, dim3(grid), dim3(kWave), 0, stream, bursts,
                           ticks, links, blob, guide, a, hist_acc, replay, replay_count, work, outE,
                           outAcc, outF);
    else
        hipLaunchKernelGGL(walk_kernel<kPassGeneral>, dim3(grid), dim3(kWave), 0, stream, bursts,
                           ticks, links, blob, guide, a, hist_acc, replay, replay_count, work, outE,
                           outAcc, outF);
    return hipGetLastError();
}

hipError_t launch_guide(const ModelBlob* blobs, int nsets, int max_states, ThrPair* thr,
                        uint8_t* guide, hipStream_t stream) {
    hipLaunchKernelGGL(guide_kernel, dim3(uint32_t(max_states) * kBinomTable, uint32_t(nsets)),
                       dim3(kWave), 0, stream, blobs, thr, guide);
    return hipGetLastError();
}

hipError_t launch_finish(const RunArgs& a, const float* obs, int32_t* hist_acc,
                         uint32_t* replay_count, uint32_t* work, ResultHeader* result,
                         hipStream_t stream) {
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(kWave * kFinishWaves), 0, stream, a, obs,
                       hist_acc, replay_count, work, result);
    return hipGetLastError();
}

hipError_t pop_walk_occupancy(int pass, int B, int* blocks_per_cu) {
    const size_t lds = sizeof(int32_t) * 2u * size_t(B);
    if (pass == kPassFast)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            blocks_per_cu, pop_walk_kernel<kPassFast, true>, kWave, lds);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        blocks_per_cu, pop_walk_kernel<kPassGeneral, false>, kWave, lds);
}

hipError_t launch_pop_walk(int pass, uint32_t grid, const BurstDesc* bursts,
                           const uint32_t* ticks, const uint2* links, const PopControl* ctl,
                           const ModelBlob* blobs, const BinomGuide& guide, const RunArgs& a,
                           int32_t* hist_acc, uint64_t* replay, uint32_t* replay_count,
                           uint32_t* work, hipStream_t stream) {
    const size_t lds = sizeof(int32_t) * 2u * size_t(a.B);
    if (pass == kPassFast && guide.thr)
        hipLaunchKernelGGL((pop_walk_kernel<kPassFast, true>), dim3(grid), dim3(kWave), lds,
                           stream, bursts, ticks, links, ctl, blobs, guide, a, hist_acc, replay,
                           replay_count, work);
    else if (pass == kPassFast)
        hipLaunchKernelGGL((pop_walk_kernel<kPassFast, false>), dim3(grid), dim3(kWave), lds,
                           stream, bursts, ticks, links, ctl, blobs, guide, a, hist_acc, replay,
                           replay_count, work);
    else
        hipLaunchKernelGGL((pop_walk_kernel<kPassGeneral, false>), dim3(grid), dim3(kWave), lds,
                           stream, bursts, ticks, links, ctl, blobs, guide, a, hist_acc, replay,
                           replay_count, work);
    return hipGetLastError();
}

hipError_t launch_pop_replay(const BurstDesc* bursts, const uint32_t* ticks, const uint2* links,
                             const PopControl* ctl, const ModelBlob* blobs, const RunArgs& a,
                             int32_t* hist_acc, const uint64_t* replay,
                             const uint32_t* replay_count, uint32_t grid, hipStream_t stream) {
    hipLaunchKernelGGL(pop_replay_kernel, dim3(grid), dim3(kWave), 0, stream, bursts, ticks,
                       links, ctl, blobs, a, hist_acc, replay, replay_count);
    return hipGetLastError();
}

hipError_t launch_pop_finish(const ModelBlob* blobs, const RunArgs& a, const float* obs,
                             int32_t* hist_acc, uint32_t* replay_count, uint32_t* work,
                             PopStatus* status, SetResult* results, int32_t* hists,
                             hipStream_t stream) {
    hipLaunchKernelGGL(pop_finish_kernel, dim3(a.nsets), dim3(kWave), 0, stream, blobs, a, obs,
                       hist_acc, replay_count, work, status, results, hists);
    return hipGetLastError();
}

hipError_t launch_primitive(int op, const void* in, void* out, int64_t n, hipStream_t stream) {
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(primitive_kernel, dim3(uint32_t(blocks)), dim3(256), 0, stream, op, in, out,
                       n);
    return hipGetLastError();
}

hipError_t microbench_shape(int op, uint32_t* grid, uint32_t* threads) {
    int per_cu = 0;
    hipError_t e = hipSuccess;
    switch (op) {
        case kMbPhilox: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, microbench_kernel<kMbPhilox>, kMbThreads, 0); break;
        case kMbWindow: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, microbench_kernel<kMbWindow>, kMbThreads, 0); break;
        case kMbSearch: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, microbench_kernel<kMbSearch>, kMbThreads, 0); break;
        case kMbPoisson: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, microbench_kernel<kMbPoisson>, kMbThreads, 0); break;
        case kMbBinom: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, microbench_kernel<kMbBinom>, kMbThreads, 0); break;
        case kMbUnit: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, microbench_kernel<kMbUnit>, kMbThreads, 0); break;
        case kMbOutcome: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, microbench_kernel<kMbOutcome>, kMbThreads, 0); break;
        default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    int dev = 0, cus = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    *grid = uint32_t(cus) * uint32_t(per_cu);
    *threads = kMbThreads;
    return hipSuccess;
}

hipError_t launch_microbench(int op, uint32_t grid, int iters, float param, const double* qn,
                             const int32_t* bound, uint64_t* out, hipStream_t stream) {
    const dim3 g(grid), b(kMbThreads);
    switch (op) {
        case kMbPhilox: hipLaunchKernelGGL(microbench_kernel<kMbPhilox>, g, b, 0, stream, iters, param, qn, bound, out); break;
        case kMbWindow: hipLaunchKernelGGL(microbench_kernel<kMbWindow>, g, b, 0, stream, iters, param, qn, bound, out); break;
        case kMbSearch: hipLaunchKernelGGL(microbench_kernel<kMbSearch>, g, b, 0, stream, iters, param, qn, bound, out); break;
        case kMbPoisson: hipLaunchKernelGGL(microbench_kernel<kMbPoisson>, g, b, 0, stream, iters, param, qn, bound, out); break;
        case kMbBinom: hipLaunchKernelGGL(microbench_kernel<kMbBinom>, g, b, 0, stream, iters, param, qn, bound, out); break;
        case kMbUnit: hipLaunchKernelGGL(microbench_kernel<kMbUnit>, g, b, 0, stream, iters, param, qn, bound, out); break;
        case kMbOutcome: hipLaunchKernelGGL(microbench_kernel<kMbOutcome>, g, b, 0, stream, iters, param, qn, bound, out); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace pda
