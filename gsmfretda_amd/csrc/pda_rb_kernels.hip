// pda_rb_kernels.hip -- gfx950 kernels of the Rao-Blackwellised (RB) PDA form.
//
//   rb_walk_kernel       the walk of every (burst, resample) unit with the
//                        keyed draws of the plain kernel (src/mc.cu:92-191)
//                        but no acceptor draw; per unit it stores the photon
//                        counts per state, F, and (for rate gradients) the
//                        score of the walk's draws, and lists the units whose
//                        acceptor total is too large for the memo; fast form,
//                        the rare units beyond it listed for
//   rb_replay_kernel     the general form over those units
//   rb_table_kernel      the binomial pmfs of every state, counts < 256
//   rb_memo_kernel       the pmf (a convolution of one binomial per state)
//                        of every count vector with a small acceptor total,
//                        one lane per vector, in a strip of LDS of its own,
//                        (the derivative passes' leave-one-trial-out pmfs
//                        are the pmfs of the vectors with one trial fewer)
//   rb_unit_kernel       every unit with a small total: its pmf from the
//                        memo spread into the E histogram as 2^-32 fixed
//                        point (forward), or the cotangent-weighted pmf and
//                        its derivatives (adjoint), or the derivatives of
//                        every bin (Jacobian); lanes over the (unit,
//                        outcome) pairs of 64 units at a time
//   rb_big_kernel        the same for the listed units, one wave per unit,
//                        lanes over acceptor totals, the pmf built in LDS
//   rb_score_mean_kernel per burst, the mean score of its R resamples
//   rb_finish_kernel     reduces the histogram rows, computes chi2 on the
//                        expected counts, resets the counters
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pda_math.h"
#include "pda_rb.h"
#include "pda_types.h"
#include "pda_walk.h"

namespace pda {

// A unit's pmf has Mt + 1 entries, Mt its acceptor total (the sum of its
// photon counts per state).  The pmfs of every count vector with
// Mt < kRbMemoCap are built once per forward (rb_memo_kernel) and the units
// look theirs up (rb_unit_kernel); the walk lists the units with larger
// totals for rb_big_kernel.  The mean Mt is about 6 on the 3-state
// workloads and 24 on the 4-state fast-dynamics one, where 97 % of the
// units have Mt < 48.
//
// rb_memo_kernel: one lane per count vector, the pmf in a strip of LDS of
// its own; kRbStride floats between the strips of consecutive lanes (odd,
// so that lanes reading the same offset hit distinct LDS banks).
constexpr int kRbStride = kRbMemoCap + 1;
constexpr int kRbMemoWaves = 1;  // waves per workgroup of rb_memo_kernel
// Entries per state of the pmf table rb_memo_kernel keeps in LDS: counts
// m < kRbMemoCap, entry (m, k) at m (m + 1) / 2 + k (the first entries of
// the state's block of the global table).
constexpr int kRbMemoTab = kRbMemoCap * (kRbMemoCap + 1) / 2;
// rb_unit_kernel: waves per workgroup, and its work counters (the
// kAuxQueues lines after the walk's, workgroup g on counter g % kAuxQueues,
// each over a contiguous share of the units: one queue per XCD).
constexpr int kRbUnitWaves = 4;
constexpr uint32_t kRbUnitCounter = kMaxQueues * kCounterStride;
// Jacobian rows in global memory, kRbJacCols doubles each: the set's n
// efficiency columns, then its n (n - 1) rates (rb_big_kernel's LDS rows
// hold the rates from column kRbMaxStates).
constexpr int kRbJacCols = kRbMaxStates + kRbMaxPairs;

constexpr int kRbFwd = 0, kRbGrad = 1, kRbJac = 2;

// What an RB walk accumulates over a unit's windows, in registers (kN
// states, known at compile time, so every index below is a constant): the
// photon counts per state and the score of the walk's draws per rate pair,
//   d/dK(j', s) log[P(j | s) K(j, s) exp(-K(j, s) tau)]
//     = [j' = j] (2 / K(j, s) - tau) - 1 / sum_{j''} K(j'', s)
// summed over the windows in float, in the order of the oracle's
// restatement: per window, the taken pair's 2 / K - tau first, then
// 1 / sum K off every pair leaving the window's state.
template <int kN, bool kScores>
struct RbAcc {
    static constexpr int kK = kN * (kN - 1);
    uint32_t cnt[kN];
    float S[kScores ? kK : 1];

    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int s = 0; s < kN; ++s) cnt[s] = 0u;
#pragma unroll
        for (int i = 0; i < (kScores ? kK : 1); ++i) S[i] = 0.0f;
    }

    __device__ __forceinline__ void absorb(const RbModel& R, const WindowInfo& wi) {
        const uint32_t m = wi.m > 0 ? uint32_t(wi.m) : 0u;
#pragma unroll
        for (int s = 0; s < kN; ++s) cnt[s] += s == wi.s ? m : 0u;
        if (kScores) {
            const int p = rate_pair(kN, wi.s, wi.j);
            const float v = R.two_over_k[wi.s][wi.j] - wi.tau;
            const float inv = R.inv_sum[wi.s];
#pragma unroll
            for (int i = 0; i < kK; ++i) {
                // pair i = rate_pair(kN, from, to): to = i / (kN - 1)
                const int to = i / (kN - 1), r = i % (kN - 1);
                const int from = r < to ? r : r + 1;
                const float t = i == p ? S[i] + v : S[i];
                S[i] = from == wi.s ? t - inv : S[i];
            }
        }
    }

    // Stores the unit: its packed counts, F and scores; a unit too long for
    // the memo joins the big list (total < kRbBigSplit, count big_count[0])
    // or the huge one (count big_count[1]): one atomic per wave, list and
    // call site, the lanes storing now ranked among themselves.
    __device__ __forceinline__ void store(uint32_t unit, float F, uint32_t lane,
                                          uint2* __restrict__ cnt_out, float* __restrict__ F_out,
                                          float* __restrict__ S_out, uint32_t* __restrict__ big,
                                          uint32_t* __restrict__ huge,
                                          uint32_t* __restrict__ big_count) const {
        uint32_t c[kRbMaxStates] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int s = 0; s < kN; ++s) c[s] = cnt[s];
        uint2 packed;  // rb_pack of the four counts (states >= n hold 0)
        packed.x = (c[0] & 0xFFFFu) | (c[1] << 16);
        packed.y = (c[2] & 0xFFFFu) | (c[3] << 16);
        cnt_out[unit] = packed;
        F_out[unit] = F;
        if (kScores) {
            // kK (2, 6 or 12) floats per unit: 8-byte aligned rows, two floats
            // a store
            float2* row = reinterpret_cast<float2*>(S_out + size_t(unit) * size_t(kK));
#pragma unroll
            for (int i = 0; i < kK / 2; ++i) row[i] = make_float2(S[2 * i], S[2 * i + 1]);
        }
        const uint32_t mt = c[0] + c[1] + c[2] + c[3];
        append(mt >= uint32_t(kRbMemoCap) && mt < uint32_t(kRbBigSplit), unit, lane, big,
               &big_count[0]);
        append(mt >= uint32_t(kRbBigSplit), unit, lane, huge, &big_count[1]);
    }

    __device__ __forceinline__ static void append(bool take, uint32_t unit, uint32_t lane,
                                                  uint32_t* __restrict__ list,
                                                  uint32_t* __restrict__ count) {
        const uint64_t mk = __ballot(take);
        if (mk) {
            const int leader = __ffsll(static_cast<long long>(mk)) - 1;
            uint32_t base = 0;
            if (int(lane) == leader) base = atomicAdd(count, uint32_t(__popcll(mk)));
            base = __shfl(base, leader);
            if (take)
                list[base + __builtin_amdgcn_mbcnt_hi(uint32_t(mk >> 32),
                                                      __builtin_amdgcn_mbcnt_lo(uint32_t(mk), 0u))] = unit;
        }
    }
};

// The walk of every unit in the fast form (32-bit window ends, background
// means <= 64: the plain walk's common case) with the work stealing of
// walk_kernel; a unit that leaves the common case is listed for
// rb_replay_kernel, which walks it again from its start in the general form
// with the same keyed draws, so every unit's stored summary is the general
// form's.
template <bool kScores, int kN>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(7))) void rb_walk_kernel(
    const BurstDesc* __restrict__ bursts, const uint32_t* __restrict__ ticks,
    const uint2* __restrict__ links, const ModelBlob* __restrict__ blob,
    const RbModel* __restrict__ rbm, const RunArgs a, uint2* __restrict__ cnt_out,
    float* __restrict__ F_out, float* __restrict__ S_out, uint32_t* __restrict__ work,
    uint32_t* __restrict__ big, uint32_t* __restrict__ huge, uint32_t* __restrict__ big_count,
    uint32_t* __restrict__ replay, uint32_t* __restrict__ replay_count) {
    __shared__ Model sM;
    __shared__ RbModel sR;
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
    RbAcc<kN, kScores> acc;

    const uint32_t queue = blockIdx.x % a.queues;
    const uint32_t q_begin = uint32_t(uint64_t(a.nunits) * queue / a.queues);
    const uint32_t q_end = uint32_t(uint64_t(a.nunits) * (queue + 1u) / a.queues);
    uint32_t pool_next = 0, pool_end = 0;
    bool busy = false;
    uint32_t u = 0;
    WalkState<true> st{};
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
                const uint32_t bl = unit_div(u, divR);
                const uint32_t b = a.begin + bl;
                d = bursts[b];
                key.r = u - bl * a.R;
                key.b = b;
                acc.clear();
                if (begin_unit<true>(sM, key, d.P, d.T, st) == kStepMore) {
                    busy = true;
                    need = false;
                } else {
                    acc.store(u, st.F, lane, cnt_out, F_out, S_out, big, huge, big_count);
                }
            }
            const uint32_t taken = uint32_t(__popcll(want));
            pool_next += taken < avail ? taken : avail;
        }
        if (!__any(busy)) break;
        if (busy) {
            WindowInfo wi;
            const int r = step_window<true, true>(sM, nullptr, key, d.P, d.stop_rel, d.T,
                                                  ticks + d.off, links + d.off, a, st, &wi);
            if (r == kStepReplay) {
                replay[atomicAdd(replay_count, 1u)] = u;
                busy = false;
            } else {
                acc.absorb(sR, wi);
                if (r != kStepMore) {
                    acc.store(u, st.F, lane, cnt_out, F_out, S_out, big, huge, big_count);
                    busy = false;
                }
            }
        }
    }
}

// The units rb_walk_kernel listed, in the general form, one unit per lane
// over the list; workgroups beyond the list's need exit at once.  Launched
// after rb_walk_kernel on the same stream.
template <bool kScores, int kN>
__global__ __launch_bounds__(kWave) void rb_replay_kernel(
    const BurstDesc* __restrict__ bursts, const uint32_t* __restrict__ ticks,
    const uint2* __restrict__ links, const ModelBlob* __restrict__ blob,
    const RbModel* __restrict__ rbm, const RunArgs a, uint2* __restrict__ cnt_out,
    float* __restrict__ F_out, float* __restrict__ S_out, uint32_t* __restrict__ big,
    uint32_t* __restrict__ huge, uint32_t* __restrict__ big_count,
    const uint32_t* __restrict__ replay, const uint32_t* __restrict__ replay_count) {
    const uint32_t count = *replay_count;
    if (blockIdx.x * kWave >= count) return;  // uniform over the workgroup
    __shared__ Model sM;
    __shared__ RbModel sR;
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
    const uint32_t first = blockIdx.x * kWave;
    const uint32_t stride = gridDim.x * kWave;
    for (uint32_t base = first; base < count; base += stride) {
        const uint32_t i = base + lane;
        if (i < count) {
            const uint32_t u = replay[i];
            const uint32_t bl = u / a.R;
            const uint32_t b = a.begin + bl;
            const BurstDesc d = bursts[b];
            const UnitKey key{a.seed, a.set, u - bl * a.R, b};
            WalkState<false> st{};
            RbAcc<kN, kScores> acc;
            acc.clear();
            int r = begin_unit<false>(sM, key, d.P, d.T, st);
            while (r == kStepMore) {
                WindowInfo wi;
                r = step_window<false, true>(sM, nullptr, key, d.P, d.stop_rel, d.T,
                                             ticks + d.off, links + d.off, a, st, &wi);
                acc.absorb(sR, wi);
            }
            acc.store(u, st.F, lane, cnt_out, F_out, S_out, big, huge, big_count);
        }
    }
}

// Sum over the wave (every lane gets the total, the same on every run).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Inclusive prefix sum / maximum over the lanes of the wave.
__device__ __forceinline__ int wave_scan_sum(int v, int lane) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int t = __shfl_up(v, d);
        if (lane >= d) v += t;
    }
    return v;
}
__device__ __forceinline__ int wave_scan_max(int v, int lane) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int t = __shfl_up(v, d);
        if (lane >= d) v = t > v ? t : v;
    }
    return v;
}

// Orders the LDS accesses of the wave's lanes: every lane's earlier LDS
// writes are visible to every lane's later LDS reads.  The lanes of a wave
// share data through LDS here without a workgroup barrier; the compiler
// treats each lane as a thread of its own and, without this fence, may keep
// a value a lane stored in a register instead of reading what another lane
// wrote over it, or move accesses across the exchange.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Spreading the elements of 64 variable-length items over the lanes: item
// l (owned by lane l) has len_l elements at positions pos_l .. pos_l +
// len_l - 1 of the wave's concatenation (pos = exclusive prefix sum of len).
// Window w holds positions 64 w .. 64 w + 63, one per lane; the owner of a
// lane's position is the last item starting at or before it: each item
// starting inside the window marks its start slot, and a running maximum
// over the slots (carried across windows) names the owner.  slot: 64 ints of
// the wave's LDS.  Calls body(j, owner) for each lane (j >= total: no
// element), with every lane of the wave present at each call (shuffles may
// be used); what the caller's lanes wrote to LDS before the call is visible
// to body, and the body's LDS accesses end before the next window's.
template <typename Body>
__device__ __forceinline__ void wave_windows(int len, int pos, int total, int lane, int* slot,
                                             Body&& body) {
    int carry = -1;
    wave_lds_sync();
    for (int wb = 0; wb < total; wb += kWave) {
        slot[lane] = -1;
        wave_lds_sync();
        if (len > 0 && pos >= wb && pos < wb + kWave) slot[pos - wb] = lane;
        wave_lds_sync();
        int own = wave_scan_max(slot[lane], lane);
        own = own > carry ? own : carry;
        carry = __shfl(own, kWave - 1);
        body(wb + lane, own);
        wave_lds_sync();
    }
}

// What the passes compute per unit, with P(a), a = 0..Mt, the unit's pmf
// and bin(a) the histogram bin of E = a / F:
//
// Forward: P(a) into the bin of a, in 2^-32 fixed point (integer sums: the
// result does not depend on the order in which units or outcomes arrive).
//
// Adjoint, for a cotangent G over the bins: h = sum_a G(bin(a)) P(a) and,
// per state s with m_s photons,
//   d h / d p_s = m_s sum_c P_s(c) (G(bin(c + 1)) - G(bin(c)))
// where P_s is the pmf with one trial of state s removed,
// P(a) = q_s P_s(a) + p_s P_s(a - 1): for a memo unit the memo's pmf of its
// count vector with m_s - 1 (built by the same convolutions); rb_big_kernel
// gets it from P by a first-order recurrence, upwards when p_s <= 1/2 and
// downwards otherwise (the direction in which it damps rounding errors).
// The rate gradient is the likelihood
// ratio R / (R - 1) sum_units h (S - Sbar), S the unit's score and Sbar the
// mean score of its burst's R resamples: the per-burst leave-one-out
// baseline, sum_r h_r S_r - (sum_r h_r)(sum_r S_r) / R, summed per unit.
//
// Jacobian: the adjoint for every indicator cotangent G_b(a) = [bin(a) = b]
// at once.  G_b(c + 1) - G_b(c) is nonzero only where c and c + 1 fall in
// different bins, so d M_b / d p_s gathers m_s P_s(c) at the unit's bin
// boundaries; the rate columns add P(a) (S - Sbar) to the row of bin(a).

// Bytes of rb_memo_kernel's dynamic LDS: the n states' pmf tables, then per
// lane one strip (the pmf).
__host__ __device__ inline size_t rb_memo_lds(int n) {
    return sizeof(float) * (size_t(n) * kRbMemoTab + size_t(kRbMemoWaves) * kWave * kRbStride);
}

// The memo: for each count vector of the list (one lane each) its pmf P
// into poolP at its offP.  (The derivative passes' leave-one-out pmfs are
// the pmfs of the vectors with one trial fewer, in the same pool.)  The list
// runs by total, the offsets follow it, and each wave copies its lanes'
// blocks out of LDS together: adjacent blocks, coalesced writes.
__global__ __launch_bounds__(kWave * kRbMemoWaves) void rb_memo_kernel(
    const RbModel* __restrict__ rbm, const float* __restrict__ pmf_tab,
    const uint4* __restrict__ list, uint32_t nvec, float* __restrict__ poolP) {
    extern __shared__ float s_f[];
    __shared__ RbModel sR;
    __shared__ int s_slot[kRbMemoWaves][kWave];
    __shared__ int s_pos[kRbMemoWaves][kWave];
    __shared__ uint32_t s_dst[kRbMemoWaves][kWave];
    const int tid = int(threadIdx.x), lane = tid % kWave, wave = tid / kWave;
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(rbm);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&sR);
        for (uint32_t i = uint32_t(tid); i < sizeof(RbModel) / 4; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    const int n = sR.n;
    float* tab = s_f;
    for (int i = tid; i < n * kRbMemoTab; i += int(blockDim.x)) {
        const int s = i / kRbMemoTab;
        tab[i] = pmf_tab[size_t(s) * kRbPmfStride + size_t(i - s * kRbMemoTab)];
    }
    __syncthreads();
    float* strips = tab + n * kRbMemoTab + size_t(wave) * kWave * kRbStride;
    float* X = strips + lane * kRbStride;
    const uint32_t v = blockIdx.x * blockDim.x + uint32_t(tid);
    const bool act = v < nvec;
    const uint4 e = act ? list[v] : make_uint4(0u, 0u, 0u, 0u);
    const uint2 c = make_uint2(e.x, e.y);
    // the pmf in X: the states' binomials convolved in state order, each
    // output summed k ascending from 0.0f (the oracle's order); outputs
    // descending, so that each overwrites an entry of X no later output reads
    int xl = 1;
    X[0] = 1.0f;
#pragma unroll
    for (int s = 0; s < kRbMaxStates; ++s) {
        const int m = s < n ? int(rb_count(c, s)) : 0;
        if (m == 0) continue;
        const float* Y = tab + s * kRbMemoTab + m * (m + 1) / 2;
        if (xl == 1) {
            for (int k = 0; k <= m; ++k) X[k] = Y[k];
        } else {
            for (int q = xl + m - 1; q >= 0; --q) {
                const int k0 = q - (xl - 1) > 0 ? q - (xl - 1) : 0;
                const int k1 = q < m ? q : m;
                float z = 0.0f;
                for (int k = k0; k <= k1; ++k) z = z + Y[k] * X[q - k];
                X[q] = z;
            }
        }
        xl += m;
    }
    int* slot = s_slot[wave];
    int* spos = s_pos[wave];
    uint32_t* sdst = s_dst[wave];
    // copy the wave's pmfs out: lane l's block of xl floats at offP
    const int len = act ? xl : 0;
    const int incl = wave_scan_sum(len, lane);
    const int pos = incl - len;
    const int total = __shfl(incl, kWave - 1);
    spos[lane] = pos;
    sdst[lane] = e.z;
    wave_windows(len, pos, total, lane, slot, [&](int j, int own) {
        if (j < total) {
            const int ai = j - spos[own];
            poolP[sdst[own] + uint32_t(ai)] = strips[own * kRbStride + ai];
        }
    });
}

// Owner of position j of the wave's concatenation: the last lane l with
// pos_l <= j (pos is nondecreasing over the lanes and pos_0 = 0; a lane with
// len 0 shares its pos with the next lane, so for j < total the last such
// lane has len > 0).  Binary search over the lanes with shuffles; *pown
// gets the owner's pos.
__device__ __forceinline__ int wave_owner(int pos, int j, int* pown) {
    int idx = 0, pbest = 0;
#pragma unroll
    for (int step = kWave / 2; step > 0; step >>= 1) {
        const int t = idx + step;
        const int p = __shfl(pos, t);
        if (p <= j) {
            idx = t;
            pbest = p;
        }
    }
    *pown = pbest;
    return idx;
}

// Per-wave LDS of rb_unit_kernel (doubles), derivative passes: S - Sbar per
// unit and rate pair of the set's n states.
__host__ __device__ inline int rb_unit_wave_doubles(int mode, int n) {
    return mode == kRbFwd ? 0 : kWave * n * (n - 1);
}
// The Jacobian's LDS row per bin: n efficiency columns, then the n (n - 1)
// rates, padded to an odd number of doubles (odd rows spread the rows'
// entries over the banks).
__host__ __device__ constexpr int rb_unit_jac_row(int n) { return (n + n * (n - 1)) | 1; }
// Dynamic LDS of rb_unit_kernel: the 8-byte head (the forward's fixed-point
// histogram, the adjoint's cotangent, the Jacobian's [B][rb_unit_jac_row]
// rows), rounded to 16 bytes, then the waves' regions.  Sized to the set's
// states: the 3-state Jacobian fits seven four-wave workgroups per CU.
__host__ __device__ inline int rb_unit_head(int B, int mode, int n) {
    return ((mode == kRbJac ? B * rb_unit_jac_row(n) : B) + 1) & ~1;
}
__host__ __device__ inline size_t rb_unit_lds(int B, int mode, int n) {
    return sizeof(double) * (size_t(rb_unit_head(B, mode, n)) +
                             size_t(kRbUnitWaves) * size_t(rb_unit_wave_doubles(mode, n)));
}

// One (unit, outcome) element of a window: its unit's lane, outcome, and
// what the pass reads for it (prefetched one window ahead).
struct RbElem {
    bool act;
    int own, pown, ai, Mi;
    float F, P;
    uint32_t cx, cy;
    float pd[kRbMaxStates];  // P_s(ai) of the unit's counted states (derivative passes)
};

// Every unit u < a.nunits with Mt < kRbMemoCap.  A wave takes a.batch
// consecutive units per grab and handles them 64 at a time: the units'
// outcomes a = 0..Mt, concatenated, are spread over the lanes in windows of
// 64, so each lane handles one (unit, outcome) pair per window whatever the
// units' lengths.  A lane finds its pair's unit by a binary search over the
// lanes' start positions and takes the unit's data from that lane by
// shuffles; the pmf entries come from the memo, each window's loads issued
// while the previous window is processed.  No barrier after the prologue:
// each wave adds to the workgroup's LDS accumulators with atomics.
template <int kMode, int kN>
__global__ __launch_bounds__(kWave * kRbUnitWaves) __attribute__((amdgpu_waves_per_eu(
    kMode == kRbFwd ? 1 : 4))) void rb_unit_kernel(
    const RbModel* __restrict__ rbm, const RunArgs a, const uint2* __restrict__ cnt_in,
    const float* __restrict__ F_in, const float* __restrict__ S_in,
    const double* __restrict__ Sbar, const double* __restrict__ g,
    const uint32_t* __restrict__ rank_offP, const uint4* __restrict__ rank_offD,
    const float* __restrict__ poolP, unsigned long long* __restrict__ hist_acc,
    double* __restrict__ grad,
    uint32_t* __restrict__ work) {
    constexpr bool kDeriv = kMode != kRbFwd;
    extern __shared__ __align__(16) double s_unit[];
    __shared__ RbModel sR;
    const int tid = int(threadIdx.x), lane = tid % kWave, wave = tid / kWave;
    const int B = a.B;
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(rbm);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&sR);
        for (uint32_t i = uint32_t(tid); i < sizeof(RbModel) / 4; i += blockDim.x) dst[i] = src[i];
    }
    const int head = rb_unit_head(B, kMode, kN);
    unsigned long long* s_h = reinterpret_cast<unsigned long long*>(s_unit);
    double* s_d = s_unit;
    for (int i = tid; i < head; i += int(blockDim.x)) {
        if (kMode == kRbGrad) s_d[i] = i < B ? g[i] : 0.0;
        else if (kMode == kRbJac) s_d[i] = 0.0;
        else s_h[i] = 0ull;
    }
    __syncthreads();
    constexpr int n = kN, K = kN * (kN - 1);  // states and rate pairs of the set
    constexpr int kRow = rb_unit_jac_row(kN);
    double* w_x = s_unit + head + wave * rb_unit_wave_doubles(kMode, kN);

    const uint32_t nq = gridDim.x < kAuxQueues ? gridDim.x : kAuxQueues;
    const uint32_t queue = blockIdx.x % nq;
    const uint32_t q_begin = uint32_t(uint64_t(a.nunits) * queue / nq);
    const uint32_t q_len = uint32_t(uint64_t(a.nunits) * (queue + 1u) / nq) - q_begin;
    uint32_t* counter = &work[kRbUnitCounter + queue * kCounterStride];
    double gp[kN] = {};  // adjoint: d h / d p_s over the lane's elements
    double gk[K] = {};   // adjoint: G P (S - Sbar) over the lane's elements
    while (true) {
        uint32_t off = 0;
        if (lane == 0) off = atomicAdd(counter, a.batch);
        off = __shfl(off, 0);
        if (off >= q_len) break;
        const uint32_t end = q_len - off < a.batch ? q_len : off + a.batch;
        for (uint32_t r0 = off; r0 < end; r0 += kWave) {
            // this lane's unit of the round
            const uint32_t o = r0 + uint32_t(lane);
            const bool valid = o < end;
            const uint32_t u = q_begin + (valid ? o : r0);
            const uint2 c = valid ? cnt_in[u] : make_uint2(0u, 0u);
            int mm[kRbMaxStates] = {0, 0, 0, 0};
            int Mt = 0;
#pragma unroll
            for (int s = 0; s < kN; ++s) {
                mm[s] = int(rb_count(c, s));
                Mt += mm[s];
            }
            // units with Mt >= kRbMemoCap: listed by the walk for rb_big_kernel
            const bool memo = valid && Mt < kRbMemoCap;
            const int len = memo ? Mt + 1 : 0;
            uint32_t offP = 0;
            uint4 offD = make_uint4(0u, 0u, 0u, 0u);
            float F = 0.0f;
            // derivative passes: the unit's score less its burst's mean, in
            // the wave's LDS
            double* w_sc = w_x + lane * K;
            if (memo) {
                const uint32_t rk = rb_rank(mm, n, kRbMemoCap - 1);
                offP = rank_offP[rk];
                F = F_in[u];
                if (kDeriv) offD = rank_offD[rk];
            }
            if (kDeriv) {
                const uint32_t bl = u / a.R;
#pragma unroll
                for (int k = 0; k < K; ++k)
                    w_sc[k] = memo ? double(S_in[size_t(u) * size_t(K) + size_t(k)]) -
                                                  Sbar[size_t(bl) * size_t(K) + size_t(k)]
                                            : 0.0;
            }
            const int incl = wave_scan_sum(len, lane);
            const int pos = incl - len;
            const int total = __shfl(incl, kWave - 1);
            if (kDeriv) wave_lds_sync();

            auto fetch = [&](int wb, RbElem& e) {
                const int j = wb + lane;
                e.act = j < total;
                e.own = wave_owner(pos, j, &e.pown);
                e.ai = j - e.pown;
                e.F = __shfl(F, e.own);
                const uint32_t op = __shfl(offP, e.own);
                e.P = e.act ? poolP[op + uint32_t(e.ai)] : 0.0f;
                if (kDeriv) {
                    e.cx = __shfl(c.x, e.own);
                    e.cy = __shfl(c.y, e.own);
                    const uint2 ci = make_uint2(e.cx, e.cy);
                    const uint32_t d0 = __shfl(offD.x, e.own), d1 = __shfl(offD.y, e.own);
                    const uint32_t d2 = __shfl(offD.z, e.own), d3 = __shfl(offD.w, e.own);
                    int Mi = 0;
#pragma unroll
                    for (int s = 0; s < kN; ++s) Mi += int(rb_count(ci, s));
                    e.Mi = Mi;
                    const bool inner = e.act && e.ai < Mi;
                    int js = 0;
#pragma unroll
                    for (int s = 0; s < kN; ++s) {
                        const bool counted = rb_count(ci, s) > 0u;
                        const uint32_t base = js == 0 ? d0 : js == 1 ? d1 : js == 2 ? d2 : d3;
                        e.pd[s] = inner && counted ? poolP[base + uint32_t(e.ai)] : 0.0f;
                        js += counted ? 1 : 0;
                    }
                }
            };
            auto process = [&](const RbElem& e, int wb) {
                const int b0 = e.act ? rb_bin(e.ai, e.F, B) : -1;
                if (kMode == kRbFwd) {
                    if (b0 >= 0) atomicAdd(&s_h[b0], (unsigned long long)rb_fixed(e.P));
                    return;
                }
                // the next outcome's bin: from the next lane, which holds it
                // unless this lane is the window's last
                int b1 = __shfl_down(b0, 1);
                const bool inner = e.act && e.ai < e.Mi;
                if (inner && lane == kWave - 1) b1 = rb_bin(e.ai + 1, e.F, B);
                const uint2 ci = make_uint2(e.cx, e.cy);
                if (kMode == kRbGrad) {
                    const double G0 = b0 >= 0 ? s_d[b0] : 0.0;
                    if (inner) {
                        const double G1 = b1 >= 0 ? s_d[b1] : 0.0;
                        const double dG = double(float(G1 - G0));
#pragma unroll
                        for (int s = 0; s < kN; ++s)
                            gp[s] += double(rb_count(ci, s)) * double(e.pd[s]) * dG;
                    }
                    // h (S - Sbar) of the unit, one element at a time (h is
                    // the sum of G(bin(a)) P(a) over the unit's outcomes)
                    if (e.act && b0 >= 0) {
                        const double hv = G0 * double(e.P);
                        const double* scu = w_x + e.own * K;
#pragma unroll
                        for (int k = 0; k < K; ++k)
                            gk[k] += hv * scu[k];
                    }
                } else {
                    if (b0 >= 0) {
                        double* row = s_d + b0 * kRow + kN;
                        const double* scu = w_x + e.own * K;
#pragma unroll
                        for (int k = 0; k < K; ++k)
                            atomicAdd(&row[k], double(e.P) * scu[k]);
                    }
                    if (inner && b0 != b1) {
#pragma unroll
                        for (int s = 0; s < kN; ++s) {
                            const uint32_t m = rb_count(ci, s);
                            if (m > 0u) {
                                const double w = double(m) * double(e.pd[s]);
                                if (b1 >= 0) atomicAdd(&s_d[b1 * kRow + s], w);
                                if (b0 >= 0) atomicAdd(&s_d[b0 * kRow + s], -w);
                            }
                        }
                    }
                }
            };
            if (total > 0) {
                RbElem cur, nxt;
                fetch(0, cur);
                for (int wb = 0; wb < total; wb += kWave) {
                    if (wb + kWave < total) fetch(wb + kWave, nxt);
                    process(cur, wb);
                    cur = nxt;
                }
            }
            if (kDeriv) wave_lds_sync();  // the round's LDS reads before the next round's writes
        }
    }
    if (kMode == kRbFwd) {
        __syncthreads();
        unsigned long long* row = hist_acc + size_t(blockIdx.x & 7u) * kMaxBins;
        for (int i = tid; i < B; i += int(blockDim.x))
            if (s_h[i]) atomicAdd(&row[i], s_h[i]);
    } else if (kMode == kRbGrad) {
        const double f = a.R > 1 ? double(a.R) / double(a.R - 1) : 0.0;
#pragma unroll
        for (int s = 0; s < kN; ++s) {
            const double v = wave_sum(gp[s]);
            if (lane == 0 && s < n) atomicAdd(&grad[s], v);
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double v = wave_sum(gk[k]);
            if (lane == 0) atomicAdd(&grad[n + k], f * v);
        }
    } else {
        // grad holds J[b][i], i < n + K (efficiencies, then rates in args
        // order), B rows of kRbJacCols
        __syncthreads();
        const double f = a.R > 1 ? double(a.R) / double(a.R - 1) : 0.0;
        const int w = n + K;
        for (int i = tid; i < B * w; i += int(blockDim.x)) {
            const int b = i / w, col = i - b * w;
            const double v = col < n ? s_d[b * kRow + col] : f * s_d[b * kRow + col];
            if (v != 0.0) atomicAdd(&grad[size_t(b) * kRbJacCols + size_t(col)], v);
        }
    }
}

// Bytes of rb_big_kernel's dynamic LDS: B 8-byte entries (the forward's
// histogram, the adjoint's cotangent) or the Jacobian's [B][kRbJacCols]
// rows, then three pmf buffers (X, Y, Z) of cap floats.
__host__ __device__ inline size_t rb_big_lds(int B, int cap, int mode, bool wide) {
    const size_t head = size_t(mode == kRbJac ? B * kRbJacCols : B);
    return sizeof(double) * head +
           (mode != kRbFwd && wide ? sizeof(double) : sizeof(float)) * 3u * size_t(cap);
}

// The units the walk listed (Mt >= kRbMemoCap), one wave per unit, lanes
// over acceptor totals; cap = the largest Mt + 1 of the units of the list.
// The forward convolves in float, in the oracle's order (bit-exact).  The
// derivative passes take the leave-one-out pmfs from P by a recurrence
// whose factor has magnitude |alpha| = p / q (or q / p): it carries P's
// rounding errors, amplified up to 1 / (1 - |alpha|) -- 3x at p = 0.4, the
// length of the chain at p = 1/2, where a P rounded to float at every
// convolution step had put 1e-3 of an efficiency gradient off at
// Mt ~ 1,000.  kWide (sets with a state of 0.4 < p < 0.6): those passes
// convolve in double (from the same float binomial tables), at twice the
// LDS.
template <int kMode, bool kWide>
__global__ __launch_bounds__(kWave) void rb_big_kernel(
    const RbModel* __restrict__ rbm, const double* __restrict__ lf,
    const float* __restrict__ pmf_tab, const RunArgs a, const uint2* __restrict__ cnt_in,
    const float* __restrict__ F_in, const float* __restrict__ S_in,
    const double* __restrict__ Sbar, const double* __restrict__ g,
    const uint32_t* __restrict__ big, const uint32_t* __restrict__ big_count,
    unsigned long long* __restrict__ hist_acc, double* __restrict__ grad, int cap) {
    extern __shared__ double s_dyn[];
    __shared__ RbModel sR;
    const uint32_t lane = threadIdx.x;
    const int B = a.B;
    const uint32_t count = *big_count;
    if (blockIdx.x >= count) return;  // whole workgroup: no barrier below is skipped by part
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(rbm);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&sR);
        for (uint32_t i = lane; i < sizeof(RbModel) / 4; i += kWave) dst[i] = src[i];
    }
    const int head = kMode == kRbJac ? B * kRbJacCols : B;
    unsigned long long* s_h = reinterpret_cast<unsigned long long*>(s_dyn);
    double* s_d = s_dyn;
    using T = typename std::conditional<kMode != kRbFwd && kWide, double, float>::type;
    T* bufs = reinterpret_cast<T*>(s_dyn + head);
    for (int i = int(lane); i < head; i += kWave) {
        if (kMode == kRbGrad) s_d[i] = g[i];
        else if (kMode == kRbJac) s_d[i] = 0.0;
        else s_h[i] = 0ull;
    }
    __syncthreads();
    const int n = sR.n, K = sR.K;
    double gp_acc = 0.0, gk_acc = 0.0;  // adjoint: lane s holds d/dp_s, lane k d/dk_k
    for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
        const uint32_t u = big[i];
        const uint2 c = cnt_in[u];
        const float F = F_in[u];
        int bX = 0, bY = 1, bZ = 2;
        int xl = 1;
        if (lane == 0) bufs[0] = T(1);
#pragma unroll
        for (int s = 0; s < kRbMaxStates; ++s) {
            const int m = s < n ? int(rb_count(c, s)) : 0;
            T* Y = bufs + bY * cap;
            for (int k = int(lane); m > 0 && k <= m; k += kWave) {
                const double y = m < kRbPmfTable
                                     ? double(pmf_tab[rb_pmf_index(s, m, k)])
                                     : rb_binom_pmf(m, k, lf, sR.p[s], sR.lp[s], sR.lq[s]);
                Y[k] = kMode == kRbFwd ? T(float(y)) : T(y);
            }
            __syncthreads();
            if (m > 0) {
                if (xl == 1) {  // X = [1]: the product is Y itself, bit for bit
                    const int t = bX;
                    bX = bY;
                    bY = t;
                    xl = m + 1;
                } else {
                    const T* X = bufs + bX * cap;
                    T* Z = bufs + bZ * cap;
                    for (int q = int(lane); q < xl + m; q += kWave)
                        Z[q] = rb_conv_at(Y, m + 1, X, xl, q);
                    const int t = bX;
                    bX = bZ;
                    bZ = t;
                    xl += m;
                }
            }
            __syncthreads();
        }
        const T* P = bufs + bX * cap;
        if (kMode == kRbFwd) {
            for (int q = int(lane); q < xl; q += kWave) {
                const int bin = rb_bin(q, F, B);
                if (bin >= 0) atomicAdd(&s_h[bin], (unsigned long long)rb_fixed(P[q]));
            }
            __syncthreads();
            continue;
        }
        const int Mt = xl - 1;
        const uint32_t bl = u / a.R;
        double sc[kRbMaxPairs];  // Jacobian: the unit's score less its burst's mean
#pragma unroll
        for (int k = 0; k < kRbMaxPairs; ++k)
            sc[k] = kMode == kRbJac && k < K
                        ? double(S_in[size_t(u) * size_t(K) + size_t(k)]) -
                              Sbar[size_t(bl) * size_t(K) + size_t(k)]
                        : 0.0;
        T* dG = bufs + bY * cap;  // adjoint: G(bin(c + 1)) - G(bin(c)), rounded to float
        double hp = 0.0;
        for (int q = int(lane); q < xl; q += kWave) {
            const int b0 = rb_bin(q, F, B);
            if (kMode == kRbJac) {
                if (b0 >= 0) {
                    const double w = double(P[q]);
                    double* row = s_d + b0 * kRbJacCols + kRbMaxStates;
#pragma unroll
                    for (int k = 0; k < kRbMaxPairs; ++k)
                        if (k < K) atomicAdd(&row[k], w * sc[k]);
                }
                continue;
            }
            const double G0 = b0 >= 0 ? s_d[b0] : 0.0;
            hp += G0 * double(P[q]);
            if (q + 1 < xl) {
                const int b1 = rb_bin(q + 1, F, B);
                dG[q] = T(float((b1 >= 0 ? s_d[b1] : 0.0) - G0));
            }
        }
        __syncthreads();
        const double h = wave_sum(hp);
#pragma unroll
        for (int s = 0; s < kRbMaxStates; ++s) {
            const int m = s < n ? int(rb_count(c, s)) : 0;
            if (m == 0) continue;  // uniform over the wave
            const double p = double(sR.p[s]), q = 1.0 - p;
            const bool up = p <= 0.5;
            const double alpha = up ? -p / q : -q / p;
            const double inv = up ? 1.0 / q : 1.0 / p;
            double carry = 0.0, part = 0.0;
            for (int j0 = 0; j0 < Mt; j0 += kWave) {
                // lanes over 64 steps of the recurrence: a scan of the
                // affine maps x -> alpha x + P(.) inv
                const int j = j0 + int(lane);
                const bool v = j < Mt;
                double A = v ? alpha : 1.0;
                double Bv = v ? double(P[up ? j : Mt - j]) * inv : 0.0;
#pragma unroll
                for (int d = 1; d < kWave; d <<= 1) {
                    const double A2 = __shfl_up(A, d);
                    const double B2 = __shfl_up(Bv, d);
                    if (int(lane) >= d) {
                        Bv = A * B2 + Bv;
                        A = A * A2;
                    }
                }
                const double x = A * carry + Bv;  // P_s(up ? j : Mt - 1 - j)
                carry = __shfl(x, kWave - 1);
                const int ci = up ? j : Mt - 1 - j;
                if (kMode == kRbJac) {
                    const int b0 = rb_bin(ci, F, B), b1 = rb_bin(ci + 1, F, B);
                    if (v && b0 != b1) {
                        const double w = double(m) * x;
                        if (b1 >= 0) atomicAdd(&s_d[b1 * kRbJacCols + s], w);
                        if (b0 >= 0) atomicAdd(&s_d[b0 * kRbJacCols + s], -w);
                    }
                } else if (v) {
                    part += x * double(dG[ci]);
                }
            }
            if (kMode == kRbGrad) {
                part = wave_sum(part) * double(m);
                if (int(lane) == s) gp_acc += part;
            }
        }
        if (kMode == kRbGrad && int(lane) < K)
            gk_acc += h * (double(S_in[size_t(u) * size_t(K) + lane]) -
                           Sbar[size_t(bl) * size_t(K) + lane]);
        __syncthreads();
    }
    if (kMode == kRbFwd) {
        unsigned long long* row = hist_acc + size_t(blockIdx.x & 7u) * kMaxBins;
        for (int i = int(lane); i < B; i += kWave)
            if (s_h[i]) atomicAdd(&row[i], s_h[i]);
    } else if (kMode == kRbGrad) {
        const double f = a.R > 1 ? double(a.R) / double(a.R - 1) : 0.0;
        if (int(lane) < n) atomicAdd(&grad[lane], gp_acc);
        if (int(lane) < K) atomicAdd(&grad[n + int(lane)], f * gk_acc);
    } else {
        const double f = a.R > 1 ? double(a.R) / double(a.R - 1) : 0.0;
        const int w = n + K;
        for (int i = int(lane); i < B * w; i += kWave) {
            const int b = i / w, col = i - b * w;
            const double v = col < n ? s_d[b * kRbJacCols + col]
                                     : f * s_d[b * kRbJacCols + kRbMaxStates + (col - n)];
            if (v != 0.0) atomicAdd(&grad[size_t(b) * kRbJacCols + size_t(col)], v);
        }
    }
}

// Per burst bl and rate k: Sbar[bl K + k] = (1 / R) sum_r S[(bl R + r) K + k],
// summed r ascending in double.  A workgroup stages the scores of a run of
// whole bursts in LDS with coalesced loads, then one thread per (burst,
// rate) sums them; bursts whose R K scores exceed the stage are summed
// straight from global memory.
constexpr uint32_t kRbSbarStage = 8192;  // floats
__global__ __launch_bounds__(256) void rb_score_mean_kernel(const RunArgs a, int K,
                                                            const float* __restrict__ S,
                                                            double* __restrict__ Sbar) {
    __shared__ float st[kRbSbarStage];
    const uint64_t nb = a.nunits / a.R;
    const uint32_t per = a.R * uint32_t(K);  // scores per burst
    const uint32_t bpb = per <= kRbSbarStage ? kRbSbarStage / per : 0u;
    if (bpb == 0) {
        for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nb * uint64_t(K);
             i += uint64_t(gridDim.x) * blockDim.x) {
            const uint64_t bl = i / uint64_t(K), k = i - bl * uint64_t(K);
            const float* src = S + bl * per + k;
            double sum = 0.0;
            for (uint32_t r = 0; r < a.R; ++r) sum += double(src[uint64_t(r) * uint64_t(K)]);
            Sbar[i] = sum / double(a.R);
        }
        return;
    }
    for (uint64_t b0 = uint64_t(blockIdx.x) * bpb; b0 < nb; b0 += uint64_t(gridDim.x) * bpb) {
        const uint32_t nbh = uint32_t(nb - b0 < bpb ? nb - b0 : bpb);
        const float* src = S + b0 * per;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nbh * per; i += blockDim.x) st[i] = src[i];
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < nbh * uint32_t(K); t += blockDim.x) {
            const uint32_t bl = t / uint32_t(K), k = t - bl * uint32_t(K);
            const float* x = st + bl * per + k;
            double sum = 0.0;
            for (uint32_t r = 0; r < a.R; ++r) sum += double(x[r * uint32_t(K)]);
            Sbar[(b0 + bl) * uint64_t(K) + k] = sum / double(a.R);
        }
    }
}

// The binomial pmfs of every state for counts m < kRbPmfTable (one thread
// per entry), read by the pmf passes instead of recomputing them per unit.
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

hipError_t rb_walk_occupancy(bool scores, int n, int* per_cu) {
#define PDA_RB_OCC(SC, N) \
    hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_walk_kernel<SC, N>, kWave, 0)
    if (scores) return n == 2 ? PDA_RB_OCC(true, 2) : n == 3 ? PDA_RB_OCC(true, 3) : PDA_RB_OCC(true, 4);
    return n == 2 ? PDA_RB_OCC(false, 2) : n == 3 ? PDA_RB_OCC(false, 3) : PDA_RB_OCC(false, 4);
#undef PDA_RB_OCC
}

template <int kMode>
static hipError_t rb_unit_occupancy_m(int n, size_t l, int* per_cu) {
    const int tu = kWave * kRbUnitWaves;
    if (n == 2) return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_unit_kernel<kMode, 2>, tu, l);
    if (n == 3) return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_unit_kernel<kMode, 3>, tu, l);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_unit_kernel<kMode, 4>, tu, l);
}

hipError_t rb_unit_occupancy(int mode, int n, int dyn_lds, int* per_cu) {
    const size_t l = size_t(dyn_lds);
    if (mode == kRbGrad) return rb_unit_occupancy_m<kRbGrad>(n, l, per_cu);
    if (mode == kRbJac) return rb_unit_occupancy_m<kRbJac>(n, l, per_cu);
    return rb_unit_occupancy_m<kRbFwd>(n, l, per_cu);
}

hipError_t rb_big_occupancy(int mode, bool wide, int dyn_lds, int* per_cu) {
    const size_t l = size_t(dyn_lds);
    if (mode == kRbGrad)
        return wide ? hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_big_kernel<kRbGrad, true>, kWave, l)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_big_kernel<kRbGrad, false>, kWave, l);
    if (mode == kRbJac)
        return wide ? hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_big_kernel<kRbJac, true>, kWave, l)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_big_kernel<kRbJac, false>, kWave, l);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rb_big_kernel<kRbFwd, false>, kWave, l);
}

template <bool kScores, int kN>
static void launch_rb_walk_n(uint32_t grid, uint32_t replay_grid, const BurstDesc* bursts,
                             const uint32_t* ticks, const uint2* links, const ModelBlob* blob,
                             const RbModel* rbm, const RunArgs& a, uint2* cnt, float* F, float* S,
                             uint32_t* work, uint32_t* big, uint32_t* huge, uint32_t* big_count,
                             uint32_t* replay, uint32_t* replay_count, hipStream_t stream) {
    hipLaunchKernelGGL((rb_walk_kernel<kScores, kN>), dim3(grid), dim3(kWave), 0, stream, bursts,
                       ticks, links, blob, rbm, a, cnt, F, S, work, big, huge, big_count, replay,
                       replay_count);
    hipLaunchKernelGGL((rb_replay_kernel<kScores, kN>), dim3(replay_grid), dim3(kWave), 0, stream,
                       bursts, ticks, links, blob, rbm, a, cnt, F, S, big, huge, big_count, replay,
                       replay_count);
}

hipError_t launch_rb_walk(bool scores, int n, uint32_t grid, uint32_t replay_grid,
                          const BurstDesc* bursts, const uint32_t* ticks, const uint2* links,
                          const ModelBlob* blob, const RbModel* rbm, const RunArgs& a, uint2* cnt,
                          float* F, float* S, uint32_t* work, uint32_t* big, uint32_t* huge,
                          uint32_t* big_count, uint32_t* replay, uint32_t* replay_count,
                          hipStream_t stream) {
#define PDA_RB_WALK(SC, N)                                                                    \
    launch_rb_walk_n<SC, N>(grid, replay_grid, bursts, ticks, links, blob, rbm, a, cnt, F, S,  \
                            work, big, huge, big_count, replay, replay_count, stream)
    if (n < 2 || n > kRbMaxStates) return hipErrorInvalidValue;
    if (scores) {
        if (n == 2) PDA_RB_WALK(true, 2);
        else if (n == 3) PDA_RB_WALK(true, 3);
        else PDA_RB_WALK(true, 4);
    } else {
        if (n == 2) PDA_RB_WALK(false, 2);
        else if (n == 3) PDA_RB_WALK(false, 3);
        else PDA_RB_WALK(false, 4);
    }
#undef PDA_RB_WALK
    return hipGetLastError();
}

int rb_unit_threads() { return kWave * kRbUnitWaves; }
size_t rb_unit_lds_bytes(int B, int mode, int n) { return rb_unit_lds(B, mode, n); }
size_t rb_big_lds_bytes(int B, int cap, int mode, bool wide) { return rb_big_lds(B, cap, mode, wide); }
size_t rb_memo_lds_bytes(int n) { return rb_memo_lds(n); }

hipError_t launch_rb_table(const RbModel* rbm, const double* lf, float* pmf_tab, int n,
                           hipStream_t stream) {
    const int total = n * kRbPmfStride;
    hipLaunchKernelGGL(rb_table_kernel, dim3(uint32_t((total + 255) / 256)), dim3(256), 0, stream,
                       rbm, lf, pmf_tab);
    return hipGetLastError();
}

hipError_t launch_rb_memo(int n, const RbModel* rbm, const float* pmf_tab, const uint4* list,
                          uint32_t nvec, float* poolP, hipStream_t stream) {
    const uint32_t threads = kWave * kRbMemoWaves;
    const dim3 grid((nvec + threads - 1) / threads);
    hipLaunchKernelGGL(rb_memo_kernel, grid, dim3(threads), rb_memo_lds(n), stream, rbm, pmf_tab,
                       list, nvec, poolP);
    return hipGetLastError();
}

template <int kMode, int kN>
static void launch_rb_unit_n(uint32_t grid, size_t lds, const RbModel* rbm, const RunArgs& a,
                             const uint2* cnt, const float* F, const float* S, const double* Sbar,
                             const double* g, const uint32_t* rank_offP, const uint4* rank_offD,
                             const float* poolP, unsigned long long* hist_acc, double* grad,
                             uint32_t* work, hipStream_t stream) {
    hipLaunchKernelGGL((rb_unit_kernel<kMode, kN>), dim3(grid), dim3(uint32_t(kWave * kRbUnitWaves)),
                       lds, stream, rbm, a, cnt, F, S, Sbar, g, rank_offP, rank_offD, poolP,
                       hist_acc, grad, work);
}

hipError_t launch_rb_unit(int mode, int n, uint32_t grid, const RbModel* rbm, const RunArgs& a,
                          const uint2* cnt, const float* F, const float* S, const double* Sbar,
                          const double* g, const uint32_t* rank_offP, const uint4* rank_offD,
                          const float* poolP, unsigned long long* hist_acc, double* grad,
                          uint32_t* work, hipStream_t stream) {
    if (n < 2 || n > kRbMaxStates) return hipErrorInvalidValue;
    const size_t lds = rb_unit_lds(a.B, mode, n);
#define PDA_RB_UNIT(M, N)                                                                      \
    launch_rb_unit_n<M, N>(grid, lds, rbm, a, cnt, F, S, Sbar, g, rank_offP, rank_offD, poolP, \
                           hist_acc, grad, work, stream)
#define PDA_RB_UNIT_M(M)            \
    if (n == 2) PDA_RB_UNIT(M, 2);  \
    else if (n == 3) PDA_RB_UNIT(M, 3); \
    else PDA_RB_UNIT(M, 4)
    if (mode == kRbGrad) {
        PDA_RB_UNIT_M(kRbGrad);
    } else if (mode == kRbJac) {
        PDA_RB_UNIT_M(kRbJac);
    } else {
        PDA_RB_UNIT_M(kRbFwd);
    }
#undef PDA_RB_UNIT_M
#undef PDA_RB_UNIT
    return hipGetLastError();
}

template <int kMode, bool kWide>
static void launch_rb_big_m(uint32_t grid, int cap, const RbModel* rbm, const double* lf,
                            const float* pmf_tab, const RunArgs& a, const uint2* cnt,
                            const float* F, const float* S, const double* Sbar, const double* g,
                            const uint32_t* big, const uint32_t* big_count,
                            unsigned long long* hist_acc, double* grad, hipStream_t stream) {
    hipLaunchKernelGGL((rb_big_kernel<kMode, kWide>), dim3(grid), dim3(kWave),
                       rb_big_lds(a.B, cap, kMode, kWide), stream, rbm, lf, pmf_tab, a, cnt, F, S,
                       Sbar, g, big, big_count, hist_acc, grad, cap);
}

hipError_t launch_rb_big(int mode, bool wide, uint32_t grid, int cap, const RbModel* rbm,
                         const double* lf, const float* pmf_tab, const RunArgs& a,
                         const uint2* cnt, const float* F, const float* S, const double* Sbar,
                         const double* g, const uint32_t* big, const uint32_t* big_count,
                         unsigned long long* hist_acc, double* grad, hipStream_t stream) {
#define PDA_RB_BIG(M, W)                                                                      \
    launch_rb_big_m<M, W>(grid, cap, rbm, lf, pmf_tab, a, cnt, F, S, Sbar, g, big, big_count, \
                          hist_acc, grad, stream)
    if (mode == kRbGrad) {
        if (wide) PDA_RB_BIG(kRbGrad, true);
        else PDA_RB_BIG(kRbGrad, false);
    } else if (mode == kRbJac) {
        if (wide) PDA_RB_BIG(kRbJac, true);
        else PDA_RB_BIG(kRbJac, false);
    } else {
        PDA_RB_BIG(kRbFwd, false);
    }
#undef PDA_RB_BIG
    return hipGetLastError();
}

hipError_t launch_rb_score_mean(const RunArgs& a, int K, const float* S, double* Sbar,
                                int num_cu, hipStream_t stream) {
    const uint64_t nb = a.nunits / a.R;
    const uint64_t per = uint64_t(a.R) * uint64_t(K);
    const uint64_t bpb = per <= kRbSbarStage ? kRbSbarStage / per : 0;
    const uint64_t want = bpb ? (nb + bpb - 1) / bpb : (nb * uint64_t(K) + 255) / 256;
    // a few workgroups per CU, each looping over its share
    const uint64_t cap = uint64_t(num_cu > 0 ? num_cu : 1) * 4u;
    const uint32_t grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(want, cap)));
    hipLaunchKernelGGL(rb_score_mean_kernel, dim3(grid), dim3(256), 0, stream, a, K, S, Sbar);
    return hipGetLastError();
}

hipError_t launch_rb_finish(const RunArgs& a, const float* obs, unsigned long long* hist_acc,
                            uint32_t* work, void* res, hipStream_t stream) {
    hipLaunchKernelGGL(rb_finish_kernel, dim3(1), dim3(kWave), 0, stream, a, obs, hist_acc, work,
                       static_cast<RbResultHeader*>(res));
    return hipGetLastError();
}

}  // namespace pda
