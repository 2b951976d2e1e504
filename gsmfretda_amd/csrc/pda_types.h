// pda_types.h -- device data layout of the Monte-Carlo PDA path.
#pragma once
#include <cstdint>

namespace pda {

constexpr int kMaxStates = 8;
constexpr int kMaxBins = 1024;
constexpr int kWave = 64;            // lanes per workgroup of the walk kernel
constexpr uint32_t kTickSentinel = 0x80000000u;  // "never closes a window"
constexpr uint32_t kIndexMask = 0x3FFFFFFFu;
constexpr int kBinomTable = 128;     // photon counts with tabulated inversion constants
constexpr uint32_t kMaxQueues = 64;     // work counters of a run
constexpr uint32_t kDefaultQueues = 32;  // four per XCD unless tuned
// Each work counter sits on its own 256-byte line: device-scope atomics on
// one line serialise, and counters sharing a line would serialise together.
constexpr uint32_t kCounterStride = 64;  // uint32 words between counters
// Counter words of a run: kMaxQueues queue counters and kAuxQueues more
// (the RB pmf pass's burst counters), each on its own line.
constexpr uint32_t kAuxQueues = 8;
constexpr size_t kWorkWords = size_t(kMaxQueues + kAuxQueues) * kCounterStride;

// One burst, original order (the range [begin, end) of mc::setBurstBd
// addresses these directly).  16 bytes, one dwordx4 load.
struct BurstDesc {
    uint32_t off;       // first photon in the photon arrays
    uint32_t P;         // photon count (istop - istart + 1)
    uint32_t stop_rel;  // stop - start, ticks
    float T;            // T_burst_duration, seconds
};

// Per photon i of a burst (local indices):
//   tick[i]   burst-relative tick; photons in neither channel carry the tick
//             of the next channel photon (kTickSentinel if none), so the
//             ticks are sorted and "first channel photon at or after w1"
//             is a lower_bound.
//   link[i].x first channel photon f >= i (P if none) | colour(f) << 30
//             (colour: 1 donor, 2 acceptor, 3 both)
//   link[i].y first photon after f carrying the other colour (P if none;
//             f itself when f carries both)
// The sticky-flag count of a window [c, close) is then
// min(close, link[c].y) - f for f = link[c].x < close (src/mc.cu:136-181).

// Model tables of one parameter set, built on the host from the args of
// mc::set_params (src/mc.cu:790-800, src/eigenhelper.cpp:4-54).
struct Model {
    int32_t n;
    int32_t any_v;                // some v[s] != 0: broadening draws needed
    int32_t init_state;           // >= 0: the initial-state draw always gives this state
    float gmx;                    // gamma - Dch2Ach
    float omdd;                   // 1 - DexDirAem
    float r0;
    // first state row of this set in its threshold table (population runs:
    // the sets' rows follow one another; 0 for a slot's single set)
    int32_t guide_base;
    float pad0;
    float p0cum[kMaxStates];                      // running sums of genMatP's inflow
    float nscum[kMaxStates][kMaxStates];          // [s][jj] running sums of P_i2j column s
    float krate[kMaxStates][kMaxStates];          // [s][j] = K(j, s), rate s -> j
    float v[kMaxStates];
    float base[kMaxStates];                       // r0 (1/E - 1)^(1/6)
    float ep0[kMaxStates];                        // broadened E when v == 0
    // Binomial(m, ep0[s]) by inversion (src/binom.cuh:135-169) on
    // pp = min(p, 1-p): pp, q = 1 - pp, and whether p > 1/2 (draw n - X).
    double pp[kMaxStates];
    double qq[kMaxStates];
    int32_t flip[kMaxStates];
};

// q^m and the inversion bound min(m, mp + 10 sqrt(mpq + 1)) per state and
// photon count m < kBinomTable, for the fast walk (v == 0).
struct BinomTable {
    double qn[kMaxStates][kBinomTable];
    int32_t bound[kMaxStates][kBinomTable];
};

// Thresholds of the fast walk's binomial inversion on the 53-bit uniform
// grid (binomial_guided and build_guide_row_host in pda_walk.h, guide_kernel):
// row (s, m) holds,
// for X = 0 .. its last entry, a pair (up, lo) of grid indices -- u >= up
// proves that the inversion goes past X, u <= lo that it stops at or before
// X -- followed by a sentinel, and a guide of kGuideBuckets entries: the
// first X that a uniform in each 1/64 of [0, 1) can end at.
constexpr int kGuideBuckets = 64;
constexpr int kGuideShift = 47;            // 53-bit grid index >> 47 = bucket
struct ThrPair {
    int64_t up, lo;
};
// Row (s, m) holds at most m + 2 pairs (X = 0 .. bound <= m, the sentinel):
// rows of a state packed one after another, m (m + 3) / 2 pairs before row m.
constexpr size_t kGuideStatePairs = size_t(kBinomTable) * (kBinomTable + 3) / 2;
constexpr size_t kGuideThrPairs = size_t(kMaxStates) * kGuideStatePairs;
constexpr size_t guide_row_offset(int s, int m) {
    return size_t(s) * kGuideStatePairs + size_t(m) * size_t(m + 3) / 2;
}
struct BinomGuide {
    const ThrPair* thr;    // kGuideThrPairs, rows at guide_row_offset
    const uint8_t* guide;  // [kMaxStates][kBinomTable][kGuideBuckets]
};
constexpr size_t kGuideBytes = size_t(kMaxStates) * kBinomTable * kGuideBuckets;

// What a stream slot uploads per parameter set.
struct ModelBlob {
    Model m;
    BinomTable tab;
};

// Arguments of one launch of the walk kernel (passed by value).
struct RunArgs {
    uint32_t begin;      // first burst
    uint32_t nunits;     // bursts * R
    uint32_t R;
    uint32_t seed, set;
    uint32_t batch;      // units a wave takes from its work counter at a time
    uint32_t queues;     // work counters in use, 1..kMaxQueues
    int32_t B, left_pad, bg_on;
    float clk_p, bg_ad, bg_dd;
    int32_t nstates;     // for the chi2 degrees of freedom
    uint32_t replay_cap; // capacity of the replay list
    uint32_t nsets;      // population runs: parameter sets
    uint32_t bps;        // population runs: work batches per set
};

// Device result of one run, followed by int32 hist[B].
struct ResultHeader {
    double chi2_64;
    float chi2_ref;
    int32_t effN;
    int64_t units;
    int32_t B;
    int32_t replays;     // units the fast pass abandoned to the replay pass
};

// Population runs: random-stream keys of the sets, uploaded with the models
// (set s draws with key {seed, set_base + s * set_stride}).
struct PopControl {
    uint32_t seed, set_base, set_stride, nsets;
    uint32_t pad[60];  // the models that follow start 256-byte aligned
};

// Population runs: what the finish kernel reports besides the sets' results.
struct PopStatus {
    uint32_t replays;   // units the fast pass abandoned
    uint32_t overflow;  // 1: more than the replay list holds (results invalid)
    uint32_t pad[2];
};

// Population runs: result of one set (its int32 histogram follows the array
// of all sets' results).
struct SetResult {
    double chi2_64;
    float chi2_ref;
    int32_t effN;
    int32_t n;
    int32_t pad;
};

}  // namespace pda
