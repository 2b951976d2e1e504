// pda_ctx.h -- host-side state of a pda_ctx (internal to the library).
//
// A context owns the resident bursts and a set of stream slots; a slot is the
// MI355X counterpart of one of the reference's CUDA streams with its
// per-stream buffers (src/mc.hpp:32-180, src/mc.cu:523-565).  Besides the
// single-set run of mc::run_kernel, a slot can hold a population of parameter
// sets evaluated by one graph-captured launch sequence.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gsmfretda_pda.h"
#include "pda_layout.h"
#include "pda_types.h"

namespace pda {

int fail(int code, const char* fmt, ...);
void set_error(const std::string& msg);
const std::string& last_error();

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::pda::fail(PDA_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                               __FILE__, __LINE__);                                        \
    } while (0)

// Launch shape of the walk kernels; 0 = chosen per run.
struct LaunchShape {
    uint32_t batch_tiles = 0;   // 64-unit tiles per work-counter grab
    uint32_t waves_per_cu = 0;  // resident one-wave workgroups per CU (grid size)
    uint32_t queues = 0;        // work counters (1..kMaxQueues)
};

// Key of a captured population graph: everything baked into its nodes.
struct PopGraphKey {
    int nsets = 0, R = 0, B = 0, left_pad = 0, pass = 0, bg_on = 0;
    int64_t b0 = 0, b1 = 0;
    uint32_t grid = 0, batch = 0, queues = 0, replay_cap = 0;
    float clk_p = 0.f, bg_ad = 0.f, bg_dd = 0.f;
    // the resident bursts the nodes read: their buffers and the load that
    // produced them (a reload may return a freed buffer's address)
    const void* bursts = nullptr;
    const void* ticks = nullptr;
    const void* links = nullptr;
    uint64_t data_gen = 0;
    const void* obs = nullptr;
    // binomial threshold rows: whether the walk reads them, the guide
    // kernel's state count and the table's buffer
    bool guided = false;
    int max_states = 0;
    const void* thr = nullptr;
    bool operator==(const PopGraphKey& o) const;
};

struct Population {
    int nsets = 0;
    std::vector<int> n;              // states per set
    std::vector<float> args;         // packed args of every set
    std::vector<int64_t> off;        // start of set s in args
    bool any_v = false;
    size_t cap_sets = 0;             // allocated sets
    uint8_t* h_in = nullptr;         // pinned: PopControl + ModelBlob[cap_sets]
    uint8_t* d_in = nullptr;
    int cap_B = 0;
    int32_t* d_hist_acc = nullptr;   // [cap_sets][8][cap_B], zero between runs
    uint8_t* d_out = nullptr;        // PopStatus + SetResult[cap_sets] + int32 hist[cap_sets][B]
    uint8_t* h_out = nullptr;        // pinned
    size_t out_bytes = 0;
    uint64_t* d_replay = nullptr;
    uint32_t* d_replay_count = nullptr;
    size_t replay_cap = 0;
    uint32_t* d_work = nullptr;      // kWorkWords: queue counters, one per line
    // binomial threshold rows of every set (fast form; set i's rows start at
    // its model's guide_base), when the table fits kPopGuideBudget
    ThrPair* d_thr = nullptr;
    uint8_t* d_guide = nullptr;
    size_t cap_guide_states = 0;
    bool guided = false;
    int max_states = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    PopGraphKey key;
    int captures = 0;
    bool ran = false;
    bool in_flight = false;
    int last_B = 0, last_R = 0, last_left_pad = 0;
    int64_t last_b0 = 0, last_b1 = 0;  // burst range of the last run
    int64_t last_units = 0;          // per set
    uint32_t last_seed = 0, last_base = 0, last_stride = 0;
    bool last_graph = false;
    hipEvent_t ev0 = nullptr, ev_mid = nullptr, ev1 = nullptr;
    bool ready = false;  // counters and events allocated
};

// Rao-Blackwellised runs of a slot: per-unit walk summaries of the last
// forward run (reused by the adjoint), histogram rows, result, gradient.
struct RbState {
    void* d_rbm = nullptr;             // RbModel
    void* h_rbm = nullptr;             // pinned
    // the walk's model when the state is one set of an RB population (the
    // single-set forward walks the slot's own d_blob)
    ModelBlob* d_blob = nullptr;
    ModelBlob* h_blob = nullptr;       // pinned
    uint2* d_cnt = nullptr;            // per unit: photon counts per state
    float* d_F = nullptr;              // per unit: F
    float* d_S = nullptr;              // per unit: scores of the rates
    size_t cap_units = 0, cap_scores = 0;
    unsigned long long* d_hist = nullptr;  // [8][kMaxBins]
    uint32_t* d_work = nullptr;        // kWorkWords: walk queues + pmf pass queues
    uint8_t* d_res = nullptr;          // header + uint64 M[kMaxBins]
    uint8_t* h_res = nullptr;          // pinned
    double* d_g = nullptr;             // cotangent per bin
    double* h_g = nullptr;             // pinned
    double* d_grad = nullptr;          // n + K
    float* d_pmf_tab = nullptr;        // binomial pmfs of the states, counts < 256
    double* d_jac = nullptr;           // [kMaxBins][n + K] rows of 16
    double* h_jac = nullptr;           // pinned
    // units whose acceptor total is too large for the memo, listed by the
    // walk: totals below kRbBigSplit (d_big) and the rest (d_huge), each of
    // capacity cap_units, with their counts d_big_count[0] and [1]
    uint32_t* d_big = nullptr;
    uint32_t* d_huge = nullptr;
    uint32_t* d_big_count = nullptr;
    // units the fast-form walk abandoned (capacity cap_units), and their
    // count, for the general-form replay
    uint32_t* d_replay = nullptr;
    uint32_t* d_replay_count = nullptr;
    double* d_Sbar = nullptr;          // per burst: mean score of its R units
    size_t cap_sbar = 0;
    // pmf memo of the last forward (layout: the context's RbMemoLayout of n
    // states): every count vector's pmf, and (with scores) its pmfs with
    // one trial of each counted state removed
    float* d_poolP = nullptr;
    size_t cap_poolP = 0;
    bool sbar_valid = false;           // d_Sbar holds the last forward's means
    double* h_grad = nullptr;          // pinned
    hipEvent_t ev0 = nullptr, ev_walk = nullptr, ev1 = nullptr, ev_a0 = nullptr, ev_a1 = nullptr;
    bool ready = false;                // every fixed-size buffer above allocated
    bool ran = false, scores = false, adj_ran = false;
    int R = 0, B = 0, n = 0, K = 0, cap = 0, left_pad = 0;
    int64_t b0 = 0, b1 = 0;
    uint32_t seed = 0, set = 0;
    int init_state = 0;
    bool rates_positive = false, p_interior = false;
    bool wide = false;  // a state with 0.4 < p < 0.6: long units' derivatives in double
};

struct Slot {
    hipStream_t stream = nullptr;
    ModelBlob* h_blob = nullptr;  // pinned staging of the model tables
    ModelBlob* d_blob = nullptr;
    // threshold rows of the fast walk's binomial draws for the uploaded set
    // (guide_kernel, enqueued with every model upload)
    ThrPair* d_thr = nullptr;     // kGuideThrPairs
    uint8_t* d_guide = nullptr;   // kGuideBytes
    bool any_v = false;           // broadening: the general pass walks every unit
    uint32_t* d_replay = nullptr;  // units the fast pass abandoned
    uint32_t* d_replay_count = nullptr;
    uint32_t* d_work = nullptr;    // kWorkWords: work counters of the run, one per line
    size_t replay_cap = 0;
    hipEvent_t model_copied = nullptr;
    bool model_pending = false;
    int n = 0;
    bool have_params = false;
    Photophysics ph;
    std::vector<float> args;
    int64_t begin = 0, end = -1;
    int32_t* d_hist_acc = nullptr;
    ResultHeader* d_result = nullptr;
    ResultHeader* h_result = nullptr;  // pinned
    int last_B = 0;
    int64_t last_units = 0;
    bool ran = false;
    float* d_E = nullptr;
    int32_t* d_acc = nullptr;
    float* d_F = nullptr;
    float* h_E = nullptr;  // pinned (the reference's hmcE[sid])
    size_t sample_cap = 0;
    bool samples_valid = false;
    hipEvent_t ev0 = nullptr, ev_mid = nullptr, ev1 = nullptr;
    Population pop;
    RbState rb;
    // RB forward: the pmf table and memo are built on rb_aux while the walk
    // runs on stream (fork after the model copy, join before the pmf pass)
    hipStream_t rb_aux = nullptr;
    hipEvent_t rb_fork = nullptr, rb_join = nullptr;
    std::vector<RbState> rbpop;  // RB population: one state per set
    int rbpop_sets = 0;          // sets of the last RB population forward
};

}  // namespace pda

namespace pda {
// Parameter-independent layout of the RB pmf memo for n states: the count
// vectors with sum < kRbMemoCap, listed by total (so the lanes of a wave
// build pmfs of one length), and per vector its offset in the pool and the
// offsets of its leave-one-out pmfs (the pmfs of the vectors with one trial
// of the 1st..4th counted state removed, which the pool holds too).
struct RbMemoLayout {
    uint32_t nvec = 0;
    uint64_t poolP = 0;               // floats
    uint4* d_list = nullptr;          // list order: counts (x, y), offP, 0
    uint32_t* d_rank_offP = nullptr;  // by rb_rank
    uint4* d_rank_offD = nullptr;     // by rb_rank: leave-one-out offsets in the pool
};
}  // namespace pda

struct pda_ctx {
    int device = 0;
    int num_cu = 256;
    std::vector<pda::Slot> slots;
    int64_t nbursts = 0, nphotons = 0;
    uint64_t data_gen = 0;          // bumped by every burst load
    pda::BurstDesc* d_bursts = nullptr;
    uint32_t* d_ticks = nullptr;
    uint2* d_links = nullptr;
    float clk_p = 0.f, bg_ad = 0.f, bg_dd = 0.f;
    int B = 0, left_pad = 2;
    float* d_obs = nullptr;
    pda::LaunchShape shape;
    std::vector<uint32_t> h_P;      // photons per burst (host copy)
    double* d_lf = nullptr;         // log k! for the RB form
    size_t lds_per_block = 65536;   // LDS a workgroup may allocate
    pda::RbMemoLayout rb_memo[5];   // by state count (2..4 used)
};

namespace pda {

Slot* slot_of(pda_ctx* ctx, int sid);
int sync_all(pda_ctx* ctx);
void range_of(const pda_ctx* ctx, const Slot& s, int64_t* b0, int64_t* b1);
void free_population(Population& p);
void drop_graph(Population& p);
// pda_noalloc.hip: looks up the device's copy of the no-alloc scratch array
// (called by pda_create with the device current)
int noalloc_setup(int device);
void free_rb(RbState& r);
void free_rb_memo(pda_ctx* ctx);

// Walk passes (pda_kernels.hip): fast form, general form over the replay
// list, general form over every unit.
constexpr int kPassFast = 0, kPassReplay = 1, kPassGeneral = 2;

// Grid and batch of a walk launch over `units` work units.
struct Launch {
    uint32_t grid = 1, batch = kWave, queues = 1;
};
Launch plan_launch(const pda_ctx* ctx, int pass, bool population, uint64_t units);

// kernels (pda_kernels.hip)
hipError_t walk_occupancy(int pass, int* blocks_per_cu);
hipError_t pop_walk_occupancy(int pass, int B, int* blocks_per_cu);
hipError_t launch_walk(int pass, uint32_t grid, const BurstDesc* bursts, const uint32_t* ticks,
                       const uint2* links, const ModelBlob* blob, const BinomGuide& guide,
                       const RunArgs& a, int32_t* hist_acc, uint32_t* replay,
                       uint32_t* replay_count, uint32_t* work, float* outE, int32_t* outAcc,
                       float* outF, hipStream_t stream);
hipError_t launch_guide(const ModelBlob* blobs, int nsets, int max_states, ThrPair* thr,
                        uint8_t* guide, hipStream_t stream);
hipError_t launch_finish(const RunArgs& a, const float* obs, int32_t* hist_acc,
                         uint32_t* replay_count, uint32_t* work, ResultHeader* result,
                         hipStream_t stream);
// guide.thr == nullptr: the fast form draws binomials by the inversion loop
hipError_t launch_pop_walk(int pass, uint32_t grid, const BurstDesc* bursts,
                           const uint32_t* ticks, const uint2* links, const PopControl* ctl,
                           const ModelBlob* blobs, const BinomGuide& guide, const RunArgs& a,
                           int32_t* hist_acc, uint64_t* replay, uint32_t* replay_count,
                           uint32_t* work, hipStream_t stream);
hipError_t launch_pop_replay(const BurstDesc* bursts, const uint32_t* ticks, const uint2* links,
                             const PopControl* ctl, const ModelBlob* blobs, const RunArgs& a,
                             int32_t* hist_acc, const uint64_t* replay,
                             const uint32_t* replay_count, uint32_t grid, hipStream_t stream);
hipError_t launch_pop_finish(const ModelBlob* blobs, const RunArgs& a, const float* obs,
                             int32_t* hist_acc, uint32_t* replay_count, uint32_t* work,
                             PopStatus* status, SetResult* results, int32_t* hists,
                             hipStream_t stream);
hipError_t launch_primitive(int op, const void* in, void* out, int64_t n, hipStream_t stream);
// event microbenchmarks of the cost model: grid at full occupancy, launch
// (out: 3 uint64 per lane -- checksum, events, inner iterations)
hipError_t microbench_shape(int op, uint32_t* grid, uint32_t* threads);
hipError_t launch_microbench(int op, uint32_t grid, int iters, float param, const double* qn,
                             const int32_t* bound, uint64_t* out, hipStream_t stream);
struct RbModel;
// rb_big_kernel, mode 0 forward, 1 adjoint, 2 Jacobian; wide: derivative
// passes in double
hipError_t rb_big_occupancy(int mode, bool wide, int dyn_lds, int* per_cu);
// rb_unit_kernel of n states (2..4), mode 0 forward, 1 adjoint, 2 Jacobian
hipError_t rb_unit_occupancy(int mode, int n, int dyn_lds, int* per_cu);
// the RB walk of n states (2..4), with or without scores
hipError_t rb_walk_occupancy(bool scores, int n, int* per_cu);
// the fast-form walk and, behind it, the general-form replay of the units it
// abandoned (replay_grid workgroups; those beyond the list's need exit)
hipError_t launch_rb_walk(bool scores, int n, uint32_t grid, uint32_t replay_grid,
                          const BurstDesc* bursts, const uint32_t* ticks, const uint2* links,
                          const ModelBlob* blob, const RbModel* rbm, const RunArgs& a, uint2* cnt,
                          float* F, float* S, uint32_t* work, uint32_t* big, uint32_t* huge,
                          uint32_t* big_count, uint32_t* replay, uint32_t* replay_count,
                          hipStream_t stream);
// dynamic LDS bytes (mode: 0 forward, 1 adjoint, 2 Jacobian)
int rb_unit_threads();
size_t rb_unit_lds_bytes(int B, int mode, int n);
size_t rb_big_lds_bytes(int B, int cap, int mode, bool wide);
size_t rb_memo_lds_bytes(int n);
hipError_t launch_rb_table(const RbModel* rbm, const double* lf, float* pmf_tab, int n,
                           hipStream_t stream);
hipError_t launch_rb_memo(int n, const RbModel* rbm, const float* pmf_tab, const uint4* list,
                          uint32_t nvec, float* poolP, hipStream_t stream);
hipError_t launch_rb_unit(int mode, int n, uint32_t grid, const RbModel* rbm, const RunArgs& a,
                          const uint2* cnt, const float* F, const float* S, const double* Sbar,
                          const double* g, const uint32_t* rank_offP, const uint4* rank_offD,
                          const float* poolP, unsigned long long* hist_acc,
                          double* grad, uint32_t* work, hipStream_t stream);
hipError_t launch_rb_big(int mode, bool wide, uint32_t grid, int cap, const RbModel* rbm,
                         const double* lf,
                         const float* pmf_tab, const RunArgs& a, const uint2* cnt, const float* F,
                         const float* S, const double* Sbar, const double* g,
                         const uint32_t* big, const uint32_t* big_count,
                         unsigned long long* hist_acc, double* grad, hipStream_t stream);
hipError_t launch_rb_score_mean(const RunArgs& a, int K, const float* S, double* Sbar,
                                int num_cu, hipStream_t stream);
hipError_t launch_rb_finish(const RunArgs& a, const float* obs, unsigned long long* hist_acc,
                            uint32_t* work, void* res, hipStream_t stream);

}  // namespace pda
