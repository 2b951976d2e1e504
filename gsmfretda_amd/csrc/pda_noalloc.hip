// pda_noalloc.hip -- pda_run_noalloc: the single-set evaluation of pda_run
// without any device allocation, for callers that cannot allocate (code
// under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
//
// pda_run's kernels use per-run scratch beyond their inputs and outputs: the
// walk's work counters, the list of units the fast pass abandons and its
// count, and the per-XCD histogram rows.  pda_run keeps these in buffers the
// context allocates (and grows the replay list to the run's size).  Here
// they live in one file-scope device array, `scratch`, of kScratchBytes on
// every GPU the library runs on, held until the process ends.  Host code
// never passes the array's name to a kernel (in host code it is a stand-in,
// not a device address): pda_create calls noalloc_setup once per device,
// with that device current, and the device address hipGetSymbolAddress
// returns is kept per device and used for every kernel argument.
//
// Every call on a GPU shares that GPU's array, so calls that overlap on one
// GPU (two slots, two contexts, two threads) race; a caller serialises them.
// Each call clears the counters and rows it relies on with a small kernel
// first (no hipMemsetAsync, so the sequence can be captured into a graph).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/gsmfretda_pda.h"
#include "pda_ctx.h"
#include "pda_types.h"

namespace pda {

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // per GPU
// layout (byte offsets, each part 256-byte aligned)
constexpr size_t kOffWork = 0;
constexpr size_t kOffCount = kOffWork + sizeof(uint32_t) * kWorkWords;
constexpr size_t kOffHist = kOffCount + 256;
constexpr size_t kOffReplay = kOffHist + sizeof(int32_t) * size_t(kMaxBins) * 8u;
constexpr size_t kReplayCapacity = (kScratchBytes - kOffReplay) / sizeof(uint32_t);
static_assert(kOffCount % 256 == 0 && kOffHist % 256 == 0 && kOffReplay % 256 == 0,
              "scratch parts are 256-byte aligned");
static_assert(PDA_NOALLOC_MAX_UNITS <= kReplayCapacity,
              "the replay list holds every unit of the largest call");

static __device__ __align__(256) uint8_t scratch[kScratchBytes];
static_assert(sizeof(pda_result) == sizeof(ResultHeader), "pda_result is the result header");

std::mutex g_mu;
std::vector<uint8_t*> g_scratch;  // device address of `scratch`, per device

uint8_t* scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0 || size_t(device) >= g_scratch.size()) return nullptr;
    return g_scratch[size_t(device)];
}

// Zeroes the run's counters, replay count and histogram rows.
__global__ __launch_bounds__(1024) void clear_scratch_kernel(uint32_t* __restrict__ work,
                                                             uint32_t* __restrict__ count,
                                                             int32_t* __restrict__ hist_acc) {
    const uint32_t t = threadIdx.x;
    if (t < kMaxQueues + kAuxQueues) work[size_t(t) * kCounterStride] = 0u;
    if (t == 0) *count = 0u;
    for (uint32_t i = t; i < uint32_t(kMaxBins) * 8u; i += blockDim.x) hist_acc[i] = 0;
}

}  // namespace

// Called by pda_create with `device` current: looks up this device's copy of
// `scratch` once.
int noalloc_setup(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0) return fail(PDA_E_ARG, "bad device %d", device);
    if (size_t(device) >= g_scratch.size()) g_scratch.resize(size_t(device) + 1, nullptr);
    if (g_scratch[size_t(device)]) return PDA_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(PDA_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                    hipGetErrorString(e));
    g_scratch[size_t(device)] = static_cast<uint8_t*>(p);
    return PDA_OK;
}

}  // namespace pda

using namespace pda;

extern "C" int pda_run_noalloc(pda_ctx* ctx, int sid, int R, uint32_t seed, uint32_t set_id,
                               void* d_result, int64_t result_bytes) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (R < 1) return fail(PDA_E_ARG, "R = %d < 1", R);
    if (!d_result) return fail(PDA_E_ARG, "null result buffer");
    if (!s->have_params) return fail(PDA_E_STATE, "stream %d: no parameters set", sid);
    if (ctx->B < 1 || !ctx->d_obs) return fail(PDA_E_STATE, "no observed histogram set");
    const int64_t need = int64_t(sizeof(pda_result)) + int64_t(sizeof(int32_t)) * ctx->B;
    if (result_bytes < need)
        return fail(PDA_E_ARG, "result buffer of %lld bytes, %d bins need %lld",
                    (long long)result_bytes, ctx->B, (long long)need);
    int64_t b0, b1;
    range_of(ctx, *s, &b0, &b1);
    const int64_t nunits = (b1 - b0) * int64_t(R);
    if (nunits > PDA_NOALLOC_MAX_UNITS)
        return fail(PDA_E_ARG,
                    "%lld bursts x R = %d is %lld units; pda_run_noalloc takes at most %d per "
                    "call (its scratch is fixed)",
                    (long long)(b1 - b0), R, (long long)nunits, PDA_NOALLOC_MAX_UNITS);
    HIP_TRY(hipSetDevice(ctx->device));
    uint8_t* base = scratch_of(ctx->device);
    if (!base) return fail(PDA_E_STATE, "scratch of device %d not set up", ctx->device);
    uint32_t* work = reinterpret_cast<uint32_t*>(base + kOffWork);
    uint32_t* count = reinterpret_cast<uint32_t*>(base + kOffCount);
    int32_t* hist_acc = reinterpret_cast<int32_t*>(base + kOffHist);
    uint32_t* replay = reinterpret_cast<uint32_t*>(base + kOffReplay);

    RunArgs a{};
    a.begin = uint32_t(b0);
    a.nunits = uint32_t(nunits);
    a.R = uint32_t(R);
    a.seed = seed;
    a.set = set_id;
    a.B = ctx->B;
    a.left_pad = ctx->left_pad;
    a.bg_on = ctx->bg_dd < 1e-4f ? 0 : 1;  // src/mc.cu:153
    a.clk_p = ctx->clk_p;
    a.bg_ad = ctx->bg_ad;
    a.bg_dd = ctx->bg_dd;
    a.nstates = s->n;
    a.replay_cap = uint32_t(kReplayCapacity);
    a.nsets = 1;
    a.bps = 0;
    const int pass = s->any_v ? kPassGeneral : kPassFast;
    const Launch L = plan_launch(ctx, pass, false, uint64_t(nunits));
    a.batch = L.batch;
    a.queues = L.queues;

    hipLaunchKernelGGL(clear_scratch_kernel, dim3(1), dim3(1024), 0, s->stream, work, count,
                       hist_acc);
    HIP_TRY(hipGetLastError());
    const BinomGuide guide{s->d_thr, s->d_guide};
    HIP_TRY(launch_walk(pass, L.grid, ctx->d_bursts, ctx->d_ticks, ctx->d_links, s->d_blob, guide,
                        a, hist_acc, replay, count, work, nullptr, nullptr, nullptr, s->stream));
    if (pass == kPassFast && nunits > 0) {
        int per_cu = 0;
        if (walk_occupancy(kPassReplay, &per_cu) != hipSuccess || per_cu < 1) per_cu = 1;
        const uint64_t want = (uint64_t(nunits) + kWave - 1) / kWave;
        const uint32_t rgrid = uint32_t(
            std::max<uint64_t>(1, std::min<uint64_t>(want, uint64_t(ctx->num_cu) * per_cu)));
        HIP_TRY(launch_walk(kPassReplay, rgrid, ctx->d_bursts, ctx->d_ticks, ctx->d_links,
                            s->d_blob, guide, a, hist_acc, replay, count, work, nullptr, nullptr,
                            nullptr, s->stream));
    }
    HIP_TRY(launch_finish(a, ctx->d_obs, hist_acc, count, work,
                          static_cast<ResultHeader*>(d_result), s->stream));
    return PDA_OK;
}
