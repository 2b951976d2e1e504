// pda_capi.cpp -- C ABI of the Monte-Carlo PDA hot path (include/gsmfretda_pda.h).
//
// Host half of the MI355X replacement for gSMFRETda's `class mc`
// (src/mc.hpp:32-180, src/mc.cu:204-829): device selection, stream slots,
// the one-time burst re-encoding and upload, per-slot parameter tables,
// launch, completion polling and result retrieval.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gsmfretda_pda.h"
#include "pda_ctx.h"
#include "pda_layout.h"
#include "pda_math.h"
#include "pda_types.h"

namespace pda {

namespace {
thread_local std::string g_err;
}  // namespace

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

void set_error(const std::string& msg) { g_err = msg; }
const std::string& last_error() { return g_err; }

Slot* slot_of(pda_ctx* ctx, int sid) {
    if (!ctx || sid < 0 || sid >= int(ctx->slots.size())) return nullptr;
    return &ctx->slots[size_t(sid)];
}

int sync_all(pda_ctx* ctx) {
    HIP_TRY(hipSetDevice(ctx->device));
    for (auto& s : ctx->slots) HIP_TRY(hipStreamSynchronize(s.stream));
    return PDA_OK;
}

void range_of(const pda_ctx* ctx, const Slot& s, int64_t* b0, int64_t* b1) {
    int64_t e = s.end;
    if (e < 0 || e > ctx->nbursts) e = ctx->nbursts;
    int64_t b = s.begin;
    if (b < 0 || b >= e) b = e;  // empty range
    *b0 = b;
    *b1 = e;
}

// Persistent grid of one-wave workgroups, as many as fit on the device (or
// the tuned number per CU); batches of units per work-counter grab: the tuned
// size, or about six grabs per wave.  A grab is a round trip to the counter
// that stalls the whole wave, but with every counter on its own line it is
// short, and smaller batches shorten the grid's tail (on 1e7-unit runs
// 256-unit batches beat both 64 and 700).
Launch plan_launch(const pda_ctx* ctx, int pass, bool population, uint64_t units) {
    int per_cu = 0;
    hipError_t e = population ? pop_walk_occupancy(pass, ctx->B, &per_cu)
                              : walk_occupancy(pass, &per_cu);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    if (ctx->shape.waves_per_cu > 0 && int(ctx->shape.waves_per_cu) < per_cu)
        per_cu = int(ctx->shape.waves_per_cu);
    const uint64_t max_grid = uint64_t(ctx->num_cu) * uint64_t(per_cu);
    uint64_t batch = uint64_t(kWave) * ctx->shape.batch_tiles;
    if (ctx->shape.batch_tiles == 0) {
        // about six grabs per wave: small batches keep the grid's tail
        // short (a wave that grabs last still walks a whole batch)
        batch = (units / (max_grid * 6u) + kWave - 1) / kWave * kWave;
        batch = std::min<uint64_t>(std::max<uint64_t>(batch, kWave), 4096);
    }
    const uint64_t batches = (units + batch - 1) / batch;
    Launch L;
    L.grid = uint32_t(std::max<uint64_t>(1, std::min(max_grid, batches)));
    L.batch = uint32_t(batch);
    const uint32_t q = ctx->shape.queues ? ctx->shape.queues : kDefaultQueues;
    L.queues = std::min<uint32_t>(q, L.grid);
    return L;
}

}  // namespace pda

using namespace pda;

namespace {

void free_data(pda_ctx* ctx) {
    (void)hipFree(ctx->d_bursts);
    (void)hipFree(ctx->d_ticks);
    (void)hipFree(ctx->d_links);
    ctx->d_bursts = nullptr;
    ctx->d_ticks = nullptr;
    ctx->d_links = nullptr;
    ctx->nbursts = ctx->nphotons = 0;
}

// Uploads the re-encoded bursts (mc::init_data_gpu, src/mc.cu:456-521).
int load_common(pda_ctx* ctx, const int64_t* start, const int64_t* stop, const int64_t* istart,
                const int32_t* phcount, int64_t nbursts, const PhotonSrc& src, int64_t nphotons,
                const float* T, float clk_p, float bg_ad, float bg_dd) {
    if (!ctx) return fail(PDA_E_ARG, "null context");
    std::vector<BurstDesc> bursts;
    std::vector<uint32_t> ticks;
    std::vector<uint2> links;
    std::string err;
    int code = encode_bursts(start, stop, istart, phcount, nbursts, src, nphotons, T, clk_p, bg_ad,
                             bg_dd, bursts, ticks, links, err);
    if (code) {
        set_error(err);
        return code;
    }
    const int64_t total = int64_t(ticks.size());
    int rc = sync_all(ctx);
    if (rc) return rc;
    // captured population graphs hold the old buffers' addresses
    for (auto& s : ctx->slots) drop_graph(s.pop);
    ++ctx->data_gen;
    free_data(ctx);
    if (nbursts > 0) {
        HIP_TRY(hipMalloc(&ctx->d_bursts, sizeof(BurstDesc) * size_t(nbursts)));
        HIP_TRY(hipMemcpy(ctx->d_bursts, bursts.data(), sizeof(BurstDesc) * size_t(nbursts),
                          hipMemcpyHostToDevice));
    }
    if (total > 0) {
        HIP_TRY(hipMalloc(&ctx->d_ticks, sizeof(uint32_t) * size_t(total)));
        HIP_TRY(hipMalloc(&ctx->d_links, sizeof(uint2) * size_t(total)));
        HIP_TRY(hipMemcpy(ctx->d_ticks, ticks.data(), sizeof(uint32_t) * size_t(total),
                          hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ctx->d_links, links.data(), sizeof(uint2) * size_t(total),
                          hipMemcpyHostToDevice));
    }
    ctx->nbursts = nbursts;
    ctx->nphotons = total;
    ctx->h_P.resize(size_t(nbursts));
    for (int64_t i = 0; i < nbursts; ++i) ctx->h_P[size_t(i)] = bursts[size_t(i)].P;
    ctx->clk_p = clk_p;
    ctx->bg_ad = bg_ad;
    ctx->bg_dd = bg_dd;
    for (auto& s : ctx->slots) {
        s.begin = 0;
        s.end = -1;
        auto reset = [](RbState& r) {
            r.ran = false;  // no derivative of a run over the previous data
        };
        reset(s.rb);
        for (auto& r : s.rbpop) reset(r);
        s.rbpop_sets = 0;
    }
    return PDA_OK;
}

int upload_model(Slot& s) {
    Model M;
    std::string err;
    int rc = build_model(s.ph, s.n, s.args.data(), int(s.args.size()), M, err);
    if (rc) {
        set_error(err);
        return rc;
    }
    if (s.model_pending) {
        HIP_TRY(hipEventSynchronize(s.model_copied));
        s.model_pending = false;
    }
    s.h_blob->m = M;
    build_binom_table(M, s.h_blob->tab);
    s.any_v = M.any_v != 0;
    HIP_TRY(hipMemcpyAsync(s.d_blob, s.h_blob, sizeof(ModelBlob), hipMemcpyHostToDevice, s.stream));
    HIP_TRY(hipEventRecord(s.model_copied, s.stream));
    if (!s.any_v) HIP_TRY(launch_guide(s.d_blob, 1, M.n, s.d_thr, s.d_guide, s.stream));
    s.model_pending = true;
    return PDA_OK;
}


}  // namespace

extern "C" {

const char* pda_last_error(void) { return last_error().c_str(); }

const char* pda_version(void) { return "gsmfretda_amd 0.1 (gfx950)"; }

int pda_create(int device, int nstreams, pda_ctx** out) {
    if (!out) return fail(PDA_E_ARG, "null out");
    *out = nullptr;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fail(PDA_E_ARG, "gpu id %d outside [0, %d)", device, ndev);
    if (nstreams == 0) nstreams = 4;
    if (nstreams < 0 || nstreams > 64) return fail(PDA_E_ARG, "nstreams %d outside [1, 64]", nstreams);
    HIP_TRY(hipSetDevice(device));
    int rc0 = noalloc_setup(device);
    if (rc0) return rc0;
    auto* ctx = new pda_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0) {
        ctx->num_cu = prop.multiProcessorCount;
        if (prop.sharedMemPerBlock > 0) ctx->lds_per_block = prop.sharedMemPerBlock;
    }
    ctx->slots.resize(size_t(nstreams));
    for (auto& s : ctx->slots) {
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void**>(&s.h_blob), sizeof(ModelBlob), hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&s.d_blob, sizeof(ModelBlob)) != hipSuccess ||
            hipMalloc(&s.d_thr, sizeof(ThrPair) * kGuideThrPairs) != hipSuccess ||
            hipMalloc(&s.d_guide, kGuideBytes) != hipSuccess ||
            hipMalloc(&s.d_replay_count, sizeof(uint32_t)) != hipSuccess ||
            hipMemset(s.d_replay_count, 0, sizeof(uint32_t)) != hipSuccess ||
            hipMalloc(&s.d_work, sizeof(uint32_t) * kWorkWords) != hipSuccess ||
            hipMemset(s.d_work, 0, sizeof(uint32_t) * kWorkWords) != hipSuccess ||
            hipEventCreateWithFlags(&s.model_copied, hipEventDisableTiming) != hipSuccess ||
            hipMalloc(&s.d_hist_acc, sizeof(int32_t) * kMaxBins * 8) != hipSuccess ||

            hipMalloc(&s.d_result, sizeof(ResultHeader) + sizeof(int32_t) * kMaxBins) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void**>(&s.h_result),
                          sizeof(ResultHeader) + sizeof(int32_t) * kMaxBins, hipHostMallocDefault) != hipSuccess ||
            hipEventCreate(&s.ev0) != hipSuccess || hipEventCreate(&s.ev_mid) != hipSuccess ||
            hipEventCreate(&s.ev1) != hipSuccess ||
            hipMemset(s.d_hist_acc, 0, sizeof(int32_t) * kMaxBins * 8) != hipSuccess ||
            hipDeviceSynchronize() != hipSuccess) {
            pda_destroy(ctx);
            return fail(PDA_E_HIP, "stream slot setup failed: %s", hipGetErrorString(hipGetLastError()));
        }
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        pda_destroy(ctx);
        return fail(PDA_E_HIP, "device sync failed");
    }
    *out = ctx;
    return PDA_OK;
}

int pda_destroy(pda_ctx* ctx) {
    if (!ctx) return PDA_OK;
    (void)hipSetDevice(ctx->device);
    for (auto& s : ctx->slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        (void)hipFree(s.d_blob);
        (void)hipFree(s.d_thr);
        (void)hipFree(s.d_guide);
        (void)hipHostFree(s.h_blob);
        (void)hipFree(s.d_replay);
        (void)hipFree(s.d_replay_count);
        (void)hipFree(s.d_work);
        (void)hipFree(s.d_hist_acc);
        (void)hipFree(s.d_result);
        (void)hipHostFree(s.h_result);
        (void)hipFree(s.d_E);
        (void)hipFree(s.d_acc);
        (void)hipFree(s.d_F);
        (void)hipHostFree(s.h_E);
        if (s.model_copied) (void)hipEventDestroy(s.model_copied);
        if (s.ev0) (void)hipEventDestroy(s.ev0);
        if (s.ev_mid) (void)hipEventDestroy(s.ev_mid);
        if (s.ev1) (void)hipEventDestroy(s.ev1);
        free_population(s.pop);
        free_rb(s.rb);
        for (auto& r : s.rbpop) free_rb(r);
        if (s.rb_aux) {
            (void)hipStreamSynchronize(s.rb_aux);
            (void)hipStreamDestroy(s.rb_aux);
        }
        if (s.rb_fork) (void)hipEventDestroy(s.rb_fork);
        if (s.rb_join) (void)hipEventDestroy(s.rb_join);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    free_data(ctx);
    (void)hipFree(ctx->d_obs);
    (void)hipFree(ctx->d_lf);
    free_rb_memo(ctx);
    delete ctx;
    return PDA_OK;
}

int pda_num_streams(const pda_ctx* ctx) { return ctx ? int(ctx->slots.size()) : 0; }

void* pda_stream(pda_ctx* ctx, int sid) {
    Slot* s = slot_of(ctx, sid);
    return s ? reinterpret_cast<void*>(s->stream) : nullptr;
}

int pda_load_bursts(pda_ctx* ctx, const int64_t* start, const int64_t* stop, const int64_t* istart,
                    const int32_t* phcount, int64_t nbursts, const int64_t* times,
                    const uint8_t* mask_ad, const uint8_t* mask_dd, int64_t nphotons,
                    const float* T, float clk_p, float bg_ad, float bg_dd) {
    PhotonSrc src{times, mask_ad, mask_dd, nullptr, nullptr};
    if (nphotons > 0 && !times) return fail(PDA_E_ARG, "null times");
    return load_common(ctx, start, stop, istart, phcount, nbursts, src, nphotons, T, clk_p, bg_ad,
                       bg_dd);
}

int pda_load_bursts_tagged(pda_ctx* ctx, const int64_t* istart, const int64_t* start,
                           const int64_t* stop, const int32_t* phcount, int64_t nbursts,
                           const int64_t* burst_ad, const int64_t* burst_dd, int64_t nphotons,
                           const float* T, float clk_p, float bg_ad, float bg_dd) {
    PhotonSrc src{nullptr, nullptr, nullptr, burst_ad, burst_dd};
    if (nphotons > 0 && (!burst_ad || !burst_dd)) return fail(PDA_E_ARG, "null burst_ad/burst_dd");
    return load_common(ctx, start, stop, istart, phcount, nbursts, src, nphotons, T, clk_p, bg_ad,
                       bg_dd);
}

int64_t pda_num_bursts(const pda_ctx* ctx) { return ctx ? ctx->nbursts : 0; }

int pda_set_observed_hist(pda_ctx* ctx, const float* ohist, int nbins, int left_pad) {
    if (!ctx || !ohist) return fail(PDA_E_ARG, "null argument");
    if (nbins < 1 || nbins > kMaxBins) return fail(PDA_E_ARG, "nbins %d outside [1, %d]", nbins, kMaxBins);
    if (left_pad < 0 || left_pad > nbins) return fail(PDA_E_ARG, "left_pad %d outside [0, nbins]", left_pad);
    int rc = sync_all(ctx);
    if (rc) return rc;
    if (!ctx->d_obs) HIP_TRY(hipMalloc(&ctx->d_obs, sizeof(float) * kMaxBins));
    HIP_TRY(hipMemcpy(ctx->d_obs, ohist, sizeof(float) * size_t(nbins), hipMemcpyHostToDevice));
    ctx->B = nbins;
    ctx->left_pad = left_pad;
    return PDA_OK;
}

int pda_set_observed(pda_ctx* ctx, const float* sgdivsr, int64_t n, int nbins, int left_pad) {
    if (!ctx || (n > 0 && !sgdivsr)) return fail(PDA_E_ARG, "null argument");
    if (nbins < 1 || nbins > kMaxBins) return fail(PDA_E_ARG, "nbins %d outside [1, %d]", nbins, kMaxBins);
    // boost axis::regular(nbins, 0, 1) with coverage::inner (src/streamWorker.cpp:33-38,62-67)
    std::vector<float> h(size_t(nbins), 0.0f);
    for (int64_t i = 0; i < n; ++i) {
        const float x = sgdivsr[i];
        if (x >= 0.0f && x < 1.0f) {
            const int64_t b = int64_t(std::floor(double(x) * double(nbins)));
            if (b >= 0 && b < nbins) h[size_t(b)] += 1.0f;
        }
    }
    return pda_set_observed_hist(ctx, h.data(), nbins, left_pad);
}

int pda_set_params(pda_ctx* ctx, int sid, int n, const float* args, int nargs) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    HIP_TRY(hipSetDevice(ctx->device));
    Model M;
    std::string err;
    int rc = build_model(s->ph, n, args, nargs, M, err);  // validate before keeping
    if (rc) {
        set_error(err);
        return rc;
    }
    s->n = n;
    s->args.assign(args, args + nargs);
    s->have_params = true;
    return upload_model(*s);
}

int pda_set_photophysics(pda_ctx* ctx, int sid, float gamma, float dexdir, float crosstalk,
                         float r0) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (!std::isfinite(gamma) || !std::isfinite(dexdir) || !std::isfinite(crosstalk) ||
        !(r0 > 0.0f) || !std::isfinite(r0))
        return fail(PDA_E_ARG, "photophysics constants must be finite, r0 > 0");
    // The corrected count (gamma - Dch2Ach) f_id + (1 - DexDirAem) f_ia
    // (src/mc.cu:143) is the binomial's trial count; with a negative
    // coefficient it goes negative, where rk_binomial (src/binom.cuh:172-188)
    // has no meaning (it returns n < 0 acceptors for p > 1/2).
    if (!(gamma - crosstalk >= 0.0f) || !(1.0f - dexdir >= 0.0f))
        return fail(PDA_E_ARG,
                    "gamma %g < crosstalk %g or DexDirAem %g > 1: the corrected photon count "
                    "would be negative", double(gamma), double(crosstalk), double(dexdir));
    s->ph.gamma = gamma;
    s->ph.dexdir = dexdir;
    s->ph.crosstalk = crosstalk;
    s->ph.r0 = r0;
    if (!s->have_params) return PDA_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    return upload_model(*s);
}

int pda_set_burst_range(pda_ctx* ctx, int sid, int64_t begin, int64_t end, int64_t* n_out) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    s->begin = begin;
    s->end = end;
    int64_t b0, b1;
    range_of(ctx, *s, &b0, &b1);
    if (n_out) *n_out = b1 - b0;
    return PDA_OK;
}

int pda_run(pda_ctx* ctx, int sid, int R, uint32_t seed, uint32_t set_id, int flags) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (R < 1) return fail(PDA_E_ARG, "R = %d < 1", R);
    if (!s->have_params) return fail(PDA_E_STATE, "stream %d: no parameters set", sid);
    if (ctx->B < 1 || !ctx->d_obs) return fail(PDA_E_STATE, "no observed histogram set");
    HIP_TRY(hipSetDevice(ctx->device));
    int64_t b0, b1;
    range_of(ctx, *s, &b0, &b1);
    const int64_t nunits = (b1 - b0) * int64_t(R);
    if (nunits > int64_t(0x7FFFFFFF) - kWave)
        return fail(PDA_E_ARG, "%lld bursts x R = %d exceeds 2^31 units per run",
                    (long long)(b1 - b0), R);
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
    // general pass for broadening, else the fast pass and a replay pass
    const int pass = s->any_v ? kPassGeneral : kPassFast;
    const Launch L = plan_launch(ctx, pass, false, uint64_t(nunits));
    const uint32_t grid = L.grid;
    a.batch = L.batch;
    a.queues = L.queues;
    a.nsets = 1;
    a.bps = 0;

    float* outE = nullptr;
    int32_t* outAcc = nullptr;
    float* outF = nullptr;
    s->samples_valid = false;
    if ((flags & PDA_RUN_SAMPLES) && nunits > 0) {
        if (size_t(nunits) > s->sample_cap) {
            HIP_TRY(hipStreamSynchronize(s->stream));
            (void)hipFree(s->d_E);
            (void)hipFree(s->d_acc);
            (void)hipFree(s->d_F);
            (void)hipHostFree(s->h_E);
            s->d_E = nullptr;
            s->d_acc = nullptr;
            s->d_F = nullptr;
            s->h_E = nullptr;
            s->sample_cap = 0;
            HIP_TRY(hipMalloc(&s->d_E, sizeof(float) * size_t(nunits)));
            HIP_TRY(hipMalloc(&s->d_acc, sizeof(int32_t) * size_t(nunits)));
            HIP_TRY(hipMalloc(&s->d_F, sizeof(float) * size_t(nunits)));
            HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&s->h_E), sizeof(float) * size_t(nunits),
                                  hipHostMallocDefault));
            s->sample_cap = size_t(nunits);
        }
        outE = s->d_E;
        outAcc = s->d_acc;
        outF = s->d_F;
    }
    if (pass == kPassFast && size_t(nunits) > s->replay_cap) {
        HIP_TRY(hipStreamSynchronize(s->stream));
        (void)hipFree(s->d_replay);
        s->d_replay = nullptr;
        s->replay_cap = 0;
        HIP_TRY(hipMalloc(&s->d_replay, sizeof(uint32_t) * size_t(nunits)));
        s->replay_cap = size_t(nunits);
    }
    a.replay_cap = uint32_t(s->replay_cap);
    HIP_TRY(hipEventRecord(s->ev0, s->stream));
    const BinomGuide guide{s->d_thr, s->d_guide};
    HIP_TRY(launch_walk(pass, grid, ctx->d_bursts, ctx->d_ticks, ctx->d_links, s->d_blob, guide, a,
                        s->d_hist_acc, s->d_replay, s->d_replay_count, s->d_work, outE, outAcc,
                        outF, s->stream));
    HIP_TRY(hipEventRecord(s->ev_mid, s->stream));
    if (pass == kPassFast && nunits > 0) {
        // the abandoned units, in the general form over the whole GPU; the
        // grid is sized for the worst case and its unneeded workgroups exit
        // at once
        int per_cu = 0;
        if (walk_occupancy(kPassReplay, &per_cu) != hipSuccess || per_cu < 1) per_cu = 1;
        const uint64_t want = (uint64_t(nunits) + kWave - 1) / kWave;
        const uint32_t rgrid =
            uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(want, uint64_t(ctx->num_cu) * per_cu)));
        HIP_TRY(launch_walk(kPassReplay, rgrid, ctx->d_bursts, ctx->d_ticks, ctx->d_links,
                            s->d_blob, guide, a, s->d_hist_acc, s->d_replay, s->d_replay_count,
                            s->d_work,
                            outE, outAcc, outF, s->stream));
    }
    HIP_TRY(launch_finish(a, ctx->d_obs, s->d_hist_acc, s->d_replay_count, s->d_work, s->d_result,
                          s->stream));
    HIP_TRY(hipEventRecord(s->ev1, s->stream));
    HIP_TRY(hipMemcpyAsync(s->h_result, s->d_result, sizeof(ResultHeader) + sizeof(int32_t) * size_t(a.B),
                           hipMemcpyDeviceToHost, s->stream));
    if (outE)
        HIP_TRY(hipMemcpyAsync(s->h_E, outE, sizeof(float) * size_t(nunits), hipMemcpyDeviceToHost,
                               s->stream));
    s->last_B = a.B;
    s->last_units = nunits;
    s->ran = true;
    s->samples_valid = outE != nullptr;
    return PDA_OK;
}

int pda_query(pda_ctx* ctx, int sid) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    const hipError_t e = hipStreamQuery(s->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return fail(PDA_E_HIP, "hipStreamQuery: %s", hipGetErrorString(e));
}

int pda_synchronize(pda_ctx* ctx, int sid) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    HIP_TRY(hipStreamSynchronize(s->stream));
    return PDA_OK;
}

int pda_get_result(pda_ctx* ctx, int sid, int32_t* hist, int nbins, double* chi2_64,
                   float* chi2_ref, int* effN) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (!s->ran) return fail(PDA_E_STATE, "stream %d has not run", sid);
    const int q = pda_query(ctx, sid);
    if (q < 0) return q;
    if (q == 0) return fail(PDA_E_STATE, "stream %d still running", sid);
    if (hist) {
        if (nbins != s->last_B) return fail(PDA_E_ARG, "nbins %d != run's %d", nbins, s->last_B);
        std::memcpy(hist, reinterpret_cast<const int32_t*>(s->h_result + 1), sizeof(int32_t) * size_t(nbins));
    }
    if (s->last_units == 0) {  // nothing ran: empty histogram, chi2 of no bins
        if (hist) std::memset(hist, 0, sizeof(int32_t) * size_t(nbins));
    }
    if (chi2_64) *chi2_64 = s->h_result->chi2_64;
    if (chi2_ref) *chi2_ref = s->h_result->chi2_ref;
    if (effN) *effN = s->h_result->effN;
    return PDA_OK;
}

int pda_run_stats(pda_ctx* ctx, int sid, int64_t* units, int32_t* replays) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (!s->ran) return fail(PDA_E_STATE, "stream %d has not run", sid);
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (units) *units = s->last_units;
    if (replays) *replays = s->last_units > 0 ? s->h_result->replays : 0;
    return PDA_OK;
}

int pda_get_samples(pda_ctx* ctx, int sid, float* E, int32_t* acc, float* F, int64_t n) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (!s->samples_valid) return fail(PDA_E_STATE, "last run kept no samples");
    if (n != s->last_units) return fail(PDA_E_ARG, "n = %lld != %lld units", (long long)n, (long long)s->last_units);
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (E) std::memcpy(E, s->h_E, sizeof(float) * size_t(n));
    if (acc) HIP_TRY(hipMemcpy(acc, s->d_acc, sizeof(int32_t) * size_t(n), hipMemcpyDeviceToHost));
    if (F) HIP_TRY(hipMemcpy(F, s->d_F, sizeof(float) * size_t(n), hipMemcpyDeviceToHost));
    return PDA_OK;
}

float* pda_samples_host(pda_ctx* ctx, int sid) {
    Slot* s = slot_of(ctx, sid);
    if (!s || !s->samples_valid) return nullptr;
    return s->h_E;
}

int pda_last_run_ms(pda_ctx* ctx, int sid, float* ms) {
    Slot* s = slot_of(ctx, sid);
    if (!s || !ms) return fail(PDA_E_ARG, "bad argument");
    if (!s->ran) return fail(PDA_E_STATE, "stream %d has not run", sid);
    HIP_TRY(hipEventSynchronize(s->ev1));
    HIP_TRY(hipEventElapsedTime(ms, s->ev0, s->ev1));
    return PDA_OK;
}

int pda_last_pass_ms(pda_ctx* ctx, int sid, float* walk_ms, float* finish_ms) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (!s->ran) return fail(PDA_E_STATE, "stream %d has not run", sid);
    HIP_TRY(hipEventSynchronize(s->ev1));
    float a = 0.f, b = 0.f;
    HIP_TRY(hipEventElapsedTime(&a, s->ev0, s->ev_mid));
    HIP_TRY(hipEventElapsedTime(&b, s->ev_mid, s->ev1));
    if (walk_ms) *walk_ms = a;
    if (finish_ms) *finish_ms = b;
    return PDA_OK;
}

int pda_debug_microbench(pda_ctx* ctx, int op, int iters, float param, float* ms,
                         int64_t* events, int64_t* inner) {
    if (!ctx || op < 0 || op > 6 || iters < 1 || iters > (1 << 24))
        return fail(PDA_E_ARG, "bad argument");
    if ((op == 4) && !(param >= 0.0f && param < float(kBinomTable)))
        return fail(PDA_E_ARG, "binomial microbenchmark: n = %g outside [0, %d)", double(param),
                    kBinomTable);
    if ((op == 3) && !(param >= 0.0f && param <= 64.0f))
        return fail(PDA_E_ARG, "Poisson microbenchmark: mean %g outside [0, 64]", double(param));
    HIP_TRY(hipSetDevice(ctx->device));
    uint32_t grid = 0, threads = 0;
    HIP_TRY(microbench_shape(op, &grid, &threads));
    const size_t lanes = size_t(grid) * threads;
    // q^m and the inversion bound of Binomial(m, 0.3), m < kBinomTable
    std::vector<double> qn(kBinomTable);
    std::vector<int32_t> bnd(kBinomTable);
    for (int m = 0; m < kBinomTable; ++m) {
        qn[size_t(m)] = std::pow(0.7, double(m));
        const double mp = 0.3 * m;
        bnd[size_t(m)] = int32_t(std::min(double(m), std::floor(mp + 10.0 * std::sqrt(mp * 0.7 + 1.0))));
    }
    uint64_t* d_out = nullptr;
    double* d_qn = nullptr;
    int32_t* d_bnd = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<uint64_t> h(lanes * 3);
    hipError_t e = hipMalloc(&d_out, sizeof(uint64_t) * 3 * lanes);
    if (e == hipSuccess) e = hipMalloc(&d_qn, sizeof(double) * kBinomTable);
    if (e == hipSuccess) e = hipMalloc(&d_bnd, sizeof(int32_t) * kBinomTable);
    if (e == hipSuccess) e = hipMemcpy(d_qn, qn.data(), sizeof(double) * kBinomTable, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_bnd, bnd.data(), sizeof(int32_t) * kBinomTable, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // one untimed launch, then the timed one
    if (e == hipSuccess) e = launch_microbench(op, grid, iters, param, d_qn, d_bnd, d_out, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) e = launch_microbench(op, grid, iters, param, d_qn, d_bnd, d_out, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e == hipSuccess) e = hipMemcpy(h.data(), d_out, sizeof(uint64_t) * 3 * lanes, hipMemcpyDeviceToHost);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d_out);
    (void)hipFree(d_qn);
    (void)hipFree(d_bnd);
    if (e != hipSuccess) return fail(PDA_E_HIP, "microbenchmark %d: %s", op, hipGetErrorString(e));
    uint64_t ev = 0, in = 0;
    for (size_t i = 0; i < lanes; ++i) {
        ev += h[3 * i + 1];
        in += h[3 * i + 2];
    }
    if (ms) *ms = t;
    if (events) *events = int64_t(ev);
    if (inner) *inner = int64_t(in);
    return PDA_OK;
}

int pda_debug_guide(pda_ctx* ctx, int sid, void* thr, int64_t thr_bytes, void* guide,
                    int64_t guide_bytes) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (!thr || !guide || thr_bytes < int64_t(sizeof(ThrPair) * kGuideThrPairs) ||
        guide_bytes < int64_t(kGuideBytes))
        return fail(PDA_E_ARG, "buffers below %zu / %zu bytes", sizeof(ThrPair) * kGuideThrPairs,
                    kGuideBytes);
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(thr, s->d_thr, sizeof(ThrPair) * kGuideThrPairs, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(guide, s->d_guide, kGuideBytes, hipMemcpyDeviceToHost));
    return PDA_OK;
}

int pda_debug_primitive(pda_ctx* ctx, int op, const void* in, void* out, int64_t n) {
    // input / output bytes per item for each op
    static const size_t in_sz[10] = {24, 4, 4, 8, 8, 4, 8, 48, 40, 4};
    static const size_t out_sz[10] = {16, 4, 4, 8, 8, 4, 4, 8, 4, 4};
    if (!ctx || !in || !out || op < 0 || op > 9 || n < 0 || n > (int64_t(1) << 28))
        return fail(PDA_E_ARG, "bad argument");
    if (n == 0) return PDA_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    void* d_in = nullptr;
    void* d_out = nullptr;
    HIP_TRY(hipMalloc(&d_in, in_sz[op] * size_t(n)));
    if (hipMalloc(&d_out, out_sz[op] * size_t(n)) != hipSuccess) {
        (void)hipFree(d_in);
        return fail(PDA_E_HIP, "hipMalloc failed");
    }
    hipError_t e = hipMemcpy(d_in, in, in_sz[op] * size_t(n), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_primitive(op, d_in, d_out, n, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, d_out, out_sz[op] * size_t(n), hipMemcpyDeviceToHost);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    if (e != hipSuccess) return fail(PDA_E_HIP, "primitive op %d: %s", op, hipGetErrorString(e));
    return PDA_OK;
}

}  // extern "C"
