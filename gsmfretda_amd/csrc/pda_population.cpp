// pda_population.cpp -- population runs, launch shapes and the autotuner.
//
// The reference evaluates one parameter set per CUDA stream and per
// mc::run_kernel call, the parameter server handing a generation's sets to
// the GPU workers one at a time (src/gpuWorker.cpp:38-88,
// src/streamWorker.cpp:83-182).  A population run evaluates a whole
// generation -- sets of uneven state counts over the same resident bursts --
// in one launch sequence that is captured once into a hipGraph and replayed:
//   H2D copy of the generation's model tables and stream keys
//   -> pop_walk_kernel (persistent grid, fast form)
//   -> pop_replay_kernel (abandoned units, general form)
//   -> pop_finish_kernel (one workgroup per set: histogram and chi2)
//   -> D2H copy of every set's histogram and chi2.
// No node of the graph is a memset: the kernels leave their accumulators and
// counters zero for the next replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gsmfretda_pda.h"
#include "pda_ctx.h"
#include "pda_layout.h"
#include "pda_types.h"

namespace pda {

bool PopGraphKey::operator==(const PopGraphKey& o) const {
    return nsets == o.nsets && R == o.R && B == o.B && left_pad == o.left_pad && pass == o.pass &&
           bg_on == o.bg_on && b0 == o.b0 && b1 == o.b1 && grid == o.grid && batch == o.batch &&
           queues == o.queues && replay_cap == o.replay_cap && clk_p == o.clk_p &&
           bg_ad == o.bg_ad && bg_dd == o.bg_dd && bursts == o.bursts && ticks == o.ticks &&
           links == o.links && data_gen == o.data_gen && obs == o.obs && guided == o.guided &&
           max_states == o.max_states && thr == o.thr;
}

void drop_graph(Population& p) {
    if (p.exec) (void)hipGraphExecDestroy(p.exec);
    if (p.graph) (void)hipGraphDestroy(p.graph);
    p.exec = nullptr;
    p.graph = nullptr;
}

void free_population(Population& p) {
    drop_graph(p);
    (void)hipHostFree(p.h_in);
    (void)hipFree(p.d_in);
    (void)hipFree(p.d_thr);
    (void)hipFree(p.d_guide);
    p.d_thr = nullptr;
    p.d_guide = nullptr;
    p.cap_guide_states = 0;
    (void)hipFree(p.d_hist_acc);
    (void)hipFree(p.d_out);
    (void)hipHostFree(p.h_out);
    (void)hipFree(p.d_replay);
    (void)hipFree(p.d_replay_count);
    (void)hipFree(p.d_work);
    if (p.ev0) (void)hipEventDestroy(p.ev0);
    if (p.ev_mid) (void)hipEventDestroy(p.ev_mid);
    if (p.ev1) (void)hipEventDestroy(p.ev1);
    p = Population();
}

}  // namespace pda

using namespace pda;

namespace {

constexpr int kMaxSets = 1 << 16;
// Largest threshold table a population keeps for the fast form's guided
// binomial draws (134 KB per state: 1 GiB covers about 8,000 states); a
// larger population draws by the inversion loop.
constexpr size_t kPopGuideBudget = size_t(1) << 30;

size_t in_bytes(size_t nsets) { return sizeof(PopControl) + sizeof(ModelBlob) * nsets; }

const ModelBlob* blobs_of(const uint8_t* base) {
    return reinterpret_cast<const ModelBlob*>(base + sizeof(PopControl));
}

size_t out_bytes(size_t nsets, int B) {
    return sizeof(PopStatus) + sizeof(SetResult) * nsets + sizeof(int32_t) * nsets * size_t(B);
}

// Buffers that depend only on the population size and the bin count.
int ensure_pop_buffers(pda_ctx* ctx, Slot& s, int B) {
    Population& p = s.pop;
    if (!p.ready) {
        // each once: a call that fails part way leaves the rest for the next
        if (!p.d_work) HIP_TRY(hipMalloc(&p.d_work, sizeof(uint32_t) * kWorkWords));
        HIP_TRY(hipMemset(p.d_work, 0, sizeof(uint32_t) * kWorkWords));
        if (!p.d_replay_count) HIP_TRY(hipMalloc(&p.d_replay_count, sizeof(uint32_t)));
        HIP_TRY(hipMemset(p.d_replay_count, 0, sizeof(uint32_t)));
        for (hipEvent_t* e : {&p.ev0, &p.ev_mid, &p.ev1})
            if (!*e) HIP_TRY(hipEventCreate(e));
        p.ready = true;
    }
    // histogram rows: [set][XCD][B] for the run's B; pda_set_population frees
    // them when the population outgrows them
    if (!p.d_hist_acc || B > p.cap_B) {
        drop_graph(p);
        (void)hipFree(p.d_hist_acc);
        p.d_hist_acc = nullptr;
        const int capB = std::max(B, p.cap_B);
        const size_t bytes = sizeof(int32_t) * p.cap_sets * 8u * size_t(capB);
        HIP_TRY(hipMalloc(&p.d_hist_acc, bytes));
        HIP_TRY(hipMemset(p.d_hist_acc, 0, bytes));
        HIP_TRY(hipDeviceSynchronize());
        p.cap_B = capB;
    }
    const size_t ob = out_bytes(p.cap_sets, p.cap_B);
    if (ob > p.out_bytes) {
        drop_graph(p);
        (void)hipFree(p.d_out);
        (void)hipHostFree(p.h_out);
        p.d_out = p.h_out = nullptr;
        p.out_bytes = 0;
        HIP_TRY(hipMalloc(&p.d_out, ob));
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&p.h_out), ob, hipHostMallocDefault));
        p.out_bytes = ob;
    }
    (void)ctx;
    return PDA_OK;
}

// Enqueues one population evaluation on the slot's stream (captured or not).
int enqueue_population(pda_ctx* ctx, Slot& s, int pass, uint32_t grid, const RunArgs& a) {
    Population& p = s.pop;
    const PopControl* ctl = reinterpret_cast<const PopControl*>(p.d_in);
    const ModelBlob* blobs = blobs_of(p.d_in);
    PopStatus* status = reinterpret_cast<PopStatus*>(p.d_out);
    SetResult* results = reinterpret_cast<SetResult*>(p.d_out + sizeof(PopStatus));
    int32_t* hists = reinterpret_cast<int32_t*>(p.d_out + sizeof(PopStatus) +
                                                sizeof(SetResult) * size_t(p.nsets));
    HIP_TRY(hipMemcpyAsync(p.d_in, p.h_in, in_bytes(size_t(p.nsets)), hipMemcpyHostToDevice,
                           s.stream));
    BinomGuide guide{nullptr, nullptr};
    if (pass == 0 && p.guided) {
        HIP_TRY(launch_guide(blobs, p.nsets, p.max_states, p.d_thr, p.d_guide, s.stream));
        guide = BinomGuide{p.d_thr, p.d_guide};
    }
    HIP_TRY(launch_pop_walk(pass, grid, ctx->d_bursts, ctx->d_ticks, ctx->d_links, ctl, blobs,
                            guide, a, p.d_hist_acc, p.d_replay, p.d_replay_count, p.d_work,
                            s.stream));
    if (pass == 0)
        HIP_TRY(launch_pop_replay(ctx->d_bursts, ctx->d_ticks, ctx->d_links, ctl, blobs, a,
                                  p.d_hist_acc, p.d_replay, p.d_replay_count,
                                  uint32_t(ctx->num_cu) * 4u, s.stream));
    HIP_TRY(launch_pop_finish(blobs, a, ctx->d_obs, p.d_hist_acc, p.d_replay_count, p.d_work,
                              status, results, hists, s.stream));
    HIP_TRY(hipMemcpyAsync(p.h_out, p.d_out, out_bytes(size_t(p.nsets), a.B),
                           hipMemcpyDeviceToHost, s.stream));
    return PDA_OK;
}

int run_population(pda_ctx* ctx, Slot& s, int R, uint32_t seed, uint32_t set_base,
                   uint32_t set_stride, int flags, bool force_general) {
    Population& p = s.pop;
    if (p.nsets < 1) return fail(PDA_E_STATE, "no population set");
    if (R < 1) return fail(PDA_E_ARG, "R = %d < 1", R);
    if (ctx->B < 1 || !ctx->d_obs) return fail(PDA_E_STATE, "no observed histogram set");
    HIP_TRY(hipSetDevice(ctx->device));
    int64_t b0, b1;
    range_of(ctx, s, &b0, &b1);
    const int64_t units = (b1 - b0) * int64_t(R);
    const int64_t total = units * int64_t(p.nsets);
    if (units > int64_t(0x7FFFFFFF) - kWave)
        return fail(PDA_E_ARG, "%lld bursts x R = %d exceeds 2^31 units per set",
                    (long long)(b1 - b0), R);
    if (total > (int64_t(1) << 40)) return fail(PDA_E_ARG, "population too large");
    int rc = ensure_pop_buffers(ctx, s, ctx->B);
    if (rc) return rc;
    const int pass = (p.any_v || force_general) ? 2 : 0;
    Launch L = plan_launch(ctx, pass, true, uint64_t(total));
    uint64_t bps = units > 0 ? (uint64_t(units) + L.batch - 1) / L.batch : 0;
    if (bps * uint64_t(p.nsets) >= (uint64_t(1) << 31)) {
        // too many batches for the 32-bit counters: larger batches
        L.batch = uint32_t(((uint64_t(units) * p.nsets >> 30) + 1) * kWave);
        bps = (uint64_t(units) + L.batch - 1) / L.batch;
    }
    // replay list: abandoned units are rare (BTPE draws, restarts, large
    // background means); an overflow is detected and the run repeated in the
    // general form
    const size_t want_cap = size_t(std::min<int64_t>(
        total, std::max<int64_t>(int64_t(1) << 20, total / 16)));
    if (pass == 0 && want_cap > p.replay_cap) {
        HIP_TRY(hipStreamSynchronize(s.stream));
        drop_graph(p);
        (void)hipFree(p.d_replay);
        p.d_replay = nullptr;
        p.replay_cap = 0;
        HIP_TRY(hipMalloc(&p.d_replay, sizeof(uint64_t) * want_cap));
        p.replay_cap = want_cap;
    }
    RunArgs a{};
    a.begin = uint32_t(b0);
    a.nunits = uint32_t(units);
    a.R = uint32_t(R);
    a.seed = seed;
    a.set = set_base;
    a.batch = L.batch;
    a.queues = L.queues;
    a.B = ctx->B;
    a.left_pad = ctx->left_pad;
    a.bg_on = ctx->bg_dd < 1e-4f ? 0 : 1;  // src/mc.cu:153
    a.clk_p = ctx->clk_p;
    a.bg_ad = ctx->bg_ad;
    a.bg_dd = ctx->bg_dd;
    a.nstates = 0;
    a.replay_cap = uint32_t(std::min<size_t>(p.replay_cap, 0xFFFFFFFFu));
    a.nsets = uint32_t(p.nsets);
    a.bps = uint32_t(bps);
    const uint32_t grid = units > 0 ? L.grid : 1u;

    // the generation's stream keys travel with its models
    if (p.in_flight) HIP_TRY(hipStreamSynchronize(s.stream));
    PopControl* hc = reinterpret_cast<PopControl*>(p.h_in);
    hc->seed = seed;
    hc->set_base = set_base;
    hc->set_stride = set_stride;
    hc->nsets = uint32_t(p.nsets);

    const bool use_graph = !(flags & PDA_RUN_NO_GRAPH);
    if (use_graph) {
        PopGraphKey k;
        k.nsets = p.nsets;
        k.R = R;
        k.B = a.B;
        k.left_pad = a.left_pad;
        k.pass = pass;
        k.bg_on = a.bg_on;
        k.b0 = b0;
        k.b1 = b1;
        k.grid = grid;
        k.batch = a.batch;
        k.queues = a.queues;
        k.replay_cap = a.replay_cap;
        k.clk_p = a.clk_p;
        k.bg_ad = a.bg_ad;
        k.bg_dd = a.bg_dd;
        k.bursts = ctx->d_bursts;
        k.ticks = ctx->d_ticks;
        k.links = ctx->d_links;
        k.data_gen = ctx->data_gen;
        k.obs = ctx->d_obs;
        k.guided = pass == 0 && p.guided;
        k.max_states = p.max_states;
        k.thr = p.d_thr;
        if (!p.exec || !(k == p.key)) {
            drop_graph(p);
            HIP_TRY(hipStreamSynchronize(s.stream));
            HIP_TRY(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_population(ctx, s, pass, grid, a);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(s.stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            HIP_TRY(ec);
            p.graph = g;
            HIP_TRY(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0));
            p.key = k;
            ++p.captures;
        }
        HIP_TRY(hipEventRecord(p.ev0, s.stream));
        HIP_TRY(hipGraphLaunch(p.exec, s.stream));
        HIP_TRY(hipEventRecord(p.ev1, s.stream));
    } else {
        HIP_TRY(hipEventRecord(p.ev0, s.stream));
        rc = enqueue_population(ctx, s, pass, grid, a);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(p.ev1, s.stream));
    }
    p.ran = true;
    p.in_flight = true;
    p.last_B = a.B;
    p.last_R = R;
    p.last_left_pad = a.left_pad;
    p.last_b0 = b0;
    p.last_b1 = b1;
    p.last_units = units;
    p.last_seed = seed;
    p.last_base = set_base;
    p.last_stride = set_stride;
    p.last_graph = use_graph;
    return PDA_OK;
}

}  // namespace

extern "C" {

int pda_set_population(pda_ctx* ctx, int sid, int nsets, const int32_t* nstates,
                       const float* args, int64_t nargs_total) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (nsets < 1 || nsets > kMaxSets) return fail(PDA_E_ARG, "nsets %d outside [1, %d]", nsets, kMaxSets);
    if (!nstates || !args) return fail(PDA_E_ARG, "null argument");
    std::vector<int64_t> off(static_cast<size_t>(nsets));
    int64_t need = 0;
    for (int i = 0; i < nsets; ++i) {
        const int n = nstates[i];
        if (n < 2 || n > kMaxStates)
            return fail(PDA_E_ARG, "set %d: n = %d outside [2, %d]", i, n, kMaxStates);
        off[size_t(i)] = need;
        need += int64_t(n) * (n + 1);
    }
    if (need != nargs_total)
        return fail(PDA_E_ARG, "args hold %lld floats, the sets need %lld",
                    (long long)nargs_total, (long long)need);
    HIP_TRY(hipSetDevice(ctx->device));
    Population& p = s->pop;
    // validate and build every set before touching the slot
    std::vector<Model> models(static_cast<size_t>(nsets));
    bool any_v = false;
    for (int i = 0; i < nsets; ++i) {
        std::string err;
        const int n = nstates[i];
        const int rc = build_model(s->ph, n, args + off[size_t(i)], n * (n + 1), models[size_t(i)], err);
        if (rc) return fail(rc, "set %d: %s", i, err.c_str());
        any_v = any_v || models[size_t(i)].any_v != 0;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    p.in_flight = false;
    if (size_t(nsets) > p.cap_sets) {
        drop_graph(p);
        (void)hipHostFree(p.h_in);
        (void)hipFree(p.d_in);
        (void)hipFree(p.d_hist_acc);
        p.h_in = p.d_in = nullptr;
        p.d_hist_acc = nullptr;
        p.cap_sets = 0;
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&p.h_in), in_bytes(size_t(nsets)),
                              hipHostMallocDefault));
        HIP_TRY(hipMalloc(&p.d_in, in_bytes(size_t(nsets))));
        p.cap_sets = size_t(nsets);
        std::memset(p.h_in, 0, in_bytes(size_t(nsets)));
    }
    // threshold rows of the fast form's binomial draws: set i's states take
    // rows guide_base .. guide_base + n_i - 1 of one table
    size_t states = 0;
    int max_states = 0;
    for (int i = 0; i < nsets; ++i) {
        models[size_t(i)].guide_base = int32_t(states);
        states += size_t(nstates[i]);
        max_states = std::max(max_states, nstates[i]);
    }
    const bool guided = !any_v && states * kGuideStatePairs * sizeof(ThrPair) <= kPopGuideBudget;
    if (guided && states > p.cap_guide_states) {
        drop_graph(p);
        (void)hipFree(p.d_thr);
        (void)hipFree(p.d_guide);
        p.d_thr = nullptr;
        p.d_guide = nullptr;
        p.cap_guide_states = 0;
        HIP_TRY(hipMalloc(&p.d_thr, states * kGuideStatePairs * sizeof(ThrPair)));
        HIP_TRY(hipMalloc(&p.d_guide, states * size_t(kBinomTable) * kGuideBuckets));
        p.cap_guide_states = states;
    }
    p.guided = guided;
    p.max_states = max_states;
    ModelBlob* hb = reinterpret_cast<ModelBlob*>(p.h_in + sizeof(PopControl));
    for (int i = 0; i < nsets; ++i) {
        hb[i].m = models[size_t(i)];
        build_binom_table(models[size_t(i)], hb[i].tab);
    }
    p.nsets = nsets;
    p.n.assign(nstates, nstates + nsets);
    p.args.assign(args, args + nargs_total);
    p.off = off;
    p.any_v = any_v;
    return PDA_OK;
}

int pda_run_population(pda_ctx* ctx, int sid, int R, uint32_t seed, uint32_t set_base,
                       uint32_t set_stride, int flags) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    return run_population(ctx, *s, R, seed, set_base, set_stride, flags, false);
}

int pda_get_population_result(pda_ctx* ctx, int sid, int32_t* hist, int nbins, double* chi2_64,
                              float* chi2_ref, int32_t* effN) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    Population& p = s->pop;
    if (!p.ran) return fail(PDA_E_STATE, "stream %d has no population run", sid);
    if (hist && nbins != p.last_B) return fail(PDA_E_ARG, "nbins %d != run's %d", nbins, p.last_B);
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    p.in_flight = false;
    const PopStatus* st = reinterpret_cast<const PopStatus*>(p.h_out);
    if (st->overflow) {
        // more abandoned units than the replay list holds: evaluate again in
        // the general form, which abandons nothing (same keyed draws) -- over
        // the same bursts and bins as the run being repeated
        int64_t b0, b1;
        range_of(ctx, *s, &b0, &b1);
        if (b0 != p.last_b0 || b1 != p.last_b1 || ctx->B != p.last_B ||
            ctx->left_pad != p.last_left_pad)
            return fail(PDA_E_STATE,
                        "stream %d: the population run overflowed its replay list and the burst "
                        "range or bins changed since; run it again", sid);
        int rc = run_population(ctx, *s, p.last_R, p.last_seed, p.last_base, p.last_stride,
                                PDA_RUN_NO_GRAPH, true);
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(s->stream));
        p.in_flight = false;
    }
    const SetResult* res = reinterpret_cast<const SetResult*>(p.h_out + sizeof(PopStatus));
    const int32_t* hs = reinterpret_cast<const int32_t*>(p.h_out + sizeof(PopStatus) +
                                                         sizeof(SetResult) * size_t(p.nsets));
    for (int i = 0; i < p.nsets; ++i) {
        if (chi2_64) chi2_64[i] = res[i].chi2_64;
        if (chi2_ref) chi2_ref[i] = res[i].chi2_ref;
        if (effN) effN[i] = res[i].effN;
    }
    if (hist) std::memcpy(hist, hs, sizeof(int32_t) * size_t(p.nsets) * size_t(nbins));
    return PDA_OK;
}

int pda_population_stats(pda_ctx* ctx, int sid, int64_t* units_per_set, int32_t* replays,
                         int32_t* graph_captures, float* ms) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    Population& p = s->pop;
    if (!p.ran) return fail(PDA_E_STATE, "stream %d has no population run", sid);
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    p.in_flight = false;
    if (units_per_set) *units_per_set = p.last_units;
    if (replays) *replays = int32_t(reinterpret_cast<const PopStatus*>(p.h_out)->replays);
    if (graph_captures) *graph_captures = p.captures;
    if (ms) HIP_TRY(hipEventElapsedTime(ms, p.ev0, p.ev1));
    return PDA_OK;
}

int pda_set_launch_shape(pda_ctx* ctx, int batch_tiles, int waves_per_cu, int queues) {
    if (!ctx) return fail(PDA_E_ARG, "null context");
    if (batch_tiles < 0 || batch_tiles > 1024)
        return fail(PDA_E_ARG, "batch_tiles %d outside [0, 1024]", batch_tiles);
    if (waves_per_cu < 0 || waves_per_cu > 64)
        return fail(PDA_E_ARG, "waves_per_cu %d outside [0, 64]", waves_per_cu);
    if (queues < 0 || queues > int(kMaxQueues))
        return fail(PDA_E_ARG, "queues %d outside [0, %u]", queues, kMaxQueues);
    int rc = sync_all(ctx);
    if (rc) return rc;
    ctx->shape.batch_tiles = uint32_t(batch_tiles);
    ctx->shape.waves_per_cu = uint32_t(waves_per_cu);
    ctx->shape.queues = uint32_t(queues);
    return PDA_OK;
}

int pda_get_launch_shape(const pda_ctx* ctx, int* batch_tiles, int* waves_per_cu, int* queues) {
    if (!ctx) return fail(PDA_E_ARG, "null context");
    if (batch_tiles) *batch_tiles = int(ctx->shape.batch_tiles);
    if (waves_per_cu) *waves_per_cu = int(ctx->shape.waves_per_cu);
    if (queues) *queues = int(ctx->shape.queues);
    return PDA_OK;
}

// Coordinate search over the launch shape on the slot's current work (its
// single set, or its population): batch size, then resident waves per CU,
// then the number of work counters; each candidate is run `reps` times after
// one untimed run and scored by its median HIP-event time.  Results do not
// depend on the shape (every unit's draws are keyed), only the time does.
int pda_autotune(pda_ctx* ctx, int sid, int population, int R, int reps, int* batch_tiles,
                 int* waves_per_cu, int* queues, float* best_ms) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (reps < 1 || reps > 100) return fail(PDA_E_ARG, "reps %d outside [1, 100]", reps);
    int rc = sync_all(ctx);
    if (rc) return rc;
    int per_cu = 0;
    const int pass = population ? (s->pop.any_v ? 2 : 0) : (s->any_v ? 2 : 0);
    HIP_TRY(population ? pop_walk_occupancy(pass, ctx->B, &per_cu) : walk_occupancy(pass, &per_cu));
    if (per_cu < 1) per_cu = 1;
    const LaunchShape saved = ctx->shape;
    auto time_shape = [&](const LaunchShape& sh, float* out) -> int {
        ctx->shape = sh;
        std::vector<float> t;
        for (int i = 0; i <= reps; ++i) {
            float ms = 0.f;
            if (population) {
                int r2 = run_population(ctx, *s, R, 0u, 0u, 1u, PDA_RUN_NO_GRAPH, false);
                if (r2) return r2;
                HIP_TRY(hipStreamSynchronize(s->stream));
                s->pop.in_flight = false;
                HIP_TRY(hipEventElapsedTime(&ms, s->pop.ev0, s->pop.ev1));
            } else {
                int r2 = pda_run(ctx, sid, R, 0u, 0u, 0);
                if (r2) return r2;
                HIP_TRY(hipStreamSynchronize(s->stream));
                HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
            }
            if (i > 0) t.push_back(ms);
        }
        std::sort(t.begin(), t.end());
        *out = t[t.size() / 2];
        return PDA_OK;
    };
    LaunchShape best;
    float best_t = 0.f;
    rc = time_shape(best, &best_t);
    if (rc) {
        ctx->shape = saved;
        return rc;
    }
    const uint32_t tiles[] = {1, 2, 3, 4, 6, 8, 16, 32, 64};
    for (uint32_t bt : tiles) {
        LaunchShape c = best;
        c.batch_tiles = bt;
        float t = 0.f;
        rc = time_shape(c, &t);
        if (rc) break;
        if (t < best_t) {
            best_t = t;
            best = c;
        }
    }
    for (int w = per_cu - 1; !rc && w >= std::max(1, per_cu - 4); --w) {
        LaunchShape c = best;
        c.waves_per_cu = uint32_t(w);
        float t = 0.f;
        rc = time_shape(c, &t);
        if (rc) break;
        if (t < best_t) {
            best_t = t;
            best = c;
        }
    }
    const uint32_t qs[] = {8, 16, 64};
    for (uint32_t q : qs) {
        if (rc) break;
        LaunchShape c = best;
        c.queues = q;
        float t = 0.f;
        rc = time_shape(c, &t);
        if (rc) break;
        if (t < best_t) {
            best_t = t;
            best = c;
        }
    }
    if (rc) {
        ctx->shape = saved;
        return rc;
    }
    ctx->shape = best;
    if (batch_tiles) *batch_tiles = int(best.batch_tiles);
    if (waves_per_cu) *waves_per_cu = int(best.waves_per_cu);
    if (queues) *queues = int(best.queues);
    if (best_ms) *best_ms = best_t;
    return PDA_OK;
}

}  // extern "C"
