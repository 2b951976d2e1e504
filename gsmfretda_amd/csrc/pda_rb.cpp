// pda_rb.cpp -- host side of the Rao-Blackwellised (RB) PDA form: forward
// (expected histogram and chi2) and adjoint (d chi2 / d args) through the C ABI.
//
// No counterpart in the reference, whose simulator draws every acceptor count
// (src/mc.cu:167-170) and is therefore not differentiable.  The forward keeps
// the reference's walk and binning and replaces the acceptor draw by its
// exact distribution; the adjoint returns, for a cotangent g_b = d chi2 /
// d M_b supplied by the caller, the derivative in the efficiencies (exact for
// the RB estimator) and in the rates (likelihood-ratio estimator with a
// per-burst leave-one-out baseline, unbiased for the derivative of the
// expected histogram).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/gsmfretda_pda.h"
#include "pda_ctx.h"
#include "pda_layout.h"
#include "pda_rb.h"
#include "pda_types.h"

namespace pda {

void free_rb(RbState& r) {
    (void)hipFree(r.d_rbm);
    (void)hipHostFree(r.h_rbm);
    (void)hipFree(r.d_blob);
    (void)hipHostFree(r.h_blob);
    (void)hipFree(r.d_cnt);
    (void)hipFree(r.d_F);
    (void)hipFree(r.d_S);
    (void)hipFree(r.d_hist);
    (void)hipFree(r.d_work);
    (void)hipFree(r.d_res);
    (void)hipHostFree(r.h_res);
    (void)hipFree(r.d_g);
    (void)hipHostFree(r.h_g);
    (void)hipFree(r.d_grad);
    (void)hipFree(r.d_pmf_tab);
    (void)hipFree(r.d_jac);
    (void)hipHostFree(r.h_jac);
    (void)hipFree(r.d_big);
    (void)hipFree(r.d_huge);
    (void)hipFree(r.d_big_count);
    (void)hipFree(r.d_replay);
    (void)hipFree(r.d_replay_count);
    (void)hipFree(r.d_Sbar);
    (void)hipFree(r.d_poolP);
    (void)hipHostFree(r.h_grad);
    for (hipEvent_t e : {r.ev0, r.ev_walk, r.ev1, r.ev_a0, r.ev_a1})
        if (e) (void)hipEventDestroy(e);
    r = RbState();
}

void free_rb_memo(pda_ctx* ctx) {
    for (RbMemoLayout& L : ctx->rb_memo) {
        (void)hipFree(L.d_list);
        (void)hipFree(L.d_rank_offP);
        (void)hipFree(L.d_rank_offD);
        L = RbMemoLayout();
    }
}

}  // namespace pda

using namespace pda;

namespace {

struct RbResultHeaderH {  // device RbResultHeader
    double chi2;
    int32_t effN;
    int32_t B;
};
constexpr size_t kRbResBytes = sizeof(RbResultHeaderH) + sizeof(uint64_t) * kMaxBins;
constexpr int kGradLen = kRbMaxStates + kRbMaxPairs;
constexpr size_t kRbCounters = kWorkWords;  // queue counters + the pmf pass's

int ensure_rb(pda_ctx* ctx, RbState& r) {
    if (!r.ready) {
        // each buffer once: a call that fails part way (out of memory) leaves
        // what it got for the next call, which allocates the rest
        if (!r.d_rbm) HIP_TRY(hipMalloc(&r.d_rbm, sizeof(RbModel)));
        if (!r.h_rbm) HIP_TRY(hipHostMalloc(&r.h_rbm, sizeof(RbModel), hipHostMallocDefault));
        if (!r.d_hist) HIP_TRY(hipMalloc(&r.d_hist, sizeof(unsigned long long) * 8u * kMaxBins));
        HIP_TRY(hipMemset(r.d_hist, 0, sizeof(unsigned long long) * 8u * kMaxBins));
        if (!r.d_work) HIP_TRY(hipMalloc(&r.d_work, sizeof(uint32_t) * kRbCounters));
        if (!r.d_big_count) HIP_TRY(hipMalloc(&r.d_big_count, 2 * sizeof(uint32_t)));
        if (!r.d_replay_count) HIP_TRY(hipMalloc(&r.d_replay_count, sizeof(uint32_t)));
        HIP_TRY(hipMemset(r.d_work, 0, sizeof(uint32_t) * kRbCounters));
        if (!r.d_res) HIP_TRY(hipMalloc(&r.d_res, kRbResBytes));
        if (!r.h_res)
            HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&r.h_res), kRbResBytes,
                                  hipHostMallocDefault));
        if (!r.d_g) HIP_TRY(hipMalloc(&r.d_g, sizeof(double) * kMaxBins));
        if (!r.h_g)
            HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&r.h_g), sizeof(double) * kMaxBins,
                                  hipHostMallocDefault));
        if (!r.d_grad) HIP_TRY(hipMalloc(&r.d_grad, sizeof(double) * kGradLen));
        if (!r.d_pmf_tab)
            HIP_TRY(hipMalloc(&r.d_pmf_tab, sizeof(float) * size_t(kRbMaxStates) * kRbPmfStride));
        if (!r.d_jac) HIP_TRY(hipMalloc(&r.d_jac, sizeof(double) * kGradLen * kMaxBins));
        if (!r.h_jac)
            HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&r.h_jac),
                                  sizeof(double) * kGradLen * kMaxBins, hipHostMallocDefault));
        if (!r.h_grad)
            HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&r.h_grad), sizeof(double) * kGradLen,
                                  hipHostMallocDefault));
        for (hipEvent_t* e : {&r.ev0, &r.ev_walk, &r.ev1, &r.ev_a0, &r.ev_a1})
            if (!*e) HIP_TRY(hipEventCreate(e));
        HIP_TRY(hipDeviceSynchronize());
        r.ready = true;
    }
    if (!ctx->d_lf) {
        std::vector<double> lf(kRbMaxCount + 1);
        log_factorials(lf.data(), kRbMaxCount);
        HIP_TRY(hipMalloc(&ctx->d_lf, sizeof(double) * lf.size()));
        HIP_TRY(hipMemcpy(ctx->d_lf, lf.data(), sizeof(double) * lf.size(), hipMemcpyHostToDevice));
    }
    return PDA_OK;
}

void build_rb_model(const Model& M, RbModel& R, bool* rates_positive, bool* p_interior) {
    std::memset(&R, 0, sizeof R);
    const int n = M.n;
    R.n = n;
    R.K = n * (n - 1);
    R.coef = std::max({M.gmx, M.omdd, 0.0f});
    *rates_positive = true;
    *p_interior = true;
    for (int s = 0; s < n; ++s) {
        R.p[s] = M.ep0[s];
        const double p = double(M.ep0[s]);
        R.lp[s] = p > 0.0 ? dlog(p) : 0.0;
        R.lq[s] = p < 1.0 ? dlog(1.0 - p) : 0.0;
        if (p > 0.0 && p < 1.0) R.dscale[s] = 1.0 / (p * (1.0 - p));
        else *p_interior = false;
        float sum = 0.0f;
        for (int j = 0; j < n; ++j)
            if (j != s) sum += M.krate[s][j];
        R.inv_sum[s] = sum > 0.0f ? 1.0f / sum : 0.0f;
        for (int j = 0; j < n; ++j) {
            if (j == s) continue;
            R.pair[s][j] = rate_pair(n, s, j);
            const float k = M.krate[s][j];
            R.two_over_k[s][j] = k > 0.0f ? 2.0f / k : 0.0f;
            if (!(k > 0.0f)) *rates_positive = false;
        }
    }
}

// The memo layout of n states (built once per context and state count):
// the count vectors with sum < kRbMemoCap in lexicographic order are ranked
// 0, 1, ... (rb_rank); the build list holds them by total, and the pools
// follow the list, so that the blocks written by one wave of rb_memo_kernel
// (64 consecutive list entries) are adjacent: P blocks of Mt + 1 floats in
// list order.  A vector's leave-one-out pmfs are the blocks of the vectors
// with one trial fewer of each counted state (Mt - 1 < kRbMemoCap: in the
// pool too); their offsets are kept per vector.
int ensure_memo_layout(pda_ctx* ctx, int n) {
    RbMemoLayout& L = ctx->rb_memo[n];
    if (L.d_list) return PDA_OK;
    const int T = kRbMemoCap - 1;
    struct Vec {
        int m[kRbMaxStates];
        int Mt;
        uint32_t rank;
    };
    std::vector<Vec> v;
    int m[kRbMaxStates] = {0, 0, 0, 0};
    // lexicographic enumeration of (m_0 .. m_{n-1}) with sum <= T
    std::function<void(int, int)> rec = [&](int i, int left) {
        if (i == n) {
            Vec e{};
            int Mt = 0;
            for (int s = 0; s < kRbMaxStates; ++s) {
                e.m[s] = s < n ? m[s] : 0;
                Mt += e.m[s];
            }
            e.Mt = Mt;
            e.rank = uint32_t(v.size());
            v.push_back(e);
            return;
        }
        for (int x = 0; x <= left; ++x) {
            m[i] = x;
            rec(i + 1, left - x);
        }
        m[i] = 0;
    };
    rec(0, T);
    for (const Vec& e : v)
        if (rb_rank(e.m, n, T) != e.rank)
            return fail(PDA_E_STATE, "RB memo: rank of a count vector disagrees with the enumeration");
    std::stable_sort(v.begin(), v.end(), [](const Vec& x, const Vec& y) { return x.Mt < y.Mt; });
    const size_t nv = v.size();
    std::vector<uint4> list(nv);
    std::vector<uint32_t> rank_offP(nv);
    std::vector<uint4> rank_offD(nv, make_uint4(0u, 0u, 0u, 0u));
    uint64_t offP = 0;
    for (size_t i = 0; i < nv; ++i) {
        uint32_t cnt[kRbMaxStates];
        for (int s = 0; s < kRbMaxStates; ++s) cnt[s] = uint32_t(v[i].m[s]);
        const uint2 c = rb_pack(cnt);
        list[i] = make_uint4(c.x, c.y, uint32_t(offP), 0u);
        rank_offP[v[i].rank] = uint32_t(offP);
        offP += uint64_t(v[i].Mt) + 1u;
    }
    // the pmf with one trial of counted state s removed is the pmf of the
    // count vector with m_s - 1 (the same convolutions, one trial fewer):
    // the derivative passes read it from the pool at that vector's offset
    for (const Vec& e : v) {
        uint32_t* d = &rank_offD[e.rank].x;
        int j = 0;  // slot: counted states before s
        for (int s = 0; s < n; ++s) {
            if (e.m[s] == 0) continue;
            int mm[kRbMaxStates];
            for (int t = 0; t < kRbMaxStates; ++t) mm[t] = e.m[t];
            mm[s] -= 1;
            d[j++] = rank_offP[rb_rank(mm, n, T)];
        }
    }
    if (offP >= (uint64_t(1) << 32)) return fail(PDA_E_STATE, "RB memo pool exceeds 2^32 entries");
    auto upload = [&](auto** dst, const auto& src) -> int {
        using T = std::remove_reference_t<decltype(src[0])>;
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(dst), sizeof(T) * src.size()));
        HIP_TRY(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
        return PDA_OK;
    };
    RbMemoLayout out;
    int rc = upload(&out.d_list, list);
    if (!rc) rc = upload(&out.d_rank_offP, rank_offP);
    if (!rc) rc = upload(&out.d_rank_offD, rank_offD);
    if (rc) {
        (void)hipFree(out.d_list);
        (void)hipFree(out.d_rank_offP);
        (void)hipFree(out.d_rank_offD);
        return rc;
    }
    out.nvec = uint32_t(nv);
    out.poolP = offP;
    L = out;
    return PDA_OK;
}

// Launch shape of one pmf pass (mode 0 forward, 1 adjoint, 2 Jacobian):
// rb_unit_kernel on every slot its LDS allows (at most enough workgroups for
// the units), about eight grabs per wave; rb_big_kernel on every slot its
// LDS allows (its workgroups beyond the listed units exit at once).
struct PassShape {
    uint32_t unit_grid = 1, batch = kWave, big_grid = 1, huge_grid = 1;
    int big_cap = 1;    // rb_big_kernel's buffers for the big list (the huge list: the run's cap)
    bool wide = false;  // rb_big_kernel's derivative buffers in double
};
int pass_shape(pda_ctx* ctx, int mode, int n, bool wide, int B, int cap, uint32_t nunits,
               PassShape& ps) {
    static const char* names[] = {"forward", "adjoint", "Jacobian"};
    constexpr size_t kStatic = sizeof(RbModel) + 64;  // the kernels' static LDS
    const size_t lds = rb_unit_lds_bytes(B, mode, n);
    if (lds + kStatic > ctx->lds_per_block)
        return fail(PDA_E_DATA, "%d bins exceed the LDS of the RB %s pass", B, names[mode]);
    int per = 0;
    HIP_TRY(rb_unit_occupancy(mode, n, int(lds), &per));
    const uint64_t threads = uint64_t(rb_unit_threads());
    const uint64_t g = std::min<uint64_t>(uint64_t(ctx->num_cu) * uint64_t(std::max(per, 1)),
                                          (uint64_t(nunits) + threads - 1) / threads);
    ps.unit_grid = uint32_t(std::max<uint64_t>(g, 1));
    uint64_t bt = uint64_t(nunits) / (uint64_t(ps.unit_grid) * (threads / kWave) * 8u);
    bt = (bt + kWave - 1) / kWave * kWave;
    ps.batch = uint32_t(std::min<uint64_t>(std::max<uint64_t>(bt, kWave), 4096));
    // double buffers where asked for and they fit, else float
    ps.wide = wide && rb_big_lds_bytes(B, cap, mode, true) + kStatic <= ctx->lds_per_block;
    const size_t blds = rb_big_lds_bytes(B, cap, mode, ps.wide);
    if (blds + kStatic > ctx->lds_per_block)
        return fail(PDA_E_DATA, "%d bins and bursts needing %d acceptor outcomes exceed the LDS "
                    "of the RB %s pass", B, cap, names[mode]);
    per = 0;
    HIP_TRY(rb_big_occupancy(mode, ps.wide, int(blds), &per));
    ps.huge_grid = uint32_t(std::max(1, ctx->num_cu * std::max(per, 1)));
    ps.big_cap = std::min(cap, kRbBigSplit);
    per = 0;
    HIP_TRY(rb_big_occupancy(mode, ps.wide, int(rb_big_lds_bytes(B, ps.big_cap, mode, ps.wide)),
                             &per));
    ps.big_grid = uint32_t(std::max(1, ctx->num_cu * std::max(per, 1)));
    return PDA_OK;
}

// Enqueues one RB forward of model M -- the walk reading M from the device
// blob d_blob -- over the slot's burst range into state r, on the slot's
// stream.  sync_model: first wait for the stream, whose last run may still be
// copying r's pinned model (a population forward waits once for all sets).
int rb_enqueue_forward(pda_ctx* ctx, Slot& s, RbState& r, const Model& M,
                       const ModelBlob* d_blob, int R, uint32_t seed, uint32_t set_id, int flags,
                       bool sync_model) {
    int rc = ensure_rb(ctx, r);
    if (rc) return rc;
    const int n = M.n;
    int64_t b0, b1;
    range_of(ctx, s, &b0, &b1);
    const int64_t units = (b1 - b0) * int64_t(R);
    if (units > int64_t(0x7FFFFFFF) - kWave)
        return fail(PDA_E_ARG, "%lld bursts x R = %d exceeds 2^31 units", (long long)(b1 - b0), R);
    // acceptor totals per unit: m_w <= floor(c (photons of the window) + 1/2)
    // with c = max(gamma - crosstalk, 1 - dexdir); at most the burst's photons
    // when c <= 1
    uint32_t maxP = 0;
    for (int64_t b = b0; b < b1; ++b) maxP = std::max(maxP, ctx->h_P[size_t(b)]);
    const float coef = std::max({M.gmx, M.omdd, 0.0f});
    const int bound = rb_total_bound(maxP, coef);
    if (bound > kRbMaxCount)
        return fail(PDA_E_DATA, "the RB form supports acceptor totals up to %d (bursts of up to "
                    "%u photons need %d)", kRbMaxCount, maxP, bound);
    const int cap = bound + 1;  // rb_big_kernel's pmf buffers
    const bool scores = (flags & PDA_RB_SCORES) != 0;
    if (rb_big_lds_bytes(ctx->B, cap, 0, false) + sizeof(RbModel) > ctx->lds_per_block)
        return fail(PDA_E_DATA, "bursts of %u photons exceed the LDS of the RB kernels", maxP);

    // buffers of the per-unit summaries and the walk's list of long units
    if (size_t(units) > r.cap_units) {
        HIP_TRY(hipStreamSynchronize(s.stream));
        (void)hipFree(r.d_cnt);
        (void)hipFree(r.d_F);
        (void)hipFree(r.d_big);
        (void)hipFree(r.d_huge);
        (void)hipFree(r.d_replay);
        r.d_cnt = nullptr;
        r.d_F = nullptr;
        r.d_big = nullptr;
        r.d_huge = nullptr;
        r.d_replay = nullptr;
        r.cap_units = 0;
        const size_t nu = size_t(std::max<int64_t>(units, 1));
        HIP_TRY(hipMalloc(&r.d_cnt, sizeof(uint2) * nu));
        HIP_TRY(hipMalloc(&r.d_F, sizeof(float) * nu));
        HIP_TRY(hipMalloc(&r.d_big, sizeof(uint32_t) * nu));
        HIP_TRY(hipMalloc(&r.d_huge, sizeof(uint32_t) * nu));
        HIP_TRY(hipMalloc(&r.d_replay, sizeof(uint32_t) * nu));
        r.cap_units = size_t(units);
    }
    const size_t nsc = scores ? size_t(units) * size_t(n * (n - 1)) : 0;
    if (nsc > r.cap_scores) {
        HIP_TRY(hipStreamSynchronize(s.stream));
        (void)hipFree(r.d_S);
        r.d_S = nullptr;
        r.cap_scores = 0;
        HIP_TRY(hipMalloc(&r.d_S, sizeof(float) * nsc));
        r.cap_scores = nsc;
    }
    const bool sbar = scores && R >= 2;  // the derivative passes need R >= 2
    const size_t nsb = sbar ? size_t(b1 - b0) * size_t(n * (n - 1)) : 0;
    if (nsb > r.cap_sbar) {
        HIP_TRY(hipStreamSynchronize(s.stream));
        (void)hipFree(r.d_Sbar);
        r.d_Sbar = nullptr;
        r.cap_sbar = 0;
        HIP_TRY(hipMalloc(&r.d_Sbar, sizeof(double) * nsb));
        r.cap_sbar = nsb;
    }
    if (sync_model) HIP_TRY(hipStreamSynchronize(s.stream));  // the pinned model may be in use
    RbModel* hm = static_cast<RbModel*>(r.h_rbm);
    bool rates_positive = false, p_interior = false;
    build_rb_model(M, *hm, &rates_positive, &p_interior);
    bool wide = false;  // a state whose leave-one-out recurrence barely damps errors
    for (int st = 0; st < n; ++st) wide = wide || (hm->p[st] > 0.4f && hm->p[st] < 0.6f);

    RunArgs a{};
    a.begin = uint32_t(b0);
    a.nunits = uint32_t(units);
    a.R = uint32_t(R);
    a.seed = seed;
    a.set = set_id;
    a.B = ctx->B;
    a.left_pad = ctx->left_pad;
    a.bg_on = ctx->bg_dd < 1e-4f ? 0 : 1;  // src/mc.cu:153
    a.clk_p = ctx->clk_p;
    a.bg_ad = ctx->bg_ad;
    a.bg_dd = ctx->bg_dd;
    a.nstates = n;
    a.nsets = 1;
    // walk grid: persistent, batches of about six grabs per wave
    int per_cu = 0;
    HIP_TRY(rb_walk_occupancy(scores, n, &per_cu));
    per_cu = std::max(per_cu, 1);
    if (ctx->shape.waves_per_cu > 0) per_cu = std::min(per_cu, int(ctx->shape.waves_per_cu));
    const uint64_t max_grid = uint64_t(ctx->num_cu) * uint64_t(per_cu);
    uint64_t batch = uint64_t(kWave) * ctx->shape.batch_tiles;
    if (ctx->shape.batch_tiles == 0) {
        batch = (uint64_t(units) / (max_grid * 6u) + kWave - 1) / kWave * kWave;
        batch = std::min<uint64_t>(std::max<uint64_t>(batch, kWave), 4096);
    }
    const uint32_t grid = uint32_t(std::max<uint64_t>(1, std::min(max_grid, (uint64_t(units) + batch - 1) / batch)));
    a.batch = uint32_t(batch);
    a.queues = std::min<uint32_t>(ctx->shape.queues ? ctx->shape.queues : kDefaultQueues, grid);
    PassShape ps;
    if (int rc = pass_shape(ctx, 0, n, false, a.B, cap, a.nunits, ps)) return rc;
    RunArgs la = a;
    la.batch = ps.batch;
    // the memo of this set: every count vector's pmf
    if (int rc = ensure_memo_layout(ctx, n)) return rc;
    const RbMemoLayout& ML = ctx->rb_memo[n];
    if (rb_memo_lds_bytes(n) + 512 > ctx->lds_per_block)
        return fail(PDA_E_STATE, "the RB memo kernel exceeds the LDS of a workgroup");
    if (ML.poolP > r.cap_poolP) {
        HIP_TRY(hipStreamSynchronize(s.stream));
        (void)hipFree(r.d_poolP);
        r.d_poolP = nullptr;
        r.cap_poolP = 0;
        HIP_TRY(hipMalloc(&r.d_poolP, sizeof(float) * ML.poolP));
        r.cap_poolP = ML.poolP;
    }

    if (!s.rb_aux) HIP_TRY(hipStreamCreateWithFlags(&s.rb_aux, hipStreamNonBlocking));
    if (!s.rb_fork) HIP_TRY(hipEventCreateWithFlags(&s.rb_fork, hipEventDisableTiming));
    if (!s.rb_join) HIP_TRY(hipEventCreateWithFlags(&s.rb_join, hipEventDisableTiming));
    HIP_TRY(hipMemsetAsync(r.d_work, 0, sizeof(uint32_t) * kRbCounters, s.stream));
    HIP_TRY(hipMemsetAsync(r.d_big_count, 0, 2 * sizeof(uint32_t), s.stream));
    HIP_TRY(hipMemsetAsync(r.d_replay_count, 0, sizeof(uint32_t), s.stream));
    HIP_TRY(hipMemcpyAsync(r.d_rbm, r.h_rbm, sizeof(RbModel), hipMemcpyHostToDevice, s.stream));
    const RbModel* d_rbm = static_cast<const RbModel*>(r.d_rbm);
    HIP_TRY(hipEventRecord(r.ev0, s.stream));
    // the pmf table and memo depend on the model only: built on the side
    // stream while the walk runs (the memo's few, long waves hide under it)
    HIP_TRY(hipEventRecord(s.rb_fork, s.stream));
    HIP_TRY(hipStreamWaitEvent(s.rb_aux, s.rb_fork, 0));
    HIP_TRY(launch_rb_table(d_rbm, ctx->d_lf, r.d_pmf_tab, n, s.rb_aux));
    HIP_TRY(launch_rb_memo(n, d_rbm, r.d_pmf_tab, ML.d_list, ML.nvec, r.d_poolP, s.rb_aux));
    HIP_TRY(hipEventRecord(s.rb_join, s.rb_aux));
    HIP_TRY(launch_rb_walk(scores, n, grid, uint32_t(ctx->num_cu) * 4u, ctx->d_bursts, ctx->d_ticks,
                           ctx->d_links, d_blob, d_rbm, a, r.d_cnt, r.d_F, r.d_S, r.d_work,
                           r.d_big, r.d_huge, r.d_big_count, r.d_replay, r.d_replay_count,
                           s.stream));
    HIP_TRY(hipEventRecord(r.ev_walk, s.stream));
    HIP_TRY(hipStreamWaitEvent(s.stream, s.rb_join, 0));
    HIP_TRY(launch_rb_unit(0, n, ps.unit_grid, d_rbm, la, r.d_cnt, r.d_F, nullptr, nullptr, nullptr,
                           ML.d_rank_offP, ML.d_rank_offD, r.d_poolP, r.d_hist,
                           nullptr, r.d_work, s.stream));
    HIP_TRY(launch_rb_big(0, false, ps.big_grid, ps.big_cap, d_rbm, ctx->d_lf, r.d_pmf_tab, a, r.d_cnt,
                          r.d_F, nullptr, nullptr, nullptr, r.d_big, r.d_big_count, r.d_hist,
                          nullptr, s.stream));
    if (cap > kRbBigSplit)
        HIP_TRY(launch_rb_big(0, false, ps.huge_grid, cap, d_rbm, ctx->d_lf, r.d_pmf_tab, a, r.d_cnt,
                              r.d_F, nullptr, nullptr, nullptr, r.d_huge, r.d_big_count + 1,
                              r.d_hist, nullptr, s.stream));
    HIP_TRY(launch_rb_finish(a, ctx->d_obs, r.d_hist, r.d_work, r.d_res, s.stream));
    HIP_TRY(hipEventRecord(r.ev1, s.stream));
    HIP_TRY(hipMemcpyAsync(r.h_res, r.d_res, sizeof(RbResultHeaderH) + sizeof(uint64_t) * size_t(a.B),
                           hipMemcpyDeviceToHost, s.stream));
    r.ran = true;
    r.adj_ran = false;
    r.scores = scores;
    r.sbar_valid = false;  // the derivative passes compute the means first
    r.R = R;
    r.B = a.B;
    r.left_pad = a.left_pad;
    r.n = n;
    r.K = n * (n - 1);
    r.cap = cap;
    r.b0 = b0;
    r.b1 = b1;
    r.seed = seed;
    r.set = set_id;
    r.init_state = M.init_state;
    r.rates_positive = rates_positive;
    r.p_interior = p_interior;
    r.wide = wide;
    return PDA_OK;
}

}  // namespace

extern "C" {

int pda_rb_forward(pda_ctx* ctx, int sid, int R, uint32_t seed, uint32_t set_id, int flags) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (R < 1) return fail(PDA_E_ARG, "R = %d < 1", R);
    if (!s->have_params) return fail(PDA_E_STATE, "stream %d: no parameters set", sid);
    if (ctx->B < 1 || !ctx->d_obs) return fail(PDA_E_STATE, "no observed histogram set");
    if (s->n > kRbMaxStates)
        return fail(PDA_E_ARG, "the RB form takes n <= %d states (n = %d)", kRbMaxStates, s->n);
    if (s->any_v)
        return fail(PDA_E_ARG, "the RB form needs v = 0 (no distance broadening) in every state");
    HIP_TRY(hipSetDevice(ctx->device));
    Model M;
    {
        std::string err;
        int rc = build_model(s->ph, s->n, s->args.data(), int(s->args.size()), M, err);
        if (rc) return fail(rc, "%s", err.c_str());
    }
    return rb_enqueue_forward(ctx, *s, s->rb, M, s->d_blob, R, seed, set_id, flags, true);
}

int pda_rb_get_result(pda_ctx* ctx, int sid, uint64_t* M_fixed, double* M, int nbins,
                      double* chi2, int32_t* effN) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    RbState& r = s->rb;
    if (!r.ran) return fail(PDA_E_STATE, "stream %d has no RB run", sid);
    if ((M_fixed || M) && nbins != r.B) return fail(PDA_E_ARG, "nbins %d != run's %d", nbins, r.B);
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    const RbResultHeaderH* h = reinterpret_cast<const RbResultHeaderH*>(r.h_res);
    const uint64_t* m = reinterpret_cast<const uint64_t*>(h + 1);
    for (int b = 0; b < r.B; ++b) {
        if (M_fixed) M_fixed[b] = m[b];
        if (M) M[b] = double(m[b]) * 0x1p-32;
    }
    if (chi2) *chi2 = h->chi2;
    if (effN) *effN = h->effN;
    return PDA_OK;
}

}  // extern "C"

namespace {

// Preconditions shared by the adjoint and the Jacobian of an RB forward run.
int check_derivative_run(const RbState& r, int sid, int nbins, int nargs) {
    if (!r.ran) return fail(PDA_E_STATE, "stream %d has no RB forward run", sid);
    if (nbins != r.B) return fail(PDA_E_ARG, "nbins %d != run's %d", nbins, r.B);
    if (nargs != r.n * (r.n + 1)) return fail(PDA_E_ARG, "nargs %d != %d", nargs, r.n * (r.n + 1));
    if (!r.scores) return fail(PDA_E_STATE, "the forward run kept no scores (PDA_RB_SCORES)");
    if (r.R < 2) return fail(PDA_E_STATE, "rate gradients need R >= 2 resamples");
    if (!r.p_interior) return fail(PDA_E_ARG, "efficiency gradients need 0 < E < 1");
    if (!r.rates_positive) return fail(PDA_E_ARG, "rate gradients need every rate > 0");
    if (r.init_state < 0)
        return fail(PDA_E_ARG, "rate gradients need the initial state fixed (inflow of state 0 > 1 s^-1)");
    return PDA_OK;
}

// Enqueues the lane and long-unit kernels of a derivative pass (mode 1
// adjoint with cotangent r.d_g, 2 Jacobian) of state r's last forward,
// accumulating into out (zeroed by the caller).
int enqueue_derivative(pda_ctx* ctx, Slot& s, RbState& r, int mode, double* out) {
    RunArgs a{};
    a.begin = uint32_t(r.b0);
    a.nunits = uint32_t((r.b1 - r.b0) * int64_t(r.R));
    a.R = uint32_t(r.R);
    a.B = r.B;
    a.left_pad = r.left_pad;
    a.nstates = r.n;
    PassShape ps;
    if (int rc = pass_shape(ctx, mode, r.n, r.wide, r.B, r.cap, a.nunits, ps)) return rc;
    RunArgs la = a;
    la.batch = ps.batch;
    const RbModel* d_rbm = static_cast<const RbModel*>(r.d_rbm);
    const RbMemoLayout& ML = ctx->rb_memo[r.n];
    const double* g = mode == 1 ? r.d_g : nullptr;
    HIP_TRY(hipMemsetAsync(r.d_work, 0, sizeof(uint32_t) * kRbCounters, s.stream));
    HIP_TRY(hipEventRecord(r.ev_a0, s.stream));
    if (!r.sbar_valid) {
        HIP_TRY(launch_rb_score_mean(a, r.K, r.d_S, r.d_Sbar, ctx->num_cu, s.stream));
        r.sbar_valid = true;
    }
    HIP_TRY(launch_rb_unit(mode, r.n, ps.unit_grid, d_rbm, la, r.d_cnt, r.d_F, r.d_S, r.d_Sbar, g,
                           ML.d_rank_offP, ML.d_rank_offD, r.d_poolP, nullptr, out,
                           r.d_work, s.stream));
    HIP_TRY(launch_rb_big(mode, ps.wide, ps.big_grid, ps.big_cap, d_rbm, ctx->d_lf, r.d_pmf_tab, a,
                          r.d_cnt, r.d_F, r.d_S, r.d_Sbar, g, r.d_big, r.d_big_count, nullptr, out,
                          s.stream));
    if (r.cap > kRbBigSplit)
        HIP_TRY(launch_rb_big(mode, ps.wide, ps.huge_grid, r.cap, d_rbm, ctx->d_lf, r.d_pmf_tab, a,
                              r.d_cnt, r.d_F, r.d_S, r.d_Sbar, g, r.d_huge, r.d_big_count + 1,
                              nullptr, out, s.stream));
    HIP_TRY(hipEventRecord(r.ev_a1, s.stream));
    return PDA_OK;
}

// Enqueues the adjoint pass of state r's last forward for the cotangent g
// (nbins entries; r's pinned buffers must be free: the caller waited for the
// stream), its gradient copied to r.h_grad.
int rb_enqueue_adjoint(pda_ctx* ctx, Slot& s, RbState& r, const double* g) {
    std::memcpy(r.h_g, g, sizeof(double) * size_t(r.B));
    HIP_TRY(hipMemsetAsync(r.d_grad, 0, sizeof(double) * kGradLen, s.stream));
    HIP_TRY(hipMemcpyAsync(r.d_g, r.h_g, sizeof(double) * size_t(r.B), hipMemcpyHostToDevice,
                           s.stream));
    if (int rc = enqueue_derivative(ctx, s, r, 1, r.d_grad)) return rc;
    HIP_TRY(hipMemcpyAsync(r.h_grad, r.d_grad, sizeof(double) * kGradLen, hipMemcpyDeviceToHost,
                           s.stream));
    HIP_TRY(hipMemsetAsync(r.d_work, 0, sizeof(uint32_t) * kRbCounters, s.stream));
    r.adj_ran = true;
    return PDA_OK;
}

// The adjoint's gradient in mc::set_params layout (after the stream finished).
void rb_collect_adjoint(const RbState& r, double* grad, int nargs) {
    const int n = r.n, K = r.K;
    for (int i = 0; i < n; ++i) grad[i] = r.h_grad[i];            // d chi2 / dE_i (dp/dE = 1)
    for (int i = 0; i < K; ++i) grad[n + i] = r.h_grad[n + i];    // rates, args order
    for (int i = n + K; i < nargs; ++i) grad[i] = 0.0;            // v (held at 0)
}

// Enqueues the Jacobian pass of state r's last forward, copied to r.h_jac.
int rb_enqueue_jacobian(pda_ctx* ctx, Slot& s, RbState& r) {
    const size_t bytes = sizeof(double) * kGradLen * size_t(r.B);
    HIP_TRY(hipMemsetAsync(r.d_jac, 0, bytes, s.stream));
    if (int rc = enqueue_derivative(ctx, s, r, 2, r.d_jac)) return rc;
    HIP_TRY(hipMemcpyAsync(r.h_jac, r.d_jac, bytes, hipMemcpyDeviceToHost, s.stream));
    HIP_TRY(hipMemsetAsync(r.d_work, 0, sizeof(uint32_t) * kRbCounters, s.stream));
    r.adj_ran = true;
    return PDA_OK;
}

// The Jacobian rows J[b * nargs + i] (after the stream finished).
void rb_collect_jacobian(const RbState& r, double* J, int nbins, int nargs) {
    const int w = r.n + r.K;
    for (int b = 0; b < nbins; ++b) {
        for (int i = 0; i < w; ++i) J[size_t(b) * size_t(nargs) + size_t(i)] = r.h_jac[size_t(b) * kGradLen + size_t(i)];
        for (int i = w; i < nargs; ++i) J[size_t(b) * size_t(nargs) + size_t(i)] = 0.0;  // v
    }
}

}  // namespace

extern "C" {

int pda_rb_adjoint(pda_ctx* ctx, int sid, const double* g, int nbins, double* grad, int nargs) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (!g || !grad) return fail(PDA_E_ARG, "null argument");
    if (int rc = check_derivative_run(s->rb, sid, nbins, nargs)) return rc;
    for (int b = 0; b < nbins; ++b)
        if (!std::isfinite(g[b])) return fail(PDA_E_ARG, "g[%d] not finite", b);
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (int rc = rb_enqueue_adjoint(ctx, *s, s->rb, g)) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    rb_collect_adjoint(s->rb, grad, nargs);
    return PDA_OK;
}

int pda_rb_jacobian(pda_ctx* ctx, int sid, double* J, int nbins, int nargs) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (!J) return fail(PDA_E_ARG, "null argument");
    if (int rc = check_derivative_run(s->rb, sid, nbins, nargs)) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (int rc = rb_enqueue_jacobian(ctx, *s, s->rb)) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    rb_collect_jacobian(s->rb, J, nbins, nargs);
    return PDA_OK;
}

// ---- populations -------------------------------------------------------------

int pda_rb_forward_population(pda_ctx* ctx, int sid, int R, uint32_t seed, uint32_t set_base,
                              uint32_t set_stride, int flags) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    Population& p = s->pop;
    if (p.nsets < 1) return fail(PDA_E_STATE, "no population set");
    if (R < 1) return fail(PDA_E_ARG, "R = %d < 1", R);
    if (ctx->B < 1 || !ctx->d_obs) return fail(PDA_E_STATE, "no observed histogram set");
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<Model> models(size_t(p.nsets));
    for (int k = 0; k < p.nsets; ++k) {
        const int n = p.n[size_t(k)];
        if (n > kRbMaxStates)
            return fail(PDA_E_ARG, "set %d: the RB form takes n <= %d states (n = %d)", k,
                        kRbMaxStates, n);
        std::string err;
        int rc = build_model(s->ph, n, p.args.data() + p.off[size_t(k)], n * (n + 1),
                             models[size_t(k)], err);
        if (rc) return fail(rc, "set %d: %s", k, err.c_str());
        if (models[size_t(k)].any_v)
            return fail(PDA_E_ARG, "set %d: the RB form needs v = 0 in every state", k);
    }
    // every set's pinned model may still be in use by the previous run
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->rbpop.size() < size_t(p.nsets)) s->rbpop.resize(size_t(p.nsets));
    for (int k = 0; k < p.nsets; ++k) {
        RbState& r = s->rbpop[size_t(k)];
        if (!r.d_blob) HIP_TRY(hipMalloc(reinterpret_cast<void**>(&r.d_blob), sizeof(ModelBlob)));
        if (!r.h_blob)
            HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&r.h_blob), sizeof(ModelBlob),
                                  hipHostMallocDefault));
        r.h_blob->m = models[size_t(k)];
        HIP_TRY(hipMemcpyAsync(r.d_blob, r.h_blob, sizeof(Model), hipMemcpyHostToDevice,
                               s->stream));
        int rc = rb_enqueue_forward(ctx, *s, r, models[size_t(k)], r.d_blob, R, seed,
                                    set_base + uint32_t(k) * set_stride, flags, false);
        if (rc) {
            s->rbpop_sets = 0;
            return fail(rc, "set %d: %s", k, last_error().c_str());
        }
    }
    s->rbpop_sets = p.nsets;
    return PDA_OK;
}

int pda_rb_get_population_result(pda_ctx* ctx, int sid, uint64_t* M_fixed, double* M, int nbins,
                                 double* chi2, int32_t* effN) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (s->rbpop_sets < 1) return fail(PDA_E_STATE, "stream %d has no RB population run", sid);
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int k = 0; k < s->rbpop_sets; ++k) {
        const RbState& r = s->rbpop[size_t(k)];
        if ((M_fixed || M) && nbins != r.B) return fail(PDA_E_ARG, "nbins %d != run's %d", nbins, r.B);
        const RbResultHeaderH* h = reinterpret_cast<const RbResultHeaderH*>(r.h_res);
        const uint64_t* m = reinterpret_cast<const uint64_t*>(h + 1);
        for (int b = 0; b < r.B; ++b) {
            if (M_fixed) M_fixed[size_t(k) * size_t(nbins) + size_t(b)] = m[b];
            if (M) M[size_t(k) * size_t(nbins) + size_t(b)] = double(m[b]) * 0x1p-32;
        }
        if (chi2) chi2[k] = h->chi2;
        if (effN) effN[k] = h->effN;
    }
    return PDA_OK;
}

int pda_rb_adjoint_population(pda_ctx* ctx, int sid, const double* g, int nbins, double* grad,
                              int nargs) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (!g || !grad) return fail(PDA_E_ARG, "null argument");
    if (s->rbpop_sets < 1) return fail(PDA_E_STATE, "stream %d has no RB population run", sid);
    const int K = s->rbpop_sets;
    for (int k = 0; k < K; ++k) {
        if (int rc = check_derivative_run(s->rbpop[size_t(k)], sid, nbins, nargs))
            return fail(rc, "set %d: %s", k, last_error().c_str());
        for (int b = 0; b < nbins; ++b)
            if (!std::isfinite(g[size_t(k) * size_t(nbins) + size_t(b)]))
                return fail(PDA_E_ARG, "set %d: g[%d] not finite", k, b);
    }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int k = 0; k < K; ++k)
        if (int rc = rb_enqueue_adjoint(ctx, *s, s->rbpop[size_t(k)], g + size_t(k) * size_t(nbins)))
            return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int k = 0; k < K; ++k)
        rb_collect_adjoint(s->rbpop[size_t(k)], grad + size_t(k) * size_t(nargs), nargs);
    return PDA_OK;
}

int pda_rb_jacobian_population(pda_ctx* ctx, int sid, const int32_t* which, double* J, int nbins,
                               int nargs) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    if (!J) return fail(PDA_E_ARG, "null argument");
    if (s->rbpop_sets < 1) return fail(PDA_E_STATE, "stream %d has no RB population run", sid);
    const int K = s->rbpop_sets;
    for (int k = 0; k < K; ++k)
        if (!which || which[k])
            if (int rc = check_derivative_run(s->rbpop[size_t(k)], sid, nbins, nargs))
                return fail(rc, "set %d: %s", k, last_error().c_str());
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int k = 0; k < K; ++k)
        if (!which || which[k])
            if (int rc = rb_enqueue_jacobian(ctx, *s, s->rbpop[size_t(k)])) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    const size_t per = size_t(nbins) * size_t(nargs);
    for (int k = 0; k < K; ++k) {
        if (!which || which[k]) rb_collect_jacobian(s->rbpop[size_t(k)], J + size_t(k) * per, nbins, nargs);
        else std::memset(J + size_t(k) * per, 0, sizeof(double) * per);
    }
    return PDA_OK;
}

int pda_rb_stats(pda_ctx* ctx, int sid, float* walk_ms, float* forward_ms, float* adjoint_ms) {
    Slot* s = slot_of(ctx, sid);
    if (!s) return fail(PDA_E_ARG, "bad stream id %d", sid);
    RbState& r = s->rb;
    if (!r.ran) return fail(PDA_E_STATE, "stream %d has no RB run", sid);
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (walk_ms) HIP_TRY(hipEventElapsedTime(walk_ms, r.ev0, r.ev_walk));
    if (forward_ms) HIP_TRY(hipEventElapsedTime(forward_ms, r.ev0, r.ev1));
    if (adjoint_ms) {
        *adjoint_ms = 0.f;
        if (r.adj_ran) HIP_TRY(hipEventElapsedTime(adjoint_ms, r.ev_a0, r.ev_a1));
    }
    return PDA_OK;
}

}  // extern "C"
