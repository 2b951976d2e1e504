// pda_layout.cpp -- see pda_layout.h.
#include "pda_layout.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "../../include/gsmfretda_pda.h"
#include "pda_math.h"

namespace pda {

namespace {
int fail_to(std::string& err, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return code;
}
}  // namespace

#define fail(code, ...) fail_to(err, code, __VA_ARGS__)

// Re-encodes the reference's burst arrays (src/loadHdf5.cpp:101-128) into the
// resident layout of pda_types.h.
int encode_bursts(const int64_t* start, const int64_t* stop, const int64_t* istart,
                  const int32_t* phcount, int64_t nbursts, const PhotonSrc& src, int64_t nphotons,
                  const float* T, float clk_p, float bg_ad, float bg_dd,
                  std::vector<BurstDesc>& bursts, std::vector<uint32_t>& ticks,
                  std::vector<uint2>& links, std::string& err) {
    if (nbursts < 0 || nbursts > int64_t(0x7FFFFFFF))
        return fail(PDA_E_ARG, "nbursts %lld out of range", (long long)nbursts);
    if (nphotons < 0) return fail(PDA_E_ARG, "nphotons < 0");
    if (nbursts > 0 && (!start || !stop || !istart || !phcount || !T))
        return fail(PDA_E_ARG, "null burst array");
    if (!(clk_p > 0.0f) || !std::isfinite(clk_p)) return fail(PDA_E_DATA, "clk_p must be > 0");
    if (!std::isfinite(bg_ad) || !std::isfinite(bg_dd) || bg_ad < 0.f || bg_dd < 0.f)
        return fail(PDA_E_DATA, "background rates must be finite and >= 0");

    bursts.assign(static_cast<size_t>(nbursts), BurstDesc{});
    int64_t total = 0;
    for (int64_t b = 0; b < nbursts; ++b) {
        const int64_t P = phcount[b];
        if (P < 0 || istart[b] < 0 || istart[b] + P > nphotons)
            return fail(PDA_E_DATA, "burst %lld: photons [%lld, +%lld) outside [0, %lld)",
                        (long long)b, (long long)istart[b], (long long)P, (long long)nphotons);
        const int64_t sr = stop[b] - start[b];
        if (sr < 0 || sr >= int64_t(kTickSentinel))
            return fail(PDA_E_DATA, "burst %lld: stop - start = %lld outside [0, 2^31)",
                        (long long)b, (long long)sr);
        if (!std::isfinite(T[b])) return fail(PDA_E_DATA, "burst %lld: T not finite", (long long)b);
        bursts[size_t(b)] = BurstDesc{uint32_t(total), uint32_t(P), uint32_t(sr), T[b]};
        total += P;
        if (total >= int64_t(0xFFFFFFFFll))
            return fail(PDA_E_DATA, "more than 2^32-1 photons in bursts");
    }
    if (total > 0 && !src.times && !(src.burst_ad && src.burst_dd))
        return fail(PDA_E_ARG, "null photon arrays");
    if (total > 0 && src.times && (!src.mask_ad || !src.mask_dd))
        return fail(PDA_E_ARG, "null photon masks");

    ticks.assign(static_cast<size_t>(total), 0u);
    links.resize(static_cast<size_t>(total));
    std::vector<uint32_t> col;
    for (int64_t b = 0; b < nbursts; ++b) {
        const BurstDesc& d = bursts[size_t(b)];
        const uint32_t P = d.P;
        if (P == 0) continue;
        const int64_t base = istart[b];
        const int64_t t0 = start[b];
        uint32_t* tk = ticks.data() + d.off;
        uint2* lk = links.data() + d.off;
        col.resize(P);
        for (uint32_t i = 0; i < P; ++i) {
            tk[i] = kTickSentinel;
            lk[i].x = P;
            lk[i].y = P;
        }
        int64_t prev = -1;
        for (uint32_t i = 0; i < P; ++i) {
            col[i] = src.colour(base + i);
            if (col[i]) {
                const int64_t t = src.tick(base + i) - t0;
                if (t < 0 || t >= int64_t(kTickSentinel))
                    return fail(PDA_E_DATA, "burst %lld photon %u: tick %lld outside [start, start+2^31)",
                                (long long)b, i, (long long)t);
                if (t < prev)
                    return fail(PDA_E_DATA, "burst %lld photon %u: channel photons out of time order",
                                (long long)b, i);
                prev = t;
                tk[i] = uint32_t(t);
            }
        }
        // backward pass: next channel photon, next photon carrying each colour
        uint32_t next_col = P, next_d = P, next_a = P;
        uint32_t next_tick = kTickSentinel;
        for (uint32_t ii = P; ii-- > 0;) {
            const uint32_t c = col[ii];
            // first photon after ii carrying the colour that ends a run of c
            uint32_t other = P;
            if (c == 1u) other = next_a;
            else if (c == 2u) other = next_d;
            else if (c == 3u) other = ii;
            if (c) {
                next_col = ii;
                next_tick = tk[ii];
                lk[ii].y = other;  // temporary: run end of this photon
            }
            if (c & 1u) next_d = ii;
            if (c & 2u) next_a = ii;
            if (!c) tk[ii] = next_tick;
            lk[ii].x = next_col;
        }
        // forward pass: fold the run end of f into every photon whose first
        // channel photon is f; entries >= ii are still untouched here
        for (uint32_t ii = 0; ii < P; ++ii) {
            const uint32_t f = lk[ii].x;
            if (f < P) {
                lk[ii].x = f | (col[f] << 30);
                lk[ii].y = lk[f].y;  // run end stored at f (f >= ii, already final)
            } else {
                lk[ii].y = P;
            }
        }
    }
    return PDA_OK;
}

// genMatK / genP_i2j / genMatP (src/eigenhelper.cpp:4-54) and the arg split
// of mc::set_params (src/mc.cu:790-800), folded into the kernel's tables.
int build_model(const Photophysics& s, int n, const float* args, int nargs, Model& M,
                std::string& err) {
    if (n < 2 || n > kMaxStates) return fail(PDA_E_ARG, "n = %d outside [2, %d]", n, kMaxStates);
    if (!args) return fail(PDA_E_ARG, "null args");
    if (nargs != n * (n + 1))
        return fail(PDA_E_ARG, "n = %d needs %d args, got %d", n, n * (n + 1), nargs);
    const float* E = args;
    const float* kargs = args + n;
    const float* v = args + n * n;
    for (int i = 0; i < n; ++i) {
        if (!(E[i] >= 0.0f && E[i] <= 1.0f)) return fail(PDA_E_ARG, "E[%d] = %g outside [0,1]", i, E[i]);
        if (!std::isfinite(v[i])) return fail(PDA_E_ARG, "v[%d] not finite", i);
    }
    for (int i = 0; i < n * (n - 1); ++i)
        if (!(kargs[i] >= 0.0f && kargs[i] <= 1e8f))
            return fail(PDA_E_ARG, "rate arg %d = %g outside [0, 1e8] s^-1", i, kargs[i]);
    float K[kMaxStates * kMaxStates];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            K[i + j * n] = i < j ? kargs[i * (n - 1) + j - 1] : (i > j ? kargs[i * (n - 1) + j] : 0.f);
    for (int j = 0; j < n; ++j)
        for (int ii = 0; ii < n; ++ii)
            if (ii != j) K[j + j * n] -= K[ii + j * n];
    std::memset(&M, 0, sizeof M);
    M.n = n;
    M.gmx = s.gamma - s.crosstalk;
    M.omdd = 1.0f - s.dexdir;
    M.r0 = s.r0;
    // P_i2j column s, without the diagonal, as running sums (drawJ_Si2Sj)
    for (int st = 0; st < n; ++st) {
        float sum = 0.0f;
        for (int ii = 0; ii < n; ++ii)
            if (ii != st) sum += K[ii + st * n];
        float run = 0.0f;
        for (int jj = 0; jj < n - 1; ++jj) {
            const int to = jj >= st ? jj + 1 : jj;
            run += K[to + st * n] / sum;
            M.nscum[st][jj] = run;
        }
        for (int j = 0; j < n; ++j) M.krate[st][j] = K[j + st * n];
    }
    // genMatP's unnormalised inflow (diagonal added then removed), running sums
    float run = 0.0f;
    for (int i = 0; i < n; ++i) {
        float acc = 0.0f;
        for (int j = 0; j < n; ++j) acc += K[i + j * n];
        acc -= K[i + i * n];
        run += acc;
        M.p0cum[i] = run;
    }
    // drawDisIdx returns 0 for every u in (0,1] once the first running sum
    // exceeds 1 (an inflow above 1 s^-1): the draw need not be made
    M.init_state = M.p0cum[0] > 1.0f ? 0 : -1;
    for (int i = 0; i < n; ++i) {
        M.v[i] = v[i];
        M.base[i] = efficiency_distance(E[i], s.r0);
        M.ep0[i] = distance_efficiency(M.base[i], s.r0);
        if (v[i] != 0.0f) M.any_v = 1;
        const double p = double(M.ep0[i]);
        M.flip[i] = p > 0.5 ? 1 : 0;
        M.pp[i] = M.flip[i] ? 1.0 - p : p;
        M.qq[i] = 1.0 - M.pp[i];
    }
    return PDA_OK;
}

// Inversion constants of binomial_inversion() (pda_math.h) per state and count.
void build_binom_table(const Model& M, BinomTable& T) {
    std::memset(&T, 0, sizeof T);
    for (int s = 0; s < M.n; ++s) {
        const double pp = M.pp[s], q = M.qq[s];
        const double lq = dlog(q);
        for (int n = 0; n < kBinomTable; ++n) {
            T.qn[s][n] = dexp(double(n) * lq);
            const double np = double(n) * pp;
            double lim = np + 10.0 * std::sqrt(np * q + 1.0);
            if (double(n) < lim) lim = double(n);
            T.bound[s][n] = int32_t(lim);
        }
    }
}

#undef fail

}  // namespace pda
