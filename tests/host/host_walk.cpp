// host_walk.cpp -- test harness: the product's per-unit walk (pda_walk.h)
// and burst re-encoding (pda_layout.cpp) compiled for the host, so the
// algorithm (binary search over re-timed ticks + channel-run links, early
// exit) can be compared with the oracle's literal photon scan without a GPU.
// Test infrastructure; the product never runs this build.
#include <cstdint>
#include <cmath>
#include <cstring>
#include <utility>
#include <string>
#include <vector>

#include "pda_layout.h"
#include "pda_walk.h"

using namespace pda;

static std::string g_err;
static int64_t replays = 0;
static int guided = 1;  // fast form draws binomials by the threshold rows

extern "C" int64_t hw_replays() { return replays; }

extern "C" void hw_set_guided(int on) { guided = on; }

// Threshold rows of every (state, count) row of a model, as the guide kernel
// builds them on the device.
static void build_guide(const Model& M, const BinomTable& tab, std::vector<ThrPair>& thr,
                        std::vector<uint8_t>& guide) {
    thr.assign(kGuideThrPairs, ThrPair{0, 0});
    guide.assign(kGuideBytes, 0);
    for (int s = 0; s < M.n; ++s)
        for (int m = 0; m < kBinomTable; ++m) {
            const size_t row = size_t(s) * kBinomTable + size_t(m);
            build_guide_row_host(M, tab, s, m, thr.data() + guide_row_offset(s, m),
                            guide.data() + row * kGuideBuckets);
        }
}

// The rows of one model as build_guide_row_host makes them (the layout of
// pda_debug_guide).
extern "C" int hw_build_guide(int n, const float* args, const float* photophysics, ThrPair* thr,
                              uint8_t* guide) {
    Photophysics ph;
    ph.gamma = photophysics[0];
    ph.dexdir = photophysics[1];
    ph.crosstalk = photophysics[2];
    ph.r0 = photophysics[3];
    Model M;
    int rc = build_model(ph, n, args, n * (n + 1), M, g_err);
    if (rc) return rc;
    BinomTable tab;
    build_binom_table(M, tab);
    std::vector<ThrPair> t;
    std::vector<uint8_t> g;
    build_guide(M, tab, t, g);
    std::memcpy(thr, t.data(), sizeof(ThrPair) * t.size());
    std::memcpy(guide, g.data(), g.size());
    return 0;
}

static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// The capped-Poisson draw as a plain inversion loop (the form of
// background_count's single chunk, src/mc.cu:153-166), for comparison.
// trace: the last k the loop reached and whether some pk was subnormal.
static int32_t g_loop_k = 0;
static bool g_loop_subnormal = false;
static float poisson_loop(float lam, float cap, float u_first, float u_cap) {
    g_loop_k = 0;
    g_loop_subnormal = false;
    if (lam > 64.0f || !(cap < 0x1p30f)) return -1.0f;
    int32_t k = 0;
    bool over = false;
    if (lam > 0.0f) {
        float pk = fexp(-lam);
        float cdf = pk;
        while (u_first > cdf) {
            ++k;
            if (float(k) > cap) {
                over = true;
                break;
            }
            pk = (pk * lam) / float(k);
            if (pk != 0.0f && pk < 0x1p-126f) g_loop_subnormal = true;
            cdf = cdf + pk;
        }
    }
    g_loop_k = k;
    if (!over && float(k) <= cap) return float(k);
    return floorf(cap * u_cap);
}

// background_single against poisson_loop: out[0] draws compared, out[1]
// results that differ (bitwise), out[2] draws whose loop went past k = 255,
// out[3] draws whose loop made a subnormal pk.  Means from [0, 64] on
// several scales, caps of 0, 1/2, 1, 3/2, 2, 5/2, 3, 10 and random, uniforms
// on the 2^-24 grid at random and at every cumulative sum of the first 8
// steps and its neighbours; then, for n_deep means spread over (0, 64], caps
// of 300 and 1000, the uniforms 1, 1 - 2^-24 and 1 - 2^-23 (past a cdf that
// stalls below them the loop runs to the cap) and every cumulative sum of the
// whole loop, until pk underflows, with its neighbours.
extern "C" int hw_poisson_check(int n_lam, int n_deep, uint64_t seed, int64_t* out) {
    uint64_t st = seed;
    auto unif = [&]() { return float(double(splitmix(st) >> 11) * 0x1p-53); };
    const float caps[] = {0.0f, 0.5f, 1.0f, 1.5f, 2.0f, 2.5f, 3.0f, 10.0f};
    out[0] = out[1] = out[2] = out[3] = 0;
    auto cmp = [&](float lam, float cap, float u1, float u2) {
        const float a = background_single(lam, cap, u1, u2);
        const float b = poisson_loop(lam, cap, u1, u2);
        ++out[0];
        out[2] += g_loop_k > 255;
        out[3] += g_loop_subnormal;
        uint32_t ba, bb;
        std::memcpy(&ba, &a, 4);
        std::memcpy(&bb, &b, 4);
        if (ba != bb) ++out[1];
    };
    for (int i = 0; i < n_deep; ++i) {
        const float lam = 64.0f * float(i + 1) / float(n_deep);
        const float u2 = uniform24(uint32_t(splitmix(st)));
        for (float cap : {300.0f, 1000.0f})
            for (float u : {1.0f, 1.0f - 0x1p-24f, 1.0f - 0x1p-23f}) cmp(lam, cap, u, u2);
        float pk = fexp(-lam), cdf = pk;
        for (int k = 1; k < 400 && pk > 0.0f; ++k) {
            for (float u : {cdf, std::nextafter(cdf, 0.0f), std::nextafter(cdf, 2.0f)})
                if (u > 0.0f && u <= 1.0f) cmp(lam, 300.0f, u, u2);
            pk = (pk * lam) / float(k);
            cdf = cdf + pk;
        }
    }
    for (int i = 0; i < n_lam; ++i) {
        const int scale = i % 4;
        float lam = scale == 0 ? unif() * 3.0f
                  : scale == 1 ? unif() * 64.0f
                  : scale == 2 ? std::ldexp(unif(), -int(splitmix(st) % 100))
                               : float(splitmix(st) % 6) * 0.5f;
        const float cap = (i / 4) % 9 < 8 ? caps[(i / 4) % 9] : unif() * 50.0f;
        const float u2 = uniform24(uint32_t(splitmix(st)));
        for (int r = 0; r < 8; ++r) cmp(lam, cap, uniform24(uint32_t(splitmix(st))), u2);
        // the uniforms at the first cumulative sums and their neighbours
        float pk = fexp(-lam), cdf = pk;
        for (int k = 1; k <= 8 && lam > 0.0f; ++k) {
            for (float u : {cdf, std::nextafter(cdf, 0.0f), std::nextafter(cdf, 2.0f)})
                if (u > 0.0f && u <= 1.0f) cmp(lam, cap, u, u2);
            pk = (pk * lam) / float(k);
            cdf = cdf + pk;
        }
    }
    return 0;
}


// unit_div against the integer division: out[0] quotients compared, out[1]
// that differ.  Every divisor 1 .. 4096 and n_rand random ones (up to
// 2^32 - 1), each at the numerators 0, 1, d - 1, d, d + 1, multiples of d and
// their neighbours near 2^31, 2^31 - 1, and n_u random ones below 2^31.
extern "C" int hw_unit_div_check(int n_rand, int n_u, uint64_t seed, int64_t* out) {
    uint64_t st = seed;
    out[0] = out[1] = 0;
    auto check = [&](uint32_t d) {
        const UnitDiv q = make_unit_div(d);
        auto one = [&](uint64_t u) {
            if (u >= (uint64_t(1) << 31)) return;
            ++out[0];
            if (unit_div(uint32_t(u), q) != uint32_t(u) / d) ++out[1];
        };
        for (uint64_t u : {uint64_t(0), uint64_t(1), uint64_t(d) - 1, uint64_t(d), uint64_t(d) + 1,
                           (uint64_t(1) << 31) - 1})
            one(u);
        const uint64_t top = ((uint64_t(1) << 31) - 1) / d * d;
        for (uint64_t k = 0; k < 4 && top >= k * d; ++k)
            for (int64_t e = -1; e <= 1; ++e) one(top - k * d + uint64_t(e));
        for (int i = 0; i < n_u; ++i) one(splitmix(st) & 0x7FFFFFFFull);
    };
    for (uint32_t d = 1; d <= 4096; ++d) check(d);
    for (int i = 0; i < n_rand; ++i) {
        const uint32_t d = uint32_t(splitmix(st) >> (32 + splitmix(st) % 32));
        if (d) check(d);
    }
    check(0xFFFFFFFFu);
    check(0x80000000u);
    check(0x80000001u);
    return 0;
}

// The guided draw against the inversion loop of binomial_table on one
// model: for every state and count 1..127, at every threshold pair +- 3 grid
// steps and at n_rand random grid indices.  out: [0] draws compared, [1]
// draws where both give an X and they differ, [2] guided -1 where the loop
// gives an X (inside a band), [3] loop -1 where the guided draw gives an X,
// [4] thresholds the loop's exact crossing lies outside (lo_k, up_k] of,
// [5] largest row length, [6] rows whose guide is not monotone, [7] random
// draws that fell inside a band.
extern "C" int hw_guide_check(int n, const float* args, const float* photophysics, int n_rand,
                              uint64_t seed, int64_t* out) {
    Photophysics ph;
    ph.gamma = photophysics[0];
    ph.dexdir = photophysics[1];
    ph.crosstalk = photophysics[2];
    ph.r0 = photophysics[3];
    Model M;
    int rc = build_model(ph, n, args, n * (n + 1), M, g_err);
    if (rc) return rc;
    BinomTable tab;
    build_binom_table(M, tab);
    std::vector<ThrPair> thr;
    std::vector<uint8_t> guide;
    build_guide(M, tab, thr, guide);
    const BinomGuide G{thr.data(), guide.data()};
    for (int i = 0; i < 8; ++i) out[i] = 0;
    auto draw_pair = [&](int s, int m, int64_t u) {
        const uint32_t hi = uint32_t(u >> 26) << 5;
        const uint32_t lo = uint32_t(u & ((int64_t(1) << 26) - 1)) << 6;
        const int64_t a = binomial_table(m, s, M, tab, hi, lo);
        const int64_t b = binomial_guided(m, s, M, G, hi, lo);
        return std::make_pair(a, b);
    };
    auto check = [&](int s, int m, int64_t u, bool rnd) {
        if (u < 0 || u > kU53Max) return;
        const auto ab = draw_pair(s, m, u);
        ++out[0];
        if (ab.first >= 0 && ab.second >= 0 && ab.first != ab.second) ++out[1];
        if (ab.first >= 0 && ab.second < 0) ++out[rnd ? 7 : 2];
        if (ab.first < 0 && ab.second >= 0) ++out[3];
    };
    // the loop's X as a function of the grid index, before the flip
    auto loop_x = [&](int s, int m, int64_t u) {
        const int64_t x = draw_pair(s, m, u).first;
        return x < 0 ? int64_t(kBinomTable + 1) : (M.flip[s] ? int64_t(m) - x : x);
    };
    uint64_t st = seed;
    for (int s = 0; s < M.n; ++s)
        for (int m = 1; m < kBinomTable; ++m) {
            const size_t row = size_t(s) * kBinomTable + size_t(m);
            const ThrPair* t = thr.data() + guide_row_offset(s, m);
            const uint8_t* g = guide.data() + row * kGuideBuckets;
            int len = 0;
            while (t[len].up != INT64_MAX) ++len;
            if (len > out[5]) out[5] = len;
            for (int j = 1; j < kGuideBuckets; ++j)
                if (g[j] < g[j - 1]) {
                    ++out[6];
                    break;
                }
            for (int k = 0; k < len; ++k) {
                for (int d = -3; d <= 3; ++d) {
                    check(s, m, t[k].up + d, false);
                    check(s, m, t[k].lo + d, false);
                }
                // exact crossing of X >= k + 1 by bisection on the grid
                // (X is non-decreasing in u): it must lie in (lo_k, up_k]
                if (t[k].up <= kU53Max && loop_x(s, m, kU53Max) >= k + 1) {
                    int64_t a = 0, b = kU53Max;  // loop_x(b) >= k + 1
                    if (loop_x(s, m, 0) >= k + 1) b = 0;
                    while (b - a > 1) {
                        const int64_t mid = a + (b - a) / 2;
                        if (loop_x(s, m, mid) >= k + 1) b = mid;
                        else a = mid;
                    }
                    if (!(b > t[k].lo && b <= t[k].up)) ++out[4];
                }
            }
            for (int i = 0; i < n_rand; ++i) check(s, m, int64_t(splitmix(st) >> 11), true);
        }
    return 0;
}

extern "C" const char* hw_error() { return g_err.c_str(); }

extern "C" int hw_run(const int64_t* start, const int64_t* stop, const int64_t* istart,
                      const int32_t* phcount, int64_t nb, const int64_t* times,
                      const uint8_t* mask_ad, const uint8_t* mask_dd, int64_t np, const float* T,
                      float clk_p, float bg_ad, float bg_dd, int n, const float* args,
                      const float* photophysics, int64_t b0, int64_t b1, int R, uint32_t seed,
                      uint32_t set, int32_t* acc, float* F, float* E, uint32_t* link_dump) {
    std::vector<BurstDesc> bursts;
    std::vector<uint32_t> ticks;
    std::vector<uint2> links;
    PhotonSrc src{times, mask_ad, mask_dd, nullptr, nullptr};
    int rc = encode_bursts(start, stop, istart, phcount, nb, src, np, T, clk_p, bg_ad, bg_dd,
                           bursts, ticks, links, g_err);
    if (rc) return rc;
    if (link_dump) {
        for (size_t i = 0; i < ticks.size(); ++i) {
            link_dump[3 * i] = ticks[i];
            link_dump[3 * i + 1] = links[i].x;
            link_dump[3 * i + 2] = links[i].y;
        }
    }
    Photophysics ph;
    ph.gamma = photophysics[0];
    ph.dexdir = photophysics[1];
    ph.crosstalk = photophysics[2];
    ph.r0 = photophysics[3];
    Model M;
    rc = build_model(ph, n, args, n * (n + 1), M, g_err);
    if (rc) return rc;
    BinomTable tab;
    build_binom_table(M, tab);
    std::vector<ThrPair> thr;
    std::vector<uint8_t> guide;
    build_guide(M, tab, thr, guide);
    const BinomGuide G{thr.data(), guide.data()};
    RunArgs a{};
    a.R = uint32_t(R);
    a.seed = seed;
    a.set = set;
    a.bg_on = bg_dd < 1e-4f ? 0 : 1;
    a.clk_p = clk_p;
    a.bg_ad = bg_ad;
    a.bg_dd = bg_dd;
    for (int64_t b = b0; b < b1; ++b) {
        const BurstDesc& d = bursts[size_t(b)];
        for (int r = 0; r < R; ++r) {
            const UnitKey key{seed, set, uint32_t(r), uint32_t(b)};
            // fast form unless the model broadens; abandoned units replay in full
            WalkOut o{0.0f, 0.0f, true};
            if (!M.any_v && guided)
                o = walk_unit<true, BinomGuide>(M, &G, key, d.P, d.stop_rel, d.T,
                                                ticks.data() + d.off, links.data() + d.off, a);
            else if (!M.any_v)
                o = walk_unit<true>(M, &tab, key, d.P, d.stop_rel, d.T, ticks.data() + d.off,
                                    links.data() + d.off, a);
            if (o.replay) {
                ++replays;
                o = walk_unit<false>(M, nullptr, key, d.P, d.stop_rel, d.T, ticks.data() + d.off,
                                     links.data() + d.off, a);
            }
            const int64_t u = (b - b0) * R + r;
            acc[u] = int32_t(o.acc);
            F[u] = o.F;
            E[u] = o.F > 0.0f ? o.acc / o.F : 0.0f;
        }
    }
    return 0;
}
