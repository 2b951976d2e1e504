// host_walk.cpp -- test harness: the product's per-unit walk (pda_walk.h)
// and burst re-encoding (pda_layout.cpp) compiled for the host, so the
// algorithm (binary search over re-timed ticks + channel-run links, early
// exit) can be compared with the oracle's literal photon scan without a GPU.
// Test infrastructure; the product never runs this build.
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "pda_layout.h"
#include "pda_walk.h"

using namespace pda;

static std::string g_err;
static int64_t replays = 0;

extern "C" int64_t hw_replays() { return replays; }

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
            if (!M.any_v)
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
