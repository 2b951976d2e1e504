// pda_layout.h -- host-side construction of the resident device layout:
// burst re-encoding (src/loadHdf5.cpp:101-128 -> pda_types.h) and the model
// tables of one parameter set (src/eigenhelper.cpp:4-54, src/mc.cu:790-800).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "pda_types.h"

namespace pda {

// Photon channel of the reference's masks: bit 0 donor (mask_dd), bit 1
// acceptor (mask_ad) -- the photons burst_data encodes as "time or 0"
// (src/loadHdf5.cpp:120-121).
struct PhotonSrc {
    const int64_t* times;      // absolute ticks (or null: use ad/dd arrays)
    const uint8_t* mask_ad;
    const uint8_t* mask_dd;
    const int64_t* burst_ad;   // reference encoding (or null)
    const int64_t* burst_dd;
    uint32_t colour(int64_t i) const {
        if (times) return (mask_dd[i] ? 1u : 0u) | (mask_ad[i] ? 2u : 0u);
        return (burst_dd[i] != 0 ? 1u : 0u) | (burst_ad[i] != 0 ? 2u : 0u);
    }
    int64_t tick(int64_t i) const {
        if (times) return times[i];
        return burst_dd[i] != 0 ? burst_dd[i] : burst_ad[i];
    }
};


struct Photophysics {
    float gamma = 0.34f, dexdir = 0.08f, crosstalk = 0.07f, r0 = 52.0f;  // src/mc.cu:89-91
};

// Validates the reference's burst arrays and re-encodes them.  Returns 0 or a
// negative PDA_E_* code with a message in err.
int encode_bursts(const int64_t* start, const int64_t* stop, const int64_t* istart,
                  const int32_t* phcount, int64_t nbursts, const PhotonSrc& src, int64_t nphotons,
                  const float* T, float clk_p, float bg_ad, float bg_dd,
                  std::vector<BurstDesc>& bursts, std::vector<uint32_t>& ticks,
                  std::vector<uint2>& links, std::string& err);

// Builds the kernel's model tables.  Returns 0 or PDA_E_ARG with err set.
int build_model(const Photophysics& ph, int n, const float* args, int nargs, Model& M,
                std::string& err);

// Tabulated inversion constants for the fast walk (needs M from build_model).
void build_binom_table(const Model& M, BinomTable& T);

}  // namespace pda
