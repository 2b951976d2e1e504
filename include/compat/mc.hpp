/*
 * compat/mc.hpp -- drop-in replacement for gSMFRETda's src/mc.hpp.
 *
 * Copy this file over src/mc.hpp in the reference tree (and stop compiling
 * src/mc.cu).  It keeps what the reference header gives its includers
 * (src/mc.hpp:1-17), because the unchanged callers rely on it:
 *   * <atomic>, <vector>, <cstdint>, <queue> and `using namespace std;`
 *     (src/streamWorker.hpp:11 declares `string* _url`, src/streamWorker.cpp
 *     uses unqualified vector<float>);
 *   * "loadHdf5.hpp" (the reference's own header, unchanged: `retype`,
 *     loadhdf5, burst_data) and "eigenhelper.hpp" when Eigen is on the
 *     include path, as it is in the reference's own build (src/main.cpp also
 *     includes it directly);
 *   * showGPUsInfo (src/mc.hpp:31, called by src/main.cpp:78-80);
 *   * class mc with the same methods and public fields, here over
 *     libgsmfretda_pda.so (gsmfretda_mc.hpp).
 * Not kept: the CUDA-only pieces -- <curand_kernel.h>, cuda_tools.hpp's
 * CUDA_CHECK_RETURN, mrImp and the rk_state struct (RNG state of the CUDA
 * kernel; the MI355X path keeps no RNG state) -- which no caller outside
 * src/mc.cu uses.  getStreams returns hipStream_t* where the reference
 * returns cudaStream_t*; src/main.cpp:132 discards the pointer.
 */
#pragma once
#include <atomic>
#include <cstdint>
#include <queue>
#include <string>
#include <vector>
using namespace std;
#if __has_include(<Eigen/Core>)
#include "eigenhelper.hpp"
#endif
#include "loadHdf5.hpp"
#define DEBUGMC 0

#include "gsmfretda_mc.hpp"
