// mc_shim.cpp -- `class mc` (include/gsmfretda_mc.hpp) over the C ABI, so the
// reference's gpuWorker / streamWorker / main (src/gpuWorker.cpp:38-88,
// src/streamWorker.cpp:39-248, src/main.cpp:129-160) link against the
// MI355X path unchanged.
#include "../../include/gsmfretda_mc.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <algorithm>

namespace {

// CUDA_CHECK_RETURN (src/cuda_tools.hpp:20-28): report and exit.
void check(int rc, const char* what) {
    if (rc < 0) {
        std::fprintf(stderr, "gsmfretda_amd: %s failed: %s\n", what, pda_last_error());
        std::exit(EXIT_FAILURE);
    }
}

}  // namespace

int showGPUsInfo(int dn, char* gpuuid, int* streamCount) {
    int ndev = 0;
    check(hipGetDeviceCount(&ndev) == hipSuccess ? 0 : -1, "hipGetDeviceCount");
    if (dn >= ndev) return ndev;
    int prev = 0;
    (void)hipGetDevice(&prev);
    const int first = dn >= 0 ? dn : 0;
    const int last = dn >= 0 ? dn + 1 : ndev;
    for (int i = first; i < last; ++i) {
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, i) == hipSuccess ? 0 : -1, "hipGetDeviceProperties");
        std::printf("Device Index: %d\n  Device UUID: ", i);
        for (int k = 0; k < 16; ++k) std::printf("%02x", (unsigned char)prop.uuid.bytes[k]);
        if (gpuuid) std::memcpy(gpuuid, prop.uuid.bytes, 16);
        std::printf("\n  Concurrent copy and kernel execution: \t %s with %d copy engine(s)\n",
                    prop.asyncEngineCount > 0 ? "Yes" : "No", prop.asyncEngineCount);
        if (streamCount) *streamCount = prop.asyncEngineCount > 0 ? prop.asyncEngineCount * 2 : 1;
        std::printf("  Device name: %s\n", prop.name);
        size_t free_b = 0, total_b = 0;
        if (hipSetDevice(i) == hipSuccess && hipMemGetInfo(&free_b, &total_b) == hipSuccess)
            std::printf("GPU %d memory: free=%zu MB, total=%zu MB\n", i, free_b >> 20, total_b >> 20);
    }
    (void)hipSetDevice(prev);
    return ndev;
}

mc::mc(int id, int _streamNum, unsigned char de, std::uintmax_t, bool _profiler)
    : devid(id), workerNum(0), hmcE(nullptr), s_n(nullptr), sz_burst(0), reSampleTimes(4),
      ctx(nullptr), streamNum(_streamNum == 0 ? 4 : _streamNum), debug(de), profiler(_profiler),
      data_gpu_inited(false), fretHistNum(200), leftPad(2), seed(1234u), common(false),
      copy_samples(true) {
    std::memset(gpuuuid, 0, sizeof gpuuuid);
    workerNum.store(streamNum);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || devid < 0 || devid >= ndev) {
        std::printf("gpu id set error!\n");  // src/mc.cu:227-230: half-built object
        return;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, devid) == hipSuccess) {
        for (int i = 0; i < 16; ++i)
            std::snprintf(gpuuuid + 2 * i, 3, "%02x", (unsigned char)prop.uuid.bytes[i]);
    }
    check(pda_create(devid, streamNum, &ctx), "pda_create");
    streams.resize(size_t(streamNum));
    for (int s = 0; s < streamNum; ++s)
        streams[size_t(s)] = reinterpret_cast<hipStream_t>(pda_stream(ctx, s));
    hmcE = new float*[size_t(streamNum)]();
    s_n = new int[size_t(streamNum)]();
    runN.assign(size_t(streamNum), 0);
    evals.assign(size_t(streamNum), 0u);
}

mc::~mc() {
    if (ctx) pda_destroy(ctx);
    delete[] hmcE;
    delete[] s_n;
}

hipStream_t* mc::getStreams(int* numStream) {
    if (numStream) *numStream = streamNum;
    return streams.data();
}

void mc::set_gpuid() {
    check(hipSetDevice(devid) == hipSuccess ? 0 : -1, "hipSetDevice");
    // src/mc.cu:207-210 starts the CUDA profiler here; on ROCm the kernels are
    // traced from outside by rocprofv3 --kernel-trace.
    if (profiler) std::printf("profiling: run under rocprofv3 --kernel-trace\n");
}

void mc::init_data_gpu(std::vector<int64_t>& istart, std::vector<int64_t>& start,
                       std::vector<int64_t>& stop, std::vector<int>& phCount, long _sz_tag,
                       int64_t* burst_ad, int64_t* burst_dd, std::vector<float>& T_burst_duration,
                       std::vector<float>& SgDivSr, float& iclk_p, float& ibg_ad_rate,
                       float& ibg_dd_rate) {
    check(ctx ? 0 : -1, "init_data_gpu (no device)");
    std::vector<int32_t> ph(phCount.begin(), phCount.end());
    check(pda_load_bursts_tagged(ctx, istart.data(), start.data(), stop.data(), ph.data(),
                                 int64_t(start.size()), burst_ad, burst_dd, int64_t(_sz_tag),
                                 T_burst_duration.data(), iclk_p, ibg_ad_rate, ibg_dd_rate),
          "pda_load_bursts_tagged");
    check(pda_set_observed(ctx, SgDivSr.data(), int64_t(SgDivSr.size()), fretHistNum, leftPad),
          "pda_set_observed");
    sz_burst = int(start.size());
    observed_.assign(SgDivSr.begin(), SgDivSr.end());
    data_gpu_inited = true;
}

bool mc::set_hist(int nbins, int pad) {
    fretHistNum = nbins;
    leftPad = pad;
    if (!data_gpu_inited) return true;
    return pda_set_observed(ctx, observed_.data(), int64_t(observed_.size()), nbins, pad) == 0;
}

bool mc::get_max_conversion_capacity(int) { return data_gpu_inited; }

void mc::free_data_gpu() {
    // Resident data is released with the context (~mc); nothing else to do.
}

int mc::set_nstates(int n, int sid) {
    check(sid >= 0 && sid < streamNum ? 0 : -1, "set_nstates (stream id)");
    const int prev = s_n[sid];
    s_n[sid] = n;
    return prev != n ? prev : n;
}

bool mc::set_params(int n, int sid, std::vector<float>& args) {
    if (!ctx) return false;
    return pda_set_params(ctx, sid, n, args.data(), int(args.size())) == 0;
}

int mc::setBurstBd(int cstart, int cstop, int sid) {
    // src/mc.cu:523-535
    int rcstart = cstart, rcstop = cstop;
    if (cstop >= sz_burst || cstop < 0) rcstop = sz_burst;
    if (cstart >= rcstop) rcstart = 0;
    int N = rcstop - rcstart;
    if (cstart == -1) N = -1;
    if (N >= 0) check(pda_set_burst_range(ctx, sid, rcstart, rcstop, nullptr), "pda_set_burst_range");
    return N;
}

void mc::set_reSampleTimes(int n) { reSampleTimes = n > 0 ? n : 1; }

void mc::init_randstate(int N, int sid) { runN[size_t(sid)] = N; }

void mc::set_params_buff(int, int N_sid, int sid) { runN[size_t(sid)] = N_sid; }

void mc::set_seed(uint32_t s, bool c) {
    seed = s;
    common = c;
}

void mc::sample_copy(bool on) { copy_samples = on; }

void mc::run_kernel(int N, int sid) {
    if (N < 0) return;
    runN[size_t(sid)] = N;
    const uint32_t set = common ? uint32_t(sid) : uint32_t(sid) + uint32_t(streamNum) * evals[size_t(sid)]++;
    check(pda_run(ctx, sid, reSampleTimes, seed, set, copy_samples ? PDA_RUN_SAMPLES : 0),
          "pda_run");
    if (copy_samples) hmcE[sid] = pda_samples_host(ctx, sid);
}

bool mc::streamQuery(int sid) {
    if (sid < 0 || sid >= streamNum || !ctx) return false;
    return pda_query(ctx, sid) == 1;
}

bool mc::get_hist_chi2(int sid, std::vector<int32_t>& hist, double& chi2_64, float& chi2_ref) {
    hist.assign(size_t(fretHistNum), 0);
    return pda_get_result(ctx, sid, hist.data(), fretHistNum, &chi2_64, &chi2_ref, nullptr) == 0;
}

void mc::get_res(int sid, int N) {
    // src/mc.cu:602-616: debug print of the first samples, max and mean
    if (!debug || !hmcE[sid]) return;
    const int nn = N * reSampleTimes;
    const float* e = hmcE[sid];
    for (int i = 0; i < std::min(10, nn); ++i) std::printf("%2.4f \t", e[i]);
    const float mx = nn > 0 ? *std::max_element(e, e + nn) : 0.0f;
    const double mean = nn > 0 ? std::accumulate(e, e + nn, 0.0) / nn : 0.0;
    std::printf("\n%d======%d\n%g,%g\n", sid, nn, mx, mean);
}
