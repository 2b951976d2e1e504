// gsmfretda_mc.hpp -- drop-in `class mc` for gSMFRETda's gpuWorker /
// streamWorker / main (src/mc.hpp:32-180), implemented over the C ABI of
// gsmfretda_pda.h on MI355X.
//
// Same method names, argument meaning and call order as the reference:
//   mc(devid, streamNum) -> init_data_gpu(...) -> get_max_conversion_capacity()
//   per parameter set: set_nstates -> set_params -> setBurstBd ->
//   init_randstate (when N changed) -> set_params_buff -> run_kernel ->
//   streamQuery until true -> read hmcE[sid][0 .. N*reSampleTimes).
// Differences, all deliberate:
//   * random numbers are stateless Philox streams keyed by (seed, stream,
//     evaluation, burst, resample); init_randstate / set_params_buff only
//     record sizes;
//   * reSampleTimes may be any positive integer (the reference needs a power
//     of two, src/mc.cu:101);
//   * setBurstBd's [cstart, cstop) is honoured by the kernel (the reference
//     always simulates bursts [0, N), src/mc.cu:101);
//   * the GPU also bins the samples and computes chi2; get_hist_chi2()
//     returns them, so a caller can skip the per-sample copy (sample_copy).
// Errors: HIP failures print and exit(EXIT_FAILURE) like CUDA_CHECK_RETURN
// (src/cuda_tools.hpp:20-28); set_params returns false on bad arguments.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <vector>

#include "gsmfretda_pda.h"

// Replaces showGPUsInfo (src/mc.cu:26-79, declared at src/mc.hpp:31): prints
// index, UUID, copy engines, name and memory of device dn (every device when
// dn < 0), writes device dn's 16 UUID bytes to gpuuid and 2 x its copy-engine
// count (1 without overlap) to streamCount when they are not null, and
// returns the number of devices.
int showGPUsInfo(int dn = -1, char* gpuuid = nullptr, int* streamCount = nullptr);

class mc {
   public:
    mc(int devid, int _streamNum = 16, unsigned char debug = 0, std::uintmax_t hdf5size = 0,
       bool profiler = false);
    ~mc();

    hipStream_t* getStreams(int* numStream = nullptr);
    int devid;
    char gpuuuid[33];
    std::atomic_int workerNum;
    float** hmcE;  // per stream: host copy of the last run's N*reSampleTimes samples
    int* s_n;
    int sz_burst;
    bool streamQuery(int sid);
    int reSampleTimes;

    int set_nstates(int n, int sid);
    void set_gpuid();
    int setBurstBd(int cstart, int cstop, int sid);
    void set_reSampleTimes(int n);
    void free_data_gpu();
    void init_randstate(int N, int sid);
    void run_kernel(int N, int sid);
    void get_res(int sid, int N);
    void set_params_buff(int oldS_n, int N_sid, int sid);
    void init_data_gpu(std::vector<int64_t>& istart, std::vector<int64_t>& start,
                       std::vector<int64_t>& stop, std::vector<int>& phCount, long _sz_tag,
                       int64_t* burst_ad, int64_t* burst_dd, std::vector<float>& T_burst_duration,
                       std::vector<float>& SgDivSr, float& iclk_p, float& ibg_ad_rate,
                       float& ibg_dd_rate);
    bool get_max_conversion_capacity(int max_stateNum = 8);
    bool set_params(int n, int sid, std::vector<float>& args);

    // -- additions without a reference counterpart --
    // Histogram binning of the on-GPU chi2 (fret_hist_num, leftPad).
    bool set_hist(int fretHistNum, int leftPad);
    // Seed of the Philox streams; common = true reuses the same streams on
    // every evaluation of a stream (common random numbers).
    void set_seed(uint32_t seed, bool common);
    // false: run_kernel skips the per-sample D2H copy (hmcE stays stale).
    void sample_copy(bool on);
    // Histogram and reduced chi2 of the stream's last completed run.
    bool get_hist_chi2(int sid, std::vector<int32_t>& hist, double& chi2_64, float& chi2_ref);
    pda_ctx* context() { return ctx; }

   private:
    pda_ctx* ctx;
    int streamNum;
    unsigned char debug;
    bool profiler;
    bool data_gpu_inited;
    int fretHistNum, leftPad;
    uint32_t seed;
    bool common;
    bool copy_samples;
    std::vector<hipStream_t> streams;
    std::vector<int> runN;
    std::vector<uint32_t> evals;
    std::vector<float> observed_;
};
