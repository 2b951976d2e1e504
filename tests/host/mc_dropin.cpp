// mc_dropin.cpp -- test harness: drives the MI355X `class mc` exactly as the
// reference's callers do -- main (src/main.cpp:129-151), one gpuWorker pass
// (src/gpuWorker.cpp:71-80) and the streamWorker result step
// (src/streamWorker.cpp:184-206) -- on burst arrays written by the test.
//
// usage: mc_dropin <input dir> <output file> <n> <R> <seed> <begin> <end> <sid>
// input dir holds raw little-endian arrays: start, stop, istart (int64),
// phcount (int32), burst_ad, burst_dd (int64), T, SgDivSr (float32),
// scalars (float32 clk_p, bg_ad, bg_dd), args (float32).
// output: N*R float mcE, then int32 hist[200], double chi2_64, float
// chi2_ref, float chi2 computed here from hmcE like streamWorker.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gsmfretda_mc.hpp"

template <class T>
static std::vector<T> load(const std::string& path) {
    std::vector<T> v;
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", path.c_str());
        std::exit(2);
    }
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    v.resize(size_t(bytes) / sizeof(T));
    if (std::fread(v.data(), sizeof(T), v.size(), f) != v.size()) std::exit(2);
    std::fclose(f);
    return v;
}

int main(int argc, char** argv) {
    if (argc != 9) return 2;
    const std::string dir = argv[1];
    const int n = std::atoi(argv[3]);
    const int R = std::atoi(argv[4]);
    const unsigned seed = unsigned(std::strtoul(argv[5], nullptr, 10));
    const int begin = std::atoi(argv[6]), end = std::atoi(argv[7]), sid = std::atoi(argv[8]);
    std::vector<int64_t> start = load<int64_t>(dir + "/start"), stop = load<int64_t>(dir + "/stop");
    std::vector<int64_t> istart = load<int64_t>(dir + "/istart");
    std::vector<int32_t> ph32 = load<int32_t>(dir + "/phcount");
    std::vector<int> phCount(ph32.begin(), ph32.end());
    std::vector<int64_t> burst_ad = load<int64_t>(dir + "/burst_ad");
    std::vector<int64_t> burst_dd = load<int64_t>(dir + "/burst_dd");
    std::vector<float> T = load<float>(dir + "/T"), SgDivSr = load<float>(dir + "/SgDivSr");
    std::vector<float> sc = load<float>(dir + "/scalars"), params = load<float>(dir + "/args");
    float clk_p = sc[0], bg_ad = sc[1], bg_dd = sc[2];

    // main (src/main.cpp:129-151)
    mc pdamc(0, 2);
    pdamc.init_data_gpu(istart, start, stop, phCount, long(burst_ad.size()), burst_ad.data(),
                        burst_dd.data(), T, SgDivSr, clk_p, bg_ad, bg_dd);
    if (!pdamc.get_max_conversion_capacity()) return 3;
    pdamc.set_seed(seed, true);
    pdamc.set_reSampleTimes(R);
    pdamc.set_gpuid();

    // one gpuWorker pass (src/gpuWorker.cpp:71-80)
    int oldS_n = pdamc.set_nstates(n, sid);
    if (!pdamc.set_params(n, sid, params)) return 4;
    int N_sid = pdamc.setBurstBd(begin, end, sid);
    pdamc.init_randstate(N_sid, sid);
    pdamc.set_params_buff(oldS_n, N_sid, sid);
    pdamc.run_kernel(N_sid, sid);
    while (!pdamc.streamQuery(sid)) {
    }

    // streamWorker result step (src/streamWorker.cpp:184-206), fretHistNum 200, leftPad 2
    const int B = 200, leftPad = 2;
    std::vector<float> mcE(pdamc.hmcE[sid], pdamc.hmcE[sid] + size_t(N_sid) * pdamc.reSampleTimes);
    std::vector<float> vMc(B, 0.0f), vO(B, 0.0f);
    for (float x : mcE)
        if (x >= 0.0f && x < 1.0f) vMc[size_t(std::floor(double(x) * B))] += 1.0f;
    for (float x : SgDivSr)
        if (x >= 0.0f && x < 1.0f) vO[size_t(std::floor(double(x) * B))] += 1.0f;
    int effN = B - leftPad;
    float chisqr = 0.0f;
    for (int b = leftPad; b < B; ++b) {
        if (vMc[b] > 0)
            chisqr += std::pow(float(vO[b] - vMc[b] / pdamc.reSampleTimes), 2) /
                      (float(vMc[b]) / pdamc.reSampleTimes);
        else
            --effN;
    }
    chisqr = chisqr / (effN - n * (n + 1));

    std::vector<int32_t> hist;
    double c64 = 0;
    float c32 = 0;
    if (!pdamc.get_hist_chi2(sid, hist, c64, c32)) return 5;
    FILE* f = std::fopen(argv[2], "wb");
    std::fwrite(mcE.data(), sizeof(float), mcE.size(), f);
    std::fwrite(hist.data(), sizeof(int32_t), hist.size(), f);
    std::fwrite(&c64, sizeof c64, 1, f);
    std::fwrite(&c32, sizeof c32, 1, f);
    std::fwrite(&chisqr, sizeof chisqr, 1, f);
    std::fclose(f);
    pdamc.workerNum--;
    return 0;
}
