// ps_loop.cpp -- test harness: the reference's worker loop against the MI355X
// `class mc`, fed by a fake in-process parameter server.
//
// The reference runs one GPU thread (gpuWorker::run, src/gpuWorker.cpp:38-88)
// and one network thread per stream (streamWorker::run,
// src/streamWorker.cpp:39-248) that talk to the pySMFRETda parameter server
// over NNG REQ/REP with protobuf messages (3rdparty/protobuf/args.proto),
// each prefixed by one letter: "c" p_cap -> p_n, "p" p_str -> p_ga,
// "k" p_sid -> (ignored), "r" res -> p_sid.  p_ga.idx = -1 means "no work
// yet", -2 "stop"; p_sid.sid = -1 after a result ends the stream.  The
// per-stream handshake is the `dataready` state machine 0 -> 1 -> 2 -> 3 -> 4
// -> 0 under a mutex and a condition variable per stream.
//
// NNG and protobuf are not available here, so this harness carries the same
// messages, encoded in the protobuf wire format by a few lines of its own, over
// an in-process REQ/REP channel, and plays the parameter server itself:
// it hands out a population of parameter sets, collects every chi2 and writes
// them out.  The worker side follows the reference's state machine and calls
// the `mc` methods in the reference's order, computing chi2 from hmcE as
// streamWorker does.
//
// usage: ps_loop <input dir> <output file> <n> <R> <seed> <streams> <nsets>
//   input dir: the arrays of mc_dropin.cpp, plus pop_args (float32, nsets x
//   n(n+1)).  output (text): "ohist b1 .. bB", then per set "set idx sid chi2
//   sent_hist" and its shist when sent.
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gsmfretda_mc.hpp"

using namespace std::chrono_literals;

// ---- protobuf wire format (proto3: default values are not written) --------
namespace wire {

struct Writer {
    std::string b;
    void varint(uint64_t v) {
        while (v >= 0x80) {
            b.push_back(char(v | 0x80));
            v >>= 7;
        }
        b.push_back(char(v));
    }
    void key(int field, int type) { varint(uint64_t(field) << 3 | uint64_t(type)); }
    void i32(int field, int32_t v) {
        if (!v) return;
        key(field, 0);
        varint(uint64_t(int64_t(v)));  // negative int32: ten-byte varint
    }
    void boolean(int field, bool v) {
        if (!v) return;
        key(field, 0);
        varint(1);
    }
    void f32(int field, float v) {
        if (v == 0.0f) return;
        key(field, 5);
        char c[4];
        std::memcpy(c, &v, 4);
        b.append(c, 4);
    }
    void str(int field, const std::string& s) {
        if (s.empty()) return;
        key(field, 2);
        varint(s.size());
        b += s;
    }
    void packed(int field, const std::vector<float>& v) {
        if (v.empty()) return;
        key(field, 2);
        varint(v.size() * 4);
        for (float x : v) {
            char c[4];
            std::memcpy(c, &x, 4);
            b.append(c, 4);
        }
    }
};

// Field values of one message: varints, fixed32s, length-delimited bytes.
struct Msg {
    std::map<int, uint64_t> var;
    std::map<int, float> f32;
    std::map<int, std::string> bytes;
    int32_t i32(int f) const { auto it = var.find(f); return it == var.end() ? 0 : int32_t(it->second); }
    float flt(int f) const { auto it = f32.find(f); return it == f32.end() ? 0.0f : it->second; }
    std::string str(int f) const { auto it = bytes.find(f); return it == bytes.end() ? "" : it->second; }
    std::vector<float> floats(int f) const {
        std::vector<float> v;
        const std::string s = str(f);
        v.resize(s.size() / 4);
        if (!v.empty()) std::memcpy(v.data(), s.data(), v.size() * 4);
        return v;
    }
};

inline Msg parse(const std::string& s) {
    Msg m;
    size_t i = 0;
    auto varint = [&](uint64_t& v) {
        v = 0;
        for (int sh = 0; i < s.size(); sh += 7) {
            const uint8_t c = uint8_t(s[i++]);
            v |= uint64_t(c & 0x7F) << sh;
            if (!(c & 0x80)) return;
        }
    };
    while (i < s.size()) {
        uint64_t k;
        varint(k);
        const int field = int(k >> 3), type = int(k & 7);
        if (type == 0) {
            uint64_t v;
            varint(v);
            m.var[field] = v;
        } else if (type == 5) {
            float f;
            std::memcpy(&f, s.data() + i, 4);
            i += 4;
            m.f32[field] = f;
        } else if (type == 2) {
            uint64_t len;
            varint(len);
            m.bytes[field] = s.substr(i, size_t(len));
            i += size_t(len);
        } else {
            std::fprintf(stderr, "unexpected wire type %d\n", type);
            std::exit(3);
        }
    }
    return m;
}

}  // namespace wire

// ---- in-process REQ/REP channel ---------------------------------------------
struct Channel {
    std::mutex m;
    std::condition_variable cv_req, cv_rep;
    std::deque<std::pair<int, std::string>> requests;
    std::map<int, std::string> replies;
    bool closed = false;

    // nn_send followed by nn_recv on a REQ socket
    std::string request(int sock, const std::string& msg) {
        std::unique_lock<std::mutex> l(m);
        requests.emplace_back(sock, msg);
        cv_req.notify_one();
        cv_rep.wait(l, [&] { return replies.count(sock) > 0; });
        std::string r = replies[sock];
        replies.erase(sock);
        return r;
    }
    bool next(std::pair<int, std::string>& req) {
        std::unique_lock<std::mutex> l(m);
        cv_req.wait(l, [&] { return !requests.empty() || closed; });
        if (requests.empty()) return false;
        req = requests.front();
        requests.pop_front();
        return true;
    }
    void reply(int sock, const std::string& msg) {
        std::lock_guard<std::mutex> l(m);
        replies[sock] = msg;
        cv_rep.notify_all();
    }
    void close() {
        std::lock_guard<std::mutex> l(m);
        closed = true;
        cv_req.notify_all();
    }
};

// ---- the fake parameter server -------------------------------------------
struct Result {
    int idx = -1;
    std::string node;
    float chi2 = 0.0f;
    bool hist = false;
    std::vector<float> shist;
};

struct FakeServer {
    int n, N;
    std::vector<std::vector<float>> pop;
    size_t next = 0, done = 0;
    float best = 3.1e32f;
    std::vector<float> ohist;
    std::vector<Result> results;

    std::string handle(const std::string& req) {
        const char kind = req[0];
        const wire::Msg m = wire::parse(req.substr(1));
        wire::Writer w;
        if (kind == 'c') {  // p_cap -> p_n
            w.i32(1, n);
        } else if (kind == 'p') {  // p_str (gpu node id, observed histogram once) -> p_ga
            if (m.i32(1) > 0) ohist = m.floats(3);
            if (next < pop.size()) {
                w.i32(1, int32_t(next));
                w.i32(2, 0);
                w.i32(3, N);
                w.f32(4, best);
                w.packed(5, pop[next]);
                ++next;
            } else {
                w.i32(1, done < pop.size() ? -1 : -2);
                w.f32(4, best);
            }
        } else if (kind == 'r') {  // res -> p_sid
            Result r;
            r.idx = m.i32(1);
            r.chi2 = m.flt(2);
            r.hist = m.i32(3) != 0;
            r.node = m.str(4);
            r.shist = m.floats(5);
            if (r.chi2 < best) best = r.chi2;
            results.push_back(r);
            ++done;
            w.i32(1, done < pop.size() ? 0 : -1);
        } else if (kind == 'k') {  // keepalive
            w.i32(1, 0);
        }
        return w.b;
    }
};

// ---- the worker side (the reference's two loops, same state machine) --------
struct Shared {
    int streams;
    std::vector<std::mutex> m;
    std::vector<std::condition_variable> cv;
    std::vector<int> dataready, s_n, ga_start, ga_stop, N;
    std::vector<std::vector<float>> params;
    explicit Shared(int s)
        : streams(s), m(size_t(s)), cv(size_t(s)), dataready(size_t(s), 0), s_n(size_t(s), 0),
          ga_start(size_t(s), 0), ga_stop(size_t(s), 0), N(size_t(s), 0), params(size_t(s)) {}
};

static std::vector<float> hist_of(const float* x, size_t n, int B) {
    std::vector<float> h(size_t(B), 0.0f);  // boost regular(B, 0, 1), inner bins
    for (size_t i = 0; i < n; ++i)
        if (x[i] >= 0.0f && x[i] < 1.0f) h[size_t(std::floor(double(x[i]) * B))] += 1.0f;
    return h;
}

// src/gpuWorker.cpp:38-88
static void gpu_loop(mc* pdamc, Shared* sh) {
    pdamc->set_gpuid();
    do {
        for (int sid = 0; sid < sh->streams; ++sid) {
            std::unique_lock<std::mutex> lck(sh->m[size_t(sid)], std::defer_lock);
            if (!lck.try_lock()) continue;
            int& dr = sh->dataready[size_t(sid)];
            if (dr == 3) {
                if (pdamc->streamQuery(sid)) {
                    dr = 4;
                    lck.unlock();
                    sh->cv[size_t(sid)].notify_one();
                }
                continue;
            }
            if (dr == 4 || dr == 0 || dr == 1) continue;
            if (sh->cv[size_t(sid)].wait_for(lck, 500ms, [&] { return dr == 2; })) {
                const int oldS_n = pdamc->set_nstates(sh->s_n[size_t(sid)], sid);
                pdamc->set_params(sh->s_n[size_t(sid)], sid, sh->params[size_t(sid)]);
                const int N_sid = pdamc->setBurstBd(sh->ga_start[size_t(sid)], sh->ga_stop[size_t(sid)], sid);
                if (N_sid != sh->N[size_t(sid)]) {
                    pdamc->init_randstate(N_sid, sid);
                    sh->N[size_t(sid)] = N_sid;
                }
                if (oldS_n != sh->s_n[size_t(sid)] || N_sid != sh->N[size_t(sid)])
                    pdamc->set_params_buff(oldS_n, N_sid, sid);
                pdamc->run_kernel(sh->N[size_t(sid)], sid);
                dr = 3;
            }
        }
    } while (pdamc->workerNum.load() > 0);
}

// src/streamWorker.cpp:39-248
static void stream_loop(mc* pdamc, Shared* sh, Channel* ch, int sid, int sz_burst, int leftPad,
                        const std::vector<float>* SgDivSr, int fretHistNum) {
    const std::string node = std::to_string(pdamc->devid) + ":" + std::to_string(sid) + ":" +
                             std::string(pdamc->gpuuuid);
    const std::vector<float> vO = hist_of(SgDivSr->data(), SgDivSr->size(), fretHistNum);
    bool sent_ohist = false, ending = false;
    float bestcs = 3.1e32f;
    int params_idx = 0;
    unsigned long keepalive = 0;
    do {
        std::unique_lock<std::mutex> lck(sh->m[size_t(sid)], std::defer_lock);
        if (!lck.try_lock()) continue;
        int& dr = sh->dataready[size_t(sid)];
        if (dr == 0) {
            wire::Writer w;
            w.i32(1, sz_burst);
            const wire::Msg r = wire::parse(ch->request(sid, "c" + w.b));
            sh->s_n[size_t(sid)] = r.i32(1);
            dr = 1;
            continue;
        }
        if (dr == 1) {
            wire::Writer w;
            w.str(2, node);
            if (!sent_ohist) {
                w.i32(1, fretHistNum);
                w.packed(3, vO);
                sent_ohist = true;
            }
            const wire::Msg ga = wire::parse(ch->request(sid, "p" + w.b));
            bestcs = ga.flt(4);
            const int pidx = ga.i32(1);
            if (pidx == -1) continue;
            if (pidx == -2) break;
            params_idx = pidx;
            sh->params[size_t(sid)] = ga.floats(5);
            sh->params[size_t(sid)].resize(size_t(sh->s_n[size_t(sid)] * (sh->s_n[size_t(sid)] + 1)));
            sh->ga_start[size_t(sid)] = ga.i32(2);
            sh->ga_stop[size_t(sid)] = ga.i32(3);
            dr = 2;
            lck.unlock();
            sh->cv[size_t(sid)].notify_one();
            continue;
        }
        if (!sh->cv[size_t(sid)].wait_for(lck, 500ms, [&] { return dr == 4; })) {
            if (++keepalive % 60 == 0) {
                wire::Writer w;
                w.i32(1, params_idx);
                ch->request(sid, "k" + w.b);
            }
            continue;
        }
        // the result: histogram of hmcE and chi2 (src/streamWorker.cpp:184-206)
        const int Nsid = sh->N[size_t(sid)];
        const std::vector<float> vM =
            hist_of(pdamc->hmcE[sid], size_t(Nsid) * size_t(pdamc->reSampleTimes), fretHistNum);
        int effN = fretHistNum - leftPad;
        float chisqr = 0.0f;
        for (int b = leftPad; b < fretHistNum; ++b) {
            if (vM[size_t(b)] > 0)
                chisqr += std::pow(float(vO[size_t(b)] - vM[size_t(b)] / pdamc->reSampleTimes), 2) /
                          (float(vM[size_t(b)]) / pdamc->reSampleTimes);
            else
                --effN;
        }
        const int sn = sh->s_n[size_t(sid)];
        chisqr = chisqr / (effN - sn * (sn + 1));
        wire::Writer w;
        w.i32(1, params_idx);
        w.f32(2, chisqr);
        const bool send_hist = chisqr < bestcs;
        w.boolean(3, send_hist);
        w.str(4, node);
        if (send_hist) {
            std::vector<float> s;
            for (float x : vM) s.push_back(x / pdamc->reSampleTimes);
            w.packed(5, s);
        }
        const wire::Msg r = wire::parse(ch->request(sid, "r" + w.b));
        if (r.i32(1) == -1) ending = true;
        dr = 0;
    } while (!ending);
    pdamc->workerNum--;
}

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
    if (argc != 8) return 2;
    const std::string dir = argv[1];
    const int n = std::atoi(argv[3]), R = std::atoi(argv[4]);
    const unsigned seed = unsigned(std::strtoul(argv[5], nullptr, 10));
    const int streams = std::atoi(argv[6]), nsets = std::atoi(argv[7]);
    std::vector<int64_t> start = load<int64_t>(dir + "/start"), stop = load<int64_t>(dir + "/stop");
    std::vector<int64_t> istart = load<int64_t>(dir + "/istart");
    std::vector<int32_t> ph32 = load<int32_t>(dir + "/phcount");
    std::vector<int> phCount(ph32.begin(), ph32.end());
    std::vector<int64_t> burst_ad = load<int64_t>(dir + "/burst_ad");
    std::vector<int64_t> burst_dd = load<int64_t>(dir + "/burst_dd");
    std::vector<float> T = load<float>(dir + "/T"), SgDivSr = load<float>(dir + "/SgDivSr");
    std::vector<float> sc = load<float>(dir + "/scalars"), pa = load<float>(dir + "/pop_args");
    float clk_p = sc[0], bg_ad = sc[1], bg_dd = sc[2];
    const int np = n * (n + 1);
    if (int(pa.size()) != nsets * np) return 4;

    // main (src/main.cpp:129-160)
    mc pdamc(0, streams);
    int sn = 0;
    pdamc.getStreams(&sn);
    pdamc.init_data_gpu(istart, start, stop, phCount, long(burst_ad.size()), burst_ad.data(),
                        burst_dd.data(), T, SgDivSr, clk_p, bg_ad, bg_dd);
    pdamc.set_seed(seed, true);
    pdamc.set_reSampleTimes(R);
    Shared sh(streams);
    if (!pdamc.get_max_conversion_capacity()) return 3;

    FakeServer server;
    server.n = n;
    server.N = pdamc.sz_burst;
    for (int i = 0; i < nsets; ++i)
        server.pop.emplace_back(pa.begin() + i * np, pa.begin() + (i + 1) * np);
    Channel ch;
    std::thread srv([&] {
        std::pair<int, std::string> req;
        while (ch.next(req)) ch.reply(req.first, server.handle(req.second));
    });
    const int fretHistNum = 200, leftPad = 2;  // cmdline default, src/main.cpp:155
    std::vector<std::thread> net;
    for (int i = 0; i < streams; ++i)
        net.emplace_back(stream_loop, &pdamc, &sh, &ch, i, pdamc.sz_burst, leftPad, &SgDivSr,
                         fretHistNum);
    std::thread gpu(gpu_loop, &pdamc, &sh);
    for (auto& t : net) t.join();
    gpu.join();
    ch.close();
    srv.join();

    FILE* f = std::fopen(argv[2], "w");
    std::fprintf(f, "ohist");
    for (float x : server.ohist) std::fprintf(f, " %.9g", x);
    std::fprintf(f, "\n");
    for (const Result& r : server.results) {
        const int sid = std::atoi(r.node.substr(r.node.find(':') + 1).c_str());
        std::fprintf(f, "set %d %d %.9g %d", r.idx, sid, double(r.chi2), r.hist ? 1 : 0);
        for (float x : r.shist) std::fprintf(f, " %.9g", x);
        std::fprintf(f, "\n");
    }
    std::fclose(f);
    return 0;
}
