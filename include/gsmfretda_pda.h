/*
 * gsmfretda_pda.h -- C ABI of the MI355X (gfx950) Monte-Carlo PDA hot path.
 *
 * Drop-in boundary for gSMFRETda's `class mc` device API (src/mc.hpp:32-180,
 * host half of src/mc.cu:204-829).  Plain pointers and sizes only.  Every
 * entry point returns 0 on success and a negative PDA_E* code on failure;
 * pda_last_error() then describes the failure (thread-local).  Host arrays
 * are copied before the call returns unless stated otherwise.
 *
 * Threading (as the reference, src/mc.cu:199-203, src/gpuWorker.cpp:39):
 * calls that mutate a stream slot must come from one thread at a time.
 */
#ifndef GSMFRETDA_PDA_H
#define GSMFRETDA_PDA_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PDA_OK 0
#define PDA_E_ARG -1     /* bad argument (size, range, null pointer)        */
#define PDA_E_STATE -2   /* call out of order (no data, no params, ...)     */
#define PDA_E_DATA -3    /* burst data violates a precondition              */
#define PDA_E_HIP -4     /* HIP runtime error                               */
#define PDA_E_NOMEM -5   /* allocation failed                               */

#define PDA_MAX_STATES 8
#define PDA_MAX_BINS 1024

/* pda_run flags */
#define PDA_RUN_SAMPLES 1 /* also keep per-sample E (and acceptor count, F)  */
#define PDA_RUN_NO_GRAPH 2 /* population runs: launch without graph capture */
#define PDA_RB_SCORES 1    /* RB forward: keep the scores rate gradients need */

typedef struct pda_ctx pda_ctx;

/* Description of the last failure on the calling thread ("" if none). */
const char *pda_last_error(void);

/* Library version string. */
const char *pda_version(void);

/* Replaces mc::mc(devid, streamNum, ...) (src/mc.cu:212-315): selects the GPU
 * and creates `nstreams` stream slots (0 = 4).  Each slot evaluates its own
 * parameter set, as the reference's per-stream buffers (src/mc.cu:235-242). */
int pda_create(int device, int nstreams, pda_ctx **out);

/* Replaces mc::~mc (src/mc.cu:618-663): waits for every slot, frees all. */
int pda_destroy(pda_ctx *ctx);

/* Number of stream slots and their hipStream_t handles (as void*), replaces
 * mc::getStreams (src/mc.cu:317-322). */
int pda_num_streams(const pda_ctx *ctx);
void *pda_stream(pda_ctx *ctx, int sid);

/* Replaces mc::init_data_gpu (src/mc.cu:456-521) fed by burst_data
 * (src/loadHdf5.cpp:101-128): the HDF5 datasets of Appendix-B layout
 * (/sub_bursts_l/start, stop, istart; photon count istop-istart+1; times,
 * mask_ad, mask_dd; T_burst_duration; clk_p; bg_ad_rate; bg_dd_rate).
 * The bursts are re-encoded once into burst-relative uint32 ticks plus
 * per-photon channel-run links and kept resident.  Preconditions (else
 * PDA_E_DATA): 0 <= istart, istart+phcount <= nphotons; every photon of a
 * channel lies in [start, start + 2^31) ticks; channel photons of a burst
 * are in non-decreasing time order; 0 <= stop-start < 2^31; T finite;
 * clk_p > 0; nbursts*R fits the run's 2^31 unit limit.  Replaces any
 * previously loaded data (waits for all slots first). */
int pda_load_bursts(pda_ctx *ctx, const int64_t *start, const int64_t *stop,
                    const int64_t *istart, const int32_t *phcount, int64_t nbursts,
                    const int64_t *times, const uint8_t *mask_ad, const uint8_t *mask_dd,
                    int64_t nphotons, const float *T, float clk_p, float bg_ad, float bg_dd);

/* The same, from the reference's "time if in channel, else 0" photon arrays
 * burst_ad / burst_dd (exactly what mc::init_data_gpu receives,
 * src/mc.hpp:158-161).  A photon with both entries 0 belongs to neither
 * channel. */
int pda_load_bursts_tagged(pda_ctx *ctx, const int64_t *istart, const int64_t *start,
                           const int64_t *stop, const int32_t *phcount, int64_t nbursts,
                           const int64_t *burst_ad, const int64_t *burst_dd, int64_t nphotons,
                           const float *T, float clk_p, float bg_ad, float bg_dd);

int64_t pda_num_bursts(const pda_ctx *ctx);

/* Observed histogram from the per-burst measured efficiencies SgDivSr
 * (src/streamWorker.cpp:62-67): nbins regular bins on [0,1), values outside
 * dropped; chi2 skips the first left_pad bins (src/main.cpp:155). */
int pda_set_observed(pda_ctx *ctx, const float *sgdivsr, int64_t n, int nbins, int left_pad);
/* Or an observed histogram directly (nbins counts). */
int pda_set_observed_hist(pda_ctx *ctx, const float *ohist, int nbins, int left_pad);

/* Replaces mc::set_params (src/mc.cu:790-829) and the matrix builders
 * genMatK / genP_i2j / genMatP (src/eigenhelper.cpp:4-54):
 * args = [E(n), K off-diagonal (n*(n-1), row-major over the target state,
 * skipping the diagonal), v(n)].  2 <= n <= PDA_MAX_STATES; nargs must be
 * n*(n+1) (the reference returns false; here PDA_E_ARG).  Unlike the
 * reference it does not synchronize the stream. */
int pda_set_params(pda_ctx *ctx, int sid, int n, const float *args, int nargs);

/* Photophysics constants, mc_kernel's default arguments (src/mc.cu:89-91):
 * gamma 0.34, DexDirAem 0.08, Dch2Ach 0.07, r0 52 until set.  PDA_E_ARG
 * unless finite, r0 > 0, gamma >= crosstalk and dexdir <= 1 (otherwise the
 * corrected photon count, the binomial's trial count, can be negative). */
int pda_set_photophysics(pda_ctx *ctx, int sid, float gamma, float dexdir, float crosstalk,
                         float r0);

/* Replaces mc::setBurstBd (src/mc.cu:523-565): bursts [begin, end) for the
 * slot; end < 0 or end > nbursts means nbursts; begin >= end means 0.
 * Writes the number of bursts to *n_out (may be NULL). */
int pda_set_burst_range(pda_ctx *ctx, int sid, int64_t begin, int64_t end, int64_t *n_out);

/* Replaces mc::run_kernel (src/mc.cu:567-587), asynchronous on the slot's
 * stream: R resamples of every burst in the range, the integer E histogram
 * and its chi2 against the observed histogram, all on the GPU.  Random
 * numbers are keyed by (seed, set_id, burst, resample). */
int pda_run(pda_ctx *ctx, int sid, int R, uint32_t seed, uint32_t set_id, int flags);

/* Replaces mc::streamQuery (src/mc.cu:595-601): 1 when the slot's last run
 * (and its result copy) finished, 0 while it runs, <0 on error. */
int pda_query(pda_ctx *ctx, int sid);
int pda_synchronize(pda_ctx *ctx, int sid);

/* Result of the slot's last completed run: int32 histogram of nbins bins,
 * reduced chi2 in FP64 and with the reference's float accumulation
 * (src/streamWorker.cpp:197-206), number of bins kept.  Any output may be
 * NULL.  Call after pda_query returned 1 or pda_synchronize. */
int pda_get_result(pda_ctx *ctx, int sid, int32_t *hist, int nbins, double *chi2_64,
                   float *chi2_ref, int *effN);

/* Facts of the slot's last run (waits for it; no reference counterpart):
 * units simulated and units the fast pass abandoned to the general-form
 * replay pass (BTPE draws, inversion restarts, background means above 64,
 * window ends past 2^31 ticks).  Outputs may be NULL. */
int pda_run_stats(pda_ctx *ctx, int sid, int64_t *units, int32_t *replays);

/* Per-sample values of the last run made with PDA_RUN_SAMPLES, in unit
 * order (burst - begin) * R + r: E (the reference's mcE / hmcE), acceptor
 * count and corrected photon count F.  Any output may be NULL. */
int pda_get_samples(pda_ctx *ctx, int sid, float *E, int32_t *acc, float *F, int64_t n);

/* Host pointer of the slot's pinned per-sample E buffer (the reference's
 * hmcE[sid], src/mc.cu:559-562); valid after a PDA_RUN_SAMPLES run
 * completed, NULL before. */
float *pda_samples_host(pda_ctx *ctx, int sid);

/* Milliseconds spent by the slot's last run on the GPU (HIP events). */
int pda_last_run_ms(pda_ctx *ctx, int sid, float *ms);
/* The same split into the walk pass over every unit and the finishing pass
 * (replayed units, histogram reduction, chi2). */
int pda_last_pass_ms(pda_ctx *ctx, int sid, float *walk_ms, float *finish_ms);

/* Launch shape of the walk kernels, single-set and population runs (no
 * reference counterpart; the reference sizes its grid with
 * cudaOccupancyMaxPotentialBlockSize, src/mc.cu:309): batch_tiles x 64 units
 * per work-counter grab (0 = sized from the run, about six grabs per
 * wave), resident one-wave workgroups per CU (0 = as many as fit), work
 * counters, each on its own 256-byte line (0 = 32, four per XCD; at most
 * 64).  Results do not depend on the shape. */
int pda_set_launch_shape(pda_ctx *ctx, int batch_tiles, int waves_per_cu, int queues);
int pda_get_launch_shape(const pda_ctx *ctx, int *batch_tiles, int *waves_per_cu, int *queues);

/* Autotuner (no reference counterpart): coordinate search over the launch
 * shape (batch size, then waves per CU, then work counters) on the slot's
 * current work -- its parameter set (population = 0) or its population
 * (population = 1) -- with R resamples, each candidate timed `reps` times
 * with HIP events after one untimed run; keeps the fastest shape in the
 * context and returns it (outputs may be NULL).  Overwrites the slot's last
 * results. */
int pda_autotune(pda_ctx *ctx, int sid, int population, int R, int reps, int *batch_tiles,
                 int *waves_per_cu, int *queues, float *best_ms);

/* Population of parameter sets (no reference counterpart: the reference
 * evaluates one set per stream and mc::run_kernel call, the parameter server
 * feeding a generation's sets to gpuWorker::run one at a time,
 * src/gpuWorker.cpp:38-88, src/streamWorker.cpp:83-182).  nstates[i] in
 * [2, 8] per set; args holds the sets' mc::set_params argument vectors
 * (src/mc.cu:790-800) back to back, sum of n(n+1) floats.  Validates every
 * set (PDA_E_ARG naming the set) before replacing the slot's population; the
 * slot's photophysics constants apply to every set. */
int pda_set_population(pda_ctx *ctx, int sid, int nsets, const int32_t *nstates,
                       const float *args, int64_t nargs_total);

/* Evaluates the slot's population asynchronously on its stream: every set
 * over the slot's burst range with R resamples, set i drawing with the key
 * (seed, set_base + i * set_stride) -- bit-identical to pda_run of that set
 * with that set id.  The launch sequence (model upload, walk, replay,
 * per-set histogram and chi2, result download) is captured into a hipGraph
 * on first use and replayed while the shape of the work stays the same;
 * flags & PDA_RUN_NO_GRAPH launches it directly. */
int pda_run_population(pda_ctx *ctx, int sid, int R, uint32_t seed, uint32_t set_base,
                       uint32_t set_stride, int flags);

/* Results of the slot's last population run (waits for it): int32
 * histograms [nsets][nbins], reduced chi2 per set in FP64 and with the
 * reference's float accumulation, bins kept per set.  Outputs may be NULL. */
int pda_get_population_result(pda_ctx *ctx, int sid, int32_t *hist, int nbins, double *chi2_64,
                              float *chi2_ref, int32_t *effN);

/* Facts of the slot's last population run (waits for it): units per set,
 * units replayed in the general form, graph captures so far, GPU
 * milliseconds of the whole sequence.  Outputs may be NULL. */
int pda_population_stats(pda_ctx *ctx, int sid, int64_t *units_per_set, int32_t *replays,
                         int32_t *graph_captures, float *ms);

/* Rao-Blackwellised (RB) form (no reference counterpart: the reference draws
 * every acceptor count, src/mc.cu:167-170, and is not differentiable).
 * Requires v = 0 in every state and n <= 4.  The walk of every unit uses the
 * keyed draws of pda_run with the same (seed, set_id); instead of the
 * acceptor draw, the exact distribution of the unit's acceptor total -- a
 * convolution of Binomial(M_s, E_s) over the states, M_s the photon counts of
 * the windows spent in state s -- is spread into the E histogram, giving the
 * expected counts M_b (accumulated exactly in 2^-32 fixed point, so results
 * do not depend on the launch shape).  chi2 is the reference's reduced chi2
 * (src/streamWorker.cpp:197-206) on M_b / R.  flags & PDA_RB_SCORES keeps,
 * per unit, the score of the walk's draws for pda_rb_adjoint.
 * Asynchronous on the slot's stream. */
int pda_rb_forward(pda_ctx *ctx, int sid, int R, uint32_t seed, uint32_t set_id, int flags);

/* Result of the slot's last RB forward run (waits for it): M_b in 2^-32
 * units and as doubles, chi2 and bins kept.  Outputs may be NULL. */
int pda_rb_get_result(pda_ctx *ctx, int sid, uint64_t *M_fixed, double *M, int nbins,
                      double *chi2, int32_t *effN);

/* Adjoint of the slot's last RB forward run (blocking): for a cotangent
 * g[b] = d chi2 / d M_b, writes grad[nargs] in mc::set_params layout
 * (src/mc.cu:790-800): d/dE_s exactly for the RB estimator, d/dK for the
 * off-diagonal rates by the likelihood ratio of the walk's state and dwell
 * draws with a per-burst leave-one-out baseline (the photon counts are step
 * functions of the dwell times, so their pathwise derivative is zero almost
 * everywhere), 0 for v.  Needs a forward run with PDA_RB_SCORES, R >= 2,
 * 0 < E < 1, rates > 0 and a fixed initial state. */
int pda_rb_adjoint(pda_ctx *ctx, int sid, const double *g, int nbins, double *grad, int nargs);

/* Jacobian of the slot's last RB forward run (blocking): J[b * nargs + i] =
 * d M_b / d args_i for every bin at once, M_b the expected counts of
 * pda_rb_get_result (summed over the R resamples), in the layout and with the
 * estimators of pda_rb_adjoint, so sum_b g[b] J[b * nargs + i] is the
 * adjoint's grad[i] for any cotangent g.  For Gauss-Newton / Levenberg-
 * Marquardt fits of the binned chi2.  Same preconditions as pda_rb_adjoint;
 * the per-bin scratch (about 170 B per bin) must fit the LDS next to the
 * pmf buffers. */
int pda_rb_jacobian(pda_ctx *ctx, int sid, double *J, int nbins, int nargs);

/* RB population (no reference counterpart; the reference evaluates one set
 * per mc::run_kernel and differentiates none): the RB forward of every set of
 * the slot's population (pda_set_population; every set n <= 4, v = 0), set k
 * with the key (seed, set_base + k * set_stride), each bit-identical to
 * pda_rb_forward of that set with that set id.  Every set keeps its own walk
 * summaries (and, with PDA_RB_SCORES, its own scores and pmf store), so
 * pda_rb_adjoint_population / pda_rb_jacobian_population differentiate all
 * of them; the passes of all sets are enqueued back to back on the slot's
 * stream and waited for once.  Asynchronous. */
int pda_rb_forward_population(pda_ctx *ctx, int sid, int R, uint32_t seed, uint32_t set_base,
                              uint32_t set_stride, int flags);

/* Results of the slot's last RB population forward (waits for it): M and M
 * in 2^-32 units as [nsets][nbins], chi2 and bins kept per set.  Outputs may
 * be NULL. */
int pda_rb_get_population_result(pda_ctx *ctx, int sid, uint64_t *M_fixed, double *M, int nbins,
                                 double *chi2, int32_t *effN);

/* Adjoint of every set of the last RB population forward (blocking): for
 * cotangents g[nsets][nbins], grad[nsets][nargs] as pda_rb_adjoint; every set
 * has the same nargs. */
int pda_rb_adjoint_population(pda_ctx *ctx, int sid, const double *g, int nbins, double *grad,
                              int nargs);

/* Jacobians of the sets of the last RB population forward (blocking):
 * J[nsets][nbins][nargs] as pda_rb_jacobian, for the sets with which[k] != 0
 * (which may be NULL: every set); the others' rows are zero. */
int pda_rb_jacobian_population(pda_ctx *ctx, int sid, const int32_t *which, double *J, int nbins,
                               int nargs);

/* GPU milliseconds of the slot's last RB runs: the walk, the whole forward
 * sequence, the last adjoint or Jacobian kernel (0 before the first).
 * Outputs may be NULL. */
int pda_rb_stats(pda_ctx *ctx, int sid, float *walk_ms, float *forward_ms, float *adjoint_ms);

/* Device-side result of pda_run_noalloc: this header, then int32 hist[nbins]
 * (the layout pda_get_result reads from). */
typedef struct pda_result {
    double chi2_64;   /* reduced chi2 in FP64                                  */
    float chi2_ref;   /* reduced chi2, the reference's float accumulation      */
    int32_t effN;     /* bins kept                                             */
    int64_t units;    /* (burst, resample) units simulated                     */
    int32_t nbins;
    int32_t replays;  /* units the fast pass abandoned to the replay pass      */
} pda_result;

/* Largest bursts x R of one pda_run_noalloc call. */
#define PDA_NOALLOC_MAX_UNITS 16000000

/* No reference counterpart: pda_run for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Same evaluation, random streams and result as pda_run of the
 * slot's parameter set over its burst range, asynchronous on the slot's
 * stream; no per-sample output and no result copy to the host: the result
 * (pda_result, then int32 hist[nbins]) is written to the caller's device
 * buffer d_result of result_bytes (at least sizeof(pda_result) + 4 * nbins).
 * The run's scratch (work counters, replay list, histogram rows) is a fixed
 * 64 MiB device array that each GPU holds until the process ends, so a call
 * takes at most PDA_NOALLOC_MAX_UNITS units (else PDA_E_ARG) and calls no
 * allocator.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (any slot or context) race and must be serialised by
 * the caller. */
int pda_run_noalloc(pda_ctx *ctx, int sid, int R, uint32_t seed, uint32_t set_id,
                    void *d_result, int64_t result_bytes);

/* Diagnostics (no reference counterpart): evaluates one primitive of the
 * shared variate library on the GPU for n inputs, so tests can compare it
 * bit for bit with the host.  op: 0 philox (in: 6 u32 per item -> 4 u32),
 * 1 flog, 2 fexp, 3 dlog, 4 dexp, 5 cos_turns, 6 gauss (2 u32 -> f32),
 * 7 binomial (in: n i32, p f64, key 4 u32, w u32, words 2 u32 -> i64),
 * 8 background_count (in: lam, cap, u1, u2 f32, key 4 u32, w, sub -> f32),
 * 9 plain copy of u32 words (a kernel of known traffic, for counter
 * calibration).
 * in/out are host arrays; the layout per op is documented in pda_capi.cpp. */
int pda_debug_primitive(pda_ctx *ctx, int op, const void *in, void *out, int64_t n);

/* Diagnostics (no reference counterpart): microbenchmark of one event of the
 * walk algorithm, the weights of the algorithmic cost model.  Every lane of a
 * full-occupancy grid repeats the event `iters` times with a loop-carried
 * dependence: op 0 a Philox4x32-10 block, 1 a dwell window's scalar work,
 * 2 a window-end binary-search step, 3 a capped-Poisson draw of mean
 * `param` (0..64), 4 a binomial inversion draw of n = `param` (0..127)
 * trials at p = 0.3 from a q^m table, 5 a unit's start and tally, 6 one
 * outcome of the Rao-Blackwellised spread (its E bin and fixed-point
 * weight; `param` shifts F).  Writes
 * the GPU milliseconds of one launch (after an untimed one), the events run
 * and the inner iterations (Poisson / inversion iterations; 0 otherwise). */
int pda_debug_microbench(pda_ctx *ctx, int op, int iters, float param, float *ms,
                         int64_t *events, int64_t *inner);

/* Diagnostics (no reference counterpart): copies slot `sid`'s threshold rows
 * of the fast walk's binomial draws, as its last parameter upload built them
 * on the GPU, to host memory: `thr` receives 8 * 8,384 pairs of int64
 * (up, lo), the rows of each state packed with m (m + 3) / 2 pairs before the
 * row of photon count m, m + 2 pairs per row (thr_bytes at least 1,073,152),
 * `guide` 8 * 128 * 64 bytes (guide_bytes at least 65,536).  Rows of states
 * beyond the set's are not written.  Blocks until the slot's stream is
 * idle. */
int pda_debug_guide(pda_ctx *ctx, int sid, void *thr, int64_t thr_bytes, void *guide,
                    int64_t guide_bytes);

#ifdef __cplusplus
}
#endif
#endif
