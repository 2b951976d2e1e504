/*
 * pda_oracle.c -- CPU oracle for the Monte-Carlo PDA hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.
 * The product never links it.  Build: oracle/Makefile (-ffp-contract=off,
 * no fast-math, so every float operation is one IEEE operation).
 *
 * Literal restatement of gSMFRETda mc_kernel (src/mc.cu:83-197).  The
 * reference draws from curand Sobol64 + a numpy-legacy xorshift; here every
 * draw comes from a keyed Philox4x32-10 block instead (key = {seed, set},
 * counter = {window, sub-stream, resample, burst}):
 *   initial state      ctr {0xFFFFFFFF, 0, r, b}   word 0
 *   window w, block A  ctr {w, 0, r, b}   A0 next state, A1 dwell time,
 *                                         A2:A3 first binomial uniform (53 bit)
 *   window w, block B  ctr {w, 1, r, b}   B0 / B1 first Poisson uniform A / D,
 *                                         B2 / B3 cap uniform A / D
 *   window w, block C  ctr {w, 2, r, b}   C0:C1 Box-Muller normal
 *   binomial extra i   ctr {w, 16+i, r, b} restart i: u53(X0,X1);
 *                                         BTPE iteration i: u53(X0,X1), u53(X2,X3)
 *   Poisson chunk c>=1 ctr {w, 0x100000 (A) / 0x200000 (D) + (c-1)/4, r, b}
 *                                         word (c-1)%4
 * The oracle draws every one of these exactly as the reference calls its
 * generators, even where the value cannot matter (n == 2 next state,
 * v == 0 broadening); the HIP kernel skips those and must still agree.
 */
#include "pda_oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).                  */
/* ------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void blockw(uint32_t seed, uint32_t set, uint32_t w, uint32_t sub, uint32_t r,
                   uint32_t b, uint32_t out[4], or_stats *st) {
    uint32_t ctr[4] = {w, sub, r, b};
    uint32_t key[2] = {seed, set};
    or_philox4x32_10(ctr, key, out);
    if (st) st->philox_blocks++;
}

/* (0,1] with 24 random bits, like curand_uniform's range. */
float or_u01f(uint32_t x) { return (float)((x >> 8) + 1u) * 0x1p-24f; }
/* [0,1) with 53 random bits, like numpy's rk_double (src/binom.cuh:15-19). */
double or_u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * 0x1p-53;
}

/* ------------------------------------------------------------------ */
/* Elementary functions from +,-,*,/ and fma only, so the host and the */
/* GPU round identically.  log: 2*atanh series around 1 after a binary */
/* exponent split; exp: Cody-Waite reduction by ln2 + Taylor series.   */
/* ------------------------------------------------------------------ */
static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint64_t d2u(double f) { uint64_t u; memcpy(&u, &f, 8); return u; }
static inline double u2d(uint64_t u) { double f; memcpy(&f, &u, 8); return f; }

#define LN2HI_F 0x1.62ep-1f
#define LN2LO_F 0x1.0bfbe8p-15f
#define LN2HI_D 0x1.62e42feep-1
#define LN2LO_D 0x1.a39ef35793c76p-33

float or_logf(float x) {
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return INFINITY;
    int e = 0;
    uint32_t b = f2u(x);
    if (b < 0x00800000u) { x = x * 0x1p25f; b = f2u(x); e = -25; }
    e += (int)(b >> 23) - 127;
    float m = u2f((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; e += 1; }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float R = z * fmaf(z, fmaf(z, fmaf(z, 0x1.c71c72p-3f, 0x1.24924ap-2f), 0x1.99999ap-2f),
                       0x1.555556p-1f);
    float hfsq = 0.5f * f * f;
    float ek = (float)e;
    return ek * LN2HI_F - ((hfsq - (s * (hfsq + R) + ek * LN2LO_F)) - f);
}

static float scale2f(float p, int k) { /* p * 2^k for k in [-160, 130] */
    int k1 = k / 2, k2 = k - k1;
    return (p * u2f((uint32_t)(k1 + 127) << 23)) * u2f((uint32_t)(k2 + 127) << 23);
}

float or_expf(float x) {
    if (x != x) return x;
    if (x > 88.72f) return INFINITY;
    if (x < -103.98f) return 0.0f;
    float kf = floorf(x * 0x1.715476p+0f + 0.5f);
    float r = (x - kf * LN2HI_F) - kf * LN2LO_F;
    float p = fmaf(r, 0x1.a01a02p-13f, 0x1.6c16c2p-10f);
    p = fmaf(r, p, 0x1.111112p-7f);
    p = fmaf(r, p, 0x1.555556p-5f);
    p = fmaf(r, p, 0x1.555556p-3f);
    p = fmaf(r, p, 0.5f);
    p = fmaf(r, p, 1.0f);
    p = fmaf(r, p, 1.0f);
    return scale2f(p, (int)kf);
}

double or_log(double x) {
    if (x != x || x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return INFINITY;
    int e = 0;
    uint64_t b = d2u(x);
    if (b < 0x0010000000000000ull) { x = x * 0x1p54; b = d2u(x); e = -54; }
    e += (int)(b >> 52) - 1023;
    double m = u2d((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 0x1.6a09e667f3bcdp+0) { m = m * 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double q = 0x1.642c8590b2164p-4;
    q = fma(z, q, 0x1.8618618618618p-4);
    q = fma(z, q, 0x1.af286bca1af28p-4);
    q = fma(z, q, 0x1.e1e1e1e1e1e1ep-4);
    q = fma(z, q, 0x1.1111111111111p-3);
    q = fma(z, q, 0x1.3b13b13b13b14p-3);
    q = fma(z, q, 0x1.745d1745d1746p-3);
    q = fma(z, q, 0x1.c71c71c71c71cp-3);
    q = fma(z, q, 0x1.2492492492492p-2);
    q = fma(z, q, 0x1.999999999999ap-2);
    q = fma(z, q, 0x1.5555555555555p-1);
    double R = z * q;
    double hfsq = 0.5 * f * f;
    double ek = (double)e;
    return ek * LN2HI_D - ((hfsq - (s * (hfsq + R) + ek * LN2LO_D)) - f);
}

static double scale2d(double p, int k) { /* p * 2^k for k in [-1100, 1030] */
    int k1 = k / 2, k2 = k - k1;
    return (p * u2d((uint64_t)(k1 + 1023) << 52)) * u2d((uint64_t)(k2 + 1023) << 52);
}

double or_exp(double x) {
    if (x != x) return x;
    if (x > 709.78) return INFINITY;
    if (x < -745.2) return 0.0;
    double kf = floor(x * 0x1.71547652b82fep+0 + 0.5);
    double r = (x - kf * LN2HI_D) - kf * LN2LO_D;
    double p = 0x1.6124613a86d09p-33;
    p = fma(r, p, 0x1.1eed8eff8d898p-29);
    p = fma(r, p, 0x1.ae64567f544e4p-26);
    p = fma(r, p, 0x1.27e4fb7789f5cp-22);
    p = fma(r, p, 0x1.71de3a556c734p-19);
    p = fma(r, p, 0x1.a01a01a01a01ap-16);
    p = fma(r, p, 0x1.a01a01a01a01ap-13);
    p = fma(r, p, 0x1.6c16c16c16c17p-10);
    p = fma(r, p, 0x1.1111111111111p-7);
    p = fma(r, p, 0x1.5555555555555p-5);
    p = fma(r, p, 0x1.5555555555555p-3);
    p = fma(r, p, 0.5);
    p = fma(r, p, 1.0);
    p = fma(r, p, 1.0);
    return scale2d(p, (int)kf);
}

/* cos(2*pi*u) for u in [0,1]: quadrant split of 4u, Taylor sin/cos on
 * |a| <= pi/4. */
float or_cos2pif(float u) {
    float y = 4.0f * u;
    float qf = floorf(y + 0.5f);
    float a = (y - qf) * 0x1.921fb6p+0f;
    float a2 = a * a;
    float sn = fmaf(a2, fmaf(a2, fmaf(a2, fmaf(a2, 0x1.71de3ap-19f, -0x1.a01a02p-13f),
                                      0x1.111112p-7f), -0x1.555556p-3f), 1.0f) * a;
    float cs = fmaf(a2, fmaf(a2, fmaf(a2, fmaf(a2, fmaf(a2, -0x1.27e4fcp-22f, 0x1.a01a02p-16f),
                                                -0x1.6c16c2p-10f), 0x1.555556p-5f), -0.5f),
                    1.0f);
    int q = ((int)qf) & 3;
    return q == 0 ? cs : q == 1 ? -sn : q == 2 ? -cs : sn;
}

/* Box-Muller normal from two stream words (replaces curand_normal in drawE,
 * src/gen_rand.cuh:63-68). */
float or_normal(uint32_t a, uint32_t b) {
    float u1 = or_u01f(a), u2 = or_u01f(b);
    return sqrtf(-2.0f * or_logf(u1)) * or_cos2pif(u2);
}

/* drawDisIdx (src/gen_rand.cuh:35-50): first i with cumulative p > u, else n-1. */
int or_categorical(int n, const float *p, float u) {
    float a = 0.0f;
    for (int i = 0; i < n; ++i) {
        a += p[i];
        if (a > u) return i;
    }
    return n - 1;
}

/* ------------------------------------------------------------------ */
/* Binomial: inversion (src/binom.cuh:135-169) for n < 128 or          */
/* min(p,1-p)*n <= 30, else BTPE (Kachitvichyanukul & Schmeiser 1988;  */
/* src/binom.cuh:21-134).  numpy's split (src/binom.cuh:172-188) uses  */
/* BTPE whenever min(p,1-p)*n > 30; both are exact samplers, and       */
/* q^n >= 2^-127 keeps the inversion start far above underflow.        */
/* ------------------------------------------------------------------ */
#define OR_MAX_RESTARTS 4096
#define OR_INV_MAX_N 128

static int64_t binom_inversion(int n, double p, double U0, uint32_t seed, uint32_t set,
                               uint32_t burst, uint32_t r, uint32_t w, or_stats *st) {
    double q = 1.0 - p;
    double qn = or_exp((double)n * or_log(q));
    double np = (double)n * p;
    double bd = np + 10.0 * sqrt(np * q + 1.0);
    if ((double)n < bd) bd = (double)n;
    int64_t bound = (int64_t)bd;
    int64_t X = 0;
    double px = qn, U = U0;
    uint32_t restart = 0;
    while (U > px) {
        X++;
        if (st) st->binom_inv_iters++;
        if (X > bound) {
            if (restart == OR_MAX_RESTARTS) return (int64_t)np;
            uint32_t x[4];
            blockw(seed, set, w, 16u + restart, r, burst, x, st);
            restart++;
            X = 0;
            px = qn;
            U = or_u53(x[0], x[1]);
        } else {
            U -= px;
            px = ((double)(n - X + 1) * p * px) / ((double)X * q);
        }
    }
    return X;
}

static double stirling_tail(double f) { /* log-gamma correction, K&S step 5.3 */
    double f2 = f * f;
    return (13860. - (462. - (132. - (99. - 140. / f2) / f2) / f2) / f2) / f / 166320.;
}

static int64_t binom_btpe(int n, double p, uint32_t seed, uint32_t set, uint32_t burst,
                          uint32_t r, uint32_t w, or_stats *st) {
    double q = 1.0 - p;
    double fm = (double)n * p + p;
    int64_t m = (int64_t)floor(fm);
    double nrq = (double)n * p * q;
    double p1 = floor(2.195 * sqrt(nrq) - 4.6 * q) + 0.5;
    double xm = (double)m + 0.5;
    double xl = xm - p1;
    double xr = xm + p1;
    double c = 0.134 + 20.5 / (15.3 + (double)m);
    double a = (fm - xl) / (fm - xl * p);
    double laml = a * (1.0 + 0.5 * a);
    a = (xr - fm) / (xr * q);
    double lamr = a * (1.0 + 0.5 * a);
    double p2 = p1 * (1.0 + 2.0 * c);
    double p3 = p2 + c / laml;
    double p4 = p3 + c / lamr;
    for (uint32_t it = 0; it < OR_MAX_RESTARTS; ++it) {
        uint32_t x[4];
        blockw(seed, set, w, 16u + it, r, burst, x, st);
        if (st) st->binom_btpe_iters++;
        double u = or_u53(x[0], x[1]) * p4;
        double v = or_u53(x[2], x[3]);
        int64_t y;
        if (u <= p1) { /* triangular region: always accepted */
            return (int64_t)floor(xm - p1 * v + u);
        } else if (u <= p2) { /* parallelograms */
            double xx = xl + (u - p1) / c;
            v = v * c + 1.0 - fabs((double)m - xx + 0.5) / p1;
            if (v > 1.0) continue;
            y = (int64_t)floor(xx);
        } else if (u <= p3) { /* left exponential tail */
            double yd = floor(xl + or_log(v) / laml);
            if (!(yd >= 0.0)) continue;
            y = (int64_t)yd;
            v = v * (u - p2) * laml;
        } else { /* right exponential tail */
            double yd = floor(xr - or_log(v) / lamr);
            if (!(yd <= (double)n)) continue;
            y = (int64_t)yd;
            v = v * (u - p3) * lamr;
        }
        int64_t k = y > m ? y - m : m - y;
        if (k <= 20 || (double)k >= nrq / 2.0 - 1.0) {
            /* explicit evaluation of f(y)/f(m) by the recurrence */
            double s = p / q;
            double aa = s * (double)(n + 1);
            double F = 1.0;
            if (m < y) {
                for (int64_t i = m + 1; i <= y; ++i) F *= (aa / (double)i - s);
            } else if (m > y) {
                for (int64_t i = y + 1; i <= m; ++i) F /= (aa / (double)i - s);
            }
            if (v > F) continue;
            return y;
        }
        /* squeeze on log(f(y)/f(m)) */
        double kd = (double)k;
        double rho = (kd / nrq) * ((kd * (kd / 3.0 + 0.625) + 0.1666666666666) / nrq + 0.5);
        double t = -kd * kd / (2.0 * nrq);
        double A = or_log(v);
        if (A < t - rho) return y;
        if (A > t + rho) continue;
        double x1 = (double)(y + 1), f1 = (double)(m + 1);
        double z = (double)(n + 1 - m), ww = (double)(n - y + 1);
        /* log f(y)/f(m) by Stirling: + w(m+1) + w(n-m+1) - w(y+1) - w(n-y+1).
         * (src/binom.cuh:118-124 adds all four terms and uses 13680 for
         * 13860; this is the series as derived.) */
        double bound = xm * or_log(f1 / x1) + ((double)(n - m) + 0.5) * or_log(z / ww) +
                       (double)(y - m) * or_log(ww * p / (x1 * q)) + stirling_tail(f1) +
                       stirling_tail(z) - stirling_tail(x1) - stirling_tail(ww);
        if (A > bound) continue;
        return y;
    }
    return m;
}

int64_t or_binomial(int n, double p, uint32_t seed, uint32_t set, uint32_t burst, uint32_t r,
                    uint32_t w, uint32_t a2, uint32_t a3, or_stats *st) {
    if (n <= 0 || !(p >= 0.0 && p <= 1.0)) return 0;
    int flip = p > 0.5;
    double pp = flip ? 1.0 - p : p;
    int64_t X;
    if (n < OR_INV_MAX_N || pp * (double)n <= 30.0)
        X = binom_inversion(n, pp, or_u53(a2, a3), seed, set, burst, r, w, st);
    else
        X = binom_btpe(n, pp, seed, set, burst, r, w, st);
    return flip ? (int64_t)n - X : X;
}

/* ------------------------------------------------------------------ */
/* draw_P_B_Tr (src/gen_rand.cuh:92-107): k ~ Poisson(lam); k if       */
/* k <= cap else floor(cap*u).  Poisson by inversion, in chunks of     */
/* mean <= 64 when lam is large; stops as soon as the count passes cap */
/* (the exact count is not needed then).                               */
/* ------------------------------------------------------------------ */
float or_bg_capped(float lam, float cap, float u_first, float u_cap, uint32_t seed,
                   uint32_t set, uint32_t burst, uint32_t r, uint32_t w, uint32_t sub_base,
                   or_stats *st) {
    int over = 0;
    int64_t total = 0;
    if (lam > 0.0f) {
        int J = lam <= 64.0f ? 1 : (int)ceilf(lam / 64.0f);
        float lc = J == 1 ? lam : lam / (float)J;
        float p0 = or_expf(-lc);
        uint32_t x[4];
        for (int c = 0; c < J && !over; ++c) {
            float u;
            if (c == 0) {
                u = u_first;
            } else {
                uint32_t q = (uint32_t)(c - 1);
                if ((q & 3u) == 0u) blockw(seed, set, w, sub_base + (q >> 2), r, burst, x, st);
                u = or_u01f(x[q & 3u]);
            }
            int64_t k = 0;
            float pk = p0, F = p0;
            while (u > F) {
                k++;
                if (st) st->poisson_iters++;
                if ((float)(total + k) > cap) { over = 1; break; }
                pk = (pk * lc) / (float)k;
                F = F + pk;
            }
            if (!over) total += k;
        }
    }
    if (!over && (float)total <= cap) return (float)total;
    return floorf(cap * u_cap);
}

/* ------------------------------------------------------------------ */
/* Parameters: arg split src/mc.cu:790-800, genMatK / genP_i2j /       */
/* genMatP src/eigenhelper.cpp:4-54.                                   */
/* ------------------------------------------------------------------ */
int or_set_params(or_params *p, int n, const float *args, int nargs, float gamma,
                  float dexdir, float crosstalk, float r0) {
    if (n < 2 || n > OR_MAXN || nargs != n * (n + 1)) return -1;
    memset(p, 0, sizeof(*p));
    p->n = n;
    const float *kargs = args + n;
    for (int i = 0; i < n; ++i) {
        p->E[i] = args[i];
        p->v[i] = args[n * n + i];
    }
    /* K(i,j) at K[i + j*n]; off-diagonal args row-major over i, skipping j == i */
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            float val = 0.0f;
            if (i < j) val = kargs[i * (n - 1) + j - 1];
            else if (i > j) val = kargs[i * (n - 1) + j];
            p->K[i + j * n] = val;
        }
    for (int j = 0; j < n; ++j)
        for (int ii = 0; ii < n; ++ii)
            if (ii != j) p->K[j + j * n] -= p->K[ii + j * n];
    for (int j = 0; j < n; ++j) {
        float sum = 0.0f;
        for (int ii = 0; ii < n; ++ii)
            if (ii != j) sum += p->K[ii + j * n];
        for (int ii = 0; ii < n; ++ii)
            p->Pi2j[ii + j * n] = ii != j ? p->K[ii + j * n] / sum : 0.0f;
    }
    for (int i = 0; i < n; ++i) {
        float acc = 0.0f;
        for (int j = 0; j < n; ++j) acc += p->K[i + j * n];
        acc -= p->K[i + i * n];
        p->p0[i] = acc;
    }
    p->gamma = gamma;
    p->dexdir = dexdir;
    p->crosstalk = crosstalk;
    p->r0 = r0;
    return 0;
}

/* drawE (src/gen_rand.cuh:63-68): r = N(0,1) v + r0 (1/E-1)^(1/6);
 * E' = 1/(1+(r/r0)^6).  The 1/6 power is exp(log(x)*(float)(1/6)); the
 * sixth power is ((q*q)*(q*q))*(q*q). */
static float broaden(float E, float r0, float v, float z) {
    float x = 1.0f / E - 1.0f;
    float base = r0 * or_expf(or_logf(x) * 0x1.555556p-3f);
    float rd = z * v + base;
    float q = rd / r0;
    float q2 = q * q;
    return 1.0f / (1.0f + (q2 * q2) * q2);
}

/* ------------------------------------------------------------------ */
/* One unit = mc_kernel body (src/mc.cu:92-196) for one resample.      */
/* ------------------------------------------------------------------ */
/* Windows of the walk as the RB form needs them: photon count per state of
 * the binomial acceptor draw each window would make, and the score of the
 * state and dwell draws, d/dK(j',s) log[P(j|s) K(j,s) exp(-K(j,s) tau)]
 * = [j'=j] (2/K(j,s) - tau) - 1/sum_j'' K(j'',s), in FP64.  Windows are
 * recorded up to the one whose scan reaches the last photon (every later
 * window counts nothing; the kernels stop there). */
typedef struct {
    int stopped;
    int32_t cnt[OR_MAXN];
    double score[OR_MAXN * (OR_MAXN - 1)];
    int32_t ntrans[OR_MAXN * (OR_MAXN - 1)]; /* windows per transition s -> j */
} rb_trace;

static int rate_pair(int n, int s, int j) { return j * (n - 1) + (s < j ? s : s - 1); }

static void unit_impl(const or_data *d, const or_params *p, int64_t burst, uint32_t r,
                      uint32_t seed, uint32_t set, int32_t *acc_out, float *F_out, float *E_out,
                      or_stats *st, rb_trace *tr);

void or_unit(const or_data *d, const or_params *p, int64_t burst, uint32_t r, uint32_t seed,
             uint32_t set, int32_t *acc_out, float *F_out, float *E_out, or_stats *st) {
    unit_impl(d, p, burst, r, seed, set, acc_out, F_out, E_out, st, NULL);
}

static void unit_impl(const or_data *d, const or_params *p, int64_t burst, uint32_t r,
                      uint32_t seed, uint32_t set, int32_t *acc_out, float *F_out, float *E_out,
                      or_stats *st, rb_trace *tr) {
    const int n = p->n;
    const int64_t P = d->phcount[burst];
    const int64_t base = d->istart[burst];
    const int64_t t0 = d->start[burst];
    const int64_t stop_rel = d->stop[burst] - t0;
    const float T = d->T[burst];
    const uint32_t b = (uint32_t)burst;
    uint32_t x[4];

    /* initial state from the unnormalised inflow rates (src/mc.cu:110) */
    blockw(seed, set, 0xFFFFFFFFu, 0, r, b, x, st);
    int s = or_categorical(n, p->p0, or_u01f(x[0]));
    /* live: the walk can still change the result (cost-model counters) */
    int live = P > 1 && T > 0.0f;
    if (st) {
        st->units++;
        /* the draw matters only when the first running sum can be <= u, i.e.
         * the inflow of state 0 is at most 1 (u is in (0, 1]) */
        if (live && !(p->p0[0] > 1.0f)) st->live_philox_blocks++;
    }

    float tsum = 0.0f, F = 0.0f, accf = 0.0f;
    int64_t w0 = 0, w1 = 0, cursor = 0;
    uint32_t w = 0;
    if (tr) {
        memset(tr, 0, sizeof(*tr));
        tr->stopped = P <= 1; /* no window of such a burst counts anything */
    }
    while (T > tsum && w < (1u << 22)) { /* src/mc.cu:117; window cap: see header */
        or_stats before;
        if (st) before = *st;
        const int64_t c0 = cursor;
        const int s0 = s;
        int32_t m_draw = 0;
        uint32_t A[4];
        blockw(seed, set, w, 0, r, b, A, st);
        if (st) st->windows++;
        /* drawJ_Si2Sj (src/gen_rand.cuh:78-89): column s without the diagonal */
        float col[OR_MAXN];
        for (int ii = 0; ii < n - 1; ++ii)
            col[ii] = p->Pi2j[s * n + (ii >= s ? ii + 1 : ii)];
        int j = or_categorical(n - 1, col, or_u01f(A[0]));
        if (j >= s) ++j;
        /* drawTau at the chosen transition's rate (src/mc.cu:122) */
        float k = p->K[j + s * n];
        float tau = or_logf(or_u01f(A[1])) / (-k);
        tsum = tsum + tau;
        if (!(tsum < T)) { /* src/mc.cu:127-133 (a NaN dwell also ends the burst) */
            w1 = stop_rel;
        } else {
            float wf = (float)w0 + tsum / d->clk_p;
            w1 = wf >= 0x1p62f ? (int64_t)1 << 62 : (int64_t)wf;
        }
        /* photon scan with sticky channel flags (src/mc.cu:136-181) */
        int sd = 0, sa = 0;
        int32_t nd = 0, na = 0;
        int64_t xdraw = 0;
        for (int64_t i = cursor; i < P; ++i) {
            const int64_t idx = base + i;
            const int64_t tt = d->times[idx] - t0;
            const int isD = d->mask_dd[idx] != 0, isA = d->mask_ad[idx] != 0;
            if (st) st->photons_visited++;
            if ((isD && tt >= w1) || (isA && tt >= w1) || i == P - 1) {
                cursor = i;
                float f = (p->gamma - p->crosstalk) * (float)nd + (1.0f - p->dexdir) * (float)na;
                F += f;
                float m;
                if (d->bg_dd < 1e-4f) {
                    m = floorf(f + 0.5f);
                } else {
                    uint32_t B[4];
                    blockw(seed, set, w, 1, r, b, B, st);
                    float dt = (float)(w1 - w0) * d->clk_p;
                    float rfa = (1.0f - p->dexdir) * (float)na;
                    float bga = or_bg_capped(dt * d->bg_ad, rfa, or_u01f(B[0]), or_u01f(B[2]),
                                             seed, set, b, r, w, 0x100000u, st);
                    float rfd = (p->gamma - p->crosstalk) * (float)nd;
                    float bgd = or_bg_capped(dt * d->bg_dd, rfd, or_u01f(B[1]), or_u01f(B[3]),
                                             seed, set, b, r, w, 0x200000u, st);
                    m = floorf(((f - bgd) - bga) + 0.5f);
                }
                uint32_t C[4];
                blockw(seed, set, w, 2, r, b, C, st);
                float de = broaden(p->E[s], p->r0, p->v[s], or_normal(C[0], C[1]));
                xdraw = or_binomial((int)m, (double)de, seed, set, b, r, w, A[2], A[3], st);
                m_draw = (int32_t)m;
                if (tr && !tr->stopped) {
                    if ((int)m > 0) tr->cnt[s] += (int)m;
                    double sum = 0.0;
                    for (int ii = 0; ii < n; ++ii)
                        if (ii != s) sum += (double)p->K[ii + s * n];
                    tr->score[rate_pair(n, s, j)] += 2.0 / (double)k - (double)tau;
                    tr->ntrans[rate_pair(n, s, j)]++;
                    for (int j2 = 0; j2 < n; ++j2)
                        if (j2 != s) tr->score[rate_pair(n, s, j2)] -= 1.0 / sum;
                    if (cursor == P - 1) tr->stopped = 1;
                }
                break;
            }
            if (isD) sd = 1;
            if (isA) sa = 1;
            if (sa && sd) continue;
            if (sa) na++;
            if (sd) nd++;
        }
        accf += (float)xdraw; /* mcE[tidx] += ai (float accumulation) */
        w0 = w1;
        s = j;
        ++w;
        if (st && live) {
            st->live_windows++;
            st->live_philox_blocks += (st->philox_blocks - before.philox_blocks) -
                                      (p->v[s0] == 0.0f ? 1 : 0);
            for (int64_t left = P - c0; left > 0; left >>= 1) st->search_steps++;
            if (m_draw > 0) st->binom_draws++;
            if (d->bg_dd >= 1e-4f) st->bg_draws += 2;
            st->live_binom_inv_iters += st->binom_inv_iters - before.binom_inv_iters;
            st->live_btpe_iters += st->binom_btpe_iters - before.binom_btpe_iters;
            st->live_poisson_iters += st->poisson_iters - before.poisson_iters;
            if (cursor == P - 1) live = 0;
        }
    }
    float E = F > 0.0f ? accf / F : 0.0f;
    if (acc_out) *acc_out = (int32_t)accf;
    if (F_out) *F_out = F;
    if (E_out) *E_out = E;
}

static void hist_add(float E, int B, int64_t *hist) {
    if (E >= 0.0f && E < 1.0f) {
        int64_t bin = (int64_t)floor((double)E * (double)B);
        if (bin >= 0 && bin < B) hist[bin]++;
    }
}

void or_run(const or_data *d, const or_params *p, int64_t b0, int64_t b1, int R, uint32_t seed,
            uint32_t set, int32_t *acc, float *F, float *E, int B, int64_t *hist,
            or_stats *st) {
    for (int64_t b = b0; b < b1; ++b)
        for (int r = 0; r < R; ++r) {
            int32_t a;
            float f, e;
            or_unit(d, p, b, (uint32_t)r, seed, set, &a, &f, &e, st);
            int64_t o = (b - b0) * R + r;
            if (acc) acc[o] = a;
            if (F) F[o] = f;
            if (E) E[o] = e;
            if (hist) hist_add(e, B, hist);
        }
}

typedef struct {
    const or_data *d;
    const or_params *p;
    int64_t b0, b1;
    int R, B;
    uint32_t seed, set;
    int64_t *hist;
} or_job;

static void *or_job_run(void *arg) {
    or_job *j = (or_job *)arg;
    or_run(j->d, j->p, j->b0, j->b1, j->R, j->seed, j->set, NULL, NULL, NULL, j->B, j->hist,
           NULL);
    return NULL;
}

void or_run_threaded(const or_data *d, const or_params *p, int64_t b0, int64_t b1, int R,
                     uint32_t seed, uint32_t set, int B, int64_t *hist, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    or_job jobs[256];
    int64_t nb = b1 - b0;
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].d = d; jobs[t].p = p; jobs[t].R = R; jobs[t].B = B;
        jobs[t].seed = seed; jobs[t].set = set;
        jobs[t].b0 = b0 + nb * t / nthreads;
        jobs[t].b1 = b0 + nb * (t + 1) / nthreads;
        jobs[t].hist = (int64_t *)calloc((size_t)B, sizeof(int64_t));
        pthread_create(&th[t], NULL, or_job_run, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t) {
        pthread_join(th[t], NULL);
        for (int i = 0; i < B; ++i) hist[i] += jobs[t].hist[i];
        free(jobs[t].hist);
    }
}

void or_hist(const float *x, int64_t n, int B, int64_t *hist) {
    for (int64_t i = 0; i < n; ++i) hist_add(x[i], B, hist);
}

/* src/streamWorker.cpp:194-206: bins b >= left_pad with M_b > 0;
 * term = (O - M/R)^2 / (M/R); chi2 /= effN - n(n+1). */
void or_chi2(const float *O, const int64_t *M, int B, int left_pad, int R, int nstates,
             float *chi2_ref, double *chi2_64, int *effN_out) {
    int effN = B - left_pad;
    float c32 = 0.0f;
    double c64 = 0.0;
    for (int b = left_pad; b < B; ++b) {
        if (M[b] > 0) {
            float Mf = (float)M[b];
            float d = O[b] - Mf / (float)R;
            c32 = (float)((double)c32 + ((double)d * (double)d) / (double)(Mf / (float)R));
            double md = (double)M[b] / (double)R;
            double dd = (double)O[b] - md;
            c64 += dd * dd / md;
        } else {
            effN--;
        }
    }
    int dof = effN - nstates * (nstates + 1);
    if (chi2_ref) *chi2_ref = c32 / (float)dof;
    if (chi2_64) *chi2_64 = c64 / (double)dof;
    if (effN_out) *effN_out = effN;
}

/* ------------------------------------------------------------------ */
/* Rao-Blackwellised form: the acceptor total of a unit given its walk  */
/* is a sum over states of Binomial(M_s, p_s); its pmf is spread into   */
/* the E histogram instead of one draw.                                 */
/* ------------------------------------------------------------------ */

#define OR_RB_MAXC 4096
static double g_lf[OR_RB_MAXC + 1];
static int g_lf_ready = 0;

static const double *log_factorials(void) {
    if (!g_lf_ready) {
        g_lf[0] = 0.0;
        for (int k = 1; k <= OR_RB_MAXC; ++k) g_lf[k] = g_lf[k - 1] + or_log((double)k);
        g_lf_ready = 1;
    }
    return g_lf;
}

const double *or_log_factorials(void) { return log_factorials(); }

/* Binomial(m, p)(k) in double: exp of log factorials and logs, this order. */
static double binom_pmf(int m, int k, float p) {
    if (p <= 0.0f) return k == 0 ? 1.0 : 0.0;
    if (p >= 1.0f) return k == m ? 1.0 : 0.0;
    const double *lf = log_factorials();
    const double lp = or_log((double)p), lq = or_log(1.0 - (double)p);
    double l = lf[m] - lf[k];
    l = l - lf[m - k];
    l = l + (double)k * lp;
    l = l + (double)(m - k) * lq;
    return or_exp(l);
}

static int rb_bin(int a, float F, int B) {
    const float E = F > 0.0f ? (float)a / F : 0.0f;
    if (!(E >= 0.0f && E < 1.0f)) return -1;
    const int b = (int)floor((double)E * (double)B);
    return b < B ? b : -1;
}

static float state_p(const or_params *p, int s) { return broaden(p->E[s], p->r0, 0.0f, 0.0f); }

/* Acceptor probability of a state without broadening: E after the distance
 * round trip of drawE (src/gen_rand.cuh:63-68) with v = 0. */
float or_state_efficiency(float E, float r0) { return broaden(E, r0, 0.0f, 0.0f); }

int or_rb_unit(const or_data *d, const or_params *p, int64_t burst, uint32_t r, uint32_t seed,
               uint32_t set, int32_t *cnt, float *F, double *score, int32_t *ntrans) {
    rb_trace tr;
    unit_impl(d, p, burst, r, seed, set, NULL, F, NULL, NULL, &tr);
    for (int s = 0; s < p->n; ++s) cnt[s] = tr.cnt[s];
    for (int i = 0; i < p->n * (p->n - 1); ++i) {
        if (score) score[i] = tr.score[i];
        if (ntrans) ntrans[i] = tr.ntrans[i];
    }
    return 0;
}

/* Float path, the kernels' operations in their order: binomial pmfs rounded
 * to float, convolutions accumulated in float (k ascending from 0.0f), each
 * outcome's probability added to its bin in 2^-32 fixed point. */
void or_rb_run(const or_data *d, const or_params *p, int64_t b0, int64_t b1, int R,
               uint32_t seed, uint32_t set, int B, uint64_t *M) {
    static float X[OR_RB_MAXC + 1], Y[OR_RB_MAXC + 1], Z[OR_RB_MAXC + 1];
    for (int64_t b = b0; b < b1; ++b)
        for (int r = 0; r < R; ++r) {
            int32_t cnt[OR_MAXN];
            float F;
            or_rb_unit(d, p, b, (uint32_t)r, seed, set, cnt, &F, NULL, NULL);
            int xl = 1;
            X[0] = 1.0f;
            for (int s = 0; s < p->n; ++s) {
                const int m = cnt[s];
                if (m == 0) continue;
                const float ps = state_p(p, s);
                for (int k = 0; k <= m; ++k) Y[k] = (float)binom_pmf(m, k, ps);
                if (xl == 1) {
                    for (int k = 0; k <= m; ++k) X[k] = Y[k];
                    xl = m + 1;
                    continue;
                }
                const int nl = xl + m;
                for (int a = 0; a < nl; ++a) {
                    int k0 = a - (xl - 1);
                    if (k0 < 0) k0 = 0;
                    const int k1 = a < m ? a : m;
                    float z = 0.0f;
                    for (int k = k0; k <= k1; ++k) z = z + Y[k] * X[a - k];
                    Z[a] = z;
                }
                for (int a = 0; a < nl; ++a) X[a] = Z[a];
                xl = nl;
            }
            for (int a = 0; a < xl; ++a) {
                const int bin = rb_bin(a, F, B);
                if (bin >= 0) M[bin] += (uint64_t)((double)X[a] * 4294967296.0 + 0.5);
            }
        }
}

/* FP64 path: expected counts M (double) and, for a cotangent g[B], the
 * gradient of sum_b g_b M_b in the args: d/dE_s by forward-mode
 * differentiation of the convolution chain (d Binomial(m,p)(k)/dp =
 * pmf (k - m p) / (p (1 - p))), d/dK by the likelihood ratio with the
 * leave-one-out baseline of each burst's R resamples. */
void or_rb_grad(const or_data *d, const or_params *p, int64_t b0, int64_t b1, int R,
                uint32_t seed, uint32_t set, int B, const double *g, double *M, double *grad) {
    static double X[OR_RB_MAXC + 1], Y[OR_RB_MAXC + 1], dY[OR_RB_MAXC + 1], T[OR_RB_MAXC + 1];
    static double D[OR_MAXN][OR_RB_MAXC + 1];
    const int n = p->n, K = n * (n - 1);
    for (int i = 0; i < n + K; ++i) grad[i] = 0.0;
    for (int i = 0; i < B; ++i) M[i] = 0.0;
    for (int64_t b = b0; b < b1; ++b) {
        double H = 0.0, Sg[OR_MAXN * (OR_MAXN - 1)] = {0}, Ac[OR_MAXN * (OR_MAXN - 1)] = {0};
        for (int r = 0; r < R; ++r) {
            int32_t cnt[OR_MAXN];
            float F;
            double S[OR_MAXN * (OR_MAXN - 1)];
            or_rb_unit(d, p, b, (uint32_t)r, seed, set, cnt, &F, S, NULL);
            int xl = 1;
            X[0] = 1.0;
            unsigned dmask = 0;
            for (int s = 0; s < n; ++s) {
                const int m = cnt[s];
                if (m == 0) continue;
                const float psf = state_p(p, s);
                const double ps = (double)psf;
                for (int k = 0; k <= m; ++k) {
                    Y[k] = binom_pmf(m, k, psf);
                    dY[k] = Y[k] * ((double)k - (double)m * ps) / (ps * (1.0 - ps));
                }
                const int nl = xl + m;
                /* D_s = X (*) dY; D_t = D_t (*) Y; X = X (*) Y */
                for (int a = 0; a < nl; ++a) {
                    int k0 = a - (xl - 1);
                    if (k0 < 0) k0 = 0;
                    const int k1 = a < m ? a : m;
                    double z = 0.0;
                    for (int k = k0; k <= k1; ++k) z += dY[k] * X[a - k];
                    T[a] = z;
                }
                for (int t = 0; t < n; ++t) {
                    if (!(dmask & (1u << t))) continue;
                    double tmp[OR_RB_MAXC + 1];
                    for (int a = 0; a < nl; ++a) {
                        int k0 = a - (xl - 1);
                        if (k0 < 0) k0 = 0;
                        const int k1 = a < m ? a : m;
                        double z = 0.0;
                        for (int k = k0; k <= k1; ++k) z += Y[k] * D[t][a - k];
                        tmp[a] = z;
                    }
                    for (int a = 0; a < nl; ++a) D[t][a] = tmp[a];
                }
                for (int a = 0; a < nl; ++a) D[s][a] = T[a];
                for (int a = 0; a < nl; ++a) {
                    int k0 = a - (xl - 1);
                    if (k0 < 0) k0 = 0;
                    const int k1 = a < m ? a : m;
                    double z = 0.0;
                    for (int k = k0; k <= k1; ++k) z += Y[k] * X[a - k];
                    T[a] = z;
                }
                for (int a = 0; a < nl; ++a) X[a] = T[a];
                xl = nl;
                dmask |= 1u << s;
            }
            double h = 0.0;
            for (int a = 0; a < xl; ++a) {
                const int bin = rb_bin(a, F, B);
                if (bin < 0) continue;
                M[bin] += X[a];
                h += g[bin] * X[a];
                for (int t = 0; t < n; ++t)
                    if (dmask & (1u << t)) grad[t] += g[bin] * D[t][a];
            }
            H += h;
            for (int i = 0; i < K; ++i) {
                Sg[i] += S[i];
                Ac[i] += h * S[i];
            }
        }
        if (R > 1)
            for (int i = 0; i < K; ++i)
                grad[n + i] += ((double)R / (double)(R - 1)) * (Ac[i] - H * Sg[i] / (double)R);
    }
}
