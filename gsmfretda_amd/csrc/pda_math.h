// pda_math.h -- random streams and variates of the Monte-Carlo PDA walk,
// shared by the gfx950 kernels and the host side of the C-ABI layer.
//
// Every function here is built from IEEE +, -, *, /, fma, sqrt, floor and
// integer operations only, and the whole library is compiled with
// -ffp-contract=off and without fast-math, so the host and the GPU produce
// identical bits for identical inputs.
//
// Random stream layout (replaces curand Sobol64 + the xorshift rk_state of
// src/gen_rand.cuh:7-25 and src/binom.cuh:7-19): one Philox4x32-10 block per
// (key, counter), key = {seed, parameter set}, counter = {window, sub-stream,
// resample, burst}:
//   initial state      {0xFFFFFFFF, 0, r, b}   word 0
//   window w           {w, 0, r, b}   w0 next state, w1 dwell, w2:w3 binomial
//   background         {w, 1, r, b}   w0/w1 Poisson A/D, w2/w3 cap uniform A/D
//   broadening normal  {w, 2, r, b}   w0:w1 Box-Muller
//   binomial extra i   {w, 16+i, r, b}
//   Poisson chunk c>=1 {w, 0x100000 (A) | 0x200000 (D) + (c-1)/4, r, b} word (c-1)%4
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstring>

#define PDA_HD __host__ __device__ __forceinline__

namespace pda {

constexpr uint32_t kInitWindow = 0xFFFFFFFFu;
constexpr uint32_t kSubMain = 0, kSubBackground = 1, kSubNormal = 2, kSubBinomExtra = 16;
constexpr uint32_t kSubPoissonA = 0x100000u, kSubPoissonD = 0x200000u;
constexpr int kMaxRestarts = 4096;

struct Block {
    uint32_t w[4];
};

// Identifies one resample of one burst under one (seed, parameter set).
struct UnitKey {
    uint32_t seed, set, r, b;
};

// a ^ b ^ c.  On gfx950 one v_bitop3_b32 (truth table 0x96 = 0xF0 ^ 0xCC ^
// 0xAA), which the compiler does not form from two XORs by itself.
PDA_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

PDA_HD Block philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                      uint32_t k1) {
    // Philox4x32-10, Salmon et al., "Parallel random numbers: as easy as
    // 1, 2, 3", SC'11: ten rounds of two 32x32->64 multiplies and a Weyl key.
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const uint64_t m0 = uint64_t(0xD2511F53u) * c0;
        const uint64_t m1 = uint64_t(0xCD9E8D57u) * c2;
        const uint32_t y0 = xor3(uint32_t(m1 >> 32), c1, k0);
        const uint32_t y2 = xor3(uint32_t(m0 >> 32), c3, k1);
        c1 = uint32_t(m1);
        c3 = uint32_t(m0);
        c0 = y0;
        c2 = y2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Block{{c0, c1, c2, c3}};
}

PDA_HD Block draw_block(const UnitKey& k, uint32_t window, uint32_t sub) {
    return philox10(window, sub, k.r, k.b, k.seed, k.set);
}

// (0,1] from the top 24 bits of a word.
PDA_HD float uniform24(uint32_t x) { return float((x >> 8) + 1u) * 0x1p-24f; }
// [0,1) from 27 + 26 bits of two words.
PDA_HD double uniform53(uint32_t hi, uint32_t lo) {
    return (double(hi >> 5) * 0x1p26 + double(lo >> 6)) * 0x1p-53;
}

PDA_HD uint32_t fbits(float f) {
#ifdef __HIP_DEVICE_COMPILE__
    return __float_as_uint(f);
#else
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
#endif
}
PDA_HD float bitsf(uint32_t u) {
#ifdef __HIP_DEVICE_COMPILE__
    return __uint_as_float(u);
#else
    float f;
    std::memcpy(&f, &u, 4);
    return f;
#endif
}
PDA_HD uint64_t dbits(double f) {
#ifdef __HIP_DEVICE_COMPILE__
    return uint64_t(__double_as_longlong(f));
#else
    uint64_t u;
    std::memcpy(&u, &f, 8);
    return u;
#endif
}
PDA_HD double bitsd(uint64_t u) {
#ifdef __HIP_DEVICE_COMPILE__
    return __longlong_as_double((long long)u);
#else
    double f;
    std::memcpy(&f, &u, 8);
    return f;
#endif
}

// ---- natural log / exp --------------------------------------------------
// log: x = 2^e m with m in (sqrt(1/2), sqrt(2)], f = m - 1, s = f/(2+f),
// log(1+f) = 2 atanh(s) = f - (f^2/2 - s (f^2/2 + R(s^2))), ln2 split hi/lo.
// exp: x = k ln2 + r (Cody-Waite), Taylor polynomial of e^r, scale by 2^k.

PDA_HD float flog(float x) {
    if (!(x >= 0.0f)) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_inff();
    if (x == __builtin_inff()) return x;
    uint32_t bits = fbits(x);
    int ex = 0;
    if (bits < 0x00800000u) {
        bits = fbits(x * 0x1p25f);
        ex = -25;
    }
    ex += int(bits >> 23) - 127;
    float m = bitsf((bits & 0x007FFFFFu) | 0x3F800000u);
    if (m > 0x1.6a09e6p+0f) {
        m *= 0.5f;
        ex += 1;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float poly = fmaf(z, 0x1.c71c72p-3f, 0x1.24924ap-2f);
    poly = fmaf(z, poly, 0x1.99999ap-2f);
    poly = fmaf(z, poly, 0x1.555556p-1f);
    const float R = z * poly;
    const float half_f2 = 0.5f * f * f;
    const float e = float(ex);
    return e * 0x1.62ep-1f - ((half_f2 - (s * (half_f2 + R) + e * 0x1.0bfbe8p-15f)) - f);
}

PDA_HD float pow2f(int k) { return bitsf(uint32_t(k + 127) << 23); }

PDA_HD float fexp(float x) {
    if (x != x) return x;
    if (x > 88.72f) return __builtin_inff();
    if (x < -103.98f) return 0.0f;
    const float kf = floorf(x * 0x1.715476p+0f + 0.5f);
    const float r = (x - kf * 0x1.62ep-1f) - kf * 0x1.0bfbe8p-15f;
    float e = fmaf(r, 0x1.a01a02p-13f, 0x1.6c16c2p-10f);
    e = fmaf(r, e, 0x1.111112p-7f);
    e = fmaf(r, e, 0x1.555556p-5f);
    e = fmaf(r, e, 0x1.555556p-3f);
    e = fmaf(r, e, 0.5f);
    e = fmaf(r, e, 1.0f);
    e = fmaf(r, e, 1.0f);
    const int k = int(kf);
    const int ka = k / 2;
    return (e * pow2f(ka)) * pow2f(k - ka);
}

PDA_HD double dlog(double x) {
    if (!(x >= 0.0)) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    uint64_t bits = dbits(x);
    int ex = 0;
    if (bits < 0x0010000000000000ull) {
        bits = dbits(x * 0x1p54);
        ex = -54;
    }
    ex += int(bits >> 52) - 1023;
    double m = bitsd((bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 0x1.6a09e667f3bcdp+0) {
        m *= 0.5;
        ex += 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    // 2/(2k+1), k = 11 .. 1
    constexpr double c[11] = {0x1.642c8590b2164p-4, 0x1.8618618618618p-4, 0x1.af286bca1af28p-4,
                              0x1.e1e1e1e1e1e1ep-4, 0x1.1111111111111p-3, 0x1.3b13b13b13b14p-3,
                              0x1.745d1745d1746p-3, 0x1.c71c71c71c71cp-3, 0x1.2492492492492p-2,
                              0x1.999999999999ap-2, 0x1.5555555555555p-1};
    double poly = c[0];
#pragma unroll
    for (int i = 1; i < 11; ++i) poly = fma(z, poly, c[i]);
    const double R = z * poly;
    const double half_f2 = 0.5 * f * f;
    const double e = double(ex);
    return e * 0x1.62e42feep-1 -
           ((half_f2 - (s * (half_f2 + R) + e * 0x1.a39ef35793c76p-33)) - f);
}

PDA_HD double pow2d(int k) { return bitsd(uint64_t(k + 1023) << 52); }

PDA_HD double dexp(double x) {
    if (x != x) return x;
    if (x > 709.78) return __builtin_inf();
    if (x < -745.2) return 0.0;
    const double kf = floor(x * 0x1.71547652b82fep+0 + 0.5);
    const double r = (x - kf * 0x1.62e42feep-1) - kf * 0x1.a39ef35793c76p-33;
    // 1/k!, k = 13 .. 2
    constexpr double c[12] = {0x1.6124613a86d09p-33, 0x1.1eed8eff8d898p-29, 0x1.ae64567f544e4p-26,
                              0x1.27e4fb7789f5cp-22, 0x1.71de3a556c734p-19, 0x1.a01a01a01a01ap-16,
                              0x1.a01a01a01a01ap-13, 0x1.6c16c16c16c17p-10, 0x1.1111111111111p-7,
                              0x1.5555555555555p-5, 0x1.5555555555555p-3, 0.5};
    double e = c[0];
#pragma unroll
    for (int i = 1; i < 12; ++i) e = fma(r, e, c[i]);
    e = fma(r, e, 1.0);
    e = fma(r, e, 1.0);
    const int k = int(kf);
    const int ka = k / 2;
    return (e * pow2d(ka)) * pow2d(k - ka);
}

// cos(2 pi u), u in [0,1]: nearest quarter turn q, a = (4u - q) pi/2.
PDA_HD float cos_turns(float u) {
    const float y = 4.0f * u;
    const float q = floorf(y + 0.5f);
    const float a = (y - q) * 0x1.921fb6p+0f;
    const float a2 = a * a;
    float sp = fmaf(a2, 0x1.71de3ap-19f, -0x1.a01a02p-13f);
    sp = fmaf(a2, sp, 0x1.111112p-7f);
    sp = fmaf(a2, sp, -0x1.555556p-3f);
    const float sn = fmaf(a2, sp, 1.0f) * a;
    float cp = fmaf(a2, -0x1.27e4fcp-22f, 0x1.a01a02p-16f);
    cp = fmaf(a2, cp, -0x1.6c16c2p-10f);
    cp = fmaf(a2, cp, 0x1.555556p-5f);
    cp = fmaf(a2, cp, -0.5f);
    const float cs = fmaf(a2, cp, 1.0f);
    switch (int(q) & 3) {
        case 0: return cs;
        case 1: return -sn;
        case 2: return -cs;
        default: return sn;
    }
}

// Standard normal by Box-Muller (stands in for curand_normal, src/gen_rand.cuh:65).
PDA_HD float gauss(uint32_t a, uint32_t b) {
    return sqrtf(-2.0f * flog(uniform24(a))) * cos_turns(uniform24(b));
}

// r0 (1/E - 1)^(1/6): the distance of efficiency E (src/gen_rand.cuh:65).
PDA_HD float efficiency_distance(float E, float r0) {
    return r0 * fexp(flog(1.0f / E - 1.0f) * 0x1.555556p-3f);
}
// 1/(1 + (r/r0)^6): the efficiency of distance r (src/gen_rand.cuh:67).
PDA_HD float distance_efficiency(float rd, float r0) {
    const float q = rd / r0;
    const float q2 = q * q;
    return 1.0f / (1.0f + (q2 * q2) * q2);
}

// ---- binomial -------------------------------------------------------------
// Inversion (src/binom.cuh:135-169) when n < kInversionMaxN or
// min(p,1-p) n <= 30, BTPE (Kachitvichyanukul & Schmeiser 1988,
// src/binom.cuh:21-134) otherwise, n - X for p > 1/2.  numpy's legacy split
// (src/binom.cuh:172-188) switches to BTPE at min(p,1-p) n > 30; both are
// exact samplers, and below 128 trials q^n >= 2^-127 stays far from
// underflow, so the cheap inversion covers every count a window of a
// typical burst produces.
constexpr int kInversionMaxN = 128;

PDA_HD int64_t binomial_inversion(int n, double p, double U, const UnitKey& key, uint32_t w) {
    const double q = 1.0 - p;
    const double qn = dexp(double(n) * dlog(q));
    const double np = double(n) * p;
    double lim = np + 10.0 * sqrt(np * q + 1.0);
    if (double(n) < lim) lim = double(n);
    const int64_t bound = int64_t(lim);
    int64_t X = 0;
    double px = qn;
    int restarts = 0;
    while (U > px) {
        ++X;
        if (X > bound) {
            if (restarts == kMaxRestarts) return int64_t(np);
            const Block x = draw_block(key, w, kSubBinomExtra + uint32_t(restarts));
            ++restarts;
            X = 0;
            px = qn;
            U = uniform53(x.w[0], x.w[1]);
        } else {
            U -= px;
            px = (double(n - X + 1) * p * px) / (double(X) * q);
        }
    }
    return X;
}

PDA_HD double stirling_omega(double f) {
    // 1/(12f) - 1/(360f^3) + 1/(1260f^5) - 1/(1680f^7) + 1/(1188f^9)
    const double f2 = f * f;
    return (13860. - (462. - (132. - (99. - 140. / f2) / f2) / f2) / f2) / f / 166320.;
}

PDA_HD int64_t binomial_btpe(int n, double p, const UnitKey& key, uint32_t w) {
    const double q = 1.0 - p;
    const double fm = double(n) * p + p;
    const int64_t m = int64_t(floor(fm));
    const double nrq = double(n) * p * q;
    const double p1 = floor(2.195 * sqrt(nrq) - 4.6 * q) + 0.5;
    const double xm = double(m) + 0.5;
    const double xl = xm - p1;
    const double xr = xm + p1;
    const double c = 0.134 + 20.5 / (15.3 + double(m));
    double a = (fm - xl) / (fm - xl * p);
    const double laml = a * (1.0 + 0.5 * a);
    a = (xr - fm) / (xr * q);
    const double lamr = a * (1.0 + 0.5 * a);
    const double p2 = p1 * (1.0 + 2.0 * c);
    const double p3 = p2 + c / laml;
    const double p4 = p3 + c / lamr;
    for (int it = 0; it < kMaxRestarts; ++it) {
        const Block x = draw_block(key, w, kSubBinomExtra + uint32_t(it));
        const double u = uniform53(x.w[0], x.w[1]) * p4;
        double v = uniform53(x.w[2], x.w[3]);
        int64_t y;
        if (u <= p1) return int64_t(floor(xm - p1 * v + u));
        if (u <= p2) {
            const double xx = xl + (u - p1) / c;
            v = v * c + 1.0 - fabs(double(m) - xx + 0.5) / p1;
            if (v > 1.0) continue;
            y = int64_t(floor(xx));
        } else if (u <= p3) {
            const double yd = floor(xl + dlog(v) / laml);
            if (!(yd >= 0.0)) continue;
            y = int64_t(yd);
            v = v * (u - p2) * laml;
        } else {
            const double yd = floor(xr - dlog(v) / lamr);
            if (!(yd <= double(n))) continue;
            y = int64_t(yd);
            v = v * (u - p3) * lamr;
        }
        const int64_t k = y > m ? y - m : m - y;
        if (k <= 20 || double(k) >= nrq / 2.0 - 1.0) {
            const double s = p / q;
            const double aa = s * double(n + 1);
            double F = 1.0;
            if (m < y) {
                for (int64_t i = m + 1; i <= y; ++i) F *= (aa / double(i) - s);
            } else if (m > y) {
                for (int64_t i = y + 1; i <= m; ++i) F /= (aa / double(i) - s);
            }
            if (v > F) continue;
            return y;
        }
        const double kd = double(k);
        const double rho = (kd / nrq) * ((kd * (kd / 3.0 + 0.625) + 0.1666666666666) / nrq + 0.5);
        const double t = -kd * kd / (2.0 * nrq);
        const double A = dlog(v);
        if (A < t - rho) return y;
        if (A > t + rho) continue;
        const double x1 = double(y + 1), f1 = double(m + 1);
        const double z = double(n + 1 - m), ww = double(n - y + 1);
        const double lim = xm * dlog(f1 / x1) + (double(n - m) + 0.5) * dlog(z / ww) +
                           double(y - m) * dlog(ww * p / (x1 * q)) + stirling_omega(f1) +
                           stirling_omega(z) - stirling_omega(x1) - stirling_omega(ww);
        if (A > lim) continue;
        return y;
    }
    return m;
}

PDA_HD int64_t binomial(int n, double p, uint32_t u_hi, uint32_t u_lo, const UnitKey& key,
                        uint32_t w) {
    if (n <= 0 || !(p >= 0.0 && p <= 1.0)) return 0;
    const bool flip = p > 0.5;
    const double pp = flip ? 1.0 - p : p;
    const int64_t X = (n < kInversionMaxN || pp * double(n) <= 30.0)
                          ? binomial_inversion(n, pp, uniform53(u_hi, u_lo), key, w)
                          : binomial_btpe(n, pp, key, w);
    return flip ? int64_t(n) - X : X;
}

// ---- background photons -----------------------------------------------------
// draw_P_B_Tr (src/gen_rand.cuh:92-107): k ~ Poisson(lam), k if k <= cap else
// floor(cap * u_cap).  Poisson by inversion in chunks of mean <= 64; the count
// is abandoned once it passes cap, where only the cap branch can follow.
PDA_HD float background_count(float lam, float cap, float u_first, float u_cap,
                              const UnitKey& key, uint32_t w, uint32_t sub_base) {
    int64_t total = 0;
    bool over = false;
    if (lam > 0.0f) {
        const int chunks = lam <= 64.0f ? 1 : int(ceilf(lam / 64.0f));
        const float lc = chunks == 1 ? lam : lam / float(chunks);
        const float p0 = fexp(-lc);
        Block extra{{0, 0, 0, 0}};
        for (int c = 0; c < chunks; ++c) {
            float u = u_first;
            if (c > 0) {
                const uint32_t q = uint32_t(c - 1);
                const uint32_t i = q & 3u;
                if (i == 0u) extra = draw_block(key, w, sub_base + (q >> 2));
                // selects, not extra.w[i]: an indexed word would put the block
                // in scratch memory, stored on every call
                u = uniform24(i == 0u ? extra.w[0]
                                      : (i == 1u ? extra.w[1] : (i == 2u ? extra.w[2] : extra.w[3])));
            }
            int64_t k = 0;
            float pk = p0, cdf = p0;
            while (u > cdf) {
                ++k;
                if (float(total + k) > cap) {
                    over = true;
                    break;
                }
                pk = (pk * lc) / float(k);
                cdf = cdf + pk;
            }
            if (over) break;
            total += k;
        }
    }
    if (!over && float(total) <= cap) return float(total);
    return floorf(cap * u_cap);
}

}  // namespace pda
