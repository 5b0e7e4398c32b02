// Device-side building blocks shared by the libnqhip kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NQ_WAVE 64

// ---------------------------------------------------------------- complex
template <typename T>
struct cplx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmk(T r, T i) { cplx<T> z; z.re = r; z.im = i; return z; }
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return cmk(a.re + b.re, a.im + b.im); }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return cmk(a.re - b.re, a.im - b.im); }
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) { return cmk(a.re * s, a.im * s); }
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) { return cmk(a.re, -a.im); }
// a += b * s (s real)
template <typename T>
__device__ __forceinline__ void cfma_r(cplx<T>& a, cplx<T> b, T s) { a.re += b.re * s; a.im += b.im * s; }

// ---------------------------------------------------------------- math
// FP64: the accurate device library.  FP32: the hardware transcendentals
// (v_exp_f32, v_log_f32, v_sin_f32, v_cos_f32); their absolute errors (~1e-7)
// are below what the FP32 form states against the FP64 oracle.
__device__ __forceinline__ double nq_exp(double x) { return exp(x); }
__device__ __forceinline__ float nq_exp(float x) { return __expf(x); }
__device__ __forceinline__ double nq_log1p(double x) { return log1p(x); }
__device__ __forceinline__ float nq_log1p(float x) { return __logf(1.0f + x); }
__device__ __forceinline__ double nq_atan2(double y, double x) { return atan2(y, x); }
// FP32: octant reduction to a = min(|x|,|y|) / max(|x|,|y|) in [0, 1] and
// atan(a) = a P(a^2), P of degree 7 fitted for this file by weighted least
// squares on [0, 1] (max abs error 6e-8 rad); atan2(0, 0) = 0
__device__ __forceinline__ float nq_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float a = (mx > 0.0f) ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
    const float s = a * a;
    float p = -0.00405455f;
    p = fmaf(p, s, 0.02186288f);
    p = fmaf(p, s, -0.05591222f);
    p = fmaf(p, s, 0.0964219f);
    p = fmaf(p, s, -0.13908626f);
    p = fmaf(p, s, 0.19946565f);
    p = fmaf(p, s, -0.3332986f);
    p = fmaf(p, s, 0.99999934f);
    float r = a * p;
    r = (ay > ax) ? 1.57079633f - r : r;
    r = (x < 0.0f) ? 3.14159265f - r : r;
    return copysignf(r, y);
}
__device__ __forceinline__ void nq_sincos(double x, double* s, double* c) { sincos(x, s, c); }
__device__ __forceinline__ void nq_sincos(float x, float* s, float* c) { *s = __sinf(x); *c = __cosf(x); }
__device__ __forceinline__ double nq_sin_only(double x) { return sin(x); }
__device__ __forceinline__ float nq_sin_only(float x) { return __sinf(x); }

// softplus f(x) = log(1 + e^x) and its derivative 1/(1 + e^-x)
// (src/Networks/activation.jl:14-18), sharing one exponential.
// e / (1 + e) is accurate for every finite e; e = inf gives the limit 1.
// FP32 quotients use the hardware reciprocal (1 ulp), not the IEEE divide
__device__ __forceinline__ double nq_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float nq_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
template <typename T>
__device__ __forceinline__ T sigmoid_of_exp(T e) {
    return __builtin_isinf(e) ? T(1) : e * nq_rcp(T(1) + e);
}
template <typename T>
__device__ __forceinline__ void softplus_and_sigmoid(T x, T* f, T* df) {
    T e = nq_exp(x);
    *f = nq_log1p(e);
    *df = sigmoid_of_exp(e);
}
template <typename T>
__device__ __forceinline__ T sigmoid(T x) { return sigmoid_of_exp(nq_exp(x)); }
template <typename T>
__device__ __forceinline__ T softplus(T x) { return nq_log1p(nq_exp(x)); }

// complex softplus and its derivative: w = e^z,
//   log(1 + w) = 1/2 log1p(e^x (2 cos y + e^x)) + i atan2(e^x sin y, 1 + e^x cos y)
//   f'(z) = w / (1 + w)
template <typename T>
__device__ __forceinline__ void csoftplus_and_sigmoid(cplx<T> z, cplx<T>* f, cplx<T>* df) {
    T ex = nq_exp(z.re), s, c;
    nq_sincos(z.im, &s, &c);
    T wr = ex * c, wi = ex * s;
    T dr = T(1) + wr;
    f->re = T(0.5) * nq_log1p(ex * (T(2) * c + ex));
    f->im = nq_atan2(wi, dr);
    // w / (1 + w) = (wr + i wi)(dr - i wi) / (dr^2 + wi^2)
    T den = nq_rcp(dr * dr + wi * wi);
    df->re = (wr * dr + wi * wi) * den;
    df->im = (wi * dr - wr * wi) * den;
}
// Re log(1 + e^z) only (what the sampler needs: log p = 2 Re log rho).
template <typename T>
__device__ __forceinline__ T csoftplus_re(cplx<T> z) {
    T ex = nq_exp(z.re), s, c;
    nq_sincos(z.im, &s, &c);
    return T(0.5) * nq_log1p(ex * (T(2) * c + ex));
}

// ---------------------------------------------------------------- NDM activations
// The NDM's f (src/Networks/activation.jl): NQ_ACT_SOFTPLUS is logℒ(x) =
// log1p(e^x) with f' = 1/(1 + e^-x) (:14-18, af_softplus); NQ_ACT_LOGCOSH is
// logℒ2(x) = log cosh x with f' = tanh x (:5-10, af_logcosh).  Complex
// arguments (the ancilla units) follow :9-10:
//   logℒ2(a + ib) = logℒ2(a) + log(cos b + i tanh(a) sin b),
// evaluated here as, with e = e^-2|a| (no overflow for any a),
//   log cosh a = |a| + log1p(e) - log 2,
//   log|cos b + i tanh a sin b| = 1/2 log1p(-sin^2 b sech^2 a), sech^2 a = 4e / (1 + e)^2,
//   arg = atan2(tanh(a) sin b, cos b),
//   tanh(a + ib) = (sgn(a) (1 - e^2) + i 2 e sin 2b) / (1 + 2 e cos 2b + e^2).
enum { NQ_ACT_SOFTPLUS = 0, NQ_ACT_LOGCOSH = 1 };

template <typename T>
__device__ __forceinline__ T nq_ln2() { return T(0.69314718055994530942); }

template <typename T>
__device__ __forceinline__ void logcosh_and_tanh(T x, T* f, T* df) {
    const T ax = fabs(x), e = nq_exp(T(-2) * ax);
    *f = ax + nq_log1p(e) - nq_ln2<T>();
    *df = copysign((T(1) - e) * nq_rcp(T(1) + e), x);
}
template <typename T>
__device__ __forceinline__ void clogcosh_and_tanh(cplx<T> z, cplx<T>* f, cplx<T>* df) {
    const T ax = fabs(z.re), e = nq_exp(T(-2) * ax);
    T s, c;
    nq_sincos(z.im, &s, &c);
    const T ope = T(1) + e, rope = nq_rcp(ope);
    const T th = (T(1) - e) * rope;                       // tanh |a|
    const T sech2 = T(4) * e * rope * rope;
    f->re = ax + nq_log1p(e) - nq_ln2<T>() + T(0.5) * nq_log1p(-s * s * sech2);
    f->im = nq_atan2(copysign(th, z.re) * s, c);
    const T iD = nq_rcp(T(1) + e * (e + T(2) * (T(2) * c * c - T(1))));
    df->re = copysign((T(1) - e) * ope * iD, z.re);
    df->im = T(4) * e * s * c * iD;
}

template <int ACT, typename T>
__device__ __forceinline__ void act_f_df(T x, T* f, T* df) {
    if constexpr (ACT == NQ_ACT_LOGCOSH) logcosh_and_tanh(x, f, df);
    else softplus_and_sigmoid(x, f, df);
}
template <int ACT, typename T>
__device__ __forceinline__ T act_f(T x) {
    if constexpr (ACT == NQ_ACT_LOGCOSH) {
        const T ax = fabs(x);
        return ax + nq_log1p(nq_exp(T(-2) * ax)) - nq_ln2<T>();
    } else {
        return softplus(x);
    }
}
template <int ACT, typename T>
__device__ __forceinline__ T act_df(T x) {
    if constexpr (ACT == NQ_ACT_LOGCOSH) {
        const T e = nq_exp(T(-2) * fabs(x));
        return copysign((T(1) - e) * nq_rcp(T(1) + e), x);
    } else {
        return sigmoid(x);
    }
}
template <int ACT, typename T>
__device__ __forceinline__ void cact_f_df(cplx<T> z, cplx<T>* f, cplx<T>* df) {
    if constexpr (ACT == NQ_ACT_LOGCOSH) clogcosh_and_tanh(z, f, df);
    else csoftplus_and_sigmoid(z, f, df);
}
template <int ACT, typename T>
__device__ __forceinline__ cplx<T> cact_f(cplx<T> z) {
    cplx<T> f, df;
    cact_f_df<ACT>(z, &f, &df);
    return f;
}
template <int ACT, typename T>
__device__ __forceinline__ cplx<T> cact_df(cplx<T> z) {
    if constexpr (ACT == NQ_ACT_LOGCOSH) {
        const T e = nq_exp(T(-2) * fabs(z.re));
        T s, c;
        nq_sincos(z.im, &s, &c);
        const T iD = nq_rcp(T(1) + e * (e + T(2) * (T(2) * c * c - T(1))));
        return cmk(copysign((T(1) - e) * (T(1) + e) * iD, z.re), T(4) * e * s * c * iD);
    } else {
        // f'(z) = w / (1 + w), w = e^z
        T s, c;
        nq_sincos(z.im, &s, &c);
        const T ex = nq_exp(z.re);
        const T wr = ex * c, wi = ex * s, dr = T(1) + wr;
        const T den = nq_rcp(dr * dr + wi * wi);
        return cmk((wr * dr + wi * wi) * den, wi * den);
    }
}
template <int ACT, typename T>
__device__ __forceinline__ T cact_re(cplx<T> z) {
    if constexpr (ACT == NQ_ACT_LOGCOSH) {
        const T ax = fabs(z.re), e = nq_exp(T(-2) * ax);
        const T s = nq_sin_only(z.im), rope = nq_rcp(T(1) + e);
        return ax + nq_log1p(e) - nq_ln2<T>() + T(0.5) * nq_log1p(-s * s * T(4) * e * rope * rope);
    } else {
        return csoftplus_re(z);
    }
}

// ---------------------------------------------------------------- exact sums
// Order-independent accumulation of doubles (the force partials of a
// workgroup, added by all its waves through LDS atomics).  Each addend v is
// split into an integer multiple of 2^-20 (hi) and a multiple of 2^-64 (lo),
// both rounded to nearest, and the two parts are summed as 64-bit integers.
// Integer addition is associative, so the total does not depend on the order
// in which lanes and waves add, and it is exact up to the 2^-65 rounding of
// each addend; intermediate wrap-around is harmless as long as the final hi
// sum stays below 2^63 (|total| < 2^43) and the lo sum below 2^63 (fewer than
// 2^20 addends per entry).  Addends must satisfy |v| < 2^30: a larger or
// non-finite addend sets *poison, and the entries read as NaN.
#define NQ_FX_MAGIC 6755399441055744.0   // 1.5 * 2^52: y + MAGIC rounds |y| < 2^51 to an integer
__device__ __forceinline__ void fx_split(double v, long long* hi, long long* lo) {
    const double t = v * 1048576.0;                              // v 2^20, exact
    const double d1 = t + NQ_FX_MAGIC;
    *hi = __double_as_longlong(d1) - __double_as_longlong(NQ_FX_MAGIC);
    const double r = t - (d1 - NQ_FX_MAGIC);                     // exact, |r| <= 1/2
    const double d2 = fma(r, 17592186044416.0, NQ_FX_MAGIC);     // r 2^44 rounded
    *lo = __double_as_longlong(d2) - __double_as_longlong(NQ_FX_MAGIC);
}
// acc: 2 words per entry; entry k at acc[2k] (hi), acc[2k + 1] (lo)
__device__ __forceinline__ void fx_add(unsigned long long* acc, int k, double v, int* poison) {
    if (!(fabs(v) < 1073741824.0)) {
        atomicOr(poison, 1);
        return;
    }
    long long hi, lo;
    fx_split(v, &hi, &lo);
    atomicAdd(acc + 2 * k, (unsigned long long)hi);
    atomicAdd(acc + 2 * k + 1, (unsigned long long)lo);
}
__device__ __forceinline__ double fx_value(const unsigned long long* acc, int k) {
    return (double)(long long)acc[2 * k] * 0x1p-20 + (double)(long long)acc[2 * k + 1] * 0x1p-64;
}
// LDS bytes of an accumulator of n entries plus its poison word (16-aligned)
__host__ __device__ inline int fx_bytes(int n) { return 16 * n + 16; }

// ---------------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11).
struct philox_out { uint32_t x[4]; };
__device__ __forceinline__ philox_out philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                    uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    philox_out o; o.x[0] = c0; o.x[1] = c1; o.x[2] = c2; o.x[3] = c3;
    return o;
}
// 53-bit uniform in [0, 1) from two words (same mapping as the host oracle).
__device__ __forceinline__ double u01_53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
__device__ __forceinline__ float u01_24(uint32_t a) { return (float)(a >> 8) * (1.0f / 16777216.0f); }

// ---------------------------------------------------------------- reductions
// xor-butterfly sums over aligned groups of `width` lanes (power of two <= 64).
__device__ __forceinline__ float group_sum(float v, int width) {
    for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double group_sum(double v, int width) {
    for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v) { return group_sum(v, 64); }

// Sum over the 64 lanes, the same value returned to every lane: DPP steps
// within each row of 16 lanes (quad_perm [1,0,3,2], [2,3,0,1], half-row
// mirror, row mirror), then the four row sums read out with v_readlane.
// All 64 lanes must be active.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float rl_f(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ __forceinline__ double rl_d(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}
__device__ __forceinline__ float wave_total(float v) {
    v += dpp_f<0xB1>(v);
    v += dpp_f<0x4E>(v);
    v += dpp_f<0x141>(v);
    v += dpp_f<0x140>(v);
    return (rl_f(v, 0) + rl_f(v, 16)) + (rl_f(v, 32) + rl_f(v, 48));
}
// Sum over aligned groups of `width` lanes (power of two <= 64), the same value
// in every lane of the group: DPP within rows of 16, then xor shuffles for
// the 32 / 64 steps.  All lanes must be active.
template <typename T>
__device__ __forceinline__ T dpp_any(T v, int step) {
    if constexpr (sizeof(T) == 8) {
        switch (step) {
            case 0: return dpp_d<0xB1>(v);
            case 1: return dpp_d<0x4E>(v);
            case 2: return dpp_d<0x141>(v);
            default: return dpp_d<0x140>(v);
        }
    } else {
        switch (step) {
            case 0: return dpp_f<0xB1>(v);
            case 1: return dpp_f<0x4E>(v);
            case 2: return dpp_f<0x141>(v);
            default: return dpp_f<0x140>(v);
        }
    }
}
template <typename T>
__device__ __forceinline__ T group_total(T v, int width) {
    if (width >= 2) v += dpp_any(v, 0);
    if (width >= 4) v += dpp_any(v, 1);
    if (width >= 8) v += dpp_any(v, 2);
    if (width >= 16) v += dpp_any(v, 3);
    if (width >= 32) v += __shfl_xor(v, 16, 64);
    if (width >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

__device__ __forceinline__ double wave_total(double v) {
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    return (rl_d(v, 0) + rl_d(v, 16)) + (rl_d(v, 32) + rl_d(v, 48));
}

// Cross-half exchanges (gfx950): permlane32 swap leaves a = (a lanes 0-31 |
// b lanes 0-31) and b = (a lanes 32-63 | b lanes 32-63); permlane16 swap
// exchanges a's odd rows with b's even rows (rows of 16 lanes), leaving
// a = (a.r0, b.r0, a.r2, b.r2) and b = (a.r1, b.r1, a.r3, b.r3).
__device__ __forceinline__ void pl32_swap(uint32_t& a, uint32_t& b) {
    const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    a = r[0];
    b = r[1];
}
__device__ __forceinline__ void pl16_swap(uint32_t& a, uint32_t& b) {
    const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
    a = r[0];
    b = r[1];
}
template <bool W32>
__device__ __forceinline__ void plswap(float& a, float& b) {
    uint32_t x = __float_as_uint(a), y = __float_as_uint(b);
    if (W32) pl32_swap(x, y); else pl16_swap(x, y);
    a = __uint_as_float(x);
    b = __uint_as_float(y);
}
template <bool W32>
__device__ __forceinline__ void plswap(double& a, double& b) {
    uint32_t al = (uint32_t)__double2loint(a), ah = (uint32_t)__double2hiint(a);
    uint32_t bl = (uint32_t)__double2loint(b), bh = (uint32_t)__double2hiint(b);
    if (W32) { pl32_swap(al, bl); pl32_swap(ah, bh); } else { pl16_swap(al, bl); pl16_swap(ah, bh); }
    a = __hiloint2double((int)ah, (int)al);
    b = __hiloint2double((int)bh, (int)bl);
}
template <int CTRL>
__device__ __forceinline__ float dpp_t(float v) { return dpp_f<CTRL>(v); }
template <int CTRL>
__device__ __forceinline__ double dpp_t(double v) { return dpp_d<CTRL>(v); }

// K = 8, 16 or 32 wave sums at once (a transposed reduction).  On return
// lane L holds the sum over all 64 lanes of v[L / (64 / K)].  Each step pairs
// lanes that differ in one lane bit and halves the values every lane holds:
// the lane with the bit clear keeps the lower index of each pair
// (i, i + n/2), its partner the upper one (bit 5: permlane32 swap, bit 4:
// permlane16 swap, bit 3: row mirror, bit 2: half-row mirror, bit 1:
// quad_perm [2,3,0,1]); the lane bits left over are summed plainly.
// About 2K instructions for K sums instead of K reductions of ~15 each.
// All 64 lanes must be active; v is clobbered.
template <int K, typename T>
__device__ __forceinline__ T wave_totals(T (&v)[K], int lane) {
    static_assert(K == 8 || K == 16 || K == 32, "8, 16 or 32 sums");
#pragma unroll
    for (int i = 0; i < K / 2; ++i) {
        T a = v[i], b = v[i + K / 2];
        plswap<true>(a, b);
        v[i] = a + b;
    }
#pragma unroll
    for (int i = 0; i < K / 4; ++i) {
        T a = v[i], b = v[i + K / 4];
        plswap<false>(a, b);
        v[i] = a + b;
    }
    const bool b3 = (lane & 8) != 0, b2 = (lane & 4) != 0, b1 = (lane & 2) != 0;
#pragma unroll
    for (int i = 0; i < K / 8; ++i) {
        const T lo = v[i] + dpp_t<0x140>(v[i]);
        const T hi = v[i + K / 8] + dpp_t<0x140>(v[i + K / 8]);
        v[i] = b3 ? hi : lo;
    }
    T r;
    if constexpr (K >= 16) {
#pragma unroll
        for (int i = 0; i < K / 16; ++i) {
            const T lo = v[i] + dpp_t<0x141>(v[i]);
            const T hi = v[i + K / 16] + dpp_t<0x141>(v[i + K / 16]);
            v[i] = b2 ? hi : lo;
        }
    } else {
        v[0] += dpp_t<0x141>(v[0]);
    }
    if constexpr (K == 32) {
        const T lo = v[0] + dpp_t<0x4E>(v[0]);
        const T hi = v[1] + dpp_t<0x4E>(v[1]);
        r = b1 ? hi : lo;
    } else {
        r = v[0] + dpp_t<0x4E>(v[0]);
    }
    return r + dpp_t<0xB1>(r);
}

// ---------------------------------------------------------------- NDM layout
// Offsets of the NDM blocks in the reference's flat order (NDM.jl:17-18):
// b_mu[N] h_mu[nh] w_mu[nh x N] u_mu[na x N] b_la[N] h_la[nh] d_la[na] w_la[nh x N] u_la[na x N]
// (matrices column-major, element (h, i) at h + nh*i).
struct ndm_dims {
    int N, nh, na;
    int act;   // NQ_ACT_SOFTPLUS or NQ_ACT_LOGCOSH
    int bmu, hmu, wmu, umu, bla, hla, dla, wla, ula, np;
    // compact gradient layout (one row per sample, "P order"):
    //   lambda block (real):  b_la h_la w_la            -> cols [0, nlam)
    //   Pi block (real part): d_la u_la u_mu            -> cols [nlam, nlam + npi)
    //   Pi block (imag part): d_la u_la u_mu            -> cols [nlam + npi, nlam + 2 npi)
    //   mu block (imag):      b_mu h_mu w_mu            -> cols [nlam + 2 npi, R)
    int nlam, npi, nmu, R;
};

static inline ndm_dims make_ndm_dims(int N, int nh, int na) {
    ndm_dims d;
    d.N = N; d.nh = nh; d.na = na;
    d.act = 0;
    int p = 0;
    d.bmu = p; p += N;
    d.hmu = p; p += nh;
    d.wmu = p; p += nh * N;
    d.umu = p; p += na * N;
    d.bla = p; p += N;
    d.hla = p; p += nh;
    d.dla = p; p += na;
    d.wla = p; p += nh * N;
    d.ula = p; p += na * N;
    d.np = p;
    d.nlam = N + nh + nh * N;
    d.npi = na + 2 * na * N;
    d.nmu = N + nh + nh * N;
    d.R = d.nlam + 2 * d.npi + d.nmu;
    return d;
}
