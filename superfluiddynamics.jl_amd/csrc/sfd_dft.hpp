// sfd_dft.hpp -- complex arithmetic and fully unrolled in-register DFTs for
// gfx950.  Every DFT here is natural-order in, natural-order out; the
// compiler turns the index permutations into register renaming.
//
// Convention (AbstractFFTs, used by the reference through FFTW, Plans.jl:847):
// forward  X[k] = sum_j x[j] exp(-2 pi i jk/n)        (SIGN = -1)
// backward x[j] = sum_k X[k] exp(+2 pi i jk/n)        (SIGN = +1, unscaled)
#pragma once
#include <hip/hip_runtime.h>

#include "sfd_f32_lifting.hpp"

namespace sfd {

template <typename R>
struct alignas(2 * sizeof(R)) cx {
    R re, im;
};

template <typename R>
__device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R>
__device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * conj(b)
template <typename R>
__device__ __forceinline__ cx<R> cmulc(cx<R> a, cx<R> b)
{
    return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <typename R>
__device__ __forceinline__ cx<R> cscale(cx<R> a, R s) { return {a.re * s, a.im * s}; }

// multiply by S*i, S = +1 or -1
template <int S, typename R>
__device__ __forceinline__ cx<R> muli(cx<R> v)
{
    if constexpr (S > 0) return {-v.im, v.re};
    else return {v.im, -v.re};
}

// cos(2 pi m / 64) for m = 0..16, rounded from 36-digit values.
__device__ constexpr double kQuarterCos[17] = {
    1.0,
    0.995184726672196886244836953109479922,
    0.980785280403230449126182236134239037,
    0.956940335732208864935797886980269969,
    0.923879532511286756128183189396788287,
    0.881921264348355029712756863660388350,
    0.831469612302545237078788377617905757,
    0.773010453362736960810906609758469801,
    0.707106781186547524400844362104849039,
    0.634393284163645498215171613225493371,
    0.555570233019602224742830813948532874,
    0.471396736825997648556387625905254378,
    0.382683432365089771728459984030398867,
    0.290284677254462367636192375817395275,
    0.195090322016128267848284868477022241,
    0.0980171403295606019941955638886418459,
    0.0,
};

template <typename R>
__device__ __forceinline__ constexpr R cos64(int m)
{
    m &= 63;
    if (m <= 16) return R(kQuarterCos[m]);
    if (m <= 32) return -R(kQuarterCos[32 - m]);
    if (m <= 48) return -R(kQuarterCos[m - 32]);
    return R(kQuarterCos[64 - m]);
}
template <typename R>
__device__ __forceinline__ constexpr R sin64(int m) { return cos64<R>(m - 16); }

// Rotation of v by theta as three lifting steps (shears), t = tan(theta/2),
// s = sin(theta); lift(lift(v, t, s), -t, -s) == v in exact arithmetic for
// any rounded t and s.
__device__ __forceinline__ cx<float> lift(cx<float> v, float t, float s)
{
    v.re = fmaf(-t, v.im, v.re);
    v.im = fmaf(s, v.re, v.im);
    v.re = fmaf(-t, v.im, v.re);
    return v;
}

// v * exp(SIGN 2 pi i e / N) for a compile-time-foldable e; N divides 64.
//
// FP64 multiplies by the correctly rounded (cos, sin).  FP32 splits
// e = q N/4 + r with |r| <= N/8: the quarter turns are exact, and the rest is
// a lifting rotation with coefficients from kLiftT/kLiftS.  The inverse
// (SIGN = +1) applies the lifting steps with negated coefficients, then the
// quarter turns, so it undoes the forward rotation of the same e exactly (in
// exact arithmetic).  The inverse DFTs mirror the forward ones, so an FP32
// forward-inverse round trip carries no systematic error from the rounding
// of its constants: a complex multiply by rounded cos/sin(pi/8) instead
// loses 5.7e-8 of norm on every pass.
template <int N, int SIGN, typename R>
__device__ __forceinline__ cx<R> twid(cx<R> v, int e)
{
    static_assert(64 % N == 0, "in-register twiddles need N | 64");
    e = ((e % N) + N) % N;
    if (e == 0) return v;
    if (2 * e == N) return {-v.re, -v.im};
    if (4 * e == N) return muli<SIGN>(v);
    if (4 * e == 3 * N) return muli<-SIGN>(v);
    if constexpr (sizeof(R) == 4) {
        const int q = (4 * e + N / 2) / N;  // nearest quarter turn
        const int m = (e - q * (N / 4)) * (64 / N) + 8;
        const float t = kLiftT[m], s = kLiftS[m];
        auto quarter = [&](cx<float> x, int sgn) {
            for (int k = 0; k < (q & 3); ++k) x = sgn < 0 ? muli<-1>(x) : muli<+1>(x);
            return x;
        };
        if constexpr (SIGN < 0) return lift(quarter(v, -1), t, s);
        else return quarter(lift(v, -t, -s), +1);
    } else {
        const int m = e * (64 / N);
        const R c = cos64<R>(m);
        const R s = R(SIGN) * sin64<R>(m);
        if ((8 * e) % N == 0) {
            // odd multiple of pi/4: |c| = |s| = sqrt(1/2)
            const R h = R(kQuarterCos[8]);
            const R sc = c > R(0) ? R(1) : R(-1);
            const R ss = s > R(0) ? R(1) : R(-1);
            cx<R> r;
            if (sc == ss) { r.re = h * (v.re - v.im); r.im = h * (v.im + v.re); }
            else          { r.re = h * (v.re + v.im); r.im = h * (v.im - v.re); }
            if (sc < R(0)) { r.re = -r.re; r.im = -r.im; }
            return r;
        }
        return {v.re * c - v.im * s, v.re * s + v.im * c};
    }
}

// In-register DFT of N points (N | 64), natural order in and out.
template <int N, int SIGN, typename R>
__device__ __forceinline__ void dft(cx<R>* v)
{
    if constexpr (N == 1) {
        return;
    } else if constexpr (N == 2) {
        cx<R> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    } else if constexpr (N == 4) {
        cx<R> s02 = v[0] + v[2], d02 = v[0] - v[2];
        cx<R> s13 = v[1] + v[3], d13 = muli<SIGN>(v[1] - v[3]);
        v[0] = s02 + s13;
        v[2] = s02 - s13;
        v[1] = d02 + d13;
        v[3] = d02 - d13;
    } else if constexpr (SIGN > 0) {
        // decimation in frequency: the mirror (adjoint) of the forward
        // decimation in time below, X[k + M q] -> x[P m + p]
        constexpr int P = (N % 16 == 0) ? 4 : 2;
        constexpr int M = N / P;
        cx<R> sub[P][M];
#pragma unroll
        for (int k = 0; k < M; ++k) {
            cx<R> z[P];
#pragma unroll
            for (int q = 0; q < P; ++q) z[q] = v[k + M * q];
            dft<P, SIGN, R>(z);
#pragma unroll
            for (int p = 0; p < P; ++p) sub[p][k] = twid<N, SIGN, R>(z[p], p * k);
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            dft<M, SIGN, R>(sub[p]);
#pragma unroll
            for (int m = 0; m < M; ++m) v[P * m + p] = sub[p][m];
        }
    } else {
        // decimation in time, N = P*M: x[P m + p] -> X[k + M q]
        constexpr int P = (N % 16 == 0) ? 4 : 2;
        constexpr int M = N / P;
        cx<R> sub[P][M];
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int m = 0; m < M; ++m) sub[p][m] = v[P * m + p];
#pragma unroll
        for (int p = 0; p < P; ++p) dft<M, SIGN, R>(sub[p]);
#pragma unroll
        for (int k = 0; k < M; ++k) {
            cx<R> z[P];
#pragma unroll
            for (int p = 0; p < P; ++p) z[p] = twid<N, SIGN, R>(sub[p][k], p * k);
            dft<P, SIGN, R>(z);
#pragma unroll
            for (int q = 0; q < P; ++q) v[k + M * q] = z[q];
        }
    }
}

// ----------------------------------------------------------- sine/cosine
// sin and cos of one argument for the nonlinear phase, by Cody-Waite
// argument reduction (x = k pi/2 + r, pi/2 split in three parts applied with
// fused multiply-adds; W. J. Cody and W. Waite, "Software Manual for the
// Elementary Functions", 1980) and Taylor series on |r| <= pi/4, truncated
// where the next term is below half an ulp.  Arguments beyond the range in
// which the three-part reduction is exact go to the library routine.
__device__ __forceinline__ void sincos_cw(float x, float* sp, float* cp)
{
    if (!(fabsf(x) <= 1.0e4f)) {
        sincosf(x, sp, cp);
        return;
    }
    const float k = rintf(x * 0.636619746685028076171875f);
    float r = fmaf(-k, 1.57079637050628662109375f, x);
    r = fmaf(-k, -4.37113882867379300296306610107421875e-8f, r);
    r = fmaf(-k, -1.7151245100058819e-15f, r);
    const float r2 = r * r;
    float ps = fmaf(r2, 2.7557319223985893e-6f, -1.984126984126984e-4f);
    ps = fmaf(r2, ps, 8.333333333333333e-3f);
    ps = fmaf(r2, ps, -1.6666666666666666e-1f);
    const float sn = fmaf(r * r2, ps, r);
    float pc = fmaf(r2, -2.755731922398589e-7f, 2.48015873015873e-5f);
    pc = fmaf(r2, pc, -1.388888888888889e-3f);
    pc = fmaf(r2, pc, 4.1666666666666664e-2f);
    pc = fmaf(r2, pc, -0.5f);
    const float cs = fmaf(r2, pc, 1.0f);
    const int q = (int)k & 3;
    const float a = (q & 1) ? cs : sn;
    const float b = (q & 1) ? sn : cs;
    *sp = (q & 2) ? -a : a;
    *cp = ((q + 1) & 2) ? -b : b;
}

// sin and cos of |y| <= pi/4 by the Taylor series of sincos_cw (no reduction).
__device__ __forceinline__ void sincos_small(float y, float* sp, float* cp)
{
    const float y2 = y * y;
    float ps = fmaf(y2, 2.7557319223985893e-6f, -1.984126984126984e-4f);
    ps = fmaf(y2, ps, 8.333333333333333e-3f);
    ps = fmaf(y2, ps, -1.6666666666666666e-1f);
    *sp = fmaf(y * y2, ps, y);
    float pc = fmaf(y2, -2.755731922398589e-7f, 2.48015873015873e-5f);
    pc = fmaf(y2, pc, -1.388888888888889e-3f);
    pc = fmaf(y2, pc, 4.1666666666666664e-2f);
    pc = fmaf(y2, pc, -0.5f);
    *cp = fmaf(y2, pc, 1.0f);
}

// x reduced to [-pi, pi] by x - k 2pi, 2pi split in three parts (Cody-Waite;
// k 2pi_hi and k 2pi_mid are exact for |k| < 2048); larger |x| reduce in FP64.
__device__ __forceinline__ float reduce_2pi(float x)
{
    if (!(fabsf(x) <= 1.0e4f)) return (float)remainder((double)x, 6.283185307179586476925);
    const float k = rintf(x * 0x1.45f306p-3f);
    float r = fmaf(-k, 0x1.92p+2f, x);
    r = fmaf(-k, 0x1.fb6p-10f, r);
    return fmaf(-k, -0x1.777a5cp-23f, r);
}

// Lifting coefficients (t, s) of the rotation by -x/2 (apply twice for the
// rotation by -x): t = tan(-x/4), s = sin(-x/2), x taken mod 2pi so |t| <= 1.
__device__ __forceinline__ void lift_half_neg(float x, float* tp, float* sp)
{
    const float r = reduce_2pi(x);
    float sq, cq;
    sincos_small(0.25f * r, &sq, &cq);
    *tp = -sq * __builtin_amdgcn_rcpf(cq);
    *sp = -2.0f * sq * cq;
}

__device__ __forceinline__ void sincos_cw(double x, double* sp, double* cp)
{
    if (!(fabs(x) <= 1.0e6)) {
        sincos(x, sp, cp);
        return;
    }
    const double k = rint(x * 0.6366197723675814);
    double r = fma(-k, 1.5707963267948966, x);
    r = fma(-k, 6.123233995736766e-17, r);
    r = fma(-k, -1.4973849048591698e-33, r);
    const double r2 = r * r;
    double ps = fma(r2, 2.8114572543455206e-15, -7.647163731819816e-13);
    ps = fma(r2, ps, 1.6059043836821613e-10);
    ps = fma(r2, ps, -2.505210838544172e-08);
    ps = fma(r2, ps, 2.7557319223985893e-06);
    ps = fma(r2, ps, -0.0001984126984126984);
    ps = fma(r2, ps, 0.008333333333333333);
    ps = fma(r2, ps, -0.16666666666666666);
    const double sn = fma(r * r2, ps, r);
    double pc = fma(r2, -1.5619206968586225e-16, 4.779477332387385e-14);
    pc = fma(r2, pc, -1.1470745597729725e-11);
    pc = fma(r2, pc, 2.08767569878681e-09);
    pc = fma(r2, pc, -2.755731922398589e-07);
    pc = fma(r2, pc, 2.48015873015873e-05);
    pc = fma(r2, pc, -0.001388888888888889);
    pc = fma(r2, pc, 0.041666666666666664);
    pc = fma(r2, pc, -0.5);
    const double cs = fma(r2, pc, 1.0);
    const int q = (int)(long long)k & 3;
    const double a = (q & 1) ? cs : sn;
    const double b = (q & 1) ? sn : cs;
    *sp = (q & 2) ? -a : a;
    *cp = ((q + 1) & 2) ? -b : b;
}

// ---------------------------------------------------------------- reductions
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace sfd
