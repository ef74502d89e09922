// sfd_sweep.hpp -- the fused one-axis sweep kernel of the split step.
//
// One launch processes every line of the field along one axis.  A line of
// N = N1*N2 points is owned by N2 lanes; lane t keeps x[N2*j1 + t]
// (j1 < N1) in registers.  A round trip (one axis of solveLapRot!,
// splitting.jl:188-201) is
//
//   load line -> N1-point DFTs -> twiddle w_N^(t s) -> LDS exchange ->
//   N2-point DFTs (spectrum X[s + N1 k2] now in registers) ->
//   multiply by the phase table exp(i(cD xi^2 + r) tau)/N ->
//   inverse N2-point DFTs -> conj twiddle -> LDS exchange ->
//   inverse N1-point DFTs -> [nonlinear phase] -> store
//
// so each sweep reads and writes the field once; no transposes in HBM.
//
// Two line flavours share the code:
//   contiguous (TS == 1): the x axis, lines are consecutive runs of N values,
//                         lane t fastest in the workgroup;
//   strided    (TS  > 1): the y or z axis, element stride S; TS adjacent
//                         x-columns sit in adjacent lanes, so every load and
//                         store is a TS*sizeof(cx) contiguous segment.
#pragma once
#include "sfd_dft.hpp"

namespace sfd {

enum SweepMode : int {
    kRoundTrip = 0,  // forward, multiply, inverse (the split-step sweep)
    kForward = 1,    // forward only, spectrum stored in natural order
    kInverse = 2,    // inverse only (normalised by 1/N), spectrum read in natural order
    kEnergy = 3,     // forward only, spectral and pointwise sums, nothing stored
};

enum SweepFlags : int {
    kFlagNL = 1,        // epilogue: psi *= exp(-i (V + beta|psi|^2) dt)
    kFlagImag = 2,      // imaginary time: NL factor exp(-(V + beta|psi|^2) dt)
    kFlagNorm = 4,      // epilogue: per-workgroup sum |psi|^2 into partial[]
    kFlagScale = 8,     // prologue: psi *= *scale (pending renormalisation)
    kFlagPointwise = 16 // energy mode: also sum |psi|^2, V|psi|^2, |psi|^4
};

struct SweepArgs {
    void* psi;              // field (in; out unless `out` is set)
    void* out;              // output for kForward/kInverse (may alias psi)
    const void* V;          // real potential (NL, pointwise energy); may be null
    const void* phase;      // multiplier table, [row * phase_rows_stride + k]
    const void* tw;         // w_N^m = exp(-2 pi i m/N), m < N
    const double* xi;       // wavenumbers along the axis (energy)
    const double* coord;    // coordinate indexing the line (energy rotation sums)
    const double* scale;    // pending scale factor (kFlagScale)
    double* partial;        // per-workgroup partial sums, [k * nblocks + block]
    long long S;            // element stride along the axis
    long long OS;           // stride of the "other" line index (strided flavour)
    long long nlines;       // contiguous flavour: number of lines
    int ncols;              // strided flavour: number of x-columns
    int row_mod;            // contiguous: phase row = line % row_mod (0: row 0)
    int row_from_col;       // strided: phase row = column (else row 0)
    int phase_row_stride;   // 0 when the multiplier does not depend on the row
    int flags;
    int coord_mod;          // contiguous: coord index = line % coord_mod
    double beta;            // NL coefficient
    double dt;              // NL time step
};

template <typename R, int N1, int N2, int TS, int G>
struct SweepShape {
    static constexpr int N = N1 * N2;
    static constexpr int NL = TS * G;          // lines per workgroup
    static constexpr int NT = NL * N2;         // threads per workgroup
    static constexpr int RB = N1 / N2;         // phase-B rows per lane
    static constexpr int PAD = (TS == 1) ? 1 : 0;
    static constexpr int N2P = N2 + PAD;
    static constexpr int LDS = NL * N2P * N1;  // complex slots
    static_assert(N1 >= N2 && N1 % N2 == 0, "N1 must be a multiple of N2");
    // slot of Y[s][t] of local line L = c + TS*g
    __device__ static __forceinline__ int slot(int c, int g, int s, int t)
    {
        return c + TS * (t + N2P * (s + N1 * g));
    }
};

template <typename R>
__device__ __forceinline__ void nl_phase(cx<R>& v, R vpot, R beta, R dt, bool imag)
{
    const R a2 = v.re * v.re + v.im * v.im;
    const R th = (vpot + beta * a2) * dt;
    if (imag) {
        R e;
        if constexpr (sizeof(R) == 4) e = expf(-th);
        else e = exp(-th);
        v.re *= e;
        v.im *= e;
    } else {
        R s, c;
        if constexpr (sizeof(R) == 4) sincosf(th, &s, &c);
        else sincos(th, &s, &c);
        // exp(-i th) * v
        const cx<R> m = {c, -s};
        v = cmul(m, v);
    }
}

template <typename R, int NT, int K>
__device__ __forceinline__ void block_partials(double (&acc)[K], double* partial, long long nblocks,
                                               long long block)
{
    __shared__ double red[K][NT / 64 > 0 ? NT / 64 : 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = wave_sum(acc[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) red[k][wave] = acc[k];
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0.0;
        for (int w = 0; w < (NT + 63) / 64; ++w) s += red[threadIdx.x][w];
        partial[threadIdx.x * nblocks + block] = s;
    }
}

template <typename R, int N1, int N2, int TS, int G, int MODE>
__global__ __launch_bounds__(TS * G * N2) void k_sweep(const SweepArgs a)
{
    using Sh = SweepShape<R, N1, N2, TS, G>;
    constexpr int N = Sh::N;
    constexpr int RB = Sh::RB;
    __shared__ cx<R> lds[Sh::LDS];

    const int tid = threadIdx.x;
    const int c = tid % TS;
    const int t = (tid / TS) % N2;
    const int g = tid / (TS * N2);
    const int L = c + TS * g;

    // ---- line geometry
    long long base;
    int row = 0;
    long long coord_idx = 0;
    bool active;
    if constexpr (TS == 1) {
        const long long line = (long long)blockIdx.x * Sh::NL + L;
        active = line < a.nlines;
        base = line * N;
        if (a.row_mod > 0) row = (int)(line % a.row_mod);
        if (a.coord_mod > 0) coord_idx = line % a.coord_mod;
    } else {
        const int col = blockIdx.x * Sh::NL + L;
        active = col < a.ncols;
        base = (long long)col + (long long)blockIdx.y * a.OS;
        if (a.row_from_col) row = col;
        coord_idx = col;
    }
    const long long S = a.S;
    const int flags = a.flags;
    const cx<R>* __restrict__ tw = static_cast<const cx<R>*>(a.tw);
    cx<R>* __restrict__ psi = static_cast<cx<R>*>(a.psi);
    const R* __restrict__ V = static_cast<const R*>(a.V);

    cx<R> v[N1];
    R vp[N1];
    double acc[5] = {0, 0, 0, 0, 0};

    if constexpr (MODE != kInverse) {
        // ---- phase A: load x[N2 j1 + t]
        if (active) {
#pragma unroll
            for (int j = 0; j < N1; ++j) v[j] = psi[base + (long long)(N2 * j + t) * S];
        } else {
#pragma unroll
            for (int j = 0; j < N1; ++j) v[j] = {R(0), R(0)};
        }
        const bool need_v = (MODE == kRoundTrip && (flags & kFlagNL)) ||
                            (MODE == kEnergy && (flags & kFlagPointwise));
        if (need_v) {
#pragma unroll
            for (int j = 0; j < N1; ++j) vp[j] = (active && V) ? V[base + (long long)(N2 * j + t) * S] : R(0);
        }
        if (MODE == kRoundTrip && (flags & kFlagScale)) {
            const R sc = R(*a.scale);
#pragma unroll
            for (int j = 0; j < N1; ++j) v[j] = cscale(v[j], sc);
        }
        if (MODE == kEnergy && (flags & kFlagPointwise)) {
#pragma unroll
            for (int j = 0; j < N1; ++j) {
                const double re = v[j].re, im = v[j].im;
                const double a2 = re * re + im * im;
                acc[2] += a2;
                acc[3] += (double)vp[j] * a2;
                acc[4] += a2 * a2;
            }
        }
        dft<N1, -1, R>(v);
#pragma unroll
        for (int s = 1; s < N1; ++s) v[s] = cmul(v[s], tw[t * s]);
#pragma unroll
        for (int s = 0; s < N1; ++s) lds[Sh::slot(c, g, s, t)] = v[s];
        __syncthreads();
    }

    // ---- phase B: rows s = t + N2 r hold X[s + N1 k2]
    cx<R> w[RB][N2];
    if constexpr (MODE != kInverse) {
#pragma unroll
        for (int r = 0; r < RB; ++r) {
#pragma unroll
            for (int q = 0; q < N2; ++q) w[r][q] = lds[Sh::slot(c, g, t + N2 * r, q)];
            dft<N2, -1, R>(w[r]);
        }
    } else {
        if (active) {
            const cx<R>* __restrict__ src = static_cast<const cx<R>*>(a.psi);
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) w[r][k2] = src[base + (long long)(t + N2 * r + N1 * k2) * S];
        } else {
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) w[r][k2] = {R(0), R(0)};
        }
    }

    if constexpr (MODE == kForward) {
        if (active) {
            cx<R>* __restrict__ dst = static_cast<cx<R>*>(a.out);
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) dst[base + (long long)(t + N2 * r + N1 * k2) * S] = w[r][k2];
        }
        return;
    }

    if constexpr (MODE == kEnergy) {
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) {
                const double xi = a.xi[t + N2 * r + N1 * k2];
                const double re = w[r][k2].re, im = w[r][k2].im;
                const double p = re * re + im * im;
                acc[0] += xi * xi * p;
                acc[1] += xi * p;
            }
        if (a.coord) acc[1] *= active ? a.coord[coord_idx] : 0.0;
        const long long nblocks = (long long)gridDim.x * gridDim.y;
        const long long block = (long long)blockIdx.y * gridDim.x + blockIdx.x;
        block_partials<R, Sh::NT, 5>(acc, a.partial, nblocks, block);
        return;
    }

    if constexpr (MODE == kRoundTrip) {
        const cx<R>* __restrict__ ph =
            static_cast<const cx<R>*>(a.phase) + (long long)row * a.phase_row_stride;
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) w[r][k2] = cmul(w[r][k2], ph[t + N2 * r + N1 * k2]);
    }

    // ---- inverse phase B, conj twiddles, exchange
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        dft<N2, +1, R>(w[r]);
        const int s = t + N2 * r;
#pragma unroll
        for (int q = 0; q < N2; ++q) {
            const cx<R> z = (q == 0 || s == 0) ? w[r][q] : cmulc(w[r][q], tw[q * s]);
            lds[Sh::slot(c, g, s, q)] = z;
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < N1; ++s) v[s] = lds[Sh::slot(c, g, s, t)];
    dft<N1, +1, R>(v);

    if constexpr (MODE == kInverse) {
        const R inv = R(1) / R(N);
#pragma unroll
        for (int j = 0; j < N1; ++j) v[j] = cscale(v[j], inv);
    }

    if constexpr (MODE == kRoundTrip) {
        if (flags & kFlagNL) {
            const R beta = R(a.beta), dt = R(a.dt);
            const bool imag = flags & kFlagImag;
#pragma unroll
            for (int j = 0; j < N1; ++j) nl_phase<R>(v[j], vp[j], beta, dt, imag);
        }
        if (flags & kFlagNorm) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < N1; ++j) {
                const double re = v[j].re, im = v[j].im;
                s += re * re + im * im;
            }
            double acc1[1] = {active ? s : 0.0};
            const long long nblocks = (long long)gridDim.x * gridDim.y;
            const long long block = (long long)blockIdx.y * gridDim.x + blockIdx.x;
            block_partials<R, Sh::NT, 1>(acc1, a.partial, nblocks, block);
        }
    }

    if (active) {
        cx<R>* __restrict__ dst = static_cast<cx<R>*>(MODE == kInverse ? a.out : a.psi);
#pragma unroll
        for (int j = 0; j < N1; ++j) dst[base + (long long)(N2 * j + t) * S] = v[j];
    }
}

}  // namespace sfd
