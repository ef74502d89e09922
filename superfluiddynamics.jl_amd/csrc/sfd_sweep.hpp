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
    kRoundTripImag = 4,  // kRoundTrip for imaginary time (FP32 sweeps: real decay tables)
    kRoundTripEnergy = 5,      // kRoundTrip that also sums the energy of the loaded state
    kRoundTripImagEnergy = 6,  // kRoundTripImag that also sums the energy of the loaded state
    kRoundTripTwice = 7,         // round trip, nonlinear phase, round trip (same multiplier)
    kRoundTripTwiceImag = 8,     // kRoundTripTwice for imaginary time (FP32 sweeps)
    kRoundTripTwiceEnergy = 9,   // kRoundTripTwice whose first round trip sums the energy
    kRoundTripTwiceEnergyX = 10, // two round trips (x axis of chained steps) whose second round trip sums
                                 // the energy of the state between them
};

constexpr bool is_round_trip(int mode)
{
    return mode == kRoundTrip || mode == kRoundTripImag || mode == kRoundTripEnergy ||
           mode == kRoundTripImagEnergy;
}
constexpr bool rt_twice(int mode)
{
    return mode == kRoundTripTwice || mode == kRoundTripTwiceImag || mode == kRoundTripTwiceEnergy ||
           mode == kRoundTripTwiceEnergyX;
}
constexpr bool rt_imag(int mode)
{
    return mode == kRoundTripImag || mode == kRoundTripImagEnergy || mode == kRoundTripTwiceImag;
}
constexpr bool rt_energy(int mode)
{
    return mode == kRoundTripEnergy || mode == kRoundTripImagEnergy || mode == kRoundTripTwiceEnergy ||
           mode == kRoundTripTwiceEnergyX;
}

enum SweepFlags : int {
    kFlagNL = 1,        // epilogue: psi *= exp(-i (V + beta|psi|^2) dt)
    kFlagImag = 2,      // imaginary time: NL factor exp(-(V + beta|psi|^2) dt)
    kFlagNorm = 4,      // epilogue: per-workgroup sum |psi|^2 into partial[]
    kFlagScale = 8,     // prologue: psi *= *scale (pending renormalisation)
    kFlagPointwise = 16, // energy sums: also sum |psi|^2, V|psi|^2, |psi|^4
    kFlagNormMid = 32    // double round trip: the norm partials of the state between its round trips
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
    long long ntiles;       // workgroup tiles of the launch
    int ntx;                // strided flavour: x-tiles per "other" index
    void* save;             // kFlagNL: the state before the nonlinear phase is also stored here
    double* epartial;       // energy round trips: per-workgroup sums [k * nblocks + block], k < 4
    int nlen;               // Bluestein sweeps: points per line (the transform runs on M >= 2 nlen - 1)
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
    } else if constexpr (sizeof(R) == 4) {
        // exp(-i th) v as two lifting rotations by -th/2.  Every lifting step
        // adds a first-order increment whose rounding is unbiased; a complex
        // multiply by (cos th, -sin th) instead drops the second-order part
        // 1 - cos th whenever it is below half an ulp, the same way at every
        // site of a near-uniform density, and so drifts the norm.
        float t, s;
        lift_half_neg(th, &t, &s);
        v = lift(lift(v, t, s), t, s);
    } else {
        R s, c;
        sincos_cw(th, &s, &c);
        // exp(-i th) * v
        const cx<R> m = {c, -s};
        v = cmul(m, v);
    }
}

// The real-time nonlinear phase of a lane's N1 FP32 sites.  When every site
// of the wavefront has |th| <= pi/2 (the usual case: th = (V + beta|psi|^2)
// dt is a small angle), each site takes one lifting rotation by -th, with
// t = -tan(th/2) (|t| <= 1) and s = -sin th from the sine and cosine of th/2
// (|th/2| <= pi/4: no 2 pi reduction), instead of two rotations by -th/2;
// otherwise every site goes through nl_phase's reduced, halved rotations.
template <int N1>
__device__ __forceinline__ void nl_phase_lane(cx<float> (&v)[N1], const float (&vp)[N1], float beta, float dt)
{
    float th[N1];
    float m = 0.0f;
#pragma unroll
    for (int j = 0; j < N1; ++j) {
        const float a2 = v[j].re * v[j].re + v[j].im * v[j].im;
        th[j] = (vp[j] + beta * a2) * dt;
        m = fmaxf(m, fabsf(th[j]));
    }
    if (__builtin_amdgcn_ballot_w64(!(m <= 1.5707963f)) == 0) {
#pragma unroll
        for (int j = 0; j < N1; ++j) {
            float sh, ch;
            sincos_small(0.5f * th[j], &sh, &ch);
            v[j] = lift(v[j], -sh * __builtin_amdgcn_rcpf(ch), -2.0f * sh * ch);
        }
    } else {
#pragma unroll
        for (int j = 0; j < N1; ++j) {
            float t, sn;
            lift_half_neg(th[j], &t, &sn);
            v[j] = lift(lift(v[j], t, sn), t, sn);
        }
    }
}

// Workgroup sums of acc[K] into partial[k * nblocks + block]; ACC adds them
// to what the slot holds (a slot per workgroup: no atomics).
template <typename R, int NT, int K, bool ACC = false>
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
        if constexpr (ACC) partial[threadIdx.x * nblocks + block] += s;
        else partial[threadIdx.x * nblocks + block] = s;
    }
}

// s_waitcnt vmcnt(N) (gfx9 encoding: vmcnt in bits 3:0 and 15:14, expcnt
// and lgkmcnt left at their no-wait maxima).
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
    static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Workgroup barrier that orders LDS only.  __syncthreads() also fences
// global memory, i.e. waits for every outstanding global load, including a
// tile prefetched for later; the sweeps exchange data through LDS alone.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Geometry of one workgroup tile: a uniform base (SGPRs) plus a 32-bit lane
// offset; element j of the lane's line sits at base + lane + j*jstride.
// Lanes past the end of the field read a clamped (valid) line and never
// store, so no load needs a predicate.
struct TileGeom {
    long long ubase, jstride, kstride;
    unsigned lane;
    int row;
    long long coord_idx;
    bool active;
};

template <int N, int N2, int TS, int NL>
__device__ __forceinline__ TileGeom tile_geom(const SweepArgs& a, long long tile, int L, int t)
{
    TileGeom g;
    g.row = 0;
    g.coord_idx = 0;
    if constexpr (TS == 1) {
        const long long line0 = tile * NL;
        const long long rem = a.nlines - line0;
        const int valid = rem < NL ? (int)rem : NL;
        const int Lc = L < valid ? L : valid - 1;
        g.active = L < valid;
        g.ubase = line0 * N;
        g.lane = (unsigned)(Lc * N + t);
        g.jstride = N2;
        g.kstride = 1;
        const long long line = line0 + Lc;
        if (a.row_mod > 0) g.row = (int)(line % a.row_mod);
        if (a.coord_mod > 0) g.coord_idx = line % a.coord_mod;
    } else {
        const long long o = tile / a.ntx;
        const int col0 = (int)(tile - o * a.ntx) * NL;
        const int rem = a.ncols - col0;
        const int valid = rem < NL ? rem : NL;
        const int Lc = L < valid ? L : valid - 1;
        g.active = L < valid;
        g.ubase = (long long)col0 + o * a.OS;
        g.lane = (unsigned)Lc + (unsigned)t * (unsigned)a.S;
        g.jstride = (long long)N2 * a.S;
        g.kstride = a.S;
        if (a.row_from_col) g.row = col0 + Lc;
        g.coord_idx = col0 + Lc;
    }
    return g;
}

template <typename T>
__device__ __forceinline__ T& lane_ref(T* base, long long off, unsigned lane)
{
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(base + off) + (size_t)lane * sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T& lane_ref(const T* base, long long off, unsigned lane)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base + off) + (size_t)lane * sizeof(T));
}

// Storage precision S (HBM) and compute precision R (registers, LDS) may
// differ: complex FP32 fields can be transformed in FP64 arithmetic.
template <typename R, typename S>
__device__ __forceinline__ cx<R> widen(cx<S> x) { return {R(x.re), R(x.im)}; }
template <typename S, typename R>
__device__ __forceinline__ cx<S> narrow(cx<R> x) { return {S(x.re), S(x.im)}; }

template <typename S, typename R, int N1>
__device__ __forceinline__ void load_line(cx<R> (&v)[N1], const cx<S>* psi, const TileGeom& g)
{
    const cx<S>* b = psi + g.ubase;
#pragma unroll
    for (int j = 0; j < N1; ++j) v[j] = widen<R>(lane_ref(b, j * g.jstride, g.lane));
}

template <typename S, typename R, int N1>
__device__ __forceinline__ void load_pot(R (&vp)[N1], const S* V, const TileGeom& g)
{
    if (V) {
        const S* b = V + g.ubase;
#pragma unroll
        for (int j = 0; j < N1; ++j) vp[j] = R(lane_ref(b, j * g.jstride, g.lane));
    } else {
#pragma unroll
        for (int j = 0; j < N1; ++j) vp[j] = R(0);
    }
}

template <typename S, typename R, int N1>
__device__ __forceinline__ void store_line(const cx<R> (&v)[N1], cx<S>* psi, const TileGeom& g)
{
    if (g.active) {
        cx<S>* b = psi + g.ubase;
#pragma unroll
        for (int j = 0; j < N1; ++j) lane_ref(b, j * g.jstride, g.lane) = narrow<S>(v[j]);
    }
}

// Inter-stage twiddles.  FP64 tables hold w = exp(-2 pi i e/n) and the
// inverse multiplies by conj(w).  FP32 tables hold the lifting coefficients
// (tan(b/2), sin(b)) of the half angle b = arg(w)/2 in (-pi/2, pi/2]: the
// forward applies that rotation twice and the inverse applies its exact
// inverse twice, so FP32 round trips keep no systematic error from the
// rounding of the tables.
template <typename R>
__device__ __forceinline__ cx<R> tw_mul(cx<R> v, cx<R> w)
{
    if constexpr (sizeof(R) == 4) return lift(lift(v, w.re, w.im), w.re, w.im);
    else return cmul(v, w);
}
template <typename R>
__device__ __forceinline__ cx<R> tw_mulc(cx<R> v, cx<R> w)
{
    if constexpr (sizeof(R) == 4) return lift(lift(v, -w.re, -w.im), -w.re, -w.im);
    else return cmulc(v, w);
}

// Phase-table row of a tile in global memory (rows differ when Omega != 0).
template <typename R>
__device__ __forceinline__ const cx<R>* row_phase(const SweepArgs& a, const TileGeom& gm)
{
    return static_cast<const cx<R>*>(a.phase) + (long long)gm.row * a.phase_row_stride;
}

// store_line without the lane predicate: lanes past the end of the field
// hold a clamped copy of the last valid line, transformed by the same
// operations, so they write bit-identical values to its addresses.  Keeps the
// number of issued stores independent of the tile, which lets the compiler
// wait for a prefetched tile without also waiting for these stores.
template <typename S, typename R, int N1>
__device__ __forceinline__ void store_line_all(const cx<R> (&v)[N1], cx<S>* psi, const TileGeom& g)
{
    cx<S>* b = psi + g.ubase;
#pragma unroll
    for (int j = 0; j < N1; ++j) lane_ref(b, j * g.jstride, g.lane) = narrow<S>(v[j]);
}

// Centered round trips (FP32 arithmetic, whole lines of N = N1*N2 points).
// Lane t holds the positions N2 j + t_c (mod N) with t_c = t - N2/2, and the
// stage-A frequencies s < N1 stand for s_c = s - N1 [s >= N1/2], so that
// every inter-stage twiddle w_N^(t_c s_c) has |t_c s_c| <= N/4: an angle in
// [-pi/2, pi/2], which one lifting rotation (tan(a/2) <= 1, 3 FMAs) applies
// exactly invertibly, where the natural index ranges need two half-angle
// lifts (6 FMAs).  Stage B then leaves (-1)^k2 X[s_c + N1 k2]: the sign is
// undone by the inverse stage B of the same round trip, since the phase
// multiply in between is diagonal in k.  Only element j = 0 of lanes
// t < N2/2 wraps to the end of the line; ct_lane0 is its lane offset.
template <typename R, int N1, int N2, int H>
constexpr bool centered_ok = sizeof(R) == 4 && H == 1 && N1 >= 2 && N2 >= 2;

template <int N, int N2>
__device__ __forceinline__ unsigned ct_lane0(const TileGeom& g, int t)
{
    const long long shift = t < N2 / 2 ? N - N2 / 2 : -(N2 / 2);
    return (unsigned)((long long)g.lane + shift * g.kstride);
}

template <typename S, typename R, int N1, int N2>
__device__ __forceinline__ void load_line_ct(cx<R> (&v)[N1], const cx<S>* psi, const TileGeom& g, unsigned lane0)
{
    const cx<S>* b = psi + g.ubase;
    v[0] = widen<R>(lane_ref(b, 0, lane0));
#pragma unroll
    for (int j = 1; j < N1; ++j) v[j] = widen<R>(lane_ref(b, j * g.jstride - (N2 / 2) * g.kstride, g.lane));
}

template <typename S, typename R, int N1, int N2>
__device__ __forceinline__ void load_pot_ct(R (&vp)[N1], const S* V, const TileGeom& g, unsigned lane0)
{
    if (V) {
        const S* b = V + g.ubase;
        vp[0] = R(lane_ref(b, 0, lane0));
#pragma unroll
        for (int j = 1; j < N1; ++j) vp[j] = R(lane_ref(b, j * g.jstride - (N2 / 2) * g.kstride, g.lane));
    } else {
#pragma unroll
        for (int j = 0; j < N1; ++j) vp[j] = R(0);
    }
}

// ALL: also the lanes past the end of the field (bit-identical copies).
template <typename S, typename R, int N1, int N2, bool ALL = false>
__device__ __forceinline__ void store_line_ct(const cx<R> (&v)[N1], cx<S>* psi, const TileGeom& g, unsigned lane0)
{
    if (ALL || g.active) {
        cx<S>* b = psi + g.ubase;
        lane_ref(b, 0, lane0) = narrow<S>(v[0]);
#pragma unroll
        for (int j = 1; j < N1; ++j) lane_ref(b, j * g.jstride - (N2 / 2) * g.kstride, g.lane) = narrow<S>(v[j]);
    }
}

// Forward transform of the tile's lines: phase A (N1-point DFTs, twiddles),
// LDS exchange, phase B (N2-point DFTs).  On return w[r][k2] holds
// X[s + N1 k2] of row s = t + N2 r.
// CT: centered index sets (see centered_ok): twA holds single-lift
// coefficients of the centered angles.
template <typename R, int N1, int N2, int TS, int G, bool CT = false>
__device__ __forceinline__ void forward_ab(cx<R> (&v)[N1], cx<R> (&w)[N1 / N2][N2], cx<R>* lds, const cx<R>* twA,
                                           int c, int g, int t)
{
    using Sh = SweepShape<R, N1, N2, TS, G>;
    dft<N1, -1, R>(v);
#pragma unroll
    for (int s = 1; s < N1; ++s) {
        if constexpr (CT) v[s] = lift(v[s], twA[s * N2].re, twA[s * N2].im);
        else v[s] = tw_mul(v[s], twA[s * N2]);
    }
#pragma unroll
    for (int s = 0; s < N1; ++s) lds[Sh::slot(c, g, s, t)] = v[s];
    lds_barrier();
#pragma unroll
    for (int r = 0; r < N1 / N2; ++r) {
#pragma unroll
        for (int q = 0; q < N2; ++q) w[r][q] = lds[Sh::slot(c, g, t + N2 * r, q)];
        dft<N2, -1, R>(w[r]);
    }
}

// Inverse transform (unscaled): inverse phase B, conjugate twiddles, LDS
// exchange, inverse phase A; on return v[j] = N * x[N2 j + t].  Each lane
// only rewrites the LDS slots it read in forward_ab, so no barrier is needed
// before the writes.  Likewise the next forward_ab on the same buffer writes
// exactly the slots this lane reads at the end of inverse_ba, so round trips
// may follow each other (in one tile or over consecutive tiles) with no
// barrier in between: every cross-lane read sits behind the barrier that
// follows the writes it needs.
// CT: centered index sets; column q = 0 stands for q_c = -N2/2 and takes a
// twiddle too.
template <typename R, int N1, int N2, int TS, int G, bool CT = false>
__device__ __forceinline__ void inverse_ba(cx<R> (&w)[N1 / N2][N2], cx<R> (&v)[N1], cx<R>* lds, const cx<R>* twB,
                                           int c, int g, int t)
{
    using Sh = SweepShape<R, N1, N2, TS, G>;
#pragma unroll
    for (int r = 0; r < N1 / N2; ++r) {
        dft<N2, +1, R>(w[r]);
        const int s = t + N2 * r;
#pragma unroll
        for (int q = 0; q < N2; ++q) {
            cx<R> z;
            if constexpr (CT) {
                const cx<R> b = twB[q * N1 + N2 * r];
                z = lift(w[r][q], -b.re, -b.im);
            } else {
                z = (q == 0) ? w[r][q] : tw_mulc(w[r][q], twB[q * N1 + N2 * r]);
            }
            lds[Sh::slot(c, g, s, q)] = z;
        }
    }
    lds_barrier();
#pragma unroll
    for (int s = 0; s < N1; ++s) v[s] = lds[Sh::slot(c, g, s, t)];
    dft<N1, +1, R>(v);
}

// Lines of N = 2M points (H = 2 variants): the line's two lanes per
// sub-transform lane t hold its even (h = 0) and odd (h = 1) samples,
// y_h[m] = x[2m + h], and forward_ab / inverse_ba run the M-point transforms
// of y_0 and y_1.  The radix-2 stage joins them across the lane pair, which
// differ in bit log2(TS) of the thread index (same wavefront, TS < 64):
//   X[k] = Y0[k] + w^k Y1[k],  X[k + M] = Y0[k] - w^k Y1[k],  w = e^(-2 pi i/N),
// so lane h ends up with X[k + M h] for its k = t + N2 r + N1 k2.  twC holds
// w^k (k < M) in the twiddle format, pre-offset by t.
template <typename R>
__device__ __forceinline__ cx<R> shfl_xor_cx(cx<R> v, int mask)
{
    return {__shfl_xor(v.re, mask), __shfl_xor(v.im, mask)};
}

template <typename R, int N1, int N2, int TS>
__device__ __forceinline__ void radix2_fwd(cx<R> (&w)[N1 / N2][N2], const cx<R>* twC, int h)
{
    static_assert(TS < 64, "the lane pair must share a wavefront");
#pragma unroll
    for (int r = 0; r < N1 / N2; ++r)
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            const cx<R> x = w[r][k2];
            const cx<R> wy = tw_mul(x, twC[N2 * r + N1 * k2]);
            const cx<R> m = h ? wy : x;  // Y0 on even lanes, w^k Y1 on odd lanes
            const cx<R> o = shfl_xor_cx(m, TS);
            w[r][k2] = h ? cx<R>{o.re - m.re, o.im - m.im} : cx<R>{m.re + o.re, m.im + o.im};
        }
}

// Inverse of radix2_fwd (unscaled): lane h gets 2 Y_h[k].
template <typename R, int N1, int N2, int TS>
__device__ __forceinline__ void radix2_inv(cx<R> (&w)[N1 / N2][N2], const cx<R>* twC, int h)
{
#pragma unroll
    for (int r = 0; r < N1 / N2; ++r)
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            const cx<R> x = w[r][k2];
            const cx<R> o = shfl_xor_cx(x, TS);
            const cx<R> d = {o.re - x.re, o.im - x.im};  // X[k] - X[k + M] on odd lanes
            const cx<R> y = tw_mulc(d, twC[N2 * r + N1 * k2]);
            w[r][k2] = h ? y : cx<R>{x.re + o.re, x.im + o.im};
        }
}

// The round trip of one tile in registers: forward, phase multiply,
// inverse, then the fused epilogue (nonlinear phase, norm partial sum).
// IMAG selects the imaginary-time form of the FP32 phase multiply and NL
// factor at compile time (FP64 arithmetic reads kFlagImag at run time).
// EN adds the energy sums of the state entering the sweep (after a pending
// scale) to eacc[4], in FP64: [0] sum xi^2 |psi_hat|^2 over the transformed
// line, and with kFlagPointwise in a contiguous (x) sweep [1] sum |psi|^2,
// [2] sum V|psi|^2, [3] sum |psi|^4 (vp must then hold V).  FP32 arithmetic
// sums a lane's N1 points in FP32 first (one FP64 add per sum and lane).
// H = 2: the line is two interleaved halves joined by the radix-2 stage; g is
// then the thread's half-line group in LDS, t its sub-transform lane and h
// its half; twC the radix-2 twiddles (unused for H = 1).
// CT: centered index sets (see centered_ok; twA, twB then point at the
// centered tables, and lane0 is the wrapped lane offset of element 0).
// SC (FP32 real time): the spectrum's factor 1/N^SC, SC = 0, 1 or 2; N is a
// power of two, so the factor is exact and may be deferred: the double round
// trip applies none in its first round trip and 1/N^2 in its second, with
// beta/N^2 in the nonlinear phase between them (beta/N^2 |N psi|^2 rounds
// to the same float as beta |psi|^2).
template <typename S, typename R, int N1, int N2, int TS, int G, bool IMAG, bool EN = false, int H = 1,
          bool CT = false, int SC = 1>
__device__ __forceinline__ void round_trip_tile(cx<R> (&v)[N1], const R (&vp)[N1], const TileGeom& gm,
                                                const SweepArgs& a, cx<R>* lds, const cx<R>* twA,
                                                const cx<R>* twB, const cx<R>* twC, const cx<R>* phase, int c,
                                                int g, int t, int h, double& norm_acc, double* eacc = nullptr,
                                                unsigned lane0 = 0)
{
    static_assert(H == 1 || H == 2, "one or two halves per line");
    static_assert(!CT || centered_ok<R, N1, N2, H>, "centered round trips: FP32 whole lines");
    constexpr int RB = N1 / N2;
    constexpr int M = N1 * N2;  // sub-transform length; the line has H * M points
    const int flags = a.flags;
    if (flags & kFlagScale) {
        const R sc = R(*a.scale);
#pragma unroll
        for (int j = 0; j < N1; ++j) v[j] = cscale(v[j], sc);
    }
    if constexpr (EN && TS == 1) {  // the x sweep (contiguous flavour) takes the pointwise sums
        if (flags & kFlagPointwise) {
            if constexpr (sizeof(R) == 4) {
                // FP32 sums over the lane's N1 sites, then one FP64 add each:
                // the double round trips that carry these sums are bound by
                // VALU issue, where FP64 arithmetic runs at half rate
                float s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
#pragma unroll
                for (int j = 0; j < N1; ++j) {
                    const float a2 = fmaf(v[j].re, v[j].re, v[j].im * v[j].im);
                    s1 += a2;
                    s2 = fmaf(vp[j], a2, s2);
                    s3 = fmaf(a2, a2, s3);
                }
                eacc[1] += s1;
                eacc[2] += s2;
                eacc[3] += s3;
            } else {
#pragma unroll
                for (int j = 0; j < N1; ++j) {
                    const double re = v[j].re, im = v[j].im;
                    const double a2 = re * re + im * im;
                    eacc[1] += a2;
                    eacc[2] += (double)vp[j] * a2;
                    eacc[3] += a2 * a2;
                }
            }
        }
    }
    cx<R> w[RB][N2];
    forward_ab<R, N1, N2, TS, G, CT>(v, w, lds, twA, c, g, t);
    if constexpr (H == 2) radix2_fwd<R, N1, N2, TS>(w, twC, h);
    // spectrum index of w[r][k2]: s + N1 k2 with s = t + N2 r; centered, it
    // is s_c + N1 k2 (mod M) = s + N1 ((k2 - d) mod N2), d = [s >= N1/2]:
    // a compile-time shift for RB >= 2 (d = [r >= RB/2]), a lane-dependent
    // one for RB = 1 (d = [t >= N2/2])
    const int dct = (CT && RB == 1 && t >= N2 / 2) ? 1 : 0;
    if constexpr (EN) {
        // xi_k = kk * xi_1 with kk the signed frequency of k (the wavenumber
        // table's own product)
        constexpr int N = H * M;
        const double dk = a.xi[1];
        if constexpr (sizeof(R) == 4) {
            // sum kk^2 |psi_hat|^2 over the lane's points in FP32 (kk^2 is
            // exact), then one FP64 add of dk^2 times it
            float s0 = 0.0f;
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) {
                    const int d = RB == 1 ? dct : (r >= RB / 2 ? 1 : 0);
                    const int k = CT ? t + N2 * r + N1 * ((k2 - d) & (N2 - 1)) : t + M * h + N2 * r + N1 * k2;
                    const float kf = (float)(k < N / 2 ? k : k - N);
                    const float p = fmaf(w[r][k2].re, w[r][k2].re, w[r][k2].im * w[r][k2].im);
                    s0 = fmaf(kf * kf, p, s0);
                }
            eacc[0] += (double)s0 * (dk * dk);
        } else {
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) {
                    const int d = RB == 1 ? dct : (r >= RB / 2 ? 1 : 0);
                    const int k = CT ? t + N2 * r + N1 * ((k2 - d) & (N2 - 1)) : t + M * h + N2 * r + N1 * k2;
                    const double x = (double)(k < N / 2 ? k : k - N) * dk;
                    const double re = w[r][k2].re, im = w[r][k2].im;
                    eacc[0] += x * x * (re * re + im * im);
                }
        }
    }
    // the phase entry of w[r][k2] at that index; for RB = 1 the k2 = 0 entry
    // of lanes with d = 1 wraps to the end of the row
    const cx<R>* __restrict__ ph = phase + t + M * h;
    const cx<R>* __restrict__ ph1 = ph - N1 * dct;                 // k2 >= 1
    const cx<R>* __restrict__ ph0 = ph + dct * N1 * (N2 - 1);      // k2 = 0
    auto phase_at = [&](int r, int k2) -> cx<R> {
        if constexpr (!CT) {
            return ph[N2 * r + N1 * k2];
        } else if constexpr (RB == 1) {
            return k2 == 0 ? ph0[0] : ph1[N1 * k2];
        } else {
            const int d = r >= RB / 2 ? 1 : 0;
            return ph[N2 * r + N1 * ((k2 - d + N2) % N2)];
        }
    };
    if constexpr (sizeof(R) == 4) {
        // FP32: X m/N with X/N exact (N a power of two).  Real-time tables
        // hold the lifting coefficients (tan(b/2), sin b) of half the phase,
        // b = phi/2 with phi in (-pi, pi], applied twice: a complex multiply
        // by rounded (cos phi, sin phi) loses 1 - cos(phi) of the low modes
        // to rounding on every pass and drifts the norm.  Imaginary-time
        // tables hold d = (m - 1)/N (real): X/N + X d.
        static_assert(SC == 1 || !IMAG, "imaginary-time tables hold the factor 1/N");
        constexpr float inv_n = SC == 2 ? 1.0f / (float(H * M) * float(H * M)) : 1.0f / float(H * M);
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) {
                const cx<R> d = phase_at(r, k2);
                const cx<R> x = w[r][k2];
                if constexpr (IMAG)
                    w[r][k2] = {fmaf(x.re, d.re, x.re * inv_n), fmaf(x.im, d.re, x.im * inv_n)};
                else if constexpr (SC == 0)
                    w[r][k2] = lift(lift(x, d.re, d.im), d.re, d.im);
                else
                    w[r][k2] = lift(lift({x.re * inv_n, x.im * inv_n}, d.re, d.im), d.re, d.im);
            }
    } else {
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) w[r][k2] = cmul(w[r][k2], ph[N2 * r + N1 * k2]);
    }
    if constexpr (H == 2) radix2_inv<R, N1, N2, TS>(w, twC, h);
    inverse_ba<R, N1, N2, TS, G, CT>(w, v, lds, twB, c, g, t);
    if (flags & kFlagNL) {
        if (a.save) {
            cx<R> u[N1];
#pragma unroll
            for (int j = 0; j < N1; ++j) u[j] = SC == 0 ? cscale(v[j], R(1) / R(H * M)) : v[j];
            if constexpr (CT) store_line_ct<S, R, N1, N2>(u, static_cast<cx<S>*>(a.save), gm, lane0);
            else store_line<S, R, N1>(u, static_cast<cx<S>*>(a.save), gm);
        }
        const R beta = R(a.beta), dt = R(a.dt);
        if constexpr (sizeof(R) == 4 && !IMAG) {
            nl_phase_lane<N1>(v, vp, beta, dt);
        } else {
            const bool imag = sizeof(R) == 4 ? IMAG : (flags & kFlagImag) != 0;
#pragma unroll
            for (int j = 0; j < N1; ++j) nl_phase<R>(v[j], vp[j], beta, dt, imag);
        }
    }
    if (flags & kFlagNorm) {
        if constexpr (sizeof(R) == 4) {  // the lane's N1 points in FP32, then one FP64 add
            float s = 0.0f;
#pragma unroll
            for (int j = 0; j < N1; ++j) s = fmaf(v[j].re, v[j].re, fmaf(v[j].im, v[j].im, s));
            if (gm.active) norm_acc += (double)s;
        } else {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < N1; ++j) {
                const double re = v[j].re, im = v[j].im;
                s += re * re + im * im;
            }
            if (gm.active) norm_acc += s;
        }
    }
}

// Twiddle entries of a line of H*M points: twA | twB (M each) | twC (M, H = 2).
template <int M, int H>
constexpr int kTwiddles = (H == 2 ? 3 : 2) * M;

// Stages the twiddles (kTwiddles<M, H> entries) and one phase row (H*M
// entries) into LDS, then waits for every lane.
// CT: the centered twiddles, which follow the natural ones in the table.
template <typename R, int M, int H, int NT, bool CT = false>
__device__ __forceinline__ void stage_tables(cx<R>* tabs, const SweepArgs& a, int tid)
{
    constexpr int NTW = kTwiddles<M, H>;
    constexpr int TOT = NTW + H * M;
    const cx<R>* gtw = static_cast<const cx<R>*>(a.tw) + (CT ? NTW : 0);
    const cx<R>* gph = static_cast<const cx<R>*>(a.phase);
#pragma unroll
    for (int k = 0; k < (TOT + NT - 1) / NT; ++k) {
        const int i = tid + k * NT;
        if (TOT % NT == 0 || i < TOT) tabs[i] = i < NTW ? gtw[i] : gph[i - NTW];
    }
    lds_barrier();
}

// One launch over all tiles of an axis, one tile per workgroup; every mode.
// W is the minimum number of wavefronts per SIMD the register allocation
// must leave room for (1 = no constraint).  H = 2: lines of 2*N1*N2 points,
// two lanes per sub-transform lane (see radix2_fwd).  TL: the round trips
// read the twiddle tables from LDS, and the phase table too when it has one
// row (no rotation term on the axis).  ZV: double round trips without a
// potential (V = 0, a null pointer), so that the nonlinear phase holds no
// potential registers (N1 fewer VGPRs).
template <typename S, typename R, int N1, int N2, int TS, int G, int MODE, int W = 1, int H = 1, bool TL = false,
          bool ZV = false>
__global__ __launch_bounds__(TS * G * N2 * H, W) void k_sweep(const SweepArgs a)
{
    using Sh = SweepShape<R, N1, N2, TS, H * G>;  // LDS lines: one per half-line
    constexpr int M = Sh::N;                      // sub-transform length
    constexpr int N = H * M;                      // line length
    constexpr int RB = Sh::RB;
    constexpr int NL = TS * G;                    // lines per workgroup
    __shared__ cx<R> lds[Sh::LDS];

    const int tid = threadIdx.x;
    const int c = tid % TS;
    const int tl = (tid / TS) % (H * N2);  // lane within the line
    const int t = tl / H;                  // sub-transform lane
    const int h = tl % H;                  // half (H = 2: even or odd samples)
    const int g = tid / (TS * N2 * H);
    const int gh = H * g + h;              // LDS line of the half-line
    const int L = c + TS * g;
    const long long tile = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    const TileGeom gm = tile_geom<N, H * N2, TS, NL>(a, tile, L, tl);
    const int flags = a.flags;
    const cx<R>* __restrict__ twA = static_cast<const cx<R>*>(a.tw) + t;
    const cx<R>* __restrict__ twB = static_cast<const cx<R>*>(a.tw) + M + t;
    const cx<R>* __restrict__ twC = static_cast<const cx<R>*>(a.tw) + 2 * M + t;
    cx<S>* __restrict__ psi = static_cast<cx<S>*>(a.psi);
    const long long nblocks = (long long)gridDim.x * gridDim.y;
    // spectrum position of w[r][k2] is k0 + N2 r + N1 k2; its natural-order
    // offset from the lane's own element is (that - tl) * kstride
    const int k0 = t + M * h;
    const int dk = k0 - tl;

    cx<R> v[N1];
    R vp[N1];

    if constexpr (is_round_trip(MODE)) {
        constexpr bool EN = rt_energy(MODE);
        constexpr bool CT = centered_ok<R, N1, N2, H>;
        const unsigned lane0 = CT ? ct_lane0<N, N2>(gm, t) : 0u;
        if constexpr (CT) load_line_ct<S, R, N1, N2>(v, psi, gm, lane0);
        else load_line<S, R, N1>(v, psi, gm);
        if ((flags & kFlagNL) || (EN && TS == 1 && (flags & kFlagPointwise))) {
            if constexpr (CT) load_pot_ct<S, R, N1, N2>(vp, static_cast<const S*>(a.V), gm, lane0);
            else load_pot<S, R, N1>(vp, static_cast<const S*>(a.V), gm);
        }
        double nacc = 0.0;
        double eacc[4] = {0, 0, 0, 0};
        if constexpr (EN || TL || (CT && (RB == 1 || TS == 1))) {
            // Energy sweeps read the twiddle and phase tables (one phase row:
            // Omega = 0) from LDS: with the energy sums' registers live the
            // compiler no longer batches the tables' global loads and would
            // wait on each in turn.  TL variants do so for the same reason
            // (FP64 lines of 512 points), and so do the centered FP32 round
            // trips with N1 = N2, whose lane-dependent phase index defeats
            // that batching too (a strided 256-point round trip at 256^3: 46
            // us from LDS, 85 us from global memory), and those on the
            // contiguous axis (512 points, 4 lines per workgroup: 453 against
            // 959 us at 512^3).  Strided centered round trips with N1 > N2
            // keep them in global memory (512 points: 513 against 666 us).
            // With several phase rows (Omega != 0) the phase is read from
            // global memory.
            constexpr int NTW = kTwiddles<M, H>;
            __shared__ cx<R> tabs[NTW + N];
            stage_tables<R, M, H, Sh::NT, CT>(tabs, a, tid);
            if (EN || a.phase_row_stride == 0)
                round_trip_tile<S, R, N1, N2, TS, H * G, rt_imag(MODE), EN, H, CT>(
                    v, vp, gm, a, lds, tabs + t, tabs + M + t, tabs + 2 * M + t, tabs + NTW, c, gh, t, h, nacc,
                    eacc, lane0);
            else
                round_trip_tile<S, R, N1, N2, TS, H * G, rt_imag(MODE), EN, H, CT>(
                    v, vp, gm, a, lds, tabs + t, tabs + M + t, tabs + 2 * M + t, row_phase<R>(a, gm), c, gh, t, h,
                    nacc, eacc, lane0);
        } else {
            constexpr int CTO = CT ? kTwiddles<M, H> : 0;  // the centered twiddles follow the natural ones
            round_trip_tile<S, R, N1, N2, TS, H * G, rt_imag(MODE), false, H, CT>(
                v, vp, gm, a, lds, twA + CTO, twB + CTO, twC, row_phase<R>(a, gm), c, gh, t, h, nacc, nullptr,
                lane0);
        }
        if (flags & kFlagNorm) {
            double acc1[1] = {nacc};
            block_partials<R, Sh::NT, 1>(acc1, a.partial, nblocks, tile);
        }
        if constexpr (EN) {
            if (!gm.active) {
#pragma unroll
                for (int k = 0; k < 4; ++k) eacc[k] = 0.0;
            }
            block_partials<R, Sh::NT, 4>(eacc, a.epartial, nblocks, tile);
        }
        if constexpr (CT) store_line_ct<S, R, N1, N2>(v, psi, gm, lane0);
        else store_line<S, R, N1>(v, psi, gm);
    } else if constexpr (rt_twice(MODE)) {
        // K_d(tau) N(dt) K_d(tau) on the lines of one axis: the innermost
        // axis of a Strang step, whose flow commutes with the other axes'
        // flows (z in 3D; y in 2D when Omega = 0), so that the two half
        // steps meet at this axis around the nonlinear phase.  One read and
        // one write of the field for both round trips; the phase table (one
        // row: no rotation term on this axis) is the same for both.
        // kRoundTripTwiceEnergy: the first round trip also sums the energy
        // of the loaded state (spectral sum only: this is never the x axis).
        // kRoundTripTwiceEnergyX: the two x round trips of chained steps,
        // the second summing the energy of the state between them (the
        // state at the step boundary; spectral x sum and, with
        // kFlagPointwise, the pointwise sums, V read for them).
        constexpr bool EN = rt_energy(MODE);
        constexpr bool EN1 = MODE == kRoundTripTwiceEnergy;
        constexpr bool EN2 = MODE == kRoundTripTwiceEnergyX;
        constexpr bool CT = centered_ok<R, N1, N2, H>;
        const unsigned lane0 = CT ? ct_lane0<N, N2>(gm, t) : 0u;
        if constexpr (CT) load_line_ct<S, R, N1, N2>(v, psi, gm, lane0);
        else load_line<S, R, N1>(v, psi, gm);
        if constexpr (ZV) {
#pragma unroll
            for (int j = 0; j < N1; ++j) vp[j] = R(0);
        } else if ((flags & kFlagNL) || (EN2 && (flags & kFlagPointwise))) {
            if constexpr (CT) load_pot_ct<S, R, N1, N2>(vp, static_cast<const S*>(a.V), gm, lane0);
            else load_pot<S, R, N1>(vp, static_cast<const S*>(a.V), gm);
        }
        double nacc = 0.0;
        double eacc[4] = {0, 0, 0, 0};
        // the norm partials after the second round trip (kFlagNorm) or, in
        // the x pass of chained imaginary-time steps, after the first
        // (kFlagNormMid: the state that ends a step, before its
        // renormalisation, which the next sweep applies)
        SweepArgs a1 = a;
        a1.flags = (flags & ~(kFlagNorm | kFlagNormMid)) | ((flags & kFlagNormMid) ? kFlagNorm : 0);
        SweepArgs a2 = a;
        a2.flags = flags & ~(kFlagNL | kFlagScale | kFlagNormMid | ((flags & kFlagNormMid) ? kFlagNorm : 0));
        a2.save = nullptr;
        // FP32 (and energy) sweeps read the tables from LDS, staged once
        // per workgroup: 104 instead of 126 VGPRs,
        // and 79 instead of 91 us at 256^3 on MI355X.  FP64 sweeps keep them
        // in global memory: the 12 KiB of staged tables would cut them from
        // 8 to 6 resident wavefronts per CU (189 instead of 168 us).
        if constexpr (sizeof(R) == 4 || EN || TL) {
            constexpr int NTW = kTwiddles<M, H>;
            __shared__ cx<R> tabs[NTW + N];
            stage_tables<R, M, H, Sh::NT, CT>(tabs, a, tid);
            // FP32 real time: the first round trip leaves N psi (no 1/N),
            // the second applies 1/N^2 (see round_trip_tile's SC); not when
            // the second sums the energy of the state between them
            constexpr int SC1 = (sizeof(R) == 4 && !rt_imag(MODE) && !EN2) ? 0 : 1;
            constexpr int SC2 = SC1 == 0 ? 2 : 1;
            if constexpr (SC1 == 0) a1.beta = a.beta / ((double)N * (double)N);
            round_trip_tile<S, R, N1, N2, TS, H * G, rt_imag(MODE), EN1, H, CT, SC1>(
                v, vp, gm, a1, lds, tabs + t, tabs + M + t, tabs + 2 * M + t, tabs + NTW, c, gh, t, h, nacc, eacc,
                lane0);
            if constexpr (EN2 && sizeof(S) != sizeof(R)) {
                // the state between the steps as a step would store it: the
                // energy sums and the next step see the field's precision
#pragma unroll
                for (int j = 0; j < N1; ++j) v[j] = widen<R>(narrow<S>(v[j]));
            }
            round_trip_tile<S, R, N1, N2, TS, H * G, rt_imag(MODE), EN2, H, CT, SC2>(
                v, vp, gm, a2, lds, tabs + t, tabs + M + t, tabs + 2 * M + t, tabs + NTW, c, gh, t, h, nacc,
                eacc, lane0);
        } else {
            const cx<R>* ph = row_phase<R>(a, gm);
            round_trip_tile<S, R, N1, N2, TS, H * G, rt_imag(MODE), false, H>(v, vp, gm, a1, lds, twA, twB, twC,
                                                                              ph, c, gh, t, h, nacc);
            round_trip_tile<S, R, N1, N2, TS, H * G, rt_imag(MODE), false, H>(v, vp, gm, a2, lds, twA, twB, twC,
                                                                              ph, c, gh, t, h, nacc);
        }
        if (flags & (kFlagNorm | kFlagNormMid)) {
            double acc1[1] = {nacc};
            block_partials<R, Sh::NT, 1>(acc1, a.partial, nblocks, tile);
        }
        if constexpr (EN) {
            if (!gm.active) {
#pragma unroll
                for (int k = 0; k < 4; ++k) eacc[k] = 0.0;
            }
            block_partials<R, Sh::NT, 4>(eacc, a.epartial, nblocks, tile);
        }
        if constexpr (CT) store_line_ct<S, R, N1, N2>(v, psi, gm, lane0);
        else store_line<S, R, N1>(v, psi, gm);
    } else if constexpr (MODE == kForward) {
        cx<R> w[RB][N2];
        load_line<S, R, N1>(v, psi, gm);
        forward_ab<R, N1, N2, TS, H * G>(v, w, lds, twA, c, gh, t);
        if constexpr (H == 2) radix2_fwd<R, N1, N2, TS>(w, twC, h);
        if (gm.active) {
            cx<S>* __restrict__ ob = static_cast<cx<S>*>(a.out) + gm.ubase;
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2)
                    lane_ref(ob, (long long)(N2 * r + N1 * k2 + dk) * gm.kstride, gm.lane) = narrow<S>(w[r][k2]);
        }
    } else if constexpr (MODE == kInverse) {
        cx<R> w[RB][N2];
        const cx<S>* ib = psi + gm.ubase;
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2)
                w[r][k2] = widen<R>(lane_ref(ib, (long long)(N2 * r + N1 * k2 + dk) * gm.kstride, gm.lane));
        if constexpr (H == 2) radix2_inv<R, N1, N2, TS>(w, twC, h);
        inverse_ba<R, N1, N2, TS, H * G>(w, v, lds, twB, c, gh, t);
        const R inv = R(1) / R(N);
#pragma unroll
        for (int j = 0; j < N1; ++j) v[j] = cscale(v[j], inv);
        store_line<S, R, N1>(v, static_cast<cx<S>*>(a.out), gm);
    } else {  // kEnergy
        double acc[5] = {0, 0, 0, 0, 0};
        load_line<S, R, N1>(v, psi, gm);
        if (flags & kFlagPointwise) {
            load_pot<S, R, N1>(vp, static_cast<const S*>(a.V), gm);
#pragma unroll
            for (int j = 0; j < N1; ++j) {
                const double re = v[j].re, im = v[j].im;
                const double a2 = re * re + im * im;
                acc[2] += a2;
                acc[3] += (double)vp[j] * a2;
                acc[4] += a2 * a2;
            }
        }
        cx<R> w[RB][N2];
        forward_ab<R, N1, N2, TS, H * G>(v, w, lds, twA, c, gh, t);
        if constexpr (H == 2) radix2_fwd<R, N1, N2, TS>(w, twC, h);
        const double* __restrict__ xi = a.xi + k0;
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) {
                const double x = xi[N2 * r + N1 * k2];
                const double re = w[r][k2].re, im = w[r][k2].im;
                const double p = re * re + im * im;
                acc[0] += x * x * p;
                acc[1] += x * p;
            }
        if (a.coord) acc[1] *= a.coord[gm.coord_idx];
        if (!gm.active) {
#pragma unroll
            for (int k = 0; k < 5; ++k) acc[k] = 0.0;
        }
        block_partials<R, Sh::NT, 5>(acc, a.partial, nblocks, tile);
    }
}

// Batched round trip over independent fields: blockIdx.z selects the
// instance, whose descriptor (field, potential, tables, strides, flags,
// pending scale, partial sums) sits in descs[]; instances whose done flag is
// set are skipped.  All instances of one launch share their extents, N1, N2,
// H and the workgroup shape; grid.x*grid.y covers the tile count.  TWICE
// runs two round trips on the lines in one pass: K(tau) N K(tau) (the folded
// Strang step of kRoundTripTwice) or, without kFlagNL, K(tau) K(tau) (the x
// flow that ends one chained step and the one that starts the next).
// Imaginary-time multipliers only (FP32: real decay tables).
template <typename S, typename R, int N1, int N2, int TS, int G, bool TWICE = false, int H = 1, bool TL = false>
__global__ __launch_bounds__(TS * G * N2 * H) void k_batch_sweep(const SweepArgs* __restrict__ descs,
                                                                 const int* __restrict__ done)
{
    using Sh = SweepShape<R, N1, N2, TS, H * G>;
    constexpr int M = Sh::N;
    constexpr int N = H * M;
    constexpr int NL = TS * G;
    __shared__ cx<R> lds[Sh::LDS];
    const int inst = blockIdx.z;
    if (done[inst]) return;
    const SweepArgs a = descs[inst];
    const long long tile = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    if (tile >= a.ntiles) return;

    const int tid = threadIdx.x;
    const int c = tid % TS;
    const int tl = (tid / TS) % (H * N2);
    const int t = tl / H;
    const int h = tl % H;
    const int g = tid / (TS * N2 * H);
    const int gh = H * g + h;
    const int L = c + TS * g;
    const TileGeom gm = tile_geom<N, H * N2, TS, NL>(a, tile, L, tl);
    const cx<R>* __restrict__ twA = static_cast<const cx<R>*>(a.tw) + t;
    const cx<R>* __restrict__ twB = static_cast<const cx<R>*>(a.tw) + M + t;
    const cx<R>* __restrict__ twC = static_cast<const cx<R>*>(a.tw) + 2 * M + t;
    cx<S>* __restrict__ psi = static_cast<cx<S>*>(a.psi);
    cx<R> v[N1];
    R vp[N1];
    load_line<S, R, N1>(v, psi, gm);
    if (a.flags & kFlagNL) load_pot<S, R, N1>(vp, static_cast<const S*>(a.V), gm);
    double nacc = 0.0;
    const cx<R>* ph = row_phase<R>(a, gm);
    if constexpr (TL) {  // tables from LDS; a phase table with rows (rotation) stays in global memory
        constexpr int NTW = kTwiddles<M, H>;
        __shared__ cx<R> tabs[NTW + N];
        stage_tables<R, M, H, Sh::NT>(tabs, a, tid);
        twA = tabs + t;
        twB = tabs + M + t;
        twC = tabs + 2 * M + t;
        if (a.phase_row_stride == 0) ph = tabs + NTW;
    }
    if constexpr (TWICE) {
        // the norm partials after the second round trip (kFlagNorm) or of
        // the state between the two (kFlagNormMid: [x x] of chained steps,
        // whose first round trip ends a step)
        const int mid = a.flags & kFlagNormMid;
        SweepArgs a1 = a;
        a1.flags = (a.flags & ~(kFlagNorm | kFlagNormMid)) | (mid ? kFlagNorm : 0);
        round_trip_tile<S, R, N1, N2, TS, H * G, true, false, H>(v, vp, gm, a1, lds, twA, twB, twC, ph, c, gh, t,
                                                                 h, nacc);
        SweepArgs a2 = a;
        a2.flags = a.flags & ~(kFlagNL | kFlagScale | kFlagNormMid);
        round_trip_tile<S, R, N1, N2, TS, H * G, true, false, H>(v, vp, gm, a2, lds, twA, twB, twC, ph, c, gh, t,
                                                                 h, nacc);
    } else {
        round_trip_tile<S, R, N1, N2, TS, H * G, true, false, H>(v, vp, gm, a, lds, twA, twB, twC, ph, c, gh, t, h,
                                                                 nacc);
    }
    if (a.flags & (kFlagNorm | kFlagNormMid)) {
        double acc1[1] = {nacc};
        block_partials<R, Sh::NT, 1>(acc1, a.partial, a.ntiles, tile);
    }
    store_line<S, R, N1>(v, psi, gm);
}

// Persistent, software-pipelined round trip: each workgroup walks tiles
// blockIdx.x, blockIdx.x + gridDim.x, ... and issues the loads of its next
// tile before transforming the current one, so HBM traffic overlaps the
// FFT arithmetic within every wavefront.  The twiddle and phase tables are
// staged in LDS first: vector-memory counters retire in order, so a table
// load from memory inside the loop would wait for the prefetched tile too.
// No nonlinear epilogue, no pending scale and a single phase row (Omega = 0
// or the z axis): other sweeps run k_sweep.
template <typename S, typename R, int N1, int N2, int TS, int G, bool IMAG>
__global__ __launch_bounds__(TS * G * N2) void k_sweep_pipe(const SweepArgs a)
{
    using Sh = SweepShape<R, N1, N2, TS, G>;
    constexpr int N = Sh::N;
    constexpr int NL = Sh::NL;
    __shared__ cx<R> lds[Sh::LDS];
    __shared__ cx<R> tabs[3 * N];  // twA | twB | phase

    const int tid = threadIdx.x;
    const int c = tid % TS;
    const int t = (tid / TS) % N2;
    const int g = tid / (TS * N2);
    const int L = c + TS * g;
    const int flags = a.flags;
    cx<S>* __restrict__ psi = static_cast<cx<S>*>(a.psi);
    long long tile = blockIdx.x;
    double nacc = 0.0;
    cx<R> v[N1];
    const R vp[N1] = {};
    constexpr bool CT = centered_ok<R, N1, N2, 1>;
    TileGeom gm = tile_geom<N, N2, TS, NL>(a, tile, L, t);
    unsigned lane0 = CT ? ct_lane0<N, N2>(gm, t) : 0u;
    if (tile < a.ntiles) {
        if constexpr (CT) load_line_ct<S, R, N1, N2>(v, psi, gm, lane0);
        else load_line<S, R, N1>(v, psi, gm);
    }
    // stage the tables after the first tile's loads, so the loop starts with
    // no tile load outstanding
    {
        // unrolled, branch-free when NT divides 3N: the compiler then knows
        // every load above has completed once the tables are written
        const cx<R>* gtw = static_cast<const cx<R>*>(a.tw) + (CT ? 2 * N : 0);
        const cx<R>* gph = static_cast<const cx<R>*>(a.phase);
#pragma unroll
        for (int k = 0; k < (3 * N + Sh::NT - 1) / Sh::NT; ++k) {
            const int i = tid + k * Sh::NT;
            if ((3 * N) % Sh::NT == 0 || i < 3 * N) tabs[i] = i < 2 * N ? gtw[i] : gph[i - 2 * N];
        }
    }
    const cx<R>* twA = tabs + t;
    const cx<R>* twB = tabs + N + t;
    const cx<R>* ph = tabs + 2 * N;
    lds_barrier();  // tables staged
    while (tile < a.ntiles) {
        const long long next = tile + gridDim.x;
        const bool more = next < a.ntiles;
        TileGeom gn;
        unsigned lane0n = 0u;
        cx<R> vn[N1];
        if (more) {
            gn = tile_geom<N, N2, TS, NL>(a, next, L, t);
            if constexpr (CT) {
                lane0n = ct_lane0<N, N2>(gn, t);
                load_line_ct<S, R, N1, N2>(vn, psi, gn, lane0n);
            } else {
                load_line<S, R, N1>(vn, psi, gn);
            }
        }
        round_trip_tile<S, R, N1, N2, TS, G, IMAG, false, 1, CT>(v, vp, gm, a, lds, twA, twB, nullptr, ph, c, g, t,
                                                                   0, nacc, nullptr, lane0);
        if constexpr (CT) store_line_ct<S, R, N1, N2, true>(v, psi, gm, lane0);
        else store_line_all<S, R, N1>(v, psi, gm);
        if (!more) break;
        // the next tile's loads were issued before these N1 stores: wait for
        // them here, with the stores still in flight
        wait_vmcnt<N1>();
#pragma unroll
        for (int j = 0; j < N1; ++j) v[j] = vn[j];
        gm = gn;
        lane0 = lane0n;
        tile = next;
    }
    if (flags & kFlagNorm) {
        double acc1[1] = {nacc};
        block_partials<R, Sh::NT, 1>(acc1, a.partial, gridDim.x, blockIdx.x);
    }
}

// ------------------------------------------------------------- Bluestein
// Lines of any length n (n not a power of two) run Bluestein's chirp
// transform on M = H * N1 * N2 >= 2n - 1 points held in registers (zero
// beyond n):
//   X_k = c_k sum_j (x_j c_j) b_{k-j},  c_m = exp(-i pi m^2/n),  b = conj(c),
// the convolution being an M-point circular one with b wrapped to
// g[m] = b_m (m < n), g[M - r] = b_r (0 < r < n), done as forward M-point
// DFT, multiply by B = DFT_M(g)/M, unscaled inverse (L. I. Bluestein, IEEE
// Trans. Audio Electroacoust. 18(4), 1970).  The inverse DFT is
// conj(DFT(conj X)).  Tables after the M-point twiddles: chirp[M] (c_p for
// p < n, zero beyond) and B[M], plain complex in both precisions.
// A lane holds the line's positions P(j) = tl + H N2 j, tl = H t + h.

// Line geometry of a Bluestein tile: as tile_geom, with n points per
// contiguous line.
template <int HN2, int TS, int NL>
__device__ __forceinline__ TileGeom tile_geom_blu(const SweepArgs& a, long long tile, int L, int tl)
{
    TileGeom g;
    g.row = 0;
    g.coord_idx = 0;
    const long long n = a.nlen;
    if constexpr (TS == 1) {
        const long long line0 = tile * NL;
        const long long rem = a.nlines - line0;
        const int valid = rem < NL ? (int)rem : NL;
        const int Lc = L < valid ? L : valid - 1;
        g.active = L < valid;
        g.ubase = line0 * n;
        g.lane = (unsigned)(Lc * n + tl);
        g.jstride = HN2;
        g.kstride = 1;
        const long long line = line0 + Lc;
        if (a.row_mod > 0) g.row = (int)(line % a.row_mod);
        if (a.coord_mod > 0) g.coord_idx = line % a.coord_mod;
    } else {
        const long long o = tile / a.ntx;
        const int col0 = (int)(tile - o * a.ntx) * NL;
        const int rem = a.ncols - col0;
        const int valid = rem < NL ? rem : NL;
        const int Lc = L < valid ? L : valid - 1;
        g.active = L < valid;
        g.ubase = (long long)col0 + o * a.OS;
        g.lane = (unsigned)Lc + (unsigned)tl * (unsigned)a.S;
        g.jstride = (long long)HN2 * a.S;
        g.kstride = a.S;
        if (a.row_from_col) g.row = col0 + Lc;
        g.coord_idx = col0 + Lc;
    }
    return g;
}

// Predicated loads and stores of the line's positions below n.
template <typename S, typename R, int N1, int HN2>
__device__ __forceinline__ void load_line_blu(cx<R> (&v)[N1], const cx<S>* psi, const TileGeom& g, int tl, int n)
{
    const cx<S>* b = psi + g.ubase;
#pragma unroll
    for (int j = 0; j < N1; ++j)
        v[j] = (tl + HN2 * j < n) ? widen<R>(lane_ref(b, j * g.jstride, g.lane)) : cx<R>{R(0), R(0)};
}
template <typename S, typename R, int N1, int HN2>
__device__ __forceinline__ void load_pot_blu(R (&vp)[N1], const S* V, const TileGeom& g, int tl, int n)
{
#pragma unroll
    for (int j = 0; j < N1; ++j) vp[j] = R(0);
    if (V) {
        const S* b = V + g.ubase;
#pragma unroll
        for (int j = 0; j < N1; ++j)
            if (tl + HN2 * j < n) vp[j] = R(lane_ref(b, j * g.jstride, g.lane));
    }
}
template <typename S, typename R, int N1, int HN2>
__device__ __forceinline__ void store_line_blu(const cx<R> (&v)[N1], cx<S>* psi, const TileGeom& g, int tl, int n)
{
    if (g.active) {
        cx<S>* b = psi + g.ubase;
#pragma unroll
        for (int j = 0; j < N1; ++j)
            if (tl + HN2 * j < n) lane_ref(b, j * g.jstride, g.lane) = narrow<S>(v[j]);
    }
}

// DFT of the lane's line in place: v holds x at positions P(j) (zero at and
// beyond n) and gets X_k at position k (zero at and beyond n).  chirp is the
// chirp table offset by tl, Bt the B table offset by the lane's first
// spectrum position t + M_sub h.
template <typename R, int N1, int N2, int TS, int G, int H>
__device__ __forceinline__ void bluestein_dft(cx<R> (&v)[N1], cx<R>* lds, const cx<R>* twA, const cx<R>* twB,
                                              const cx<R>* twC, const cx<R>* chirp, const cx<R>* Bt, int c, int g,
                                              int t, int h)
{
    constexpr int HN2 = H * N2;
    constexpr int RB = N1 / N2;
#pragma unroll
    for (int j = 0; j < N1; ++j) v[j] = cmul(v[j], chirp[HN2 * j]);
    cx<R> w[RB][N2];
    forward_ab<R, N1, N2, TS, G>(v, w, lds, twA, c, g, t);
    if constexpr (H == 2) radix2_fwd<R, N1, N2, TS>(w, twC, h);
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) w[r][k2] = cmul(w[r][k2], Bt[N2 * r + N1 * k2]);
    if constexpr (H == 2) radix2_inv<R, N1, N2, TS>(w, twC, h);
    inverse_ba<R, N1, N2, TS, G>(w, v, lds, twB, c, g, t);
#pragma unroll
    for (int j = 0; j < N1; ++j) v[j] = cmul(v[j], chirp[HN2 * j]);
}

template <typename R, int N1>
__device__ __forceinline__ void conj_all(cx<R> (&v)[N1])
{
#pragma unroll
    for (int j = 0; j < N1; ++j) v[j].im = -v[j].im;
}

// Round trip of a Bluestein tile: DFT, the phase multiplier (indexed by the
// natural position k, n entries per row), inverse DFT, then the epilogue of
// round_trip_tile (nonlinear phase, saved state, norm partials).
template <typename S, typename R, int N1, int N2, int TS, int G, int H, bool IMAG>
__device__ __forceinline__ void round_trip_blu(cx<R> (&v)[N1], const R (&vp)[N1], const TileGeom& gm,
                                               const SweepArgs& a, cx<R>* lds, const cx<R>* twA, const cx<R>* twB,
                                               const cx<R>* twC, const cx<R>* chirp, const cx<R>* Bt,
                                               const cx<R>* phase, int c, int g, int t, int h, int tl,
                                               double& norm_acc)
{
    constexpr int HN2 = H * N2;
    const int n = a.nlen;
    const int flags = a.flags;
    if (flags & kFlagScale) {
        const R sc = R(*a.scale);
#pragma unroll
        for (int j = 0; j < N1; ++j) v[j] = cscale(v[j], sc);
    }
    bluestein_dft<R, N1, N2, TS, G, H>(v, lds, twA, twB, twC, chirp, Bt, c, g, t, h);
    const R inv_n = R(1) / R(n);
#pragma unroll
    for (int j = 0; j < N1; ++j) {
        const int p = tl + HN2 * j;
        const cx<R> d = phase[p < n ? p : 0];
        const cx<R> x = v[j];
        if constexpr (sizeof(R) == 4) {
            // FP32 tables as in round_trip_tile: lifting coefficients of half
            // the phase (real time), (m - 1)/n (imaginary time); 1/n rounded
            if constexpr (IMAG) v[j] = {fmaf(x.re, d.re, x.re * inv_n), fmaf(x.im, d.re, x.im * inv_n)};
            else v[j] = lift(lift({x.re * inv_n, x.im * inv_n}, d.re, d.im), d.re, d.im);
        } else {
            v[j] = cmul(x, d);
        }
    }
    conj_all<R, N1>(v);
    bluestein_dft<R, N1, N2, TS, G, H>(v, lds, twA, twB, twC, chirp, Bt, c, g, t, h);
    conj_all<R, N1>(v);
    if (flags & kFlagNL) {
        if (a.save) store_line_blu<S, R, N1, HN2>(v, static_cast<cx<S>*>(a.save), gm, tl, n);
        const R beta = R(a.beta), dt = R(a.dt);
        if constexpr (sizeof(R) == 4 && !IMAG) {
            nl_phase_lane<N1>(v, vp, beta, dt);
        } else {
            const bool imag = sizeof(R) == 4 ? IMAG : (flags & kFlagImag) != 0;
#pragma unroll
            for (int j = 0; j < N1; ++j) nl_phase<R>(v[j], vp[j], beta, dt, imag);
        }
    }
    if (flags & kFlagNorm) {
        if constexpr (sizeof(R) == 4) {  // the lane's N1 points in FP32, then one FP64 add
            float s = 0.0f;
#pragma unroll
            for (int j = 0; j < N1; ++j) s = fmaf(v[j].re, v[j].re, fmaf(v[j].im, v[j].im, s));
            if (gm.active) norm_acc += (double)s;
        } else {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < N1; ++j) {
                const double re = v[j].re, im = v[j].im;
                s += re * re + im * im;
            }
            if (gm.active) norm_acc += s;
        }
    }
}

// One launch over all tiles of an axis whose lines have n = a.nlen points
// (not a power of two), by Bluestein transforms on M = H N1 N2 points.
// Modes: round trips (single and [d N d]), forward, inverse, energy sums.
template <typename S, typename R, int N1, int N2, int TS, int G, int MODE, int H = 1>
__global__ __launch_bounds__(TS * G * N2 * H) void k_sweep_blu(const SweepArgs a)
{
    using Sh = SweepShape<R, N1, N2, TS, H * G>;
    constexpr int M = Sh::N;  // sub-transform length; the transform has H * M points
    constexpr int RB = Sh::RB;
    constexpr int NL = TS * G;
    constexpr int HN2 = H * N2;
    constexpr int NTW = kTwiddles<M, H>;
    __shared__ cx<R> lds[Sh::LDS];
    const int tid = threadIdx.x;
    const int c = tid % TS;
    const int tl = (tid / TS) % HN2;
    const int t = tl / H;
    const int h = tl % H;
    const int g = tid / (TS * HN2);
    const int gh = H * g + h;
    const int L = c + TS * g;
    const long long tile = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    const TileGeom gm = tile_geom_blu<HN2, TS, NL>(a, tile, L, tl);
    const int n = a.nlen;
    const cx<R>* tw = static_cast<const cx<R>*>(a.tw);
    const cx<R>* twA = tw + t;
    const cx<R>* twB = tw + M + t;
    const cx<R>* twC = tw + 2 * M + t;
    const cx<R>* chirp = tw + NTW + tl;
    const cx<R>* Bt = tw + NTW + H * M + t + M * h;
    cx<S>* __restrict__ psi = static_cast<cx<S>*>(a.psi);
    const long long nblocks = (long long)gridDim.x * gridDim.y;
    cx<R> v[N1];
    R vp[N1];
    if constexpr (is_round_trip(MODE) || rt_twice(MODE)) {
        static_assert(!rt_energy(MODE), "no energy round trips on Bluestein lines");
        load_line_blu<S, R, N1, HN2>(v, psi, gm, tl, n);
        if (a.flags & kFlagNL) load_pot_blu<S, R, N1, HN2>(vp, static_cast<const S*>(a.V), gm, tl, n);
        else
#pragma unroll
            for (int j = 0; j < N1; ++j) vp[j] = R(0);
        const cx<R>* ph = static_cast<const cx<R>*>(a.phase) + (long long)gm.row * a.phase_row_stride;
        double nacc = 0.0;
        if constexpr (rt_twice(MODE)) {
            SweepArgs a1 = a;
            a1.flags = a.flags & ~kFlagNorm;
            SweepArgs a2 = a;
            a2.flags = a.flags & ~(kFlagNL | kFlagScale);
            a2.save = nullptr;
            round_trip_blu<S, R, N1, N2, TS, H * G, H, rt_imag(MODE)>(v, vp, gm, a1, lds, twA, twB, twC, chirp, Bt,
                                                                     ph, c, gh, t, h, tl, nacc);
            round_trip_blu<S, R, N1, N2, TS, H * G, H, rt_imag(MODE)>(v, vp, gm, a2, lds, twA, twB, twC, chirp, Bt,
                                                                     ph, c, gh, t, h, tl, nacc);
        } else {
            round_trip_blu<S, R, N1, N2, TS, H * G, H, rt_imag(MODE)>(v, vp, gm, a, lds, twA, twB, twC, chirp, Bt,
                                                                     ph, c, gh, t, h, tl, nacc);
        }
        if (a.flags & kFlagNorm) {
            double acc1[1] = {nacc};
            block_partials<R, Sh::NT, 1>(acc1, a.partial, nblocks, tile);
        }
        store_line_blu<S, R, N1, HN2>(v, psi, gm, tl, n);
    } else if constexpr (MODE == kForward) {
        load_line_blu<S, R, N1, HN2>(v, psi, gm, tl, n);
        bluestein_dft<R, N1, N2, TS, H * G, H>(v, lds, twA, twB, twC, chirp, Bt, c, gh, t, h);
        store_line_blu<S, R, N1, HN2>(v, static_cast<cx<S>*>(a.out), gm, tl, n);
    } else if constexpr (MODE == kInverse) {
        load_line_blu<S, R, N1, HN2>(v, psi, gm, tl, n);
        conj_all<R, N1>(v);
        bluestein_dft<R, N1, N2, TS, H * G, H>(v, lds, twA, twB, twC, chirp, Bt, c, gh, t, h);
        const R inv = R(1) / R(n);
#pragma unroll
        for (int j = 0; j < N1; ++j) v[j] = {v[j].re * inv, -v[j].im * inv};
        store_line_blu<S, R, N1, HN2>(v, static_cast<cx<S>*>(a.out), gm, tl, n);
    } else if constexpr (MODE == kEnergy) {
        double acc[5] = {0, 0, 0, 0, 0};
        load_line_blu<S, R, N1, HN2>(v, psi, gm, tl, n);
        if (a.flags & kFlagPointwise) {
            load_pot_blu<S, R, N1, HN2>(vp, static_cast<const S*>(a.V), gm, tl, n);
#pragma unroll
            for (int j = 0; j < N1; ++j) {
                const double re = v[j].re, im = v[j].im;
                const double a2 = re * re + im * im;
                acc[2] += a2;
                acc[3] += (double)vp[j] * a2;
                acc[4] += a2 * a2;
            }
        }
        bluestein_dft<R, N1, N2, TS, H * G, H>(v, lds, twA, twB, twC, chirp, Bt, c, gh, t, h);
#pragma unroll
        for (int j = 0; j < N1; ++j) {
            const int p = tl + HN2 * j;
            const double x = p < n ? a.xi[p] : 0.0;
            const double re = v[j].re, im = v[j].im;
            const double pw = re * re + im * im;
            acc[0] += x * x * pw;
            acc[1] += x * pw;
        }
        if (a.coord) acc[1] *= a.coord[gm.coord_idx];
        if (!gm.active) {
#pragma unroll
            for (int k = 0; k < 5; ++k) acc[k] = 0.0;
        }
        block_partials<R, Sh::NT, 5>(acc, a.partial, nblocks, tile);
    }
    (void)RB;
}

}  // namespace sfd
