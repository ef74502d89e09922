// sfd.hip -- runtime and C ABI of libsfdhip.so (declared in include/sfd.h).
//
// The context replaces the reference's Grid + FFT Plan pair (Grids.jl:64-129,
// Plans.jl:958-999) for one field shape: it owns the FP64 geometry
// (coordinates, wavenumbers), the twiddle tables, a cache of phase-multiplier
// tables keyed by (axis, tau, coeffD, Omega, mode) and the reduction
// workspace.  Fields and potentials stay owned by the caller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>

#include "../../include/sfd.h"
#include "sfd_variants.hpp"
#include "sfd_batch.hpp"
#include "sfd_adjoint.hpp"

using namespace sfd;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define SFD_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(e_ == hipErrorOutOfMemory ? SFD_ERR_OOM : SFD_ERR_HIP,         \
                        "%s failed: %s", #call, hipGetErrorString(e_));                \
    } while (0)

#define SFD_TRY(...)              \
    do {                          \
        int r_ = (__VA_ARGS__);   \
        if (r_ != SFD_OK) return r_; \
    } while (0)

// ----------------------------------------------------------- small kernels
template <typename S>
__global__ void k_norm2(const cx<S>* __restrict__ psi, long long n, double* partial)
{
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const cx<S> v = psi[i];
        const double re = v.re, im = v.im;
        s += re * re + im * im;
    }
    double acc[1] = {s};
    block_partials<S, 256, 1>(acc, partial, gridDim.x, blockIdx.x);
}

// out[k] = sum_b partial[k*nb + b] for k < K, in a fixed order
// (deterministic: the order depends only on nb and the workgroup size).
// Each lane adds a strided subset with four independent accumulators, so its
// loads are in flight together; waves then combine through LDS.  Runs on
// one workgroup of kFinalizeThreads.
constexpr int kFinalizeThreads = 1024;

__device__ __forceinline__ void finalize_rows(const double* __restrict__ partial, long long nb, int K, double* out)
{
    __shared__ double red[kFinalizeThreads / 64];
    const int tid = threadIdx.x;
    const long long bs = blockDim.x;
    for (int k = 0; k < K; ++k) {
        const double* __restrict__ p = partial + k * nb;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        long long b = tid;
        for (; b + 3 * bs < nb; b += 4 * bs) {
            s0 += p[b];
            s1 += p[b + bs];
            s2 += p[b + 2 * bs];
            s3 += p[b + 3 * bs];
        }
        for (; b < nb; b += bs) s0 += p[b];
        const double s = wave_sum((s0 + s1) + (s2 + s3));
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < (int)(bs + 63) / 64; ++w) t += red[w];
            out[k] = t;
        }
        __syncthreads();
    }
}

// finalize_rows; if scale_out != null also *scale_out = 1 / sqrt(out[0] * dv).
__global__ __launch_bounds__(kFinalizeThreads) void k_finalize(const double* __restrict__ partial, long long nb,
                                                              int K, double* out, double dv, double* scale_out)
{
    finalize_rows(partial, nb, K, out);
    if (scale_out && threadIdx.x == 0) scale_out[0] = 1.0 / sqrt(out[0] * dv);
}

// Up to three independent finalize_rows in one launch (the per-axis energy
// sums of a step): one workgroup per row, rows of set 0 first.  Launch
// with sum K[i] workgroups.
struct FinalizeSets {
    const double* partial[3];
    long long nb[3];
    int K[3];
    double* out[3];
    int rows() const { return K[0] + K[1] + K[2]; }
};

__global__ __launch_bounds__(kFinalizeThreads) void k_finalize_sets(const FinalizeSets f)
{
    int i = 0, k = blockIdx.x;
    while (k >= f.K[i]) k -= f.K[i++];
    finalize_rows(f.partial[i] + k * f.nb[i], f.nb[i], 1, f.out[i] + k);
}

void launch_finalize(hipStream_t s, const double* partial, long long nb, int K, double* out, double dv = 0.0,
                     double* scale_out = nullptr)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinalizeThreads), 0, s, partial, nb, K, out, dv, scale_out);
}

// psi *= s (divide = 0) or psi /= s (divide = 1) in FP64, rounded once to
// the storage precision; s is read from device memory when s_dev is not null.
template <typename S>
__global__ void k_scale(cx<S>* __restrict__ psi, long long n, double s_host, const double* s_dev, int divide)
{
    const double s = s_dev ? *s_dev : s_host;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const cx<S> v = psi[i];
        double re = v.re, im = v.im;
        if (divide) { re /= s; im /= s; }
        else        { re *= s; im *= s; }
        psi[i] = {S(re), S(im)};
    }
}

// psi[i] = exp(i theta_i), theta_i from an integer hash of i: unit-modulus
// pseudo-random data for timing (sweep timings depend on the data: zeros
// run ~15 % faster than a condensate-like field on MI355X).
template <typename S>
__global__ void k_fill_phases(cx<S>* __restrict__ psi, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        unsigned h = (unsigned)i * 2654435761u;
        h ^= h >> 15;
        h *= 2246822519u;
        h ^= h >> 13;
        const float th = (float)(h >> 8) * (6.28318530718f / 16777216.0f);
        float sn, cs;
        sincosf(th, &sn, &cs);
        psi[i] = {S(cs), S(sn)};
    }
}

// Pointwise nonlinear flow (solveNL!, splitting.jl:205-209): storage S,
// arithmetic R.
template <typename S, typename R>
__global__ void k_nl(cx<S>* __restrict__ psi, const S* __restrict__ V, long long n, R beta, R dt, int imag)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        cx<R> v = widen<R>(psi[i]);
        nl_phase<R>(v, V ? R(V[i]) : R(0), beta, dt, imag != 0);
        psi[i] = narrow<S>(v);
    }
}

// ------------------------------------------------------------ host helpers
double lin_coord(double a, double b, int n, int i)
{
    // LinRange(a, b, n+1)[i+1] = (1 - t) a + t b with t = i/n (Grids.jl:78)
    const double t = (double)i / (double)n;
    return (1.0 - t) * a + t * b;
}

double fft_freq(int n, double dx, int k)
{
    // fftfreq(n, 2pi/dx)[k+1] (Plans.jl:829-830,967-969)
    const double twopi = 6.283185307179586;
    const double mult = (twopi / dx) / (double)n;
    const int kk = (k < (n + 1) / 2) ? k : k - n;
    return (double)kk * mult;
}

struct PhaseEntry {
    int axis;
    double tau, cD, Om;
    int imag;
    void* dev;
    int rows;
};

}  // namespace

struct sfd_ctx {
    int ndim = 0;
    int n[3] = {1, 1, 1};
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1}, d[3] = {1, 1, 1};
    long long nsite = 1;
    int prec = SFD_F64;
    int device = 0;
    hipStream_t stream = nullptr;
    // per axis: launch variant of k_sweep and its twiddle table
    int var[3] = {0, 0, 0};
    VariantInfo vi[3] = {};
    void* tw[3] = {nullptr, nullptr, nullptr};
    int tuned = 0;
    int pipe_occ[3] = {0, 0, 0};  // resident workgroups per CU (persistent variants)
    int ncu = 1;
    long long last_nblocks = 0;   // workgroups of the last sweep launched (its partial count)
    double* xi[3] = {nullptr, nullptr, nullptr};
    double* coord[3] = {nullptr, nullptr, nullptr};
    double* partial = nullptr;
    long long partial_cap = 0;  // doubles
    double* epartial = nullptr;  // energy round trips: per-workgroup sums
    long long epartial_cap = 0;
    double* ehist = nullptr;     // fused step + energy: 8 FP64 sums per step
    long long ehist_cap = 0;
    double* scal = nullptr;     // [0] pending scale, [1..31] sums
    std::vector<PhaseEntry> phases;
    bool pin_phases = false;    // no eviction: two streams read the cache

    size_t csize() const { return prec == SFD_F64 ? 16 : 8; }
    double dvol() const
    {
        double v = 1.0;
        for (int a = 0; a < ndim; ++a) v *= d[a];
        return v;
    }
};

namespace {

// Twiddles in the layout k_sweep reads for a line of n = h*m points split
// as m = n1*n2 (h = 1: m = n), with w_m = exp(-2 pi i / m):
//   [0, m):    twA[s][t] = w_m^(s t), s < n1, t < n2
//   [m, 2m):   twB[q][s] = w_m^(q s), q < n2, s < n1
//   [2m, 3m):  h = 2 only, the radix-2 stage's twC[k] = w_n^k, k < m
// FP64 entries are (cos, sin) of the angle a; FP32 entries are the lifting
// coefficients (tan(b/2), sin b) of b = a/2 with a taken in (-pi, pi].
template <typename R>
void fill_twiddles(std::vector<cx<R>>& w, int n, int n1, int n2, int h)
{
    const int m = n1 * n2;
    w.resize((size_t)(h == 2 ? 3 : 2) * m);
    const long double pi = 3.14159265358979323846264338327950288L;
    // entry i = exp(-2 pi i e / d)
    auto put = [&](size_t i, long long e, long long d) {
        e %= d;
        const long double a = -2.0L * pi * (long double)e / (long double)d;
        if constexpr (sizeof(R) == 4) {
            // lifting coefficients of the half angle b in (-pi/2, pi/2]
            const long double b = 0.5L * (2 * e > d ? a + 2.0L * pi : a);
            w[i].re = (R)tanl(0.5L * b);
            w[i].im = (R)sinl(b);
        } else {
            w[i].re = (R)cosl(a);
            w[i].im = (R)sinl(a);
        }
    };
    for (int s = 0; s < n1; ++s)
        for (int t = 0; t < n2; ++t) put((size_t)s * n2 + t, (long long)t * s, m);
    for (int q = 0; q < n2; ++q)
        for (int s = 0; s < n1; ++s) put((size_t)m + (size_t)q * n1 + s, (long long)s * q, m);
    if (h == 2)
        for (int k = 0; k < m; ++k) put((size_t)2 * m + k, k, n);
}

int upload_sync(void* dst, const void* src, size_t bytes)
{
    SFD_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SFD_OK;
}

// One imaginary-time multiplier exp(arg)/n: itself in FP64 tables, and
// expm1(arg)/n in FP32 tables, whose sweeps form X/n + X expm1(arg)/n.
template <typename R>
cx<R> decay_entry(double arg, double inv_n)
{
    if constexpr (sizeof(R) == 4) return {(R)(std::expm1(arg) * inv_n), (R)0};
    return {(R)(std::exp(arg) * inv_n), (R)0};
}

// Phase multiplier of one axis (splitting.jl:190,195,200), divided by n:
//   x: exp(s (cD xi^2 - Om y xi) tau)   rows indexed by y when Om != 0
//   y: exp(s (cD xi^2 + Om x xi) tau)   rows indexed by x when Om != 0
//   z: exp(s  cD xi^2 tau)
// s = i (real time) or 1 (imaginary time).  Formed in FP64 with the
// reference's operation order, rounded once to the field precision.
template <typename P>
int build_phase(sfd_ctx* c, int axis, double tau, double cD, double Om, int imag, void** out, int* rows_out)
{
    using R = typename P::R;
    using S = typename P::S;
    (void)sizeof(R);
    (void)sizeof(S);
    for (auto& p : c->phases)
        if (p.axis == axis && p.tau == tau && p.cD == cD && p.Om == Om && p.imag == imag) {
            *out = p.dev;
            *rows_out = p.rows;
            return SFD_OK;
        }
    const int n = c->n[axis];
    const bool rot = (Om != 0.0) && axis < 2 && c->ndim >= 2;
    const int oax = axis == 0 ? 1 : 0;
    const int rows = rot ? c->n[oax] : 1;
    std::vector<cx<R>> h((size_t)rows * n);
    std::vector<double> args(sizeof(R) == 4 && !imag ? h.size() : 0);
    const double inv_n = 1.0 / (double)n;
    for (int r = 0; r < rows; ++r) {
        const double co = rot ? lin_coord(c->lo[oax], c->hi[oax], c->n[oax], r) : 0.0;
        for (int k = 0; k < n; ++k) {
            const double xi = fft_freq(n, c->d[axis], k);
            double arg;
            if (axis == 0) arg = (cD * (xi * xi) - (Om * co) * xi) * tau;
            else if (axis == 1) arg = (cD * (xi * xi) + (Om * co) * xi) * tau;
            else arg = (cD * (xi * xi)) * tau;
            if (imag) h[(size_t)r * n + k] = decay_entry<R>(arg, inv_n);
            else if (sizeof(R) == 4) args[(size_t)r * n + k] = std::remainder(arg, 2.0 * M_PI);
            else h[(size_t)r * n + k] = {(R)(std::cos(arg) * inv_n), (R)(std::sin(arg) * inv_n)};
        }
    }
    // FP32 real time: lifting coefficients (tan(b/2), sin b) of the half
    // phase b = phi/2, phi in (-pi, pi]; the sweep applies them twice
    for (size_t i = 0; i < args.size(); ++i) {
        const double b = 0.5 * args[i];
        h[i] = {(R)std::tan(0.5 * b), (R)std::sin(b)};
    }
    void* dev = nullptr;
    SFD_HIP(hipMalloc(&dev, h.size() * sizeof(cx<R>)));
    int rc = upload_sync(dev, h.data(), h.size() * sizeof(cx<R>));
    if (rc != SFD_OK) { (void)hipFree(dev); return rc; }
    if (c->phases.size() >= 32 && !c->pin_phases) {  // bounded cache: drop the oldest entry
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(c->phases.front().dev);
        c->phases.erase(c->phases.begin());
    }
    c->phases.push_back({axis, tau, cD, Om, imag, dev, rows});
    *out = dev;
    *rows_out = rows;
    return SFD_OK;
}

// Grows a device workspace of the context to `need` doubles (contents lost).
int ensure_buf(sfd_ctx* c, double** buf, long long* cap, long long need)
{
    if (need <= *cap) return SFD_OK;
    if (*buf) {
        SFD_HIP(hipStreamSynchronize(c->stream));
        SFD_HIP(hipFree(*buf));
        *buf = nullptr;
        *cap = 0;
    }
    SFD_HIP(hipMalloc(buf, (size_t)need * sizeof(double)));
    *cap = need;
    return SFD_OK;
}

int ensure_partial(sfd_ctx* c, long long need) { return ensure_buf(c, &c->partial, &c->partial_cap, need); }

// Launch geometry of one sweep under the axis's current variant: returns
// grid/block and fills the line fields of the argument block.
struct Launch {
    dim3 grid, block;
    long long nblocks;
};

Launch sweep_geometry(const sfd_ctx* c, int axis, SweepArgs& a, bool persistent)
{
    const VariantInfo& v = c->vi[axis];
    const int NL = v.ts * v.g;
    Launch L;
    const long long nx = c->n[0], ny = c->n[1], nz = c->n[2];
    if (axis == 0) {
        a.S = 1;
        a.nlines = ny * nz;
        a.ntiles = (a.nlines + NL - 1) / NL;
        L.grid = dim3((unsigned)a.ntiles, 1, 1);
    } else {
        a.ncols = (int)nx;
        if (axis == 1) { a.S = nx; a.OS = nx * ny; }
        else           { a.S = nx * ny; a.OS = nx; }
        const long long others = axis == 1 ? nz : ny;
        a.ntx = (int)((nx + NL - 1) / NL);
        a.ntiles = (long long)a.ntx * others;
        L.grid = dim3((unsigned)a.ntx, (unsigned)others, 1);
    }
    if (persistent) {
        // pipelined variant: as many workgroups as stay resident
        const long long resident = (long long)c->pipe_occ[axis] * c->ncu;
        L.grid = dim3((unsigned)(a.ntiles < resident ? a.ntiles : resident), 1, 1);
    }
    L.block = dim3(NL * v.n2 * v.h, 1, 1);
    L.nblocks = (long long)L.grid.x * L.grid.y;
    return L;
}

template <typename P, int MODE>
int launch_sweep(sfd_ctx* c, int axis, SweepArgs a)
{
    using R = typename P::R;
    using S = typename P::S;
    (void)sizeof(R);
    (void)sizeof(S);
    // The pipelined kernel of a pipelined variant takes round trips without a
    // nonlinear epilogue, pending scale or energy sums and with one phase
    // row; other sweeps run k_sweep.
    const bool persistent = c->vi[axis].pipe && (MODE == kRoundTrip || MODE == kRoundTripImag) &&
                            !(a.flags & (kFlagNL | kFlagScale)) && a.phase_row_stride == 0;
    Launch L = sweep_geometry(c, axis, a, persistent);
    a.tw = c->tw[axis];
    if (!launch_variant<P, MODE>(c->n[axis], axis == 0, c->var[axis], L.grid, L.block, c->stream, a, persistent))
        return fail(SFD_ERR_SHAPE, "no sweep variant %d for line length %d", c->var[axis], c->n[axis]);
    SFD_HIP(hipGetLastError());
    c->last_nblocks = L.nblocks;
    return SFD_OK;
}

// Workgroups of a non-persistent sweep: an upper bound on the partial sums
// any sweep of the axis writes.
long long sweep_nblocks(const sfd_ctx* c, int axis)
{
    SweepArgs a{};
    return sweep_geometry(c, axis, a, false).nblocks;
}

// Select launch variant `idx` for `axis` and (re)build the twiddle table of
// its line split.
template <typename P>
int set_variant(sfd_ctx* c, int axis, int idx)
{
    using R = typename P::R;
    using S = typename P::S;
    (void)sizeof(R);
    (void)sizeof(S);
    const int n = c->n[axis];
    VariantInfo vi;
    if (!variant_info<P>(n, axis == 0, idx, &vi))
        return fail(SFD_ERR_ARG, "axis %d: variant %d does not exist (%d variants)", axis, idx,
                    n_variants<P>(n, axis == 0));
    c->pipe_occ[axis] = vi.pipe ? variant_occupancy<P>(n, axis == 0, idx) : 0;
    if (c->tw[axis] && c->vi[axis].n1 == vi.n1 && c->vi[axis].n2 == vi.n2 && c->vi[axis].h == vi.h) {
        c->var[axis] = idx;
        c->vi[axis] = vi;
        return SFD_OK;
    }
    std::vector<cx<R>> w;
    fill_twiddles<R>(w, n, vi.n1, vi.n2, vi.h);
    if (c->tw[axis]) {
        SFD_HIP(hipStreamSynchronize(c->stream));
        SFD_HIP(hipFree(c->tw[axis]));
        c->tw[axis] = nullptr;
    }
    SFD_HIP(hipMalloc(&c->tw[axis], w.size() * sizeof(cx<R>)));
    SFD_TRY(upload_sync(c->tw[axis], w.data(), w.size() * sizeof(cx<R>)));
    c->var[axis] = idx;
    c->vi[axis] = vi;
    return SFD_OK;
}

template <typename P>
int base_args(sfd_ctx* c, int axis, SweepArgs& a)
{
    using R = typename P::R;
    using S = typename P::S;
    (void)sizeof(R);
    (void)sizeof(S);
    std::memset(&a, 0, sizeof a);
    return SFD_OK;
}

// One linear round trip along `axis` over tau, with optional fused flags.
template <typename P>
// One round trip of an axis.  With `epartial` (room for 4 * sweep_nblocks
// doubles), the sweep also writes the energy sums of the state it loads
// there, 4 per workgroup (kRoundTripEnergy; kFlagPointwise in `flags` adds
// the pointwise sums).
// With `twice`, the sweep is K(tau) N K(tau) on this axis (kRoundTripTwice):
// the nonlinear phase of `flags` between two round trips.
int round_trip(sfd_ctx* c, void* psi, const void* V, int axis, double tau, double cD, double Om, int imag,
               int flags, double beta, double dt, void* save = nullptr, double* epartial = nullptr,
               bool twice = false)
{
    using R = typename P::R;
    using S = typename P::S;
    (void)sizeof(R);
    (void)sizeof(S);
    SweepArgs a;
    base_args<P>(c, axis, a);
    void* ph = nullptr;
    int rows = 1;
    SFD_TRY(build_phase<P>(c, axis, tau, cD, Om, imag, &ph, &rows));
    a.psi = psi;
    a.V = V;
    a.phase = ph;
    a.phase_row_stride = rows > 1 ? c->n[axis] : 0;
    if (axis == 0) a.row_mod = rows > 1 ? c->n[1] : 0;
    else a.row_from_col = (axis == 1 && rows > 1) ? 1 : 0;
    a.flags = flags | (imag ? kFlagImag : 0);
    a.beta = beta;
    a.dt = dt;
    a.scale = c->scal;
    a.save = save;
    if (flags & kFlagNorm) {
        SFD_TRY(ensure_partial(c, 5 * sweep_nblocks(c, axis)));
        a.partial = c->partial;
    }
    if (twice) {
        if (epartial) {
            if (imag) return fail(SFD_ERR_ARG, "no energy sums in an imaginary-time double round trip");
            a.epartial = epartial;
            a.xi = c->xi[axis];
            return launch_sweep<P, kRoundTripTwiceEnergy>(c, axis, a);
        }
        if constexpr (sizeof(R) == 4)
            if (imag) return launch_sweep<P, kRoundTripTwiceImag>(c, axis, a);
        return launch_sweep<P, kRoundTripTwice>(c, axis, a);
    }
    if (epartial) {
        a.epartial = epartial;
        a.xi = c->xi[axis];
        if constexpr (sizeof(R) == 4)
            if (imag) return launch_sweep<P, kRoundTripImagEnergy>(c, axis, a);
        return launch_sweep<P, kRoundTripEnergy>(c, axis, a);
    }
    if constexpr (sizeof(R) == 4)
        if (imag) return launch_sweep<P, kRoundTripImag>(c, axis, a);
    return launch_sweep<P, kRoundTrip>(c, axis, a);
}

int check_ctx(sfd_ctx* c)
{
    if (!c) return fail(SFD_ERR_NULL, "null context");
    SFD_HIP(hipSetDevice(c->device));
    return SFD_OK;
}

long long grid_blocks(long long n)
{
    long long b = (n + 255) / 256;
    return b < 2048 ? (b < 1 ? 1 : b) : 2048;
}

template <typename P>
int norm2_impl(sfd_ctx* c, const void* psi, double* dev_out)
{
    using R = typename P::R;
    using S = typename P::S;
    (void)sizeof(R);
    (void)sizeof(S);
    const long long nb = grid_blocks(c->nsite);
    SFD_TRY(ensure_partial(c, nb));
    hipLaunchKernelGGL(k_norm2<S>, dim3((unsigned)nb), dim3(256), 0, c->stream,
                       static_cast<const cx<S>*>(psi), c->nsite, c->partial);
    SFD_HIP(hipGetLastError());
    launch_finalize(c->stream, c->partial, nb, 1, dev_out);
    SFD_HIP(hipGetLastError());
    return SFD_OK;
}

template <typename P>
// nsteps split steps.  With `ehist` (device, 8 doubles per step), the first
// half-step sweeps of step s also sum the energy of the state they load and
// write the sums of that state (the state before step s) to ehist[8 s ..]:
// [0] sum xi_x^2 |psi_hat_x|^2, [1] sum |psi|^2, [2] sum V|psi|^2,
// [3] sum |psi|^4, [3 + d] sum xi_d^2 |psi_hat_d|^2 for d = 1, 2.  The y and
// z sums are taken after the x (and y) flow, which leaves them unchanged
// only in real time with Omega = 0 (the caller's condition).
//
// Strang steps run the axes as x, y, [z N z], x, y (fold): the innermost
// axis's flow commutes with the others (its multiplier depends on its own
// wavenumber only; with rotation the x and y flows also depend on y and x,
// so this holds for z in 3D, and for y in 2D only when Omega = 0), so the
// second half step K_z K_y K_x may apply z first, and the two z flows meet
// around the nonlinear phase in one double round-trip sweep (2D - 1 sweeps
// per step instead of 2D).  `fold = false` keeps the x, y, z, N, x, y, z
// order of splitting.jl:149-154.
int split_impl(sfd_ctx* c, void* psi, const void* V, double cD, double beta, double Om, double dt,
               int64_t nsteps, int order, int mode, double* ehist = nullptr, bool fold = true)
{
    using R = typename P::R;
    using S = typename P::S;
    (void)sizeof(R);
    (void)sizeof(S);
    const int imag = mode == SFD_IMAG_TIME;
    const int D = c->ndim;
    const double tau = order == 2 ? dt * 0.5 : dt;
    const int nlflags = kFlagNL | (imag ? kFlagImag : 0);
    // Reduction workspace for the per-step renormalisation (its last sweep
    // is on axis D - 2 when folded, else on the last axis).
    if (imag)
        SFD_TRY(ensure_partial(c, 5 * std::max(sweep_nblocks(c, D >= 2 ? D - 2 : 0), sweep_nblocks(c, D - 1))));
    // Energy partials: one region per axis, all finalized after the step's
    // first half in one launch.
    double* epart[3] = {nullptr, nullptr, nullptr};
    if (ehist) {
        long long off[3], tot = 0;
        for (int ax = 0; ax < D; ++ax) {
            off[ax] = tot;
            tot += 4 * sweep_nblocks(c, ax);
        }
        SFD_TRY(ensure_buf(c, &c->epartial, &c->epartial_cap, tot));
        for (int ax = 0; ax < D; ++ax) epart[ax] = c->epartial + off[ax];
    }
    bool pending = false;
    fold = fold && order == 2 && D >= 2 && (Om == 0.0 || D == 3);
    for (int64_t s = 0; s < nsteps && fold; ++s) {
        FinalizeSets fs{};
        auto record = [&](int ax) {
            fs.partial[ax] = epart[ax];
            fs.nb[ax] = c->last_nblocks;
            fs.K[ax] = ax == 0 ? 4 : 1;
            fs.out[ax] = ehist + 8 * s + (ax == 0 ? 0 : 3 + ax);
        };
        for (int ax = 0; ax < D - 1; ++ax) {
            int flags = (ax == 0 && pending) ? kFlagScale : 0;
            if (ehist && ax == 0) flags |= kFlagPointwise;
            SFD_TRY(round_trip<P>(c, psi, V, ax, tau, cD, Om, imag, flags, beta, dt, nullptr, epart[ax]));
            if (ehist) record(ax);
        }
        SFD_TRY(round_trip<P>(c, psi, V, D - 1, tau, cD, Om, imag, nlflags, beta, dt, nullptr, epart[D - 1], true));
        if (ehist) {
            record(D - 1);
            hipLaunchKernelGGL(k_finalize_sets, dim3(fs.rows()), dim3(kFinalizeThreads), 0, c->stream, fs);
            SFD_HIP(hipGetLastError());
        }
        for (int ax = 0; ax < D - 1; ++ax)
            SFD_TRY(round_trip<P>(c, psi, V, ax, tau, cD, Om, imag, (ax == D - 2 && imag) ? kFlagNorm : 0, beta,
                                  dt));
        if (imag) {
            launch_finalize(c->stream, c->partial, c->last_nblocks, 1, c->scal + 1, c->dvol(), c->scal);
            SFD_HIP(hipGetLastError());
            pending = true;
        }
    }
    for (int64_t s = 0; s < nsteps && !fold; ++s) {
        FinalizeSets fs{};
        for (int h = 0; h < order; ++h) {
            for (int ax = 0; ax < D; ++ax) {
                int flags = 0;
                if (ax == 0 && h == 0 && pending) flags |= kFlagScale;
                if (ax == D - 1 && h == 0) flags |= nlflags;
                if (ax == D - 1 && h == order - 1 && imag) flags |= kFlagNorm;
                const bool en = ehist && h == 0;
                if (en && ax == 0) flags |= kFlagPointwise;
                SFD_TRY(round_trip<P>(c, psi, V, ax, tau, cD, Om, imag, flags, beta, dt, nullptr,
                                      en ? epart[ax] : nullptr));
                if (en) {
                    fs.partial[ax] = epart[ax];
                    fs.nb[ax] = c->last_nblocks;
                    fs.K[ax] = ax == 0 ? 4 : 1;
                    fs.out[ax] = ehist + 8 * s + (ax == 0 ? 0 : 3 + ax);
                }
            }
            if (ehist && h == 0) {
                hipLaunchKernelGGL(k_finalize_sets, dim3(fs.rows()), dim3(kFinalizeThreads), 0, c->stream, fs);
                SFD_HIP(hipGetLastError());
            }
        }
        if (imag) {
            const long long nb = c->last_nblocks;  // partials of the norm sweep just launched
            launch_finalize(c->stream, c->partial, nb, 1, c->scal + 1, c->dvol(), c->scal);
            SFD_HIP(hipGetLastError());
            pending = true;
        }
    }
    if (pending) {
        const long long nb = grid_blocks(c->nsite);
        hipLaunchKernelGGL(k_scale<S>, dim3((unsigned)nb), dim3(256), 0, c->stream, static_cast<cx<S>*>(psi),
                           c->nsite, 0.0, (const double*)c->scal, 0);
        SFD_HIP(hipGetLastError());
    }
    return SFD_OK;
}

// (EOm, ED, Eb, E, mu) from the Parseval and pointwise sums of a state:
// kin[d] = sum xi_d^2 |psi_hat_d|^2, rot[d] = sum_lines coord sum xi_d |psi_hat_d|^2
// (d = x, y), pw = (sum |psi|^2, sum V|psi|^2, sum |psi|^4).
// Parseval: sum_sites |d_ax psi|^2 = (1/n_ax) sum xi^2 |psi_hat|^2, and
// sum_sites Re(i conj(psi) y d_x psi) = -(1/n_x) sum_lines y sum xi |psi_hat|^2.
void combine_energy(const sfd_ctx* c, const double* kin, const double* rot, const double* pw, double cD,
                    double beta, double Om, double* out5)
{
    double k = 0.0;
    for (int ax = 0; ax < c->ndim; ++ax) k += kin[ax] / (double)c->n[ax];
    double r = 0.0;
    if (c->ndim >= 2) r = -rot[0] / (double)c->n[0] + rot[1] / (double)c->n[1];
    const double dv = c->dvol();
    const double EOm = Om * r * dv;
    const double ED = (-cD * k + pw[1]) * dv;
    const double Eb = 0.5 * beta * pw[2] * dv;
    out5[0] = EOm;
    out5[1] = ED;
    out5[2] = Eb;
    out5[3] = -EOm + ED + Eb;
    out5[4] = -EOm + ED + Eb * 2.0;
}

template <typename P>
int energy_impl(sfd_ctx* c, const void* psi, const void* V, double cD, double beta, double Om, double* out5)
{
    using R = typename P::R;
    using S = typename P::S;
    (void)sizeof(R);
    (void)sizeof(S);
    double sums[3][5];
    long long off[3], tot = 0;
    for (int ax = 0; ax < c->ndim; ++ax) {
        off[ax] = tot;
        tot += 5 * sweep_nblocks(c, ax);
    }
    SFD_TRY(ensure_partial(c, tot));
    FinalizeSets fs{};
    for (int ax = 0; ax < c->ndim; ++ax) {
        SweepArgs a;
        base_args<P>(c, ax, a);
        a.psi = const_cast<void*>(psi);
        a.V = V;
        a.xi = c->xi[ax];
        a.flags = ax == 0 ? kFlagPointwise : 0;
        if (ax == 0 && c->ndim >= 2) { a.coord = c->coord[1]; a.coord_mod = c->n[1]; }
        if (ax == 1) a.coord = c->coord[0];
        a.partial = c->partial + off[ax];
        SFD_TRY(launch_sweep<P, kEnergy>(c, ax, a));
        fs.partial[ax] = a.partial;
        fs.nb[ax] = c->last_nblocks;
        fs.K[ax] = 5;
        fs.out[ax] = c->scal + 8 + 5 * ax;
    }
    hipLaunchKernelGGL(k_finalize_sets, dim3(fs.rows()), dim3(kFinalizeThreads), 0, c->stream, fs);
    SFD_HIP(hipGetLastError());
    SFD_HIP(hipMemcpyAsync(sums, c->scal + 8, 5 * c->ndim * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SFD_HIP(hipStreamSynchronize(c->stream));
    const double kin[3] = {sums[0][0], c->ndim > 1 ? sums[1][0] : 0.0, c->ndim > 2 ? sums[2][0] : 0.0};
    const double rot[2] = {sums[0][1], c->ndim > 1 ? sums[1][1] : 0.0};
    const double pw[3] = {sums[0][2], sums[0][3], sums[0][4]};
    combine_energy(c, kin, rot, pw, cD, beta, Om, out5);
    return SFD_OK;
}

// solve!'s loop: nsteps of (energy of the state, split step).  In real time
// with Omega = 0 the energy comes out of each step's first-half sweeps
// (split_impl with a history): the y and z sums are taken after the x (and
// y) flow, which is unitary along its axis and commutes with the y/z
// transforms, so the sums are the state's own.  With rotation the x flow's
// phase depends on y, and in imaginary time the flows damp; each step is
// then preceded by a separate energy evaluation.  energies: host, 5 per step.
template <typename P>
int split_energy_impl(sfd_ctx* c, void* psi, const void* V, double cD, double beta, double Om, double dt,
                      int64_t nsteps, int order, int mode, double* energies)
{
    if (Om != 0.0 || mode != SFD_REAL_TIME) {
        for (int64_t s = 0; s < nsteps; ++s) {
            SFD_TRY(energy_impl<P>(c, psi, V, cD, beta, Om, energies + 5 * s));
            SFD_TRY(split_impl<P>(c, psi, V, cD, beta, Om, dt, 1, order, mode));
        }
        return SFD_OK;
    }
    constexpr int64_t kChunk = 4096;  // steps per history buffer (256 KiB)
    std::vector<double> h;
    for (int64_t s0 = 0; s0 < nsteps; s0 += kChunk) {
        const int64_t m = std::min(kChunk, nsteps - s0);
        SFD_TRY(ensure_buf(c, &c->ehist, &c->ehist_cap, 8 * m));
        SFD_TRY(split_impl<P>(c, psi, V, cD, beta, Om, dt, m, order, mode, c->ehist));
        h.resize(8 * (size_t)m);
        SFD_HIP(hipMemcpyAsync(h.data(), c->ehist, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        SFD_HIP(hipStreamSynchronize(c->stream));
        for (int64_t s = 0; s < m; ++s) {
            const double* r = h.data() + 8 * s;
            const double kin[3] = {r[0], r[4], r[5]};
            const double rot[2] = {0.0, 0.0};
            const double pw[3] = {r[1], r[2], r[3]};
            combine_energy(c, kin, rot, pw, cD, beta, Om, energies + 5 * (s0 + s));
        }
    }
    return SFD_OK;
}

template <typename P>
int fft_impl(sfd_ctx* c, const void* in, void* out, int axis, int dir)
{
    using R = typename P::R;
    using S = typename P::S;
    (void)sizeof(R);
    (void)sizeof(S);
    SweepArgs a;
    base_args<P>(c, axis, a);
    a.psi = const_cast<void*>(in);
    a.out = out;
    if (dir == SFD_FORWARD) return launch_sweep<P, kForward>(c, axis, a);
    return launch_sweep<P, kInverse>(c, axis, a);
}

// One real-time Strang step that also stores the state before the
// nonlinear phase (psi1 = K psi) in `save`.  Where split_impl folds the step
// (D >= 2, and Omega = 0 or D = 3) this runs the same launches, the double
// round trip storing psi1, so that recomputed states equal the forward
// pass's bit for bit.
template <typename P>
int strang_step_save(sfd_ctx* c, void* psi, const void* V, double cD, double beta, double Om, double dt, void* save)
{
    const int D = c->ndim;
    const double tau = 0.5 * dt;
    if (D >= 2 && (Om == 0.0 || D == 3)) {
        for (int ax = 0; ax < D - 1; ++ax) SFD_TRY(round_trip<P>(c, psi, V, ax, tau, cD, Om, 0, 0, beta, dt));
        SFD_TRY(round_trip<P>(c, psi, V, D - 1, tau, cD, Om, 0, kFlagNL, beta, dt, save, nullptr, true));
        for (int ax = 0; ax < D - 1; ++ax) SFD_TRY(round_trip<P>(c, psi, V, ax, tau, cD, Om, 0, 0, beta, dt));
        return SFD_OK;
    }
    for (int h = 0; h < 2; ++h)
        for (int ax = 0; ax < D; ++ax) {
            const bool nl = (h == 0 && ax == D - 1);
            SFD_TRY(round_trip<P>(c, psi, V, ax, tau, cD, Om, 0, nl ? kFlagNL : 0, beta, dt, nl ? save : nullptr));
        }
    return SFD_OK;
}

// K^dagger: round trips over -tau in reverse axis order.
template <typename P>
int lap_rot_adjoint(sfd_ctx* c, void* G, double cD, double Om, double tau)
{
    for (int ax = c->ndim - 1; ax >= 0; --ax) SFD_TRY(round_trip<P>(c, G, nullptr, ax, -tau, cD, Om, 0, 0, 0.0, 0.0));
    return SFD_OK;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

// Side streams (default priority) beside the context's, with the events
// that order them: ready[b] / freed[b] per saved-state buffer b.  On
// destruction it gives the context back its stream and phase-cache
// eviction, and waits for every stream, so that buffers freed after it are
// idle.
struct SideStreams {
    static constexpr int kMax = 2;
    sfd_ctx* c = nullptr;
    int n = 0;
    hipStream_t main = nullptr, side[kMax] = {nullptr, nullptr};
    hipEvent_t start = nullptr, ready[kMax + 1] = {}, freed[kMax + 1] = {};
    int init(sfd_ctx* ctx, int nside)
    {
        c = ctx;
        main = ctx->stream;
        SFD_HIP(hipEventCreateWithFlags(&start, hipEventDisableTiming));
        for (int b = 0; b <= nside; ++b) {
            SFD_HIP(hipEventCreateWithFlags(&ready[b], hipEventDisableTiming));
            SFD_HIP(hipEventCreateWithFlags(&freed[b], hipEventDisableTiming));
        }
        for (; n < nside; ++n) SFD_HIP(hipStreamCreateWithFlags(&side[n], hipStreamNonBlocking));
        return SFD_OK;
    }
    ~SideStreams()
    {
        if (!c) return;
        c->stream = main;
        c->pin_phases = false;
        for (int k = 0; k < n; ++k) {
            (void)hipStreamSynchronize(side[k]);
            (void)hipStreamDestroy(side[k]);
        }
        (void)hipStreamSynchronize(main);
        if (start) (void)hipEventDestroy(start);
        for (int b = 0; b <= kMax; ++b) {
            if (ready[b]) (void)hipEventDestroy(ready[b]);
            if (freed[b]) (void)hipEventDestroy(freed[b]);
        }
    }
};

// Gradient of L = 1 - |<target|psi_N>|^2 through nsteps real-time Strang
// steps by the reverse sweep, recomputing forward states from checkpoints
// taken every `every` steps (sqrt(nsteps) when every <= 0): memory is
// nsteps/every + (S + 1) every + S + 2 fields with S side streams (2 for
// fields of at most 16 MiB, else 1), forward work twice the plain run.
template <typename P>
int adjoint_impl(sfd_ctx* c, const void* psi0, const void* V, const void* target, double cD, double beta, double Om,
                 double dt, int64_t N, int64_t every, double* gradV, double* grad_beta, double* loss, void* psi_final)
{
    using S = typename P::S;
    using R = typename P::R;
    const long long ns = c->nsite;
    const size_t fb = (size_t)ns * sizeof(cx<S>);
    if (every <= 0) every = (int64_t)std::ceil(std::sqrt((double)(N > 0 ? N : 1)));
    const int64_t nck = N > 0 ? (N + every - 1) / every : 0;
    const long long nb = grid_blocks(ns);
    // Small fields leave most CUs idle in every sweep: two side streams then
    // recompute two segments at once.
    const int nside = fb <= ((size_t)16 << 20) ? 2 : 1;
    const int nbuf = nside + 1;
    DevBuf psi, G, ck, seg, part, acc;
    SFD_HIP(hipMalloc(&psi.p, fb * (size_t)nside));
    SFD_HIP(hipMalloc(&G.p, fb));
    if (nck > 0) SFD_HIP(hipMalloc(&ck.p, fb * (size_t)nck));
    SFD_HIP(hipMalloc(&seg.p, nbuf * fb * (size_t)(every > 0 ? every : 1)));
    SFD_HIP(hipMalloc(&part.p, 2 * (size_t)nb * sizeof(double)));
    SFD_HIP(hipMalloc(&acc.p, 4 * sizeof(double)));
    auto fld = [&](void* base, int64_t k) { return static_cast<void*>(static_cast<char*>(base) + fb * (size_t)k); };
    double* accd = static_cast<double*>(acc.p);
    double* partd = static_cast<double*>(part.p);

    // forward pass with checkpoints
    SFD_HIP(hipMemcpyAsync(psi.p, psi0, fb, hipMemcpyDeviceToDevice, c->stream));
    for (int64_t k = 0; k < nck; ++k) {
        SFD_HIP(hipMemcpyAsync(fld(ck.p, k), psi.p, fb, hipMemcpyDeviceToDevice, c->stream));
        const int64_t m = std::min<int64_t>(every, N - k * every);
        SFD_TRY(split_impl<P>(c, psi.p, V, cD, beta, Om, dt, m, 2, SFD_REAL_TIME, nullptr));
    }
    if (psi_final) SFD_HIP(hipMemcpyAsync(psi_final, psi.p, fb, hipMemcpyDeviceToDevice, c->stream));

    // loss and terminal gradient
    hipLaunchKernelGGL(k_inner<S>, dim3((unsigned)nb), dim3(256), 0, c->stream, static_cast<const cx<S>*>(target),
                       static_cast<const cx<S>*>(psi.p), ns, partd);
    SFD_HIP(hipGetLastError());
    launch_finalize(c->stream, partd, nb, 2, accd);
    SFD_HIP(hipGetLastError());
    double cs[2];
    SFD_HIP(hipMemcpyAsync(cs, accd, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SFD_HIP(hipStreamSynchronize(c->stream));
    const double dv = c->dvol();
    const double cr = cs[0] * dv, ci = cs[1] * dv;
    if (loss) *loss = 1.0 - (cr * cr + ci * ci);
    hipLaunchKernelGGL(k_terminal<S>, dim3((unsigned)nb), dim3(256), 0, c->stream, static_cast<cx<S>*>(G.p),
                       static_cast<const cx<S>*>(target), ns, (const double*)accd, dv);
    SFD_HIP(hipGetLastError());
    if (gradV) SFD_HIP(hipMemsetAsync(gradV, 0, (size_t)ns * sizeof(double), c->stream));
    SFD_HIP(hipMemsetAsync(accd + 2, 0, sizeof(double), c->stream));
    // dL/dbeta: per-workgroup sums over all reversed steps, reduced once
    SFD_HIP(hipMemsetAsync(partd, 0, (size_t)nb * sizeof(double), c->stream));

    // reverse sweep, segment by segment from the last checkpoint.  With
    // Omega = 0 the axis flows commute, so the trailing K^dagger(tau) of one
    // reversed step and the leading one of the step before it run as one
    // K^dagger(2 tau); the sequence of operations on G does not depend on
    // the checkpoint interval.
    const double tau = 0.5 * dt;
    const bool merge = Om == 0.0;
    // Segment i (checkpoint k = nck - 1 - i) is recomputed on side stream
    // i % nside, up to nside segments ahead of the one reversed on the
    // context's stream; its saved states go to buffer i % nbuf of `seg`.
    // Every phase table the streams read is built first, and the cache
    // evicts nothing while they run.
    for (int ax = 0; ax < c->ndim; ++ax) {
        void* ph = nullptr;
        int rows = 0;
        SFD_TRY(build_phase<P>(c, ax, tau, cD, Om, 0, &ph, &rows));
        SFD_TRY(build_phase<P>(c, ax, -tau, cD, Om, 0, &ph, &rows));
        if (merge) SFD_TRY(build_phase<P>(c, ax, -2.0 * tau, cD, Om, 0, &ph, &rows));
    }
    SideStreams ss;
    SFD_TRY(ss.init(c, nside));
    auto seg_len = [&](int64_t i) { return std::min<int64_t>(every, N - (nck - 1 - i) * every); };
    auto saved = [&](int64_t i, int64_t j) { return fld(seg.p, (i % nbuf) * every + j); };
    auto recompute = [&](int64_t i) -> int {
        hipStream_t st = ss.side[i % nside];
        void* work = fld(psi.p, i % nside);
        if (i >= nbuf) SFD_HIP(hipStreamWaitEvent(st, ss.freed[i % nbuf], 0));  // segment i - nbuf reversed
        c->stream = st;
        SFD_HIP(hipMemcpyAsync(work, fld(ck.p, nck - 1 - i), fb, hipMemcpyDeviceToDevice, st));
        for (int64_t j = 0; j < seg_len(i); ++j)
            SFD_TRY(strang_step_save<P>(c, work, V, cD, beta, Om, dt, saved(i, j)));
        c->stream = ss.main;
        SFD_HIP(hipEventRecord(ss.ready[i % nbuf], st));
        return SFD_OK;
    };
    int64_t queued = 0;  // segments whose recompute is enqueued
    if (nck > 0) {
        SFD_HIP(hipEventRecord(ss.start, ss.main));  // checkpoints written, psi free
        for (int k = 0; k < nside; ++k) SFD_HIP(hipStreamWaitEvent(ss.side[k], ss.start, 0));
        c->pin_phases = true;
    }
    bool started = false;
    for (int64_t i = 0; i < nck; ++i) {
        for (; queued < nck && queued <= i + nside; ++queued) SFD_TRY(recompute(queued));
        SFD_HIP(hipStreamWaitEvent(ss.main, ss.ready[i % nbuf], 0));
        for (int64_t j = seg_len(i) - 1; j >= 0; --j) {
            SFD_TRY(lap_rot_adjoint<P>(c, G.p, cD, Om, (merge && started) ? 2.0 * tau : tau));
            started = true;
            hipLaunchKernelGGL((k_nl_adjoint<S, R>), dim3((unsigned)nb), dim3(256), 0, c->stream,
                               static_cast<cx<S>*>(G.p), static_cast<const cx<S>*>(saved(i, j)),
                               static_cast<const S*>(V), ns, (R)beta, (R)dt, gradV, partd);
            SFD_HIP(hipGetLastError());
            if (!merge) SFD_TRY(lap_rot_adjoint<P>(c, G.p, cD, Om, tau));
        }
        SFD_HIP(hipEventRecord(ss.freed[i % nbuf], ss.main));
    }
    c->pin_phases = false;
    if (merge && started) SFD_TRY(lap_rot_adjoint<P>(c, G.p, cD, Om, tau));
    hipLaunchKernelGGL(k_accumulate, dim3(1), dim3(256), 0, c->stream, (const double*)partd, nb, accd + 2);
    SFD_HIP(hipGetLastError());
    double gb = 0.0;
    SFD_HIP(hipMemcpyAsync(&gb, accd + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SFD_HIP(hipStreamSynchronize(c->stream));
    if (grad_beta) *grad_beta = gb;
    return SFD_OK;
}

// Autotuner: for every axis with more than one compiled variant, time the
// identity round trip (tau = 0) of each variant on a scratch field and keep
// the fastest.  Each candidate gets 2 warm-up launches, then the best of 3
// timed batches of 8 launches (HIP events on the context's stream).
template <typename P>
int autotune_impl(sfd_ctx* c, void* scratch, float* times_out)
{
    using R = typename P::R;
    using S = typename P::S;
    (void)sizeof(R);
    (void)sizeof(S);
    void* buf = scratch;
    bool own = false;
    if (!buf) {
        SFD_HIP(hipMalloc(&buf, (size_t)c->nsite * sizeof(cx<S>)));
        own = true;
    }
    int rc = SFD_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(SFD_ERR_HIP, "hipEventCreate failed");
    }
    if (rc == SFD_OK) {
        hipLaunchKernelGGL(k_fill_phases<S>, dim3((unsigned)grid_blocks(c->nsite)), dim3(256), 0, c->stream,
                           static_cast<cx<S>*>(buf), c->nsite);
        if (hipGetLastError() != hipSuccess) rc = fail(SFD_ERR_HIP, "scratch fill failed");
    }
    for (int ax = 0; rc == SFD_OK && ax < c->ndim; ++ax) {
        const int nv = n_variants<P>(c->n[ax], ax == 0);
        int best = 0;
        float best_ms = 1e30f;
        for (int v = 0; rc == SFD_OK && v < nv; ++v) {
            if ((rc = set_variant<P>(c, ax, v)) != SFD_OK) break;
            float vbest = 1e30f;
            for (int w = 0; rc == SFD_OK && w < 2; ++w)
                rc = round_trip<P>(c, buf, nullptr, ax, 0.0, -0.5, 0.0, 0, 0, 0.0, 0.0);
            for (int rep = 0; rc == SFD_OK && rep < 3; ++rep) {
                (void)hipEventRecord(e0, c->stream);
                for (int k = 0; rc == SFD_OK && k < 8; ++k)
                    rc = round_trip<P>(c, buf, nullptr, ax, 0.0, -0.5, 0.0, 0, 0, 0.0, 0.0);
                (void)hipEventRecord(e1, c->stream);
                if (hipEventSynchronize(e1) != hipSuccess) { rc = fail(SFD_ERR_HIP, "event sync failed"); break; }
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (ms / 8 < vbest) vbest = ms / 8;
            }
            if (times_out) times_out[ax * 8 + (v < 8 ? v : 7)] = vbest;
            if (vbest < best_ms) { best_ms = vbest; best = v; }
        }
        if (rc == SFD_OK) rc = set_variant<P>(c, ax, best);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (own) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(buf);
    }
    if (rc == SFD_OK) c->tuned = 1;
    return rc;
}

template <typename P>
int init_tables(sfd_ctx* c)
{
    using R = typename P::R;
    using S = typename P::S;
    (void)sizeof(R);
    (void)sizeof(S);
    for (int ax = 0; ax < c->ndim; ++ax) {
        const int n = c->n[ax];
        SFD_TRY(set_variant<P>(c, ax, 0));
        std::vector<double> xi(n), co(n);
        for (int i = 0; i < n; ++i) {
            xi[i] = fft_freq(n, c->d[ax], i);
            co[i] = lin_coord(c->lo[ax], c->hi[ax], n, i);
        }
        SFD_HIP(hipMalloc(&c->xi[ax], n * sizeof(double)));
        SFD_HIP(hipMalloc(&c->coord[ax], n * sizeof(double)));
        SFD_TRY(upload_sync(c->xi[ax], xi.data(), n * sizeof(double)));
        SFD_TRY(upload_sync(c->coord[ax], co.data(), n * sizeof(double)));
    }
    SFD_HIP(hipMalloc(&c->scal, 32 * sizeof(double)));
    std::vector<double> one(32, 0.0);
    one[0] = 1.0;
    SFD_TRY(upload_sync(c->scal, one.data(), 32 * sizeof(double)));
    return SFD_OK;
}

void release(sfd_ctx* c)
{
    for (int ax = 0; ax < 3; ++ax) {
        if (c->tw[ax]) (void)hipFree(c->tw[ax]);
        if (c->xi[ax]) (void)hipFree(c->xi[ax]);
        if (c->coord[ax]) (void)hipFree(c->coord[ax]);
    }
    for (auto& p : c->phases) (void)hipFree(p.dev);
    if (c->partial) (void)hipFree(c->partial);
    if (c->epartial) (void)hipFree(c->epartial);
    if (c->ehist) (void)hipFree(c->ehist);
    if (c->scal) (void)hipFree(c->scal);
}

}  // namespace

// ------------------------------------------------------------ batch solver
struct sfd_batch {
    struct Inst {
        int n[2];
        double lo[2], hi[2], d[2];
        long long nsite;
        int group;
        int user;  // index in the caller's ordering
    };
    struct Group {
        int n[2];
        int first = 0, count = 0;
        VariantInfo vi[2] = {};
        int var[2] = {0, 0};  // launch variant per axis (sfd_batch_autotune)
        void* tw[2] = {nullptr, nullptr};
        long long ntiles[2] = {0, 0};
    };
    int ninst = 0;
    int prec = SFD_F32;
    int device = 0;
    hipStream_t user_stream = nullptr;
    hipStream_t stream = nullptr;  // own stream (graph capture needs a non-default stream)
    hipEvent_t ev = nullptr;
    // one stream per size group: the groups' chains of launches are
    // independent, so they fork from `stream` and join it again each chunk
    std::vector<hipStream_t> gstream;
    std::vector<hipEvent_t> gdone;
    hipEvent_t fork = nullptr;
    std::vector<Inst> inst;        // internal (group-major) order
    std::vector<Group> groups;
    BatchInstDev* d_inst = nullptr;
    double* d_scale = nullptr;
    double* d_snap_scale = nullptr;
    double* d_dpsi = nullptr;
    int* d_done = nullptr;
    long long* d_steps = nullptr;
    double* d_partial = nullptr;
    double* d_dpart = nullptr;
    void* d_snap = nullptr;
    SweepArgs* d_desc = nullptr;   // [4 kinds][ninst]
    std::vector<void*> phase;      // per instance, per axis
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_k = 0;
    static constexpr int kDiffBlocks = 64;
    // sweeps of one imaginary-time Strang step: x, [y N y], x
    static constexpr int kKinds = 3;
    static int kind_axis(int kind) { return kind == 1 ? 1 : 0; }
};

namespace {

// Launch variant `idx` for one axis of a size group: its twiddles (rebuilt
// when the line split changes) and tile count.
template <typename P>
int batch_group_variant(sfd_batch* b, sfd_batch::Group& G, int ax, int idx)
{
    using R = typename P::R;
    VariantInfo vi;
    if (!variant_info<P>(G.n[ax], ax == 0, idx, &vi))
        return fail(SFD_ERR_SHAPE, "no sweep variant %d for line length %d", idx, G.n[ax]);
    if (!G.tw[ax] || G.vi[ax].n1 != vi.n1 || G.vi[ax].n2 != vi.n2 || G.vi[ax].h != vi.h) {
        std::vector<cx<R>> w;
        fill_twiddles<R>(w, G.n[ax], vi.n1, vi.n2, vi.h);
        if (G.tw[ax]) {
            SFD_HIP(hipStreamSynchronize(b->stream));
            SFD_HIP(hipFree(G.tw[ax]));
            G.tw[ax] = nullptr;
        }
        SFD_HIP(hipMalloc(&G.tw[ax], w.size() * sizeof(cx<R>)));
        SFD_TRY(upload_sync(G.tw[ax], w.data(), w.size() * sizeof(cx<R>)));
    }
    G.vi[ax] = vi;
    G.var[ax] = idx;
    const int NL = vi.ts * vi.g;
    G.ntiles[ax] = ax == 0 ? (G.n[1] + NL - 1) / NL : (G.n[0] + NL - 1) / NL;
    return SFD_OK;
}

// Line geometry of instance `in` along `ax` under its group's variant.
void batch_geom(SweepArgs& a, const sfd_batch::Inst& in, const sfd_batch::Group& G, int ax)
{
    if (ax == 0) {
        a.S = 1;
        a.nlines = in.n[1];
    } else {
        a.S = in.n[0];
        a.OS = (long long)in.n[0] * in.n[1];
        a.ncols = in.n[0];
        a.ntx = (int)G.ntiles[1];
    }
    a.ntiles = G.ntiles[ax];
}

template <typename P>
int batch_setup(sfd_batch* b)
{
    using S = typename P::S;
    // groups of equal extents, internal order group-major
    std::vector<sfd_batch::Inst> sorted;
    for (auto& in : b->inst) {
        int gidx = -1;
        for (size_t g = 0; g < b->groups.size(); ++g)
            if (b->groups[g].n[0] == in.n[0] && b->groups[g].n[1] == in.n[1]) gidx = (int)g;
        if (gidx < 0) {
            sfd_batch::Group G;
            G.n[0] = in.n[0];
            G.n[1] = in.n[1];
            b->groups.push_back(G);
            gidx = (int)b->groups.size() - 1;
        }
        in.group = gidx;
    }
    for (size_t g = 0; g < b->groups.size(); ++g) {
        b->groups[g].first = (int)sorted.size();
        for (auto& in : b->inst)
            if (in.group == (int)g) sorted.push_back(in);
        b->groups[g].count = (int)sorted.size() - b->groups[g].first;
    }
    b->inst = sorted;
    const int ni = b->ninst;
    long long parts = 0, sites = 0;
    for (auto& G : b->groups) {
        for (int ax = 0; ax < 2; ++ax) SFD_TRY(batch_group_variant<P>(b, G, ax, 0));
        // norm partials come from the last x sweep (one per tile of G.n[1]
        // x lines): room for the variant with the most tiles, so the
        // autotuner may switch variants
        long long most = 0;
        for (int v = 0; v < n_variants<P>(G.n[0], true); ++v) {
            VariantInfo vi;
            if (variant_info<P>(G.n[0], true, v, &vi)) {
                const int NL = vi.ts * vi.g;
                most = std::max(most, (long long)((G.n[1] + NL - 1) / NL));
            }
        }
        parts += most * G.count;
    }
    for (auto& in : b->inst) sites += in.nsite;
    SFD_HIP(hipMalloc(&b->d_inst, ni * sizeof(BatchInstDev)));
    SFD_HIP(hipMalloc(&b->d_scale, ni * sizeof(double)));
    SFD_HIP(hipMalloc(&b->d_snap_scale, ni * sizeof(double)));
    SFD_HIP(hipMalloc(&b->d_dpsi, ni * sizeof(double)));
    SFD_HIP(hipMalloc(&b->d_done, ni * sizeof(int)));
    SFD_HIP(hipMalloc(&b->d_steps, ni * sizeof(long long)));
    SFD_HIP(hipMalloc(&b->d_partial, parts * sizeof(double)));
    SFD_HIP(hipMalloc(&b->d_dpart, (size_t)ni * sfd_batch::kDiffBlocks * sizeof(double)));
    SFD_HIP(hipMalloc(&b->d_snap, (size_t)sites * sizeof(cx<S>)));
    SFD_HIP(hipMalloc(&b->d_desc, sfd_batch::kKinds * (size_t)ni * sizeof(SweepArgs)));
    b->phase.assign(2 * (size_t)ni, nullptr);
    SFD_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    for (size_t g = 0; g < b->groups.size(); ++g) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        SFD_HIP(hipStreamCreate(&st));
        b->gstream.push_back(st);
        SFD_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        b->gdone.push_back(ev);
    }
    return SFD_OK;
}

// Descriptors and phase tables of one solve (fields, potentials, beta, dt).
template <typename P>
int batch_bind(sfd_batch* b, void* const* psi, const void* const* V, const double* beta, double cD, double dt)
{
    using R = typename P::R;
    using S = typename P::S;
    const int ni = b->ninst;
    std::vector<BatchInstDev> hi(ni);
    std::vector<SweepArgs> desc(sfd_batch::kKinds * (size_t)ni);
    long long poff = 0, soff = 0;
    const double tau = 0.5 * dt;
    for (int i = 0; i < ni; ++i) {
        const auto& in = b->inst[i];
        const auto& G = b->groups[in.group];
        if (!psi[in.user]) return fail(SFD_ERR_NULL, "batch: field %d is null", in.user);
        hi[i].psi = psi[in.user];
        hi[i].snap = static_cast<char*>(b->d_snap) + soff * sizeof(cx<S>);
        hi[i].partial = b->d_partial + poff;
        hi[i].nsite = in.nsite;
        hi[i].nparts = G.ntiles[0];
        hi[i].dv = in.d[0] * in.d[1];
        // imaginary-time multipliers exp(cD xi^2 tau)/n of both axes
        for (int ax = 0; ax < 2; ++ax) {
            const int n = in.n[ax];
            std::vector<cx<R>> h(n);
            for (int k = 0; k < n; ++k) {
                const double xi = fft_freq(n, in.d[ax], k);
                h[k] = decay_entry<R>((cD * (xi * xi)) * tau, 1.0 / (double)n);
            }
            void*& dev = b->phase[2 * (size_t)i + ax];
            if (!dev) SFD_HIP(hipMalloc(&dev, n * sizeof(cx<R>)));
            SFD_TRY(upload_sync(dev, h.data(), n * sizeof(cx<R>)));
        }
        for (int k = 0; k < sfd_batch::kKinds; ++k) {
            const int ax = sfd_batch::kind_axis(k);
            SweepArgs a;
            std::memset(&a, 0, sizeof a);
            a.psi = psi[in.user];
            a.V = V ? V[in.user] : nullptr;
            a.phase = b->phase[2 * (size_t)i + ax];
            a.tw = G.tw[ax];
            a.scale = b->d_scale + i;
            a.beta = beta[in.user];
            a.dt = dt;
            batch_geom(a, in, G, ax);
            // x (rescale), [y, nonlinear, y], x (+ norm partials): the
            // folded Strang step of split_impl
            a.flags = kFlagImag;
            if (k == 0) a.flags |= kFlagScale;
            if (k == 1) a.flags |= kFlagNL;
            if (k == 2) {
                a.flags |= kFlagNorm;
                a.partial = b->d_partial + poff;
            }
            desc[(size_t)k * ni + i] = a;
        }
        poff += G.ntiles[0];
        soff += in.nsite;
    }
    SFD_TRY(upload_sync(b->d_inst, hi.data(), hi.size() * sizeof(BatchInstDev)));
    SFD_TRY(upload_sync(b->d_desc, desc.data(), desc.size() * sizeof(SweepArgs)));
    return SFD_OK;
}

// Enqueue k imaginary-time Strang steps of every unfinished instance; the
// last step is a convergence check when `check` is set.  Each size group
// runs its k steps on its own stream (forked from and joined back to the
// batch stream), so the groups' small launches overlap.
template <typename P>
int batch_enqueue(sfd_batch* b, int k, bool check, double tol)
{
    using S = typename P::S;
    const int ni = b->ninst;
    constexpr int KB = sfd_batch::kDiffBlocks;
    SFD_HIP(hipEventRecord(b->fork, b->stream));
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        const auto& G = b->groups[gi];
        hipStream_t st = b->gstream[gi];
        SFD_HIP(hipStreamWaitEvent(st, b->fork, 0));
        const int f = G.first;
        for (int s = 0; s < k; ++s) {
            const bool chk = check && s == k - 1;
            if (chk) {
                hipLaunchKernelGGL(k_batch_snapshot<S>, dim3(KB, G.count), dim3(256), 0, st, b->d_inst + f,
                                   b->d_scale + f, b->d_snap_scale + f, b->d_done + f);
                SFD_HIP(hipGetLastError());
            }
            for (int kind = 0; kind < sfd_batch::kKinds; ++kind) {
                const int ax = sfd_batch::kind_axis(kind);
                const VariantInfo& vi = G.vi[ax];
                dim3 grid((unsigned)G.ntiles[ax], 1, (unsigned)G.count);
                dim3 block(vi.ts * vi.g * vi.n2 * vi.h);
                if (!launch_batch<P>(G.n[ax], ax == 0, G.var[ax], grid, block, st,
                                     b->d_desc + (size_t)kind * ni + f, b->d_done + f, kind == 1))
                    return fail(SFD_ERR_SHAPE, "batch: no kernel for line length %d", G.n[ax]);
                SFD_HIP(hipGetLastError());
            }
            hipLaunchKernelGGL(k_batch_finalize, dim3(G.count), dim3(256), 0, st, b->d_inst + f, b->d_scale + f,
                               b->d_done + f, b->d_steps + f);
            SFD_HIP(hipGetLastError());
            if (chk) {
                hipLaunchKernelGGL(k_batch_diff<S>, dim3(KB, G.count), dim3(256), 0, st, b->d_inst + f,
                                   b->d_scale + f, b->d_snap_scale + f, b->d_done + f, b->d_dpart + (size_t)f * KB);
                SFD_HIP(hipGetLastError());
                hipLaunchKernelGGL(k_batch_diff_finalize, dim3(G.count), dim3(256), 0, st,
                                   b->d_dpart + (size_t)f * KB, KB, b->d_inst + f, b->d_dpsi + f, b->d_done + f,
                                   tol);
                SFD_HIP(hipGetLastError());
            }
        }
        SFD_HIP(hipEventRecord(b->gdone[gi], st));
        SFD_HIP(hipStreamWaitEvent(b->stream, b->gdone[gi], 0));
    }
    return SFD_OK;
}

void batch_drop_graph(sfd_batch* b);

// Batch autotuner: for every size group and axis with more than one compiled
// variant, time the identity round trip (tau = 0) of k_batch_sweep (the
// double round trip on y, as in a step) over the
// group's instances on the snapshot buffer filled with unit-modulus
// pseudo-random data (the fields are not touched), 2 warm-up launches then
// the best of 3 batches of 8, and keep the
// fastest.  times[16 g + 8 ax + v] = ms per launch of variant v (-1 if none).
template <typename P>
int batch_autotune_impl(sfd_batch* b, float* times)
{
    using R = typename P::R;
    using S = typename P::S;
    const int ni = b->ninst;
    if (times)
        for (int i = 0; i < 16 * ni; ++i) times[i] = -1.0f;
    batch_drop_graph(b);
    long long sites = 0;
    for (auto& in : b->inst) sites += in.nsite;
    SFD_HIP(hipEventRecord(b->ev, b->user_stream));
    SFD_HIP(hipStreamWaitEvent(b->stream, b->ev, 0));
    hipLaunchKernelGGL(k_fill_phases<S>, dim3((unsigned)grid_blocks(sites)), dim3(256), 0, b->stream,
                       static_cast<cx<S>*>(b->d_snap), sites);
    SFD_HIP(hipGetLastError());
    SFD_HIP(hipMemsetAsync(b->d_done, 0, ni * sizeof(int), b->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<void*> tmp;  // identity phase tables
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(b->stream);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        for (void* p : tmp) (void)hipFree(p);
    };
    auto run = [&]() -> int {
        SFD_HIP(hipEventCreate(&e0));
        SFD_HIP(hipEventCreate(&e1));
        for (size_t gi = 0; gi < b->groups.size(); ++gi) {
            auto& G = b->groups[gi];
            for (int ax = 0; ax < 2; ++ax) {
                const int nv = n_variants<P>(G.n[ax], ax == 0);
                if (nv <= 1) continue;
                const int n = G.n[ax];
                std::vector<cx<R>> h(n, decay_entry<R>(0.0, 1.0 / (double)n));
                void* ph = nullptr;
                SFD_HIP(hipMalloc(&ph, n * sizeof(cx<R>)));
                tmp.push_back(ph);
                SFD_TRY(upload_sync(ph, h.data(), n * sizeof(cx<R>)));
                float best = 1e30f;
                int best_v = 0;
                for (int v = 0; v < nv && v < 8; ++v) {
                    SFD_TRY(batch_group_variant<P>(b, G, ax, v));
                    std::vector<SweepArgs> desc(G.count);
                    long long soff = 0;
                    for (int i = 0; i < G.first; ++i) soff += b->inst[i].nsite;
                    for (int k = 0; k < G.count; ++k) {
                        const auto& in = b->inst[G.first + k];
                        SweepArgs& a = desc[k];
                        std::memset(&a, 0, sizeof a);
                        a.psi = static_cast<char*>(b->d_snap) + soff * sizeof(cx<S>);
                        a.phase = ph;
                        a.tw = G.tw[ax];
                        a.flags = kFlagImag;
                        batch_geom(a, in, G, ax);
                        soff += in.nsite;
                    }
                    SFD_TRY(upload_sync(b->d_desc, desc.data(), desc.size() * sizeof(SweepArgs)));
                    const VariantInfo& vi = G.vi[ax];
                    dim3 grid((unsigned)G.ntiles[ax], 1, (unsigned)G.count);
                    dim3 block(vi.ts * vi.g * vi.n2 * vi.h);
                    // the y sweep of a step is the double round trip
                    auto launch = [&]() {
                        return launch_batch<P>(n, ax == 0, v, grid, block, b->stream, b->d_desc, b->d_done, ax == 1);
                    };
                    for (int w = 0; w < 2; ++w)
                        if (!launch()) return fail(SFD_ERR_SHAPE, "batch autotune: no kernel %d for n = %d", v, n);
                    SFD_HIP(hipGetLastError());
                    float t_best = 1e30f;
                    for (int rep = 0; rep < 3; ++rep) {
                        SFD_HIP(hipEventRecord(e0, b->stream));
                        for (int k = 0; k < 8; ++k) launch();
                        SFD_HIP(hipEventRecord(e1, b->stream));
                        SFD_HIP(hipEventSynchronize(e1));
                        float ms = 0.0f;
                        SFD_HIP(hipEventElapsedTime(&ms, e0, e1));
                        t_best = std::min(t_best, ms / 8.0f);
                    }
                    if (times) times[16 * gi + 8 * ax + v] = t_best;
                    if (t_best < best) {
                        best = t_best;
                        best_v = v;
                    }
                }
                SFD_TRY(batch_group_variant<P>(b, G, ax, best_v));
            }
        }
        return SFD_OK;
    };
    const int rc = run();
    cleanup();
    return rc;
}

void batch_drop_graph(sfd_batch* b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->graph_k = 0;
}

template <typename P>
int batch_run(sfd_batch* b, void* const* psi, const void* const* V, const double* beta, double cD, double dt,
              int64_t max_steps, double tol, int check_every, int use_graph, int64_t* steps_out, double* dpsi_out)
{
    using S = typename P::S;
    const int ni = b->ninst;
    SFD_TRY(batch_bind<P>(b, psi, V, beta, cD, dt));
    // order after the caller's pending work
    SFD_HIP(hipEventRecord(b->ev, b->user_stream));
    SFD_HIP(hipStreamWaitEvent(b->stream, b->ev, 0));
    hipLaunchKernelGGL(k_batch_reset, dim3((ni + 255) / 256), dim3(256), 0, b->stream, b->d_scale, b->d_done,
                       b->d_steps, b->d_dpsi, ni, 1);
    SFD_HIP(hipGetLastError());
    batch_drop_graph(b);
    std::vector<int> done(ni, 0);
    int64_t left = max_steps;
    while (left > 0) {
        const int k = (int)(left < check_every ? left : check_every);
        if (use_graph && k == check_every && k > 0) {
            if (!b->gexec) {
                SFD_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                int rc = batch_enqueue<P>(b, k, true, tol);
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamEndCapture(b->stream, &g);
                if (rc != SFD_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (e != hipSuccess) return fail(SFD_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
                b->graph = g;
                SFD_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
                b->graph_k = k;
            }
            SFD_HIP(hipGraphLaunch(b->gexec, b->stream));
        } else {
            SFD_TRY(batch_enqueue<P>(b, k, true, tol));
        }
        left -= k;
        SFD_HIP(hipMemcpyAsync(done.data(), b->d_done, ni * sizeof(int), hipMemcpyDeviceToHost, b->stream));
        SFD_HIP(hipStreamSynchronize(b->stream));
        bool all = true;
        for (int i = 0; i < ni; ++i) all = all && done[i];
        if (all) break;
    }
    // apply the pending renormalisation, leave normalised fields
    hipLaunchKernelGGL(k_batch_apply_scale<S>, dim3(sfd_batch::kDiffBlocks, ni), dim3(256), 0, b->stream, b->d_inst,
                       b->d_scale);
    SFD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_batch_reset, dim3((ni + 255) / 256), dim3(256), 0, b->stream, b->d_scale, b->d_done,
                       b->d_steps, b->d_dpsi, ni, 0);
    SFD_HIP(hipGetLastError());
    std::vector<long long> st(ni);
    std::vector<double> dp(ni);
    SFD_HIP(hipMemcpyAsync(st.data(), b->d_steps, ni * sizeof(long long), hipMemcpyDeviceToHost, b->stream));
    SFD_HIP(hipMemcpyAsync(dp.data(), b->d_dpsi, ni * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    SFD_HIP(hipStreamSynchronize(b->stream));
    for (int i = 0; i < ni; ++i) {
        if (steps_out) steps_out[b->inst[i].user] = st[i];
        if (dpsi_out) dpsi_out[b->inst[i].user] = dp[i];
    }
    return SFD_OK;
}

void batch_release(sfd_batch* b)
{
    batch_drop_graph(b);
    for (auto& G : b->groups)
        for (int ax = 0; ax < 2; ++ax)
            if (G.tw[ax]) (void)hipFree(G.tw[ax]);
    for (void* p : b->phase)
        if (p) (void)hipFree(p);
    void* ptrs[] = {b->d_inst, b->d_scale, b->d_snap_scale, b->d_dpsi, b->d_done, b->d_steps,
                    b->d_partial, b->d_dpart, b->d_snap, b->d_desc};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (b->ev) (void)hipEventDestroy(b->ev);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (hipEvent_t e : b->gdone) (void)hipEventDestroy(e);
    for (hipStream_t st : b->gstream) (void)hipStreamDestroy(st);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

}  // namespace

// Call FN<precision traits>(args...) for the context's precision.
#define SFD_BY_PREC(ctx, FN, ...)                                         \
    ((ctx)->prec == SFD_F32 ? FN<PF32>(__VA_ARGS__)                       \
                            : (ctx)->prec == SFD_F64 ? FN<PF64>(__VA_ARGS__) : FN<PF32X>(__VA_ARGS__))

// ===================================================================== C ABI
extern "C" {

const char* sfd_last_error(void) { return g_last_error.c_str(); }

int sfd_version(void) { return 100; }

int sfd_ctx_create(sfd_ctx** out, int ndim, const int64_t* n, const double* box_min, const double* box_max,
                   int precision, void* stream)
{
    g_last_error.clear();
    if (!out || !n || !box_min || !box_max) return fail(SFD_ERR_NULL, "sfd_ctx_create: null argument");
    *out = nullptr;
    if (ndim < 2 || ndim > 3) return fail(SFD_ERR_SHAPE, "ndim must be 2 or 3 (got %d)", ndim);
    if (precision != SFD_F32 && precision != SFD_F64 && precision != SFD_F32X)
        return fail(SFD_ERR_ARG, "precision must be SFD_F32, SFD_F64 or SFD_F32X (got %d)", precision);
    for (int a = 0; a < ndim; ++a) {
        const int64_t m = n[a];
        if (m < 2 || m > 1024 || (m & (m - 1)))
            return fail(SFD_ERR_SHAPE, "n[%d] = %lld: line lengths must be powers of two in [2, 1024]", a,
                        (long long)m);
        if (!(box_max[a] > box_min[a]) || !std::isfinite(box_min[a]) || !std::isfinite(box_max[a]))
            return fail(SFD_ERR_ARG, "axis %d: bounds must satisfy min < max (Grids.jl:69-76)", a);
    }
    auto* c = new sfd_ctx();
    c->ndim = ndim;
    c->prec = precision;
    c->stream = static_cast<hipStream_t>(stream);
    for (int a = 0; a < ndim; ++a) {
        c->n[a] = (int)n[a];
        c->lo[a] = box_min[a];
        c->hi[a] = box_max[a];
        c->d[a] = (box_max[a] - box_min[a]) / (double)n[a];
        c->nsite *= n[a];
    }
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) {
        delete c;
        return fail(SFD_ERR_HIP, "hipGetDevice failed: %s", hipGetErrorString(e));
    }
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
        c->ncu < 1)
        c->ncu = 1;
    int rc = SFD_BY_PREC(c, init_tables, c);
    if (rc != SFD_OK) {
        release(c);
        delete c;
        return rc;
    }
    *out = c;
    return SFD_OK;
}

int sfd_ctx_destroy(sfd_ctx* ctx)
{
    if (!ctx) return SFD_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    release(ctx);
    delete ctx;
    return SFD_OK;
}

int sfd_ctx_set_stream(sfd_ctx* ctx, void* stream)
{
    SFD_TRY(check_ctx(ctx));
    ctx->stream = static_cast<hipStream_t>(stream);
    return SFD_OK;
}

int sfd_split_step(sfd_ctx* ctx, void* psi, const void* V, double coeffD, double beta, double Omega, double dt,
                   int64_t nsteps, int order, int mode)
{
    SFD_TRY(check_ctx(ctx));
    if (!psi) return fail(SFD_ERR_NULL, "sfd_split_step: psi is null");
    if (order != 1 && order != 2) return fail(SFD_ERR_ARG, "order must be 1 or 2 (got %d)", order);
    if (mode != SFD_REAL_TIME && mode != SFD_IMAG_TIME) return fail(SFD_ERR_ARG, "mode must be 0 or 1");
    if (nsteps < 0) return fail(SFD_ERR_ARG, "nsteps must be >= 0");
    if (!std::isfinite(coeffD) || !std::isfinite(beta) || !std::isfinite(Omega) || !std::isfinite(dt))
        return fail(SFD_ERR_ARG, "non-finite scalar argument");
    return SFD_BY_PREC(ctx, split_impl, ctx, psi, V, coeffD, beta, Omega, dt, nsteps, order, mode);
}

int sfd_split_step_energy(sfd_ctx* ctx, void* psi, const void* V, double coeffD, double beta, double Omega,
                          double dt, int64_t nsteps, int order, int mode, double* energies)
{
    SFD_TRY(check_ctx(ctx));
    if (!psi) return fail(SFD_ERR_NULL, "sfd_split_step_energy: psi is null");
    if (!energies && nsteps > 0) return fail(SFD_ERR_NULL, "sfd_split_step_energy: energies is null");
    if (order != 1 && order != 2) return fail(SFD_ERR_ARG, "order must be 1 or 2 (got %d)", order);
    if (mode != SFD_REAL_TIME && mode != SFD_IMAG_TIME) return fail(SFD_ERR_ARG, "mode must be 0 or 1");
    if (nsteps < 0) return fail(SFD_ERR_ARG, "nsteps must be >= 0");
    if (!std::isfinite(coeffD) || !std::isfinite(beta) || !std::isfinite(Omega) || !std::isfinite(dt))
        return fail(SFD_ERR_ARG, "non-finite scalar argument");
    return SFD_BY_PREC(ctx, split_energy_impl, ctx, psi, V, coeffD, beta, Omega, dt, nsteps, order, mode,
                       energies);
}

int sfd_lap_rot_axis(sfd_ctx* ctx, void* psi, int axis, double coeffD, double Omega, double tau, int mode)
{
    SFD_TRY(check_ctx(ctx));
    if (!psi) return fail(SFD_ERR_NULL, "sfd_lap_rot_axis: psi is null");
    if (axis < 0 || axis >= ctx->ndim) return fail(SFD_ERR_ARG, "axis %d out of range", axis);
    if (mode != SFD_REAL_TIME && mode != SFD_IMAG_TIME) return fail(SFD_ERR_ARG, "mode must be 0 or 1");
    if (!std::isfinite(coeffD) || !std::isfinite(Omega) || !std::isfinite(tau))
        return fail(SFD_ERR_ARG, "non-finite scalar argument");
    const int imag = mode == SFD_IMAG_TIME;
    return SFD_BY_PREC(ctx, round_trip, ctx, psi, nullptr, axis, tau, coeffD, Omega, imag, 0, 0.0, 0.0);
}

int sfd_lap_rot(sfd_ctx* ctx, void* psi, double coeffD, double Omega, double tau, int mode)
{
    SFD_TRY(check_ctx(ctx));
    for (int ax = 0; ax < ctx->ndim; ++ax) SFD_TRY(sfd_lap_rot_axis(ctx, psi, ax, coeffD, Omega, tau, mode));
    return SFD_OK;
}

int sfd_solve_nl(sfd_ctx* ctx, void* psi, const void* V, double beta, double dt, int mode)
{
    SFD_TRY(check_ctx(ctx));
    if (!psi) return fail(SFD_ERR_NULL, "sfd_solve_nl: psi is null");
    if (mode != SFD_REAL_TIME && mode != SFD_IMAG_TIME) return fail(SFD_ERR_ARG, "mode must be 0 or 1");
    if (!std::isfinite(beta) || !std::isfinite(dt)) return fail(SFD_ERR_ARG, "non-finite scalar argument");
    const long long nb = grid_blocks(ctx->nsite);
    const int imag = mode == SFD_IMAG_TIME;
    if (ctx->prec == SFD_F32)
        hipLaunchKernelGGL((k_nl<float, float>), dim3((unsigned)nb), dim3(256), 0, ctx->stream,
                           static_cast<cx<float>*>(psi), static_cast<const float*>(V), ctx->nsite, (float)beta,
                           (float)dt, imag);
    else if (ctx->prec == SFD_F32X)
        hipLaunchKernelGGL((k_nl<float, double>), dim3((unsigned)nb), dim3(256), 0, ctx->stream,
                           static_cast<cx<float>*>(psi), static_cast<const float*>(V), ctx->nsite, beta, dt, imag);
    else
        hipLaunchKernelGGL((k_nl<double, double>), dim3((unsigned)nb), dim3(256), 0, ctx->stream,
                           static_cast<cx<double>*>(psi), static_cast<const double*>(V), ctx->nsite, beta, dt, imag);
    SFD_HIP(hipGetLastError());
    return SFD_OK;
}

int sfd_norm2(sfd_ctx* ctx, const void* psi, double* out_sum)
{
    SFD_TRY(check_ctx(ctx));
    if (!psi || !out_sum) return fail(SFD_ERR_NULL, "sfd_norm2: null argument");
    int rc = SFD_BY_PREC(ctx, norm2_impl, ctx, psi, ctx->scal + 1);
    if (rc != SFD_OK) return rc;
    SFD_HIP(hipMemcpyAsync(out_sum, ctx->scal + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SFD_HIP(hipStreamSynchronize(ctx->stream));
    return SFD_OK;
}

int sfd_normalize(sfd_ctx* ctx, void* psi, double* out_norm)
{
    double s = 0.0;
    SFD_TRY(sfd_norm2(ctx, psi, &s));
    // norm = sqrt(sum) * sqrt(prod dx) (Fields.jl:172-180)
    double nrm = std::sqrt(s) * std::sqrt(ctx->dvol());
    if (out_norm) *out_norm = nrm;
    if (!(nrm > 0.0) || !std::isfinite(nrm)) return fail(SFD_ERR_ARG, "normalize: norm is %g", nrm);
    const long long nb = grid_blocks(ctx->nsite);
    if (ctx->prec != SFD_F64)
        hipLaunchKernelGGL(k_scale<float>, dim3((unsigned)nb), dim3(256), 0, ctx->stream, static_cast<cx<float>*>(psi),
                           ctx->nsite, nrm, (const double*)nullptr, 1);
    else
        hipLaunchKernelGGL(k_scale<double>, dim3((unsigned)nb), dim3(256), 0, ctx->stream,
                           static_cast<cx<double>*>(psi), ctx->nsite, nrm, (const double*)nullptr, 1);
    SFD_HIP(hipGetLastError());
    SFD_HIP(hipStreamSynchronize(ctx->stream));
    return SFD_OK;
}

int sfd_energy(sfd_ctx* ctx, const void* psi, const void* V, double coeffD, double beta, double Omega, double* out5)
{
    SFD_TRY(check_ctx(ctx));
    if (!psi || !out5) return fail(SFD_ERR_NULL, "sfd_energy: null argument");
    return SFD_BY_PREC(ctx, energy_impl, ctx, psi, V, coeffD, beta, Omega, out5);
}

int sfd_fft_axis(sfd_ctx* ctx, const void* in, void* out, int axis, int direction)
{
    SFD_TRY(check_ctx(ctx));
    if (!in || !out) return fail(SFD_ERR_NULL, "sfd_fft_axis: null argument");
    if (axis < 0 || axis >= ctx->ndim) return fail(SFD_ERR_ARG, "axis %d out of range", axis);
    if (direction != SFD_FORWARD && direction != SFD_BACKWARD) return fail(SFD_ERR_ARG, "direction must be -1 or 1");
    return SFD_BY_PREC(ctx, fft_impl, ctx, in, out, axis, direction);
}

int sfd_batch_create(sfd_batch** out, int ninst, const int64_t* n, const double* box, int precision, void* stream)
{
    g_last_error.clear();
    if (!out || !n || !box) return fail(SFD_ERR_NULL, "sfd_batch_create: null argument");
    *out = nullptr;
    if (ninst < 1) return fail(SFD_ERR_ARG, "sfd_batch_create: ninst must be >= 1");
    if (precision != SFD_F32 && precision != SFD_F64 && precision != SFD_F32X)
        return fail(SFD_ERR_ARG, "precision must be SFD_F32, SFD_F64 or SFD_F32X (got %d)", precision);
    auto* b = new sfd_batch();
    b->ninst = ninst;
    b->prec = precision;
    b->user_stream = static_cast<hipStream_t>(stream);
    for (int i = 0; i < ninst; ++i) {
        sfd_batch::Inst in;
        in.nsite = 1;
        for (int a = 0; a < 2; ++a) {
            const int64_t m = n[2 * i + a];
            if (m < 2 || m > 1024 || (m & (m - 1))) {
                delete b;
                return fail(SFD_ERR_SHAPE, "instance %d axis %d: n = %lld is not a power of two in [2, 1024]", i, a,
                            (long long)m);
            }
            in.n[a] = (int)m;
            in.lo[a] = box[4 * i + 2 * a];
            in.hi[a] = box[4 * i + 2 * a + 1];
            if (!(in.hi[a] > in.lo[a])) {
                delete b;
                return fail(SFD_ERR_ARG, "instance %d axis %d: bounds must satisfy min < max", i, a);
            }
            in.d[a] = (in.hi[a] - in.lo[a]) / (double)m;
            in.nsite *= m;
        }
        in.group = -1;
        in.user = i;
        b->inst.push_back(in);
    }
    hipError_t e = hipGetDevice(&b->device);
    if (e == hipSuccess) e = hipStreamCreate(&b->stream);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        batch_release(b);
        delete b;
        return fail(SFD_ERR_HIP, "sfd_batch_create: %s", hipGetErrorString(e));
    }
    int rc = SFD_BY_PREC(b, batch_setup, b);
    if (rc != SFD_OK) {
        batch_release(b);
        delete b;
        return rc;
    }
    *out = b;
    return SFD_OK;
}

int sfd_batch_autotune(sfd_batch* b, float* times_ms)
{
    g_last_error.clear();
    if (!b) return fail(SFD_ERR_NULL, "sfd_batch_autotune: null batch");
    SFD_HIP(hipSetDevice(b->device));
    return SFD_BY_PREC(b, batch_autotune_impl, b, times_ms);
}

int sfd_batch_destroy(sfd_batch* b)
{
    if (!b) return SFD_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    batch_release(b);
    delete b;
    return SFD_OK;
}

int sfd_batch_imag_run(sfd_batch* b, void* const* psi, const void* const* V, const double* beta, double coeffD,
                       double dt, int64_t max_steps, double tol, int check_every, int use_graph, int64_t* steps_out,
                       double* dpsi_out)
{
    if (!b) return fail(SFD_ERR_NULL, "null batch");
    if (!psi || !beta) return fail(SFD_ERR_NULL, "sfd_batch_imag_run: null argument");
    if (max_steps < 0 || check_every < 1) return fail(SFD_ERR_ARG, "max_steps >= 0 and check_every >= 1 required");
    if (!std::isfinite(coeffD) || !std::isfinite(dt) || !std::isfinite(tol))
        return fail(SFD_ERR_ARG, "non-finite scalar argument");
    SFD_HIP(hipSetDevice(b->device));
    return SFD_BY_PREC(b, batch_run, b, psi, V, beta, coeffD, dt, max_steps, tol, check_every, use_graph,
                       steps_out, dpsi_out);
}

int sfd_adjoint(sfd_ctx* ctx, const void* psi0, const void* V, const void* target, double coeffD, double beta,
                double Omega, double dt, int64_t nsteps, int64_t checkpoint_every, double* gradV, double* grad_beta,
                double* loss, void* psi_final)
{
    SFD_TRY(check_ctx(ctx));
    if (!psi0 || !target) return fail(SFD_ERR_NULL, "sfd_adjoint: null field");
    if (nsteps < 0) return fail(SFD_ERR_ARG, "nsteps must be >= 0");
    if (!std::isfinite(coeffD) || !std::isfinite(beta) || !std::isfinite(Omega) || !std::isfinite(dt))
        return fail(SFD_ERR_ARG, "non-finite scalar argument");
    return SFD_BY_PREC(ctx, adjoint_impl, ctx, psi0, V, target, coeffD, beta, Omega, dt, nsteps, checkpoint_every,
                       gradV, grad_beta, loss, psi_final);
}

int sfd_inner(sfd_ctx* ctx, const void* a, const void* b, double* out2)
{
    SFD_TRY(check_ctx(ctx));
    if (!a || !b || !out2) return fail(SFD_ERR_NULL, "sfd_inner: null argument");
    const long long nb = grid_blocks(ctx->nsite);
    SFD_TRY(ensure_partial(ctx, 2 * nb));
    if (ctx->prec == SFD_F64)
        hipLaunchKernelGGL(k_inner<double>, dim3((unsigned)nb), dim3(256), 0, ctx->stream,
                           static_cast<const cx<double>*>(a), static_cast<const cx<double>*>(b), ctx->nsite,
                           ctx->partial);
    else
        hipLaunchKernelGGL(k_inner<float>, dim3((unsigned)nb), dim3(256), 0, ctx->stream,
                           static_cast<const cx<float>*>(a), static_cast<const cx<float>*>(b), ctx->nsite,
                           ctx->partial);
    SFD_HIP(hipGetLastError());
    launch_finalize(ctx->stream, ctx->partial, nb, 2, ctx->scal + 4);
    SFD_HIP(hipGetLastError());
    SFD_HIP(hipMemcpyAsync(out2, ctx->scal + 4, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SFD_HIP(hipStreamSynchronize(ctx->stream));
    const double dv = ctx->dvol();
    out2[0] *= dv;
    out2[1] *= dv;
    return SFD_OK;
}

int sfd_project(sfd_ctx* ctx, const double* field, const double* basis, int nbasis, double* out)
{
    SFD_TRY(check_ctx(ctx));
    if (!field || !basis || !out) return fail(SFD_ERR_NULL, "sfd_project: null argument");
    if (nbasis < 1 || nbasis > 65535) return fail(SFD_ERR_ARG, "nbasis must be in [1, 65535]");
    const long long nb = 64;
    DevBuf part, res;
    SFD_HIP(hipMalloc(&part.p, (size_t)nb * nbasis * sizeof(double)));
    SFD_HIP(hipMalloc(&res.p, (size_t)nbasis * sizeof(double)));
    hipLaunchKernelGGL(k_project, dim3((unsigned)nb, (unsigned)nbasis), dim3(256), 0, ctx->stream, field, basis,
                       ctx->nsite, static_cast<double*>(part.p));
    SFD_HIP(hipGetLastError());
    for (int j = 0; j < nbasis; ++j) {
        launch_finalize(ctx->stream, static_cast<const double*>(part.p) + (size_t)j * nb, nb, 1,
                        static_cast<double*>(res.p) + j);
        SFD_HIP(hipGetLastError());
    }
    SFD_HIP(hipMemcpyAsync(out, res.p, (size_t)nbasis * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SFD_HIP(hipStreamSynchronize(ctx->stream));
    return SFD_OK;
}

int sfd_autotune(sfd_ctx* ctx, void* scratch, float* times_ms)
{
    SFD_TRY(check_ctx(ctx));
    if (times_ms)
        for (int i = 0; i < 24; ++i) times_ms[i] = -1.f;
    return SFD_BY_PREC(ctx, autotune_impl, ctx, scratch, times_ms);
}

int sfd_ctx_get_variant(sfd_ctx* ctx, int axis, int* index, int* count, int* shape7)
{
    if (!ctx || !index || !count) return fail(SFD_ERR_NULL, "sfd_ctx_get_variant: null argument");
    if (axis < 0 || axis >= ctx->ndim) return fail(SFD_ERR_ARG, "axis %d out of range", axis);
    *index = ctx->var[axis];
    *count = SFD_BY_PREC(ctx, n_variants, ctx->n[axis], axis == 0);
    if (shape7) {
        shape7[0] = ctx->vi[axis].n1;
        shape7[1] = ctx->vi[axis].n2;
        shape7[2] = ctx->vi[axis].ts;
        shape7[3] = ctx->vi[axis].g;
        shape7[4] = ctx->vi[axis].pipe;
        shape7[5] = ctx->vi[axis].waves;
        shape7[6] = ctx->vi[axis].h;
    }
    return SFD_OK;
}

int sfd_ctx_set_variant(sfd_ctx* ctx, int axis, int index)
{
    SFD_TRY(check_ctx(ctx));
    if (axis < 0 || axis >= ctx->ndim) return fail(SFD_ERR_ARG, "axis %d out of range", axis);
    return SFD_BY_PREC(ctx, set_variant, ctx, axis, index);
}

int sfd_alloc(sfd_ctx* ctx, size_t bytes, void** out)
{
    SFD_TRY(check_ctx(ctx));
    if (!out) return fail(SFD_ERR_NULL, "sfd_alloc: null out");
    *out = nullptr;
    SFD_HIP(hipMalloc(out, bytes ? bytes : 1));
    return SFD_OK;
}

int sfd_free(sfd_ctx* ctx, void* ptr)
{
    SFD_TRY(check_ctx(ctx));
    if (ptr) {
        SFD_HIP(hipStreamSynchronize(ctx->stream));
        SFD_HIP(hipFree(ptr));
    }
    return SFD_OK;
}

int sfd_upload(sfd_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    SFD_TRY(check_ctx(ctx));
    if (!dst || !src) return fail(SFD_ERR_NULL, "sfd_upload: null argument");
    SFD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    SFD_HIP(hipStreamSynchronize(ctx->stream));
    return SFD_OK;
}

int sfd_download(sfd_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    SFD_TRY(check_ctx(ctx));
    if (!dst || !src) return fail(SFD_ERR_NULL, "sfd_download: null argument");
    SFD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    SFD_HIP(hipStreamSynchronize(ctx->stream));
    return SFD_OK;
}

int sfd_synchronize(sfd_ctx* ctx)
{
    SFD_TRY(check_ctx(ctx));
    SFD_HIP(hipStreamSynchronize(ctx->stream));
    return SFD_OK;
}

}  // extern "C"
