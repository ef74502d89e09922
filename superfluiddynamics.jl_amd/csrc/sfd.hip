// sfd.hip -- runtime and C ABI of libsfdhip.so (declared in include/sfd.h).
//
// The context replaces the reference's Grid + FFT Plan pair (Grids.jl:64-129,
// Plans.jl:958-999) for one field shape: it owns the FP64 geometry
// (coordinates, wavenumbers), the twiddle tables, a cache of phase-multiplier
// tables keyed by (axis, tau, coeffD, Omega, mode) and the reduction
// workspace.  Fields and potentials stay owned by the caller.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sfd.h"
#include "sfd_sweep.hpp"

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

// ------------------------------------------------------------ launch shapes
constexpr int lg2(int n) { return n <= 1 ? 0 : 1 + lg2(n / 2); }
constexpr int kLdsBudget = 65536;

template <int N>
struct Split {
    // N = N1 * N2 with N1 >= N2, both powers of two, N1 <= 32.
    static constexpr int N2 = 1 << (lg2(N) / 2);
    static constexpr int N1 = N / N2;
};

template <typename R, int N>
struct Cfg {
    static constexpr int N1 = Split<N>::N1;
    static constexpr int N2 = Split<N>::N2;
    static constexpr int CS = (int)sizeof(cx<R>);
    static constexpr int gc(int g)
    {
        return (g > 1 && (g * N2 > 256 || g * (N2 + 1) * N1 * CS > kLdsBudget)) ? gc(g / 2) : g;
    }
    static constexpr int G_contig = gc(256);
    static constexpr int ts(int s) { return (s > 1 && s * N * CS > kLdsBudget) ? ts(s / 2) : s; }
    static constexpr int TS = ts(16);
    static constexpr int gs(int g)
    {
        return (g < 64 && TS * g * N2 < 64 && 2 * TS * g * N * CS <= kLdsBudget) ? gs(2 * g) : g;
    }
    static constexpr int G_strided = gs(1);
};

// ----------------------------------------------------------- small kernels
template <typename R>
__global__ void k_norm2(const cx<R>* __restrict__ psi, long long n, double* partial)
{
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const cx<R> v = psi[i];
        const double re = v.re, im = v.im;
        s += re * re + im * im;
    }
    double acc[1] = {s};
    block_partials<R, 256, 1>(acc, partial, gridDim.x, blockIdx.x);
}

// out[k] = sum_b partial[k*nb + b] in a fixed order (deterministic).
// If scale_out != null: *scale_out = 1 / sqrt(out[0] * dv).
__global__ void k_finalize(const double* __restrict__ partial, long long nb, int K, double* out,
                           double dv, double* scale_out)
{
    __shared__ double red[256];
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (long long b = threadIdx.x; b < nb; b += blockDim.x) s += partial[k * nb + b];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = red[0];
        __syncthreads();
    }
    if (scale_out && threadIdx.x == 0) scale_out[0] = 1.0 / sqrt(out[0] * dv);
}

// psi *= s (divide = 0) or psi /= s (divide = 1); s is read from device
// memory when s_dev is not null.
template <typename R>
__global__ void k_scale(cx<R>* __restrict__ psi, long long n, double s_host, const double* s_dev, int divide)
{
    const double s = s_dev ? *s_dev : s_host;
    const R f = R(s);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        cx<R> v = psi[i];
        if (divide) { v.re /= f; v.im /= f; }
        else        { v.re *= f; v.im *= f; }
        psi[i] = v;
    }
}

// Pointwise nonlinear flow (solveNL!, splitting.jl:205-209).
template <typename R>
__global__ void k_nl(cx<R>* __restrict__ psi, const R* __restrict__ V, long long n, R beta, R dt, int imag)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        cx<R> v = psi[i];
        nl_phase<R>(v, V ? V[i] : R(0), beta, dt, imag != 0);
        psi[i] = v;
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
    void* tw[3] = {nullptr, nullptr, nullptr};
    double* xi[3] = {nullptr, nullptr, nullptr};
    double* coord[3] = {nullptr, nullptr, nullptr};
    double* partial = nullptr;
    long long partial_cap = 0;  // doubles
    double* scal = nullptr;     // [0] pending scale, [1..31] sums
    std::vector<PhaseEntry> phases;

    size_t csize() const { return prec == SFD_F32 ? 8 : 16; }
    double dvol() const
    {
        double v = 1.0;
        for (int a = 0; a < ndim; ++a) v *= d[a];
        return v;
    }
};

namespace {

template <typename R>
void fill_twiddles(std::vector<cx<R>>& w, int n)
{
    w.resize(n);
    const long double pi = 3.14159265358979323846264338327950288L;
    for (int m = 0; m < n; ++m) {
        const long double a = -2.0L * pi * (long double)m / (long double)n;
        w[m].re = (R)cosl(a);
        w[m].im = (R)sinl(a);
    }
}

int upload_sync(void* dst, const void* src, size_t bytes)
{
    SFD_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SFD_OK;
}

// Phase multiplier of one axis (splitting.jl:190,195,200), divided by n:
//   x: exp(s (cD xi^2 - Om y xi) tau)   rows indexed by y when Om != 0
//   y: exp(s (cD xi^2 + Om x xi) tau)   rows indexed by x when Om != 0
//   z: exp(s  cD xi^2 tau)
// s = i (real time) or 1 (imaginary time).  Formed in FP64 with the
// reference's operation order, rounded once to the field precision.
template <typename R>
int build_phase(sfd_ctx* c, int axis, double tau, double cD, double Om, int imag, void** out, int* rows_out)
{
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
    const double inv_n = 1.0 / (double)n;
    for (int r = 0; r < rows; ++r) {
        const double co = rot ? lin_coord(c->lo[oax], c->hi[oax], c->n[oax], r) : 0.0;
        for (int k = 0; k < n; ++k) {
            const double xi = fft_freq(n, c->d[axis], k);
            double arg;
            if (axis == 0) arg = (cD * (xi * xi) - (Om * co) * xi) * tau;
            else if (axis == 1) arg = (cD * (xi * xi) + (Om * co) * xi) * tau;
            else arg = (cD * (xi * xi)) * tau;
            double mr, mi;
            if (imag) { mr = std::exp(arg); mi = 0.0; }
            else      { mr = std::cos(arg); mi = std::sin(arg); }
            h[(size_t)r * n + k].re = (R)(mr * inv_n);
            h[(size_t)r * n + k].im = (R)(mi * inv_n);
        }
    }
    void* dev = nullptr;
    SFD_HIP(hipMalloc(&dev, h.size() * sizeof(cx<R>)));
    int rc = upload_sync(dev, h.data(), h.size() * sizeof(cx<R>));
    if (rc != SFD_OK) { (void)hipFree(dev); return rc; }
    if (c->phases.size() >= 32) {  // bounded cache: drop the oldest entry
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(c->phases.front().dev);
        c->phases.erase(c->phases.begin());
    }
    c->phases.push_back({axis, tau, cD, Om, imag, dev, rows});
    *out = dev;
    *rows_out = rows;
    return SFD_OK;
}

int ensure_partial(sfd_ctx* c, long long need)
{
    if (need <= c->partial_cap) return SFD_OK;
    if (c->partial) {
        SFD_HIP(hipStreamSynchronize(c->stream));
        SFD_HIP(hipFree(c->partial));
        c->partial = nullptr;
        c->partial_cap = 0;
    }
    SFD_HIP(hipMalloc(&c->partial, (size_t)need * sizeof(double)));
    c->partial_cap = need;
    return SFD_OK;
}

// Launch geometry of one sweep: returns grid/block and fills the line fields
// of the argument block.
struct Launch {
    dim3 grid, block;
    long long nblocks;
};

template <typename R, int N>
Launch sweep_geometry(const sfd_ctx* c, int axis, SweepArgs& a)
{
    using C = Cfg<R, N>;
    Launch L;
    const long long nx = c->n[0], ny = c->n[1], nz = c->n[2];
    if (axis == 0) {
        constexpr int NL = C::G_contig;
        a.S = 1;
        a.nlines = ny * nz;
        L.grid = dim3((unsigned)((a.nlines + NL - 1) / NL), 1, 1);
        L.block = dim3(NL * C::N2, 1, 1);
    } else {
        constexpr int NL = C::TS * C::G_strided;
        a.ncols = (int)nx;
        if (axis == 1) { a.S = nx; a.OS = nx * ny; }
        else           { a.S = nx * ny; a.OS = nx; }
        const long long others = axis == 1 ? nz : ny;
        L.grid = dim3((unsigned)((nx + NL - 1) / NL), (unsigned)others, 1);
        L.block = dim3(NL * C::N2, 1, 1);
    }
    L.nblocks = (long long)L.grid.x * L.grid.y;
    return L;
}

template <typename R, int N, int MODE>
int launch_n(sfd_ctx* c, int axis, SweepArgs a)
{
    using C = Cfg<R, N>;
    Launch L = sweep_geometry<R, N>(c, axis, a);
    if (axis == 0)
        hipLaunchKernelGGL((k_sweep<R, C::N1, C::N2, 1, C::G_contig, MODE>), L.grid, L.block, 0, c->stream, a);
    else
        hipLaunchKernelGGL((k_sweep<R, C::N1, C::N2, C::TS, C::G_strided, MODE>), L.grid, L.block, 0, c->stream,
                           a);
    SFD_HIP(hipGetLastError());
    return SFD_OK;
}

template <typename R, int MODE>
int launch_sweep(sfd_ctx* c, int axis, const SweepArgs& a)
{
    switch (c->n[axis]) {
        case 2: return launch_n<R, 2, MODE>(c, axis, a);
        case 4: return launch_n<R, 4, MODE>(c, axis, a);
        case 8: return launch_n<R, 8, MODE>(c, axis, a);
        case 16: return launch_n<R, 16, MODE>(c, axis, a);
        case 32: return launch_n<R, 32, MODE>(c, axis, a);
        case 64: return launch_n<R, 64, MODE>(c, axis, a);
        case 128: return launch_n<R, 128, MODE>(c, axis, a);
        case 256: return launch_n<R, 256, MODE>(c, axis, a);
        case 512: return launch_n<R, 512, MODE>(c, axis, a);
        case 1024: return launch_n<R, 1024, MODE>(c, axis, a);
        default: return fail(SFD_ERR_SHAPE, "unsupported line length %d", c->n[axis]);
    }
}

template <typename R, int N, int MODE>
long long nblocks_n(const sfd_ctx* c, int axis)
{
    SweepArgs a{};
    return sweep_geometry<R, N>(c, axis, a).nblocks;
}

template <typename R>
long long sweep_nblocks(const sfd_ctx* c, int axis)
{
    switch (c->n[axis]) {
        case 2: return nblocks_n<R, 2, 0>(c, axis);
        case 4: return nblocks_n<R, 4, 0>(c, axis);
        case 8: return nblocks_n<R, 8, 0>(c, axis);
        case 16: return nblocks_n<R, 16, 0>(c, axis);
        case 32: return nblocks_n<R, 32, 0>(c, axis);
        case 64: return nblocks_n<R, 64, 0>(c, axis);
        case 128: return nblocks_n<R, 128, 0>(c, axis);
        case 256: return nblocks_n<R, 256, 0>(c, axis);
        case 512: return nblocks_n<R, 512, 0>(c, axis);
        case 1024: return nblocks_n<R, 1024, 0>(c, axis);
        default: return 0;
    }
}

template <typename R>
int base_args(sfd_ctx* c, int axis, SweepArgs& a)
{
    std::memset(&a, 0, sizeof a);
    a.tw = c->tw[axis];
    return SFD_OK;
}

// One linear round trip along `axis` over tau, with optional fused flags.
template <typename R>
int round_trip(sfd_ctx* c, void* psi, const void* V, int axis, double tau, double cD, double Om, int imag,
               int flags, double beta, double dt)
{
    SweepArgs a;
    base_args<R>(c, axis, a);
    void* ph = nullptr;
    int rows = 1;
    SFD_TRY(build_phase<R>(c, axis, tau, cD, Om, imag, &ph, &rows));
    a.psi = psi;
    a.V = V;
    a.phase = ph;
    a.phase_row_stride = rows > 1 ? c->n[axis] : 0;
    if (axis == 0) a.row_mod = rows > 1 ? c->n[1] : 0;
    else a.row_from_col = (axis == 1 && rows > 1) ? 1 : 0;
    a.flags = flags;
    a.beta = beta;
    a.dt = dt;
    a.scale = c->scal;
    if (flags & kFlagNorm) {
        SFD_TRY(ensure_partial(c, 5 * sweep_nblocks<R>(c, axis)));
        a.partial = c->partial;
    }
    return launch_sweep<R, kRoundTrip>(c, axis, a);
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

template <typename R>
int norm2_impl(sfd_ctx* c, const void* psi, double* dev_out)
{
    const long long nb = grid_blocks(c->nsite);
    SFD_TRY(ensure_partial(c, nb));
    hipLaunchKernelGGL(k_norm2<R>, dim3((unsigned)nb), dim3(256), 0, c->stream,
                       static_cast<const cx<R>*>(psi), c->nsite, c->partial);
    SFD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, c->stream, c->partial, nb, 1, dev_out, 0.0,
                       (double*)nullptr);
    SFD_HIP(hipGetLastError());
    return SFD_OK;
}

template <typename R>
int split_impl(sfd_ctx* c, void* psi, const void* V, double cD, double beta, double Om, double dt,
               int64_t nsteps, int order, int mode)
{
    const int imag = mode == SFD_IMAG_TIME;
    const int D = c->ndim;
    const double tau = order == 2 ? dt * 0.5 : dt;
    const int nlflags = kFlagNL | (imag ? kFlagImag : 0);
    // Reduction workspace for the per-step renormalisation.
    if (imag) SFD_TRY(ensure_partial(c, 5 * sweep_nblocks<R>(c, D - 1)));
    bool pending = false;
    for (int64_t s = 0; s < nsteps; ++s) {
        for (int h = 0; h < order; ++h) {
            for (int ax = 0; ax < D; ++ax) {
                int flags = 0;
                if (ax == 0 && h == 0 && pending) flags |= kFlagScale;
                if (ax == D - 1 && h == 0) flags |= nlflags;
                if (ax == D - 1 && h == order - 1 && imag) flags |= kFlagNorm;
                SFD_TRY(round_trip<R>(c, psi, V, ax, tau, cD, Om, imag, flags, beta, dt));
            }
        }
        if (imag) {
            const long long nb = sweep_nblocks<R>(c, D - 1);
            hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, c->stream, c->partial, nb, 1, c->scal + 1,
                               c->dvol(), c->scal);
            SFD_HIP(hipGetLastError());
            pending = true;
        }
    }
    if (pending) {
        const long long nb = grid_blocks(c->nsite);
        hipLaunchKernelGGL(k_scale<R>, dim3((unsigned)nb), dim3(256), 0, c->stream, static_cast<cx<R>*>(psi),
                           c->nsite, 0.0, (const double*)c->scal, 0);
        SFD_HIP(hipGetLastError());
    }
    return SFD_OK;
}

template <typename R>
int energy_impl(sfd_ctx* c, const void* psi, const void* V, double cD, double beta, double Om, double* out5)
{
    double sums[3][5];
    for (int ax = 0; ax < c->ndim; ++ax) {
        SweepArgs a;
        base_args<R>(c, ax, a);
        a.psi = const_cast<void*>(psi);
        a.V = V;
        a.xi = c->xi[ax];
        a.flags = ax == 0 ? kFlagPointwise : 0;
        if (ax == 0 && c->ndim >= 2) { a.coord = c->coord[1]; a.coord_mod = c->n[1]; }
        if (ax == 1) a.coord = c->coord[0];
        const long long nb = sweep_nblocks<R>(c, ax);
        SFD_TRY(ensure_partial(c, 5 * nb));
        a.partial = c->partial;
        SFD_TRY(launch_sweep<R, kEnergy>(c, ax, a));
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, c->stream, c->partial, nb, 5, c->scal + 8, 0.0,
                           (double*)nullptr);
        SFD_HIP(hipGetLastError());
        SFD_HIP(hipMemcpyAsync(sums[ax], c->scal + 8, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        SFD_HIP(hipStreamSynchronize(c->stream));
    }
    // Parseval: sum_sites |d_ax psi|^2 = (1/n_ax) sum xi^2 |psi_hat|^2, and
    // sum_sites Re(i conj(psi) y d_x psi) = -(1/n_x) sum_lines y sum xi |psi_hat|^2.
    double kin = 0.0;
    for (int ax = 0; ax < c->ndim; ++ax) kin += sums[ax][0] / (double)c->n[ax];
    double rot = 0.0;
    if (c->ndim >= 2) rot = -sums[0][1] / (double)c->n[0] + sums[1][1] / (double)c->n[1];
    const double dv = c->dvol();
    const double EOm = Om * rot * dv;
    const double ED = (-cD * kin + sums[0][3]) * dv;
    const double Eb = 0.5 * beta * sums[0][4] * dv;
    out5[0] = EOm;
    out5[1] = ED;
    out5[2] = Eb;
    out5[3] = -EOm + ED + Eb;
    out5[4] = -EOm + ED + Eb * 2.0;
    return SFD_OK;
}

template <typename R>
int fft_impl(sfd_ctx* c, const void* in, void* out, int axis, int dir)
{
    SweepArgs a;
    base_args<R>(c, axis, a);
    a.psi = const_cast<void*>(in);
    a.out = out;
    if (dir == SFD_FORWARD) return launch_sweep<R, kForward>(c, axis, a);
    return launch_sweep<R, kInverse>(c, axis, a);
}

template <typename R>
int init_tables(sfd_ctx* c)
{
    for (int ax = 0; ax < c->ndim; ++ax) {
        const int n = c->n[ax];
        std::vector<cx<R>> w;
        fill_twiddles<R>(w, n);
        SFD_HIP(hipMalloc(&c->tw[ax], n * sizeof(cx<R>)));
        SFD_TRY(upload_sync(c->tw[ax], w.data(), n * sizeof(cx<R>)));
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
    if (c->scal) (void)hipFree(c->scal);
}

}  // namespace

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
    if (precision != SFD_F32 && precision != SFD_F64)
        return fail(SFD_ERR_ARG, "precision must be SFD_F32 or SFD_F64 (got %d)", precision);
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
    int rc = precision == SFD_F32 ? init_tables<float>(c) : init_tables<double>(c);
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
    if (ctx->prec == SFD_F32) return split_impl<float>(ctx, psi, V, coeffD, beta, Omega, dt, nsteps, order, mode);
    return split_impl<double>(ctx, psi, V, coeffD, beta, Omega, dt, nsteps, order, mode);
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
    if (ctx->prec == SFD_F32) return round_trip<float>(ctx, psi, nullptr, axis, tau, coeffD, Omega, imag, 0, 0.0, 0.0);
    return round_trip<double>(ctx, psi, nullptr, axis, tau, coeffD, Omega, imag, 0, 0.0, 0.0);
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
        hipLaunchKernelGGL(k_nl<float>, dim3((unsigned)nb), dim3(256), 0, ctx->stream, static_cast<cx<float>*>(psi),
                           static_cast<const float*>(V), ctx->nsite, (float)beta, (float)dt, imag);
    else
        hipLaunchKernelGGL(k_nl<double>, dim3((unsigned)nb), dim3(256), 0, ctx->stream,
                           static_cast<cx<double>*>(psi), static_cast<const double*>(V), ctx->nsite, beta, dt, imag);
    SFD_HIP(hipGetLastError());
    return SFD_OK;
}

int sfd_norm2(sfd_ctx* ctx, const void* psi, double* out_sum)
{
    SFD_TRY(check_ctx(ctx));
    if (!psi || !out_sum) return fail(SFD_ERR_NULL, "sfd_norm2: null argument");
    int rc = ctx->prec == SFD_F32 ? norm2_impl<float>(ctx, psi, ctx->scal + 1)
                                  : norm2_impl<double>(ctx, psi, ctx->scal + 1);
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
    if (ctx->prec == SFD_F32)
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
    if (ctx->prec == SFD_F32) return energy_impl<float>(ctx, psi, V, coeffD, beta, Omega, out5);
    return energy_impl<double>(ctx, psi, V, coeffD, beta, Omega, out5);
}

int sfd_fft_axis(sfd_ctx* ctx, const void* in, void* out, int axis, int direction)
{
    SFD_TRY(check_ctx(ctx));
    if (!in || !out) return fail(SFD_ERR_NULL, "sfd_fft_axis: null argument");
    if (axis < 0 || axis >= ctx->ndim) return fail(SFD_ERR_ARG, "axis %d out of range", axis);
    if (direction != SFD_FORWARD && direction != SFD_BACKWARD) return fail(SFD_ERR_ARG, "direction must be -1 or 1");
    if (ctx->prec == SFD_F32) return fft_impl<float>(ctx, in, out, axis, direction);
    return fft_impl<double>(ctx, in, out, axis, direction);
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
