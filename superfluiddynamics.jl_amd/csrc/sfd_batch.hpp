// sfd_batch.hpp -- device kernels of the batched imaginary-time solver:
// many independent 2D and 3D fields of uneven sizes advanced together, each
// with its own FP64 renormalisation factor, step count, convergence flag and
// (for FP32 storage) escalation to FP64 near its FP32 rounding floor.
#pragma once
#include "sfd_sweep.hpp"

namespace sfd {

struct BatchInstDev {
    void* psi;          // field (storage precision)
    void* snap;         // snapshot of the field at the last check step
    double* partial;    // per-tile sums of |psi|^2 of the last sweep of a step
    long long nsite;
    long long nparts;
    double dv;          // cell volume
};

// Per instance: pending renormalisation 1/||psi|| from the partial sums of
// the step's last sweep; counts the step.  One workgroup per instance,
// fixed summation order.
static __global__ void k_batch_finalize(const BatchInstDev* __restrict__ inst, double* __restrict__ scale,
                                 const int* __restrict__ done, long long* __restrict__ steps)
{
    const int i = blockIdx.x;
    if (done[i]) return;
    const BatchInstDev d = inst[i];
    __shared__ double red[256];
    double s = 0.0;
    for (long long b = threadIdx.x; b < d.nparts; b += blockDim.x) s += d.partial[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        scale[i] = 1.0 / sqrt(red[0] * d.dv);
        steps[i] += 1;
    }
}

// snap <- psi and snap_scale <- scale (the state before a check step).
template <typename S>
__global__ void k_batch_snapshot(const BatchInstDev* __restrict__ inst, const double* __restrict__ scale,
                                 double* __restrict__ snap_scale, const int* __restrict__ done)
{
    const int i = blockIdx.y;
    if (done[i]) return;
    const BatchInstDev d = inst[i];
    const cx<S>* __restrict__ src = static_cast<const cx<S>*>(d.psi);
    cx<S>* __restrict__ dst = static_cast<cx<S>*>(d.snap);
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < d.nsite;
         k += (long long)gridDim.x * blockDim.x)
        dst[k] = src[k];
    if (blockIdx.x == 0 && threadIdx.x == 0) snap_scale[i] = scale[i];
}

// Partial sums of |s1 psi - s0 snap|^2 (normalised states one step apart).
template <typename S>
__global__ void k_batch_diff(const BatchInstDev* __restrict__ inst, const double* __restrict__ scale,
                             const double* __restrict__ snap_scale, const int* __restrict__ done,
                             double* __restrict__ dpart)
{
    const int i = blockIdx.y;
    double acc = 0.0;
    if (!done[i]) {
        const BatchInstDev d = inst[i];
        const double s1 = scale[i], s0 = snap_scale[i];
        const cx<S>* __restrict__ a = static_cast<const cx<S>*>(d.psi);
        const cx<S>* __restrict__ b = static_cast<const cx<S>*>(d.snap);
        for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < d.nsite;
             k += (long long)gridDim.x * blockDim.x) {
            const double re = s1 * (double)a[k].re - s0 * (double)b[k].re;
            const double im = s1 * (double)a[k].im - s0 * (double)b[k].im;
            acc += re * re + im * im;
        }
    }
    double acc1[1] = {acc};
    block_partials<double, 256, 1>(acc1, dpart + (long long)i * gridDim.x, gridDim.x, blockIdx.x);
}

// Thresholds of the escalation from FP32 storage to FP64 (see
// k_batch_diff_finalize): the rounding of FP32 states and arithmetic keeps
// ||Delta psi|| near 1e-7, so an instance goes on in FP64 once its increment
// is below kEscAbs, or below kEscStallMax while shrinking by less than
// kEscRatio per check (stalled at its FP32 floor).
constexpr double kEscAbs = 1e-6;
constexpr double kEscStallMax = 1e-4;
constexpr double kEscRatio = 0.8;

// ||Delta psi|| = sqrt(sum * dv) of the last check step.  Without
// escalation an instance below tol is done.  With escalation (FP32 levels
// of a batch that finishes in FP64) an instance is done at tol only when
// tol >= kEscAbs, and otherwise is flagged for FP64 (esc = 1, done here)
// once its increment is near the FP32 floor: below kEscAbs, or below
// kEscStallMax without shrinking by kEscRatio since the previous check.
static __global__ void k_batch_diff_finalize(const double* __restrict__ dpart, int nb,
                                             const BatchInstDev* __restrict__ inst, double* __restrict__ dpsi,
                                             double* __restrict__ dprev, int* __restrict__ done, double tol,
                                             int* __restrict__ esc, int escalate)
{
    const int i = blockIdx.x;
    if (done[i]) return;
    __shared__ double red[256];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += dpart[(long long)i * nb + b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double dn = sqrt(red[0] * inst[i].dv);
        const double prev = dprev[i];
        dpsi[i] = dn;
        dprev[i] = dn;
        if (!escalate) {
            if (dn < tol) done[i] = 1;
        } else if (dn < tol && tol >= kEscAbs) {
            done[i] = 1;
        } else if (dn < kEscAbs || (dn < kEscStallMax && prev >= 0.0 && dn > kEscRatio * prev)) {
            esc[i] = 1;
            done[i] = 1;
        }
    }
}

// Per-workgroup sums of psi^2 (re^2 - im^2, 2 re im) into
// sq[(2 i + k) KB + b], for every instance (esc null) or for those flagged
// for escalation at this check (esc = 1).
template <typename S>
__global__ void k_batch_sq_partials(const BatchInstDev* __restrict__ lo, const int* __restrict__ esc,
                                    double* __restrict__ sq)
{
    const int i = blockIdx.y;
    double acc[2] = {0.0, 0.0};
    if (!esc || esc[i] == 1) {
        const cx<S>* __restrict__ src = static_cast<const cx<S>*>(lo[i].psi);
        for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < lo[i].nsite;
             k += (long long)gridDim.x * blockDim.x) {
            const double re = src[k].re, im = src[k].im;
            acc[0] += re * re - im * im;
            acc[1] += 2.0 * re * im;
        }
    }
    block_partials<double, 256, 2>(acc, sq + (long long)i * 2 * gridDim.x, gridDim.x, blockIdx.x);
}

// The dominant phase arg(sum psi^2)/2 of instance i from its KB = nsq
// partials, taken within pi/2 of `near` (the phase is defined modulo pi).
__device__ __forceinline__ double dominant_phase(const double* __restrict__ sq, int i, int nsq, double near)
{
    double a = 0.0, b = 0.0;  // every workgroup sums the partials in the same order
    for (int k = 0; k < nsq; ++k) {
        a += sq[((long long)i * 2 + 0) * nsq + k];
        b += sq[((long long)i * 2 + 1) * nsq + k];
    }
    double phi = 0.5 * atan2(b, a);
    const double pi = 3.141592653589793;
    while (phi - near > 0.5 * pi) phi -= pi;
    while (phi - near < -0.5 * pi) phi += pi;
    return phi;
}

// Start of a solve: phi0[i], the dominant phase of each initial state.
static __global__ void k_batch_phase0(const double* __restrict__ sq, int nsq, double* __restrict__ phi0, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) phi0[i] = dominant_phase(sq, i, nsq, 0.0);
}

// Escalation, part 2: the FP64 field of every flagged instance becomes
// scale * its FP32-level field, widened exactly.  For an instance with a
// real Hamiltonian (real[i]: no rotation), whose ground state is real up to
// a global phase, the state is also put back on the dominant phase phi0 of
// its initial state: psi <- e^{i phi0} Re(e^{-i phi} psi), phi its current
// dominant phase.  The imaginary-time flow keeps e^{i phi0} (real state)
// so, and only the rounding of the FP32 steps moved the state off it: a
// random walk of the global phase, and phase-gradient components that decay
// at the slow rate ~1/(2<x^2>) and would hold the FP64 stage for thousands
// of steps.
template <typename S>
__global__ void k_batch_escalate(const BatchInstDev* __restrict__ lo, const BatchInstDev* __restrict__ hi,
                                 const double* __restrict__ scale, const int* __restrict__ esc,
                                 const int* __restrict__ real, const double* __restrict__ sq, int nsq,
                                 const double* __restrict__ phi0)
{
    const int i = blockIdx.y;
    if (esc[i] != 1) return;
    const bool proj = real[i] != 0;
    double c = 1.0, sn = 0.0, c0 = 1.0, s0 = 0.0;
    if (proj) {
        const double phi = dominant_phase(sq, i, nsq, phi0[i]);
        c = cos(phi);
        sn = sin(phi);
        c0 = cos(phi0[i]);
        s0 = sin(phi0[i]);
    }
    const double s = scale[i];
    const cx<S>* __restrict__ src = static_cast<const cx<S>*>(lo[i].psi);
    cx<double>* __restrict__ dst = static_cast<cx<double>*>(hi[i].psi);
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < lo[i].nsite;
         k += (long long)gridDim.x * blockDim.x) {
        const double re = s * (double)src[k].re, im = s * (double)src[k].im;
        if (proj) {
            const double t = re * c + im * sn;
            dst[k] = {t * c0, t * s0};
        } else {
            dst[k] = {re, im};
        }
    }
}

// Escalation, part 2 (after part 1 on the same stream): the flagged
// instances continue at the FP64 level with no pending scale.
static __global__ void k_batch_escalate_flags(double* __restrict__ scale, int* __restrict__ esc,
                                              int* __restrict__ done_hi, double* __restrict__ dprev, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || esc[i] != 1) return;
    esc[i] = 2;
    scale[i] = 1.0;
    dprev[i] = -1.0;
    done_hi[i] = 0;
}

// End of a solve: the final state of instance i (its FP64 field when it was
// escalated, esc = 2, else its field) times the pending scale, rounded once
// into the caller's field, and copied in FP64 to out64[i] when non-null.
template <typename S>
__global__ void k_batch_finish(const BatchInstDev* __restrict__ lo, const BatchInstDev* __restrict__ hi,
                               const double* __restrict__ scale, const int* __restrict__ esc,
                               void* const* __restrict__ out64)
{
    const int i = blockIdx.y;
    const double s = scale[i];
    cx<S>* __restrict__ p = static_cast<cx<S>*>(lo[i].psi);
    cx<double>* __restrict__ o = out64 ? static_cast<cx<double>*>(out64[i]) : nullptr;
    const bool up = hi && esc[i] == 2;
    const cx<double>* __restrict__ h = up ? static_cast<const cx<double>*>(hi[i].psi) : nullptr;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < lo[i].nsite;
         k += (long long)gridDim.x * blockDim.x) {
        const cx<double> v = up ? cx<double>{s * h[k].re, s * h[k].im}
                                : cx<double>{s * (double)p[k].re, s * (double)p[k].im};
        p[k] = {S(v.re), S(v.im)};
        if (o) o[k] = v;
    }
}

// Start of a solve: scale = 1, done = 0 at the batch's level and 1 at the
// FP64 level, steps = 0, no escalation, no previous increment.  A kernel,
// not memsets, so that the sequence could be captured in a graph.
static __global__ void k_batch_reset(double* scale, int* done_lo, int* done_hi, long long* steps, double* dpsi,
                                     double* dprev, int* esc, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    scale[i] = 1.0;
    done_lo[i] = 0;
    if (done_hi) done_hi[i] = 1;
    steps[i] = 0;
    dpsi[i] = -1.0;
    dprev[i] = -1.0;
    esc[i] = 0;
}

}  // namespace sfd
