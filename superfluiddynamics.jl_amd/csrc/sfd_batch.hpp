// sfd_batch.hpp -- device kernels of the batched imaginary-time solver:
// many independent 2D fields of uneven sizes advanced together, each with its
// own FP64 renormalisation factor, step count and convergence flag.
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

// ||Delta psi|| = sqrt(sum * dv); an instance below tol is done.
static __global__ void k_batch_diff_finalize(const double* __restrict__ dpart, int nb,
                                      const BatchInstDev* __restrict__ inst, double* __restrict__ dpsi,
                                      int* __restrict__ done, double tol)
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
        dpsi[i] = dn;
        if (dn < tol) done[i] = 1;
    }
}

// psi *= scale (the pending renormalisation), in FP64, rounded once.
template <typename S>
__global__ void k_batch_apply_scale(const BatchInstDev* __restrict__ inst, const double* __restrict__ scale)
{
    const int i = blockIdx.y;
    const BatchInstDev d = inst[i];
    const double s = scale[i];
    cx<S>* __restrict__ p = static_cast<cx<S>*>(d.psi);
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < d.nsite;
         k += (long long)gridDim.x * blockDim.x)
        p[k] = {S(s * (double)p[k].re), S(s * (double)p[k].im)};
}

// scale[i] = 1, done[i] = 0, steps[i] = 0 (start of a solve); kernel, not a
// memset, so that the sequence can be captured in a graph.
static __global__ void k_batch_reset(double* scale, int* done, long long* steps, double* dpsi, int n, int full)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    scale[i] = 1.0;
    if (full) {
        done[i] = 0;
        steps[i] = 0;
        dpsi[i] = -1.0;
    }
}

}  // namespace sfd
