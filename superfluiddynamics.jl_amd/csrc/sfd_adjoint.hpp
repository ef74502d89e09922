// sfd_adjoint.hpp -- device kernels of the reverse (adjoint) split step.
//
// Gradient convention: for a real loss L and a complex field psi = u + i v,
// G = dL/du + i dL/dv, so that dL = Re(conj(G) dpsi).  The Strang step
// psi1 = K psi_n, psi2 = exp(-i th) psi1, th = (V + beta|psi1|^2) dt,
// psi_{n+1} = K psi2 (splitting.jl:149-209) is reversed as
//   G <- K^dagger G          (round trips over -tau, axes in reverse order)
//   w = exp(+i th) G, q = Im(conj(w) psi1)
//   G <- w + 2 beta dt q psi1,  dL/dV += q dt,  dL/dbeta += q dt |psi1|^2
//   G <- K^dagger G.
#pragma once
#include "sfd_sweep.hpp"

namespace sfd {

// Reverse of the nonlinear phase at every site; dL/dV accumulates into
// gradV (FP64), dL/dbeta into per-workgroup partials that persist over the
// reverse sweep (each workgroup adds to its own slot).
template <typename S, typename R>
__global__ void k_nl_adjoint(cx<S>* __restrict__ G, const cx<S>* __restrict__ psi1, const S* __restrict__ V,
                             long long n, R beta, R dt, double* __restrict__ gradV, double* __restrict__ partial)
{
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const cx<R> p = widen<R>(psi1[i]);
        const cx<R> g = widen<R>(G[i]);
        const R a2 = p.re * p.re + p.im * p.im;
        const R th = ((V ? R(V[i]) : R(0)) + beta * a2) * dt;
        R s, c;
        sincos_cw(th, &s, &c);
        const cx<R> w = {c * g.re - s * g.im, c * g.im + s * g.re};
        const R q = w.re * p.im - w.im * p.re;
        const R f = R(2) * beta * dt * q;
        G[i] = narrow<S>(cx<R>{w.re + f * p.re, w.im + f * p.im});
        if (gradV) gradV[i] += (double)q * (double)dt;
        acc += (double)q * (double)dt * (double)a2;
    }
    double acc1[1] = {acc};
    block_partials<R, 256, 1, true>(acc1, partial, gridDim.x, blockIdx.x);
}

// Partial sums of <a|b> = sum conj(a) b (re and im, FP64).
template <typename S>
__global__ void k_inner(const cx<S>* __restrict__ a, const cx<S>* __restrict__ b, long long n,
                        double* __restrict__ partial)
{
    double acc[2] = {0.0, 0.0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const double ar = a[i].re, ai = a[i].im, br = b[i].re, bi = b[i].im;
        acc[0] += ar * br + ai * bi;
        acc[1] += ar * bi - ai * br;
    }
    block_partials<double, 256, 2>(acc, partial, gridDim.x, blockIdx.x);
}

// Terminal gradient of L = 1 - |c|^2, c = <target|psi> = dv sum conj(target) psi:
// G = -2 c dv target, with cdev = sum conj(target) psi (re, im).
template <typename S>
__global__ void k_terminal(cx<S>* __restrict__ G, const cx<S>* __restrict__ target, long long n,
                           const double* __restrict__ cdev, double dv)
{
    const double cr = cdev[0] * dv, ci = cdev[1] * dv;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const double tr = target[i].re, ti = target[i].im;
        G[i] = {S(-2.0 * dv * (cr * tr - ci * ti)), S(-2.0 * dv * (cr * ti + ci * tr))};
    }
}

// acc[0] += sum of nb partials (fixed order).
static __global__ void k_accumulate(const double* __restrict__ partial, long long nb, double* acc)
{
    __shared__ double red[256];
    double s = 0.0;
    for (long long b = threadIdx.x; b < nb; b += blockDim.x) s += partial[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] += red[0];
}

// out[j] partials of sum_x f[x] basis[j*n + x]: blockIdx.y = j.
static __global__ void k_project(const double* __restrict__ f, const double* __restrict__ basis, long long n,
                          double* __restrict__ partial)
{
    const double* b = basis + (long long)blockIdx.y * n;
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        acc += f[i] * b[i];
    double acc1[1] = {acc};
    block_partials<double, 256, 1>(acc1, partial + (long long)blockIdx.y * gridDim.x, gridDim.x, blockIdx.x);
}

}  // namespace sfd
