// sfd_impl.hpp -- the precision-dependent routines of libsfdhip.so, templated
// on the precision traits P (PF32, PF64, PF32X of sfd_variants.hpp):
// phase tables, sweep launches, split steps, energies, directional FFTs, the
// adjoint, the autotuners and the batched imaginary-time solver.  Each of
// sfd_f32.hip, sfd_f64.hip and sfd_f32x.hip instantiates them for one
// precision through make_ops<P>().
#pragma once
#include "sfd_runtime.hpp"

namespace sfd {

// Twiddles in the layout k_sweep reads for a line of n = h*m points split
// as m = n1*n2 (h = 1: m = n), with w_m = exp(-2 pi i / m):
//   [0, m):    twA[s][t] = w_m^(s t), s < n1, t < n2
//   [m, 2m):   twB[q][s] = w_m^(q s), q < n2, s < n1
//   [2m, 3m):  h = 2 only, the radix-2 stage's twC[k] = w_n^k, k < m
// FP64 entries are (cos, sin) of the angle a; FP32 entries are the lifting
// coefficients (tan(b/2), sin b) of b = a/2 with a taken in (-pi, pi].
// With `centered` (FP32, h = 1, n1, n2 >= 2: the centered round trips of
// the sweep kernels), [2m, 4m) repeats twA and twB for the centered indices
// t_c = t - n2/2, q_c = q - n2/2, s_c = s - n1 [s >= n1/2], as the lifting
// coefficients (tan(a/2), sin a) of the angle a itself, |a| <= pi/2.
template <typename R>
void fill_twiddles(std::vector<cx<R>>& w, int n, int n1, int n2, int h, bool centered)
{
    const int m = n1 * n2;
    const bool ct = centered && sizeof(R) == 4 && h == 1 && n1 >= 2 && n2 >= 2;
    w.resize((size_t)(ct ? 4 : (h == 2 ? 3 : 2)) * m);
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
    if (ct) {
        // exp(-2 pi i e / m) for a signed e with |e| <= m/4
        auto put1 = [&](size_t i, long long e) {
            const long double a = -2.0L * pi * (long double)e / (long double)m;
            w[i].re = (R)tanl(0.5L * a);
            w[i].im = (R)sinl(a);
        };
        auto sc = [&](int s) { return s >= n1 / 2 ? s - n1 : s; };
        for (int s = 0; s < n1; ++s)
            for (int t = 0; t < n2; ++t) put1((size_t)2 * m + (size_t)s * n2 + t, (long long)(t - n2 / 2) * sc(s));
        for (int q = 0; q < n2; ++q)
            for (int s = 0; s < n1; ++s)
                put1((size_t)3 * m + (size_t)q * n1 + s, (long long)sc(s) * (q - n2 / 2));
    }
}

// Bluestein tables of a line of n points on m >= 2n - 1 (appended to the
// m-point twiddles): chirp[p] = exp(-i pi p^2/n) for p < n, zero up to m,
// then B = DFT_m(g)/m of the wrapped conjugate chirp g[q] = conj(chirp[q])
// (q < n), g[m - r] = conj(chirp[r]) (0 < r < n).  Angles from p^2 mod 2n
// (exact integers) in long double; B by a direct sum in long double.
template <typename R>
void fill_bluestein(std::vector<cx<R>>& w, int n, int m)
{
    const long double pi = 3.14159265358979323846264338327950288L;
    std::vector<long double> cr(n), ci(n);
    for (int p = 0; p < n; ++p) {
        const long long q = ((long long)p * p) % (2LL * n);
        const long double a = -pi * (long double)q / (long double)n;
        cr[p] = cosl(a);
        ci[p] = sinl(a);
    }
    const size_t base = w.size();
    w.resize(base + 2 * (size_t)m);
    for (int p = 0; p < m; ++p) w[base + p] = p < n ? cx<R>{(R)cr[p], (R)ci[p]} : cx<R>{(R)0, (R)0};
    std::vector<long double> tr(m), ti(m);  // exp(-2 pi i k/m)
    for (int k = 0; k < m; ++k) {
        tr[k] = cosl(-2.0L * pi * (long double)k / (long double)m);
        ti[k] = sinl(-2.0L * pi * (long double)k / (long double)m);
    }
    std::vector<int> idx;
    std::vector<long double> gr, gi;
    for (int q = 0; q < n; ++q) { idx.push_back(q); gr.push_back(cr[q]); gi.push_back(-ci[q]); }
    for (int r = 1; r < n; ++r) { idx.push_back(m - r); gr.push_back(cr[r]); gi.push_back(-ci[r]); }
    for (int k = 0; k < m; ++k) {
        long double sr = 0.0L, si = 0.0L;
        for (size_t e = 0; e < idx.size(); ++e) {
            const int t = (int)(((long long)k * idx[e]) % m);
            sr += gr[e] * tr[t] - gi[e] * ti[t];
            si += gr[e] * ti[t] + gi[e] * tr[t];
        }
        w[base + m + k] = {(R)(sr / (long double)m), (R)(si / (long double)m)};
    }
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
// s = i (real time) or 1 (imaginary time).  Formed on the device in FP64
// with the reference's operation order from the context's own FP64
// wavenumber and coordinate tables, rounded once to the table precision R:
//   FP64 tables, real time:  (cos arg, sin arg) / n
//   FP32 tables, real time:  lifting coefficients (tan(b/2), sin b) of the
//                            half phase b = phi/2, phi = arg reduced to
//                            (-pi, pi]; the sweep applies them twice
//   imaginary time:          exp(arg)/n (FP64), expm1(arg)/n (FP32: the
//                            sweep forms X/n + X expm1(arg)/n)
template <typename R>
__global__ void k_build_phase(cx<R>* __restrict__ out, const double* __restrict__ xi, const double* __restrict__ co,
                              int n, int rows, int axis, double tau, double cD, double Om, int imag)
{
    const long long tot = (long long)rows * n;
    const double inv_n = 1.0 / (double)n;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
         i += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(i / n), k = (int)(i - (long long)r * n);
        const double x = xi[k];
        const double c = rows > 1 ? co[r] : 0.0;
        double arg;
        if (axis == 0) arg = (cD * (x * x) - (Om * c) * x) * tau;
        else if (axis == 1) arg = (cD * (x * x) + (Om * c) * x) * tau;
        else arg = (cD * (x * x)) * tau;
        cx<R> e;
        if (imag) {
            if constexpr (sizeof(R) == 4) e = {(R)(expm1(arg) * inv_n), (R)0};
            else e = {(R)(exp(arg) * inv_n), (R)0};
        } else if constexpr (sizeof(R) == 4) {
            const double b = 0.5 * remainder(arg, 6.283185307179586476925);
            e = {(R)tan(0.5 * b), (R)sin(b)};
        } else {
            double sn, cs;
            sincos(arg, &sn, &cs);
            e = {(R)(cs * inv_n), (R)(sn * inv_n)};
        }
        out[i] = e;
    }
}

// Enqueue the build of a phase table into `dev` (rows * n[axis] entries of
// cx<R>) on `stream`.
template <typename P>
int launch_build_phase(const sfd_ctx* c, hipStream_t stream, void* dev, int axis, double tau, double cD, double Om,
                       int imag, int rows)
{
    using R = typename P::R;
    const int n = c->n[axis];
    const long long tot = (long long)rows * n;
    const long long nb = std::min<long long>((tot + 255) / 256, 1024);
    hipLaunchKernelGGL(k_build_phase<R>, dim3((unsigned)nb), dim3(256), 0, stream, static_cast<cx<R>*>(dev),
                       (const double*)c->xi[axis], (const double*)c->coord[axis == 0 ? 1 : 0], n, rows, axis, tau,
                       cD, Om, imag);
    SFD_HIP(hipGetLastError());
    return SFD_OK;
}

// Phase table of (axis, tau, coeffD, Omega, mode) from the context's cache
// (least recently used first out, at most 32 entries unless pinned), built
// on the context's stream on a miss.  A full cache rebuilds its least
// recently used entry in place: the sweeps that read the old table were
// enqueued earlier on the same stream, so the build kernel runs after them
// and no host synchronisation or allocation is needed.
template <typename P>
int build_phase(sfd_ctx* c, int axis, double tau, double cD, double Om, int imag, void** out, int* rows_out)
{
    using R = typename P::R;
    // a hit moves the entry to the back
    for (size_t i = 0; i < c->phases.size(); ++i) {
        const PhaseEntry p = c->phases[i];
        if (p.axis == axis && p.tau == tau && p.cD == cD && p.Om == Om && p.imag == imag) {
            c->phases.erase(c->phases.begin() + i);
            c->phases.push_back(p);
            *out = p.dev;
            *rows_out = p.rows;
            return SFD_OK;
        }
    }
    const int rows = phase_rows(c, axis, Om);
    const size_t bytes = (size_t)rows * c->n[axis] * sizeof(cx<R>);
    void* dev = nullptr;
    size_t cap = bytes;
    if (c->phases.size() >= 32 && !c->pin_phases) {
        PhaseEntry lru = c->phases.front();
        c->phases.erase(c->phases.begin());
        if (lru.bytes >= bytes) {
            dev = lru.dev;
            cap = lru.bytes;
        } else {  // a larger table (rotation rows): the old buffer may still be read
            SFD_HIP(hipStreamSynchronize(c->stream));
            (void)hipFree(lru.dev);
        }
    }
    if (!dev) SFD_HIP(hipMalloc(&dev, bytes));
    int rc = launch_build_phase<P>(c, c->stream, dev, axis, tau, cD, Om, imag, rows);
    if (rc != SFD_OK) {
        SFD_HIP(hipStreamSynchronize(c->stream));
        (void)hipFree(dev);
        return rc;
    }
    c->phases.push_back({axis, tau, cD, Om, imag, dev, rows, cap});
    *out = dev;
    *rows_out = rows;
    return SFD_OK;
}

template <typename P, int MODE>
int launch_sweep(sfd_ctx* c, int axis, SweepArgs a)
{
    // The pipelined kernel of a pipelined variant takes round trips with one
    // phase row, without a pending scale or energy sums, and with a
    // nonlinear epilogue only at V = 0 that stores no pre-phase state: its
    // loop issues no global memory operation but the tiles' loads and stores
    // (it waits for the prefetched tile by counting them).  Other sweeps run
    // k_sweep.
    const bool persistent = c->vi[axis].pipe && (MODE == kRoundTrip || MODE == kRoundTripImag) &&
                            !(a.flags & kFlagScale) && !((a.flags & kFlagNL) && (a.V || a.save)) &&
                            a.phase_row_stride == 0;
    const bool blu = c->blu[axis];
    constexpr bool twice = rt_twice(MODE);  // the double round trips' own variant
    const int var = twice ? c->var2[axis] : c->var[axis];
    Launch L = sweep_geometry(c, axis, a, persistent && !blu, twice);
    a.tw = twice ? c->tw2[axis] : c->tw[axis];
    a.nlen = c->n[axis];
    if (c->time_sweeps)
        SFD_TRY(timer_begin(c, rt_twice(MODE) ? (axis == 0 ? 5 : 3) : (is_round_trip(MODE) ? axis : 4)));
    bool ok = false;
    if constexpr (rt_energy(MODE)) {
        if (blu) return fail(SFD_ERR_SHAPE, "no energy sums fused into Bluestein sweeps (n = %d)", c->n[axis]);
        ok = launch_variant<P, MODE>(c->n[axis], axis == 0, var, L.grid, L.block, c->stream, a, persistent);
    } else {
        ok = blu ? launch_blu<P, MODE>(c->m[axis], axis == 0, L.grid, L.block, c->stream, a)
                 : launch_variant<P, MODE>(c->n[axis], axis == 0, var, L.grid, L.block, c->stream, a, persistent);
    }
    if (!ok) return fail(SFD_ERR_SHAPE, "no sweep variant %d for line length %d", var, c->n[axis]);
    SFD_HIP(hipGetLastError());
    if (c->time_sweeps) SFD_TRY(timer_end(c));
    c->last_nblocks = L.nblocks;
    return SFD_OK;
}

// Select launch variant `idx` for `axis` and (re)build the twiddle table of
// its line split.
template <typename P>
int axis_variants(const sfd_ctx* c, int axis)
{
    return c->blu[axis] ? 1 : n_variants<P>(c->n[axis], axis == 0);
}

// `which`: 1 the axis's single round trips and other sweeps, 2 its double
// round trips ([d N d], [x x]), 3 both.  Each keeps its own twiddle table.
template <typename P>
int set_variant(sfd_ctx* c, int axis, int idx, int which)
{
    using R = typename P::R;
    const int n = c->m[axis];  // the transform length (n itself unless Bluestein)
    VariantInfo vi;
    if ((c->blu[axis] && idx != 0) || !variant_info<P>(n, axis == 0, idx, &vi))
        return fail(SFD_ERR_ARG, "axis %d: variant %d does not exist (%d variants)", axis, idx,
                    axis_variants<P>(c, axis));
    if (which < 1 || which > 3) return fail(SFD_ERR_ARG, "set_variant: which = %d", which);
    if (which & 1) c->pipe_occ[axis] = (vi.pipe && !c->blu[axis]) ? variant_occupancy<P>(n, axis == 0, idx) : 0;
    if (c->blu[axis]) vi.pipe = 0;
    for (int k = 0; k < 2; ++k) {
        if (!(which & (1 << k))) continue;
        void*& tw = k ? c->tw2[axis] : c->tw[axis];
        VariantInfo& cur = k ? c->vi2[axis] : c->vi[axis];
        if (!(tw && cur.n1 == vi.n1 && cur.n2 == vi.n2 && cur.h == vi.h)) {
            std::vector<cx<R>> w;
            fill_twiddles<R>(w, n, vi.n1, vi.n2, vi.h, !c->blu[axis]);
            if (c->blu[axis]) fill_bluestein<R>(w, c->n[axis], n);
            if (tw) {
                SFD_HIP(hipStreamSynchronize(c->stream));
                SFD_HIP(hipFree(tw));
                tw = nullptr;
            }
            SFD_HIP(hipMalloc(&tw, w.size() * sizeof(cx<R>)));
            SFD_TRY(upload_sync(tw, w.data(), w.size() * sizeof(cx<R>)));
        }
        (k ? c->var2[axis] : c->var[axis]) = idx;
        cur = vi;
    }
    return SFD_OK;
}

// One linear round trip along `axis` over tau, with optional fused flags.
// With `epartial` (room for 4 * sweep_nblocks doubles), the sweep also
// writes the energy sums of the state it loads there, 4 per workgroup
// (kRoundTripEnergy; kFlagPointwise in `flags` adds the pointwise sums).
// With `twice`, the sweep is K(tau) N K(tau) on this axis (kRoundTripTwice):
// the nonlinear phase of `flags` between two round trips.  With `ws`, the
// phase table, scale and partials come from it (its tables are for its tau
// and, where built, for dt).
template <typename P>
int round_trip(sfd_ctx* c, void* psi, const void* V, int axis, double tau, double cD, double Om, int imag,
               int flags, double beta, double dt, void* save = nullptr, double* epartial = nullptr,
               bool twice = false, const Workspace* ws = nullptr)
{
    using R = typename P::R;
    SweepArgs a;
    std::memset(&a, 0, sizeof a);
    void* ph = nullptr;
    int rows = 1;
    if (ws) {
        ph = tau == ws->tau ? ws->phase[axis] : ws->phase_dt[axis];
        if (!ph) return fail(SFD_ERR_ARG, "workspace: no phase table for tau = %g on axis %d", tau, axis);
        rows = ws->rows[axis];
    } else {
        SFD_TRY(build_phase<P>(c, axis, tau, cD, Om, imag, &ph, &rows));
    }
    a.psi = psi;
    a.V = V;
    a.phase = ph;
    a.phase_row_stride = rows > 1 ? c->n[axis] : 0;
    if (axis == 0) a.row_mod = rows > 1 ? c->n[1] : 0;
    else a.row_from_col = (axis == 1 && rows > 1) ? 1 : 0;
    a.flags = flags | (imag ? kFlagImag : 0);
    a.beta = beta;
    a.dt = dt;
    a.scale = ws ? ws->scal : c->scal;
    a.save = save;
    if (flags & (kFlagNorm | kFlagNormMid)) {
        if (ws) {
            a.partial = ws->partial;
        } else {
            SFD_TRY(ensure_partial(c, 5 * sweep_nblocks(c, axis)));
            a.partial = c->partial;
        }
    }
    if (twice) {
        if (epartial) {
            if (imag) return fail(SFD_ERR_ARG, "no energy sums in an imaginary-time double round trip");
            a.epartial = epartial;
            a.xi = c->xi[axis];
            // x: the two round trips of chained steps, the energy of the
            // state between them; innermost axis: [d N d], the energy of the
            // state it loads
            if (axis == 0) return launch_sweep<P, kRoundTripTwiceEnergyX>(c, axis, a);
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

template <typename P>
int norm2_impl(sfd_ctx* c, const void* psi, double* dev_out)
{
    using S = typename P::S;
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
// per step instead of 2D).  `fold = kFoldNone` keeps the x, y, z, N, x, y, z
// order of splitting.jl:149-154.
//
// With `fold = kFoldChain`, Omega = 0 and several steps, every flow commutes
// with every other.  In real time the half-step flows that meet between two
// nonlinear phases are one flow of dt per axis, K(tau) N K(dt) N ... N
// K(tau), with the nonlinear phase in the epilogue of the x round trip: D
// sweeps per step, each a single round trip.  Where the state
// between two steps is needed (the norm in imaginary time, the energy sums
// of solve!), each step runs its second half in reverse, x, y, [z N z], y,
// x, and the x flow ending step s and the one starting step s + 1 meet in
// one double round trip [x x] (two round trips of tau each, in one read and
// one write of the field, the state between them in registers): x, y,
// [z N z], y, [x x], y, [z N z], ..., y, x, i.e. 2D - 2 sweeps per step
// instead of 2D - 1 (2D: x, [y N y], [x x], [y N y], ..., x).
int split_impl(sfd_ctx* c, void* psi, const void* V, double cD, double beta, double Om, double dt,
               int64_t nsteps, int order, int mode, double* ehist = nullptr, int fold = kFoldInner,
               const Workspace* ws = nullptr)
{
    using S = typename P::S;
    const int imag = mode == SFD_IMAG_TIME;
    const int D = c->ndim;
    const double tau = order == 2 ? dt * 0.5 : dt;
    const int nlflags = kFlagNL | (imag ? kFlagImag : 0);
    // Reduction workspace for the per-step renormalisation (its sweep is on
    // axis D - 2 when folded, on the last axis when not, on x when chained).
    if (imag && !ws) {
        long long nb = 0;
        for (int ax = 0; ax < D; ++ax) nb = std::max(nb, sweep_nblocks(c, ax));
        SFD_TRY(ensure_partial(c, 5 * nb));
    }
    double* scal = ws ? ws->scal : c->scal;
    double* npart = ws ? ws->partial : c->partial;
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
    if (!(order == 2 && D >= 2 && (Om == 0.0 || D == 3))) fold = kFoldNone;
    if (fold == kFoldChain && Om == 0.0 && nsteps >= 2 && !imag && !ehist) {
        // Real time without energy sums: the kinetic flow between two
        // nonlinear phases is K(dt) = x(dt) y(dt) z(dt), one round trip per
        // axis in any order, and the phase is the epilogue of the x round
        // trip (the contiguous sweep, which hides its cost best): y(tau),
        // z(tau), x(tau) + N, then y(dt), z(dt), x(dt) + N per further step,
        // then y(tau), z(tau), x(tau).
        for (int64_t s = 0; s <= nsteps; ++s) {
            const double t = (s == 0 || s == nsteps) ? tau : dt;
            for (int k = 1; k <= D; ++k) {
                const int ax = k % D;  // y, z, then x with the nonlinear phase
                SFD_TRY(round_trip<P>(c, psi, V, ax, t, cD, Om, imag, (s < nsteps && ax == 0) ? nlflags : 0,
                                      beta, dt, nullptr, nullptr, false, ws));
            }
        }
        nsteps = 0;
    }
    if (fold == kFoldChain && nsteps) {
        if (Om == 0.0 && nsteps >= 2 && !c->blu[0] && !(imag && ehist)) {
            // With `ehist`, the energy of the state before step s comes out of
            // the sweeps that first transform it: x from the opening x round
            // trip (s = 0) or from the second round trip of [x x] (the state
            // between steps s - 1 and s, with the pointwise sums), y from the
            // y round trip that follows, z from the first round trip of
            // [z N z]; one finalize per step.
            // In imaginary time [x x] takes the norm partials of the state
            // between its round trips (the end of step s, kFlagNormMid); the
            // renormalisation they give is applied by the next sweep (a
            // scalar commutes with the x flow of step s + 1), before the next
            // nonlinear phase.
            FinalizeSets fs{};
            auto record = [&](int ax, int64_t s) {
                fs.partial[ax] = epart[ax];
                fs.nb[ax] = c->last_nblocks;
                fs.K[ax] = ax == 0 ? 4 : 1;
                fs.out[ax] = ehist + 8 * s + (ax == 0 ? 0 : 3 + ax);
            };
            auto renorm = [&]() -> int {
                launch_finalize(c->stream, npart, c->last_nblocks, 1, scal + 1, c->dvol(), scal);
                SFD_HIP(hipGetLastError());
                pending = true;
                return SFD_OK;
            };
            const int pw = ehist ? kFlagPointwise : 0;
            // first sweep after a renormalisation applies it
            auto take_scale = [&]() -> int {
                const int f = pending ? kFlagScale : 0;
                pending = false;
                return f;
            };
            // x, y (first half of step 0)
            for (int ax = 0; ax < D - 1; ++ax) {
                SFD_TRY(round_trip<P>(c, psi, V, ax, tau, cD, Om, imag, ax == 0 ? pw : 0, beta, dt, nullptr,
                                      epart[ax], false, ws));
                if (ehist) record(ax, 0);
            }
            for (int64_t s = 0; s < nsteps; ++s) {
                SFD_TRY(round_trip<P>(c, psi, V, D - 1, tau, cD, Om, imag, nlflags | (D == 2 ? take_scale() : 0),
                                      beta, dt, nullptr, epart[D - 1], true, ws));
                if (ehist) {
                    record(D - 1, s);
                    launch_finalize_sets(c->stream, fs);
                    SFD_HIP(hipGetLastError());
                }
                // second half in reverse down to y
                for (int ax = D - 2; ax >= 1; --ax)
                    SFD_TRY(round_trip<P>(c, psi, V, ax, tau, cD, Om, imag, 0, beta, dt, nullptr, nullptr, false, ws));
                if (s + 1 < nsteps) {
                    // x of step s and x of step s + 1: no nonlinear phase (no
                    // potential, unless the energy sums need it)
                    SFD_TRY(round_trip<P>(c, psi, ehist ? V : nullptr, 0, tau, cD, Om, imag,
                                          pw | (imag ? kFlagNormMid : 0), beta, dt, nullptr, epart[0], true, ws));
                    if (ehist) record(0, s + 1);
                    if (imag) SFD_TRY(renorm());
                    for (int ax = 1; ax < D - 1; ++ax) {
                        SFD_TRY(round_trip<P>(c, psi, V, ax, tau, cD, Om, imag, ax == 1 ? take_scale() : 0, beta, dt,
                                              nullptr, epart[ax], false, ws));
                        if (ehist) record(ax, s + 1);
                    }
                } else {
                    SFD_TRY(round_trip<P>(c, psi, V, 0, tau, cD, Om, imag, imag ? kFlagNorm : 0, beta, dt, nullptr,
                                          nullptr, false, ws));
                    if (imag) SFD_TRY(renorm());
                }
            }
            nsteps = 0;  // done: only the closing renormalisation below remains
        } else {
            fold = kFoldInner;
        }
    }
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
            SFD_TRY(round_trip<P>(c, psi, V, ax, tau, cD, Om, imag, flags, beta, dt, nullptr, epart[ax], false, ws));
            if (ehist) record(ax);
        }
        SFD_TRY(round_trip<P>(c, psi, V, D - 1, tau, cD, Om, imag, nlflags, beta, dt, nullptr, epart[D - 1], true,
                              ws));
        if (ehist) {
            record(D - 1);
            launch_finalize_sets(c->stream, fs);
            SFD_HIP(hipGetLastError());
        }
        for (int ax = 0; ax < D - 1; ++ax)
            SFD_TRY(round_trip<P>(c, psi, V, ax, tau, cD, Om, imag, (ax == D - 2 && imag) ? kFlagNorm : 0, beta,
                                  dt, nullptr, nullptr, false, ws));
        if (imag) {
            launch_finalize(c->stream, npart, c->last_nblocks, 1, scal + 1, c->dvol(), scal);
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
                                      en ? epart[ax] : nullptr, false, ws));
                if (en) {
                    fs.partial[ax] = epart[ax];
                    fs.nb[ax] = c->last_nblocks;
                    fs.K[ax] = ax == 0 ? 4 : 1;
                    fs.out[ax] = ehist + 8 * s + (ax == 0 ? 0 : 3 + ax);
                }
            }
            if (ehist && h == 0) {
                launch_finalize_sets(c->stream, fs);
                SFD_HIP(hipGetLastError());
            }
        }
        if (imag) {
            const long long nb = c->last_nblocks;  // partials of the norm sweep just launched
            launch_finalize(c->stream, npart, nb, 1, scal + 1, c->dvol(), scal);
            SFD_HIP(hipGetLastError());
            pending = true;
        }
    }
    if (pending) {
        const long long nb = grid_blocks(c->nsite);
        hipLaunchKernelGGL(k_scale<S>, dim3((unsigned)nb), dim3(256), 0, c->stream, static_cast<cx<S>*>(psi),
                           c->nsite, 0.0, (const double*)scal, 0);
        SFD_HIP(hipGetLastError());
    }
    return SFD_OK;
}

template <typename P>
int energy_impl(sfd_ctx* c, const void* psi, const void* V, double cD, double beta, double Om, double* out5)
{
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
        std::memset(&a, 0, sizeof a);
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
    launch_finalize_sets(c->stream, fs);
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
// transforms, so the sums are the state's own.  The steps are chained
// (kFoldChain): the x sums of the state between two steps come out of the
// second round trip of the [x x] pass that joins them.  With rotation the x
// flow's phase depends on y, and in imaginary time the flows damp; each step
// is then preceded by a separate energy evaluation.  energies: host, 5 per
// step.
template <typename P>
int split_energy_impl(sfd_ctx* c, void* psi, const void* V, double cD, double beta, double Om, double dt,
                      int64_t nsteps, int order, int mode, double* energies)
{
    if (Om != 0.0 || mode != SFD_REAL_TIME || c->blu[0] || c->blu[1] || c->blu[2]) {
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
        SFD_TRY(split_impl<P>(c, psi, V, cD, beta, Om, dt, m, order, mode, c->ehist, kFoldChain));
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
    SweepArgs a;
    std::memset(&a, 0, sizeof a);
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

// Gradient of L = 1 - |<target|psi_N>|^2 through nsteps real-time Strang
// steps by the reverse sweep, recomputing forward states from checkpoints
// taken every `every` steps (sqrt(nsteps) when every <= 0, at most nsteps):
// with K = ceil(nsteps/every) checkpoints and S side streams (2 for fields of
// at most 16 MiB, else 1; never more than K), memory is
// K + min(S + 1, K) every + S + 2 fields, forward work twice the plain run.
template <typename P>
int adjoint_impl(sfd_ctx* c, const void* psi0, const void* V, const void* target, double cD, double beta, double Om,
                 double dt, int64_t N, int64_t every, double* gradV, double* grad_beta, double* loss, void* psi_final)
{
    using S = typename P::S;
    using R = typename P::R;
    const long long ns = c->nsite;
    const size_t fb = (size_t)ns * sizeof(cx<S>);
    if (every <= 0) every = (int64_t)std::ceil(std::sqrt((double)(N > 0 ? N : 1)));
    every = std::min<int64_t>(every, N > 0 ? N : 1);  // one segment at most covers the run
    const int64_t nck = N > 0 ? (N + every - 1) / every : 0;
    const long long nb = grid_blocks(ns);
    // Small fields leave most CUs idle in every sweep: two side streams then
    // recompute two segments at once.  No more streams than segments, and no
    // more saved-state buffers than segments.
    int nside = adjoint_side_streams(fb);
    nside = (int)std::max<int64_t>(1, std::min<int64_t>(nside, nck));
    const int nbuf = (int)std::max<int64_t>(1, std::min<int64_t>(nside + 1, nck));
    size_t ck_bytes = 0, seg_fields = 0, seg_bytes = 0, psi_bytes = 0;
    if (__builtin_mul_overflow(fb, (size_t)nck, &ck_bytes) ||
        __builtin_mul_overflow((size_t)every, (size_t)nbuf, &seg_fields) ||
        __builtin_mul_overflow(fb, seg_fields, &seg_bytes) || __builtin_mul_overflow(fb, (size_t)nside, &psi_bytes))
        return fail(SFD_ERR_ARG, "sfd_adjoint: checkpoint storage for %lld steps every %lld overflows",
                    (long long)N, (long long)every);
    DevBuf psi, G, ck, seg, part, acc;
    SFD_HIP(hipMalloc(&psi.p, psi_bytes));
    SFD_HIP(hipMalloc(&G.p, fb));
    if (nck > 0) SFD_HIP(hipMalloc(&ck.p, ck_bytes));
    SFD_HIP(hipMalloc(&seg.p, seg_bytes));
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
    // evicts nothing from then on (the side streams' destructor lifts the pin).
    SideStreams ss;
    SFD_TRY(ss.init(c, nside));
    c->pin_phases = true;
    for (int ax = 0; ax < c->ndim; ++ax) {
        void* ph = nullptr;
        int rows = 0;
        SFD_TRY(build_phase<P>(c, ax, tau, cD, Om, 0, &ph, &rows));
        SFD_TRY(build_phase<P>(c, ax, -tau, cD, Om, 0, &ph, &rows));
        if (merge) SFD_TRY(build_phase<P>(c, ax, -2.0 * tau, cD, Om, 0, &ph, &rows));
    }
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
    using S = typename P::S;
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
        const int nv = axis_variants<P>(c, ax);
        int best[2] = {0, 0};
        float best_ms[2] = {1e30f, 1e30f};
        // Chained real-time steps run single round trips, x's with the
        // nonlinear phase in its epilogue.  The innermost axis of a
        // one-step call, of imaginary-time and of energy-loop steps also runs
        // the double round trip [d N d] (z in 3D; y in 2D without rotation),
        // and the x axis of their chained steps the double round trip [x x]:
        // the single and the double round trips each get their fastest
        // variant.
        const bool folds = c->ndim >= 2 && (ax == c->ndim - 1 || (ax == 0 && !c->blu[0]));
        const int single_flags = (c->ndim >= 2 && ax == 0) ? kFlagNL : 0;
        for (int v = 0; rc == SFD_OK && v < nv; ++v) {
            if ((rc = set_variant<P>(c, ax, v, 3)) != SFD_OK) break;
            for (int twice = 0; rc == SFD_OK && twice <= (folds ? 1 : 0); ++twice) {
                auto launch = [&]() {
                    return twice ? round_trip<P>(c, buf, nullptr, ax, 0.0, -0.5, 0.0, 0, ax == 0 ? 0 : kFlagNL, 0.0,
                                                 0.0, nullptr, nullptr, true)
                                 : round_trip<P>(c, buf, nullptr, ax, 0.0, -0.5, 0.0, 0, single_flags, 0.0, 0.0);
                };
                float vbest = 1e30f;
                for (int w = 0; rc == SFD_OK && w < 2; ++w) rc = launch();
                for (int rep = 0; rc == SFD_OK && rep < 3; ++rep) {
                    (void)hipEventRecord(e0, c->stream);
                    for (int k = 0; rc == SFD_OK && k < 8; ++k) rc = launch();
                    (void)hipEventRecord(e1, c->stream);
                    if (hipEventSynchronize(e1) != hipSuccess) {
                        rc = fail(SFD_ERR_HIP, "event sync failed");
                        break;
                    }
                    float ms = 0.f;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    if (ms / 8 < vbest) vbest = ms / 8;
                }
                if (!twice && times_out) times_out[ax * 8 + (v < 8 ? v : 7)] = vbest;
                if (vbest < best_ms[twice]) {
                    best_ms[twice] = vbest;
                    best[twice] = v;
                }
            }
        }
        if (rc == SFD_OK) rc = set_variant<P>(c, ax, best[0], 1);
        if (rc == SFD_OK) rc = set_variant<P>(c, ax, folds ? best[1] : best[0], 2);
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
    for (int ax = 0; ax < c->ndim; ++ax) {
        const int n = c->n[ax];
        SFD_TRY(set_variant<P>(c, ax, 0, 3));
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

// ------------------------------------------------------------ batch levels
// A batch advances its instances at level 0 (the batch's precision) and, for
// FP32 storage with escalation, at level 1 (FP64) once they reach their
// FP32 floor.  The routines below run one level with the kernels of
// precision P; the chunk loop, graphs and escalation policy are in sfd.hip.

// FP32 potential widened to the FP64 level's copy.
template <typename S>
__global__ void k_widen_real(const S* __restrict__ src, double* __restrict__ dst, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = (double)src[i];
}

// Launch variant `idx` for one axis of a size group at one level: its
// twiddles (rebuilt when the line split changes) and tile counts.
template <typename P>
int batch_level_variant(sfd_batch* b, int L, int g, int ax, int idx)
{
    using R = typename P::R;
    auto& G = b->groups[g];
    VariantInfo vi;
    if (!variant_info<P>(G.n[ax], ax == 0, idx, &vi))
        return fail(SFD_ERR_SHAPE, "batch: no sweep variant %d for line length %d", idx, G.n[ax]);
    if (!G.tw[L][ax] || G.vi[L][ax].n1 != vi.n1 || G.vi[L][ax].n2 != vi.n2 || G.vi[L][ax].h != vi.h) {
        std::vector<cx<R>> w;
        fill_twiddles<R>(w, G.n[ax], vi.n1, vi.n2, vi.h, false);
        if (G.tw[L][ax]) {
            SFD_HIP(hipStreamSynchronize(b->stream));
            SFD_HIP(hipFree(G.tw[L][ax]));
            G.tw[L][ax] = nullptr;
        }
        SFD_HIP(hipMalloc(&G.tw[L][ax], w.size() * sizeof(cx<R>)));
        SFD_TRY(upload_sync(G.tw[L][ax], w.data(), w.size() * sizeof(cx<R>)));
    }
    G.vi[L][ax] = vi;
    G.var[L][ax] = idx;
    const long long NL = (long long)vi.ts * vi.g;
    const long long nx = G.n[0], ny = G.n[1], nz = G.ndim == 3 ? G.n[2] : 1;
    if (ax == 0) {
        G.ntx[L][ax] = 0;
        G.ntiles[L][ax] = (ny * nz + NL - 1) / NL;
    } else {
        G.ntx[L][ax] = (int)((nx + NL - 1) / NL);
        G.ntiles[L][ax] = (long long)G.ntx[L][ax] * (ax == 1 ? nz : ny);
    }
    return SFD_OK;
}

// Tiles of the most-tiled variant of axis ax for a group: the room for the
// norm partials of a sweep on that axis whatever the autotuner picks.
template <typename P>
long long batch_most_tiles(const sfd_batch::Group& G, int ax)
{
    long long most = 0;
    const long long nx = G.n[0], ny = G.n[1], nz = G.ndim == 3 ? G.n[2] : 1;
    for (int v = 0; v < n_variants<P>(G.n[ax], ax == 0); ++v) {
        VariantInfo vi;
        if (!variant_info<P>(G.n[ax], ax == 0, v, &vi)) continue;
        const long long NL = (long long)vi.ts * vi.g;
        const long long t = ax == 0 ? (ny * nz + NL - 1) / NL : ((nx + NL - 1) / NL) * (ax == 1 ? nz : ny);
        most = std::max(most, t);
    }
    return most;
}

// Variants, twiddles and device buffers of one level (the norm partials
// have room for the norm sweep on x or y under any variant).
template <typename P>
int batch_level_setup(sfd_batch* b, int L)
{
    using S = typename P::S;
    auto& lv = b->lv[L];
    const int ni = b->ninst;
    for (size_t g = 0; g < b->groups.size(); ++g)
        for (int ax = 0; ax < b->groups[g].ndim; ++ax) SFD_TRY(batch_level_variant<P>(b, L, (int)g, ax, 0));
    lv.poff.assign(ni, 0);
    long long parts = 0;
    for (int i = 0; i < ni; ++i) {
        const auto& G = b->groups[b->inst[i].group];
        lv.poff[i] = parts;
        parts += std::max(batch_most_tiles<P>(G, 0), batch_most_tiles<P>(G, 1));
    }
    SFD_HIP(hipMalloc(&lv.d_inst, ni * sizeof(BatchInstDev)));
    SFD_HIP(hipMalloc(&lv.d_desc, (size_t)sfd_batch::kMaxKinds * ni * sizeof(SweepArgs)));
    SFD_HIP(hipMalloc(&lv.d_done, ni * sizeof(int)));
    SFD_HIP(hipMalloc(&lv.d_partial, (size_t)parts * sizeof(double)));
    SFD_HIP(hipMalloc(&lv.d_snap, (size_t)b->sites * sizeof(cx<S>)));
    if (L == 1) {
        SFD_HIP(hipMalloc(&lv.d_field, (size_t)b->sites * sizeof(cx<S>)));
        SFD_HIP(hipMalloc(&lv.d_V, (size_t)b->sites * sizeof(S)));
    }
    lv.phase.assign(3 * (size_t)ni, nullptr);
    lv.phase_bytes.assign(3 * (size_t)ni, 0);
    return SFD_OK;
}

// Descriptors, phase tables and instance records of one solve at one level.
// Level 0 works on the caller's fields and potentials; level 1 on its own
// FP64 fields and FP64 copies of the potentials.
template <typename P>
int batch_level_bind(sfd_batch* b, int L, void* const* psi, const void* const* V, const double* beta,
                     const double* Omega, double cD, double dt)
{
    using R = typename P::R;
    using S = typename P::S;
    auto& lv = b->lv[L];
    const int ni = b->ninst;
    std::vector<BatchInstDev> hi(ni);
    std::vector<SweepArgs> desc((size_t)sfd_batch::kMaxKinds * ni);
    const double tau = 0.5 * dt;
    for (int i = 0; i < ni; ++i) {
        const auto& in = b->inst[i];
        const auto& G = b->groups[in.group];
        const double Om = Omega ? Omega[in.user] : 0.0;
        const void* Vi = V ? V[in.user] : nullptr;
        void* field = psi[in.user];
        if (L == 1) {
            field = static_cast<char*>(lv.d_field) + (size_t)in.soff * sizeof(cx<S>);
            if (Vi) {
                double* v64 = static_cast<double*>(lv.d_V) + in.soff;
                hipLaunchKernelGGL(k_widen_real<float>, dim3((unsigned)grid_blocks(in.nsite)), dim3(256), 0,
                                   b->stream, static_cast<const float*>(Vi), v64, in.nsite);
                SFD_HIP(hipGetLastError());
                Vi = v64;
            }
        }
        hi[i].psi = field;
        hi[i].snap = static_cast<char*>(lv.d_snap) + (size_t)in.soff * sizeof(cx<S>);
        hi[i].partial = lv.d_partial + lv.poff[i];
        hi[i].nsite = in.nsite;
        hi[i].dv = in.d[0] * in.d[1] * (in.ndim == 3 ? in.d[2] : 1.0);
        int rows[3] = {1, 1, 1};
        for (int ax = 0; ax < in.ndim; ++ax) {
            rows[ax] = (Om != 0.0 && ax < 2) ? in.n[ax == 0 ? 1 : 0] : 1;
            const size_t bytes = (size_t)rows[ax] * in.n[ax] * sizeof(cx<R>);
            void*& dev = lv.phase[3 * (size_t)i + ax];
            if (dev && lv.phase_bytes[3 * (size_t)i + ax] != bytes) {
                SFD_HIP(hipStreamSynchronize(b->stream));
                SFD_HIP(hipFree(dev));
                dev = nullptr;
            }
            if (!dev) {
                SFD_HIP(hipMalloc(&dev, bytes));
                lv.phase_bytes[3 * (size_t)i + ax] = bytes;
            }
            const long long tot = (long long)rows[ax] * in.n[ax];
            hipLaunchKernelGGL(k_build_phase<R>, dim3((unsigned)std::min<long long>((tot + 255) / 256, 1024)),
                               dim3(256), 0, b->stream, static_cast<cx<R>*>(dev), (const double*)in.geo[ax],
                               (const double*)(in.geo[ax == 0 ? 1 : 0] + in.n[ax == 0 ? 1 : 0]), in.n[ax], rows[ax],
                               ax, tau, cD, Om, 1);
            SFD_HIP(hipGetLastError());
        }
        int norm_axis = 0;
        for (size_t k = 0; k < G.kinds.size(); ++k) {
            const auto& kd = G.kinds[k];
            const int ax = kd.axis;
            SweepArgs a;
            std::memset(&a, 0, sizeof a);
            a.psi = field;
            a.V = Vi;
            a.phase = lv.phase[3 * (size_t)i + ax];
            a.phase_row_stride = rows[ax] > 1 ? in.n[ax] : 0;
            if (ax == 0) a.row_mod = rows[ax] > 1 ? in.n[1] : 0;
            else a.row_from_col = (ax == 1 && rows[ax] > 1) ? 1 : 0;
            a.tw = G.tw[L][ax];
            a.scale = b->d_scale + i;
            a.beta = beta[in.user];
            a.dt = dt;
            a.flags = kFlagImag | kd.flags;
            if (kd.flags & (kFlagNorm | kFlagNormMid)) {  // every norm sweep of a group is on one axis
                a.partial = hi[i].partial;
                norm_axis = ax;
            }
            batch_geom(a, in, G, L, ax);
            desc[k * (size_t)ni + i] = a;
        }
        hi[i].nparts = G.ntiles[L][norm_axis];
    }
    SFD_HIP(hipMemcpyAsync(lv.d_inst, hi.data(), hi.size() * sizeof(BatchInstDev), hipMemcpyHostToDevice, b->stream));
    SFD_HIP(hipMemcpyAsync(lv.d_desc, desc.data(), desc.size() * sizeof(SweepArgs), hipMemcpyHostToDevice,
                           b->stream));
    SFD_HIP(hipStreamSynchronize(b->stream));  // the host vectors go out of scope
    return SFD_OK;
}

// The sweeps (and renormalisations, sfd_batch::kFinalize) of one sequence
// of group g at level L on stream st, over its unfinished instances.
template <typename P>
int batch_level_run(sfd_batch* b, int L, int g, hipStream_t st, const std::vector<int>& seq)
{
    const auto& G = b->groups[g];
    const auto& lv = b->lv[L];
    const int ni = b->ninst, f = G.first;
    for (const int k : seq) {
        if (k == sfd_batch::kFinalize) {
            hipLaunchKernelGGL(k_batch_finalize, dim3(G.count), dim3(256), 0, st, lv.d_inst + f, b->d_scale + f,
                               lv.d_done + f, b->d_steps + f);
            SFD_HIP(hipGetLastError());
            continue;
        }
        const int ax = G.kinds[k].axis;
        const VariantInfo& vi = G.vi[L][ax];
        dim3 grid((unsigned)G.ntiles[L][ax], 1, (unsigned)G.count);
        dim3 block(vi.ts * vi.g * vi.n2 * vi.h);
        if (!launch_batch<P>(G.n[ax], ax == 0, G.var[L][ax], grid, block, st, lv.d_desc + k * (size_t)ni + f,
                             lv.d_done + f, G.kinds[k].twice))
            return fail(SFD_ERR_SHAPE, "batch: no kernel for line length %d", G.n[ax]);
        SFD_HIP(hipGetLastError());
    }
    return SFD_OK;
}

// k imaginary-time Strang steps of the unfinished instances of group g at
// level L on stream st, the last one a convergence check: the snapshot
// before it and ||Delta psi|| after it (with the escalation policy of
// k_batch_diff_finalize when `escalate`).  Each step ends with the
// per-instance renormalisation factor, which the next sweep applies.  A
// group that chains runs the k - 1 steps before the check as one chain
// (open, body k - 2 times, close): the x flow that ends a step and the one
// that starts the next are one [x x] pass, whose first round trip takes the
// norm of the state that ends the step; the check step is never chained, as
// the snapshot needs the state before it in memory.
template <typename P>
int batch_level_steps(sfd_batch* b, int L, int g, hipStream_t st, int k, double tol, bool escalate)
{
    using S = typename P::S;
    const auto& G = b->groups[g];
    const auto& lv = b->lv[L];
    const int f = G.first;
    constexpr int KB = sfd_batch::kDiffBlocks;
    if (G.chain && k >= 3) {
        SFD_TRY(batch_level_run<P>(b, L, g, st, G.open));
        for (int s = 1; s < k - 1; ++s) SFD_TRY(batch_level_run<P>(b, L, g, st, G.body));
        SFD_TRY(batch_level_run<P>(b, L, g, st, G.close));
    } else {
        for (int s = 0; s < k - 1; ++s) SFD_TRY(batch_level_run<P>(b, L, g, st, G.step));
    }
    hipLaunchKernelGGL(k_batch_snapshot<S>, dim3(KB, G.count), dim3(256), 0, st, lv.d_inst + f, b->d_scale + f,
                       b->d_snap_scale + f, lv.d_done + f);
    SFD_HIP(hipGetLastError());
    SFD_TRY(batch_level_run<P>(b, L, g, st, G.step));
    hipLaunchKernelGGL(k_batch_diff<S>, dim3(KB, G.count), dim3(256), 0, st, lv.d_inst + f, b->d_scale + f,
                       b->d_snap_scale + f, lv.d_done + f, b->d_dpart + (size_t)f * KB);
    SFD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_batch_diff_finalize, dim3(G.count), dim3(256), 0, st, b->d_dpart + (size_t)f * KB, KB,
                       lv.d_inst + f, b->d_dpsi + f, b->d_dprev + f, lv.d_done + f, tol, b->d_esc + f,
                       escalate ? 1 : 0);
    SFD_HIP(hipGetLastError());
    return SFD_OK;
}

// Escalation of the instances of group g flagged at this check (level 0 in
// precision P to the FP64 level).
template <typename P>
int batch_escalate(sfd_batch* b, int g, hipStream_t st)
{
    using S = typename P::S;
    const auto& G = b->groups[g];
    const int f = G.first;
    constexpr int KB = sfd_batch::kDiffBlocks;
    double* sq = b->d_sq + (size_t)f * 2 * KB;
    hipLaunchKernelGGL(k_batch_sq_partials<S>, dim3(KB, G.count), dim3(256), 0, st, b->lv[0].d_inst + f,
                       b->d_esc + f, sq);
    SFD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_batch_escalate<S>, dim3(KB, G.count), dim3(256), 0, st, b->lv[0].d_inst + f,
                       b->lv[1].d_inst + f, b->d_scale + f, b->d_esc + f, b->d_real + f, (const double*)sq, KB,
                       (const double*)(b->d_phi0 + f));
    SFD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_batch_escalate_flags, dim3((G.count + 255) / 256), dim3(256), 0, st, b->d_scale + f,
                       b->d_esc + f, b->lv[1].d_done + f, b->d_dprev + f, G.count);
    SFD_HIP(hipGetLastError());
    return SFD_OK;
}

// Start of an escalating solve (level 0 in precision P): the dominant phase
// of every initial state, the phase the escalation restores.
template <typename P>
int batch_phase0(sfd_batch* b)
{
    using S = typename P::S;
    constexpr int KB = sfd_batch::kDiffBlocks;
    hipLaunchKernelGGL(k_batch_sq_partials<S>, dim3(KB, b->ninst), dim3(256), 0, b->stream, b->lv[0].d_inst,
                       (const int*)nullptr, b->d_sq);
    SFD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_batch_phase0, dim3((b->ninst + 255) / 256), dim3(256), 0, b->stream, (const double*)b->d_sq,
                       KB, b->d_phi0, b->ninst);
    SFD_HIP(hipGetLastError());
    return SFD_OK;
}

// End of a solve (level 0 in precision P): every instance's final state
// times its pending scale into the caller's field, and in FP64 to the
// caller's outputs.
template <typename P>
int batch_finish(sfd_batch* b)
{
    using S = typename P::S;
    const BatchInstDev* hi = b->lv[1].prec ? b->lv[1].d_inst : nullptr;
    hipLaunchKernelGGL(k_batch_finish<S>, dim3(sfd_batch::kDiffBlocks, b->ninst), dim3(256), 0, b->stream,
                       b->lv[0].d_inst, hi, b->d_scale, b->d_esc, (void* const*)b->d_out64);
    SFD_HIP(hipGetLastError());
    return SFD_OK;
}

// Batch autotuner of level L: for every size group and axis with more than
// one compiled variant, time the identity round trip (tau = 0) of
// k_batch_sweep over the group's instances on the level's snapshot buffer
// filled with unit-modulus pseudo-random data (the fields are not touched),
// as the double round trip on the axes that run as one in a solve: the axis
// whose flows the step folds (y in 2D, z in 3D) and x ([x x] of chained
// steps); 2 warm-up launches, then the best of 3 batches of 8, and keep the
// fastest.  times[24 g + 8 ax + v] = ms per launch (-1 if none).
template <typename P>
int batch_autotune(sfd_batch* b, int L, float* times)
{
    using R = typename P::R;
    using S = typename P::S;
    auto& lv = b->lv[L];
    const int ni = b->ninst;
    if (times)
        for (int i = 0; i < 24 * ni; ++i) times[i] = -1.0f;
    batch_drop_graphs(b);
    SFD_HIP(hipEventRecord(b->ev, b->user_stream));
    SFD_HIP(hipStreamWaitEvent(b->stream, b->ev, 0));
    hipLaunchKernelGGL(k_fill_phases<S>, dim3((unsigned)grid_blocks(b->sites)), dim3(256), 0, b->stream,
                       static_cast<cx<S>*>(lv.d_snap), b->sites);
    SFD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_batch_reset, dim3((ni + 255) / 256), dim3(256), 0, b->stream, b->d_scale, lv.d_done,
                       (int*)nullptr, b->d_steps, b->d_dpsi, b->d_dprev, b->d_esc, ni);
    SFD_HIP(hipGetLastError());
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<void*> tmp;  // identity phase tables
    std::vector<SweepArgs> desc;
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
            const int fold_axis = G.ndim == 2 ? 1 : 2;
            for (int ax = 0; ax < G.ndim; ++ax) {
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
                    SFD_TRY(batch_level_variant<P>(b, L, (int)gi, ax, v));
                    desc.assign(G.count, SweepArgs{});
                    for (int k = 0; k < G.count; ++k) {
                        const auto& in = b->inst[G.first + k];
                        SweepArgs& a = desc[k];
                        std::memset(&a, 0, sizeof a);
                        a.psi = static_cast<char*>(lv.d_snap) + (size_t)in.soff * sizeof(cx<S>);
                        a.phase = ph;
                        a.tw = G.tw[L][ax];
                        a.flags = kFlagImag;
                        batch_geom(a, in, G, L, ax);
                    }
                    SFD_HIP(hipMemcpyAsync(lv.d_desc, desc.data(), desc.size() * sizeof(SweepArgs),
                                           hipMemcpyHostToDevice, b->stream));
                    const VariantInfo& vi = G.vi[L][ax];
                    dim3 grid((unsigned)G.ntiles[L][ax], 1, (unsigned)G.count);
                    dim3 block(vi.ts * vi.g * vi.n2 * vi.h);
                    auto launch = [&]() {
                        return launch_batch<P>(n, ax == 0, v, grid, block, b->stream, lv.d_desc, lv.d_done,
                                               ax == fold_axis || ax == 0);
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
                    if (times) times[24 * gi + 8 * ax + v] = t_best;
                    if (t_best < best) {
                        best = t_best;
                        best_v = v;
                    }
                }
                SFD_TRY(batch_level_variant<P>(b, L, (int)gi, ax, best_v));
            }
        }
        return SFD_OK;
    };
    const int rc = run();
    cleanup();
    return rc;
}

// Wrappers with the signatures of PrecOps.
template <typename P>
int round_trip_axis(sfd_ctx* c, void* psi, int axis, double tau, double cD, double Om, int imag)
{
    return round_trip<P>(c, psi, nullptr, axis, tau, cD, Om, imag, 0, 0.0, 0.0);
}

template <typename P>
int split_op(sfd_ctx* c, void* psi, const void* V, double cD, double beta, double Om, double dt, int64_t nsteps,
             int order, int mode, double* ehist, int fold, const Workspace* ws)
{
    return split_impl<P>(c, psi, V, cD, beta, Om, dt, nsteps, order, mode, ehist, fold, ws);
}

template <typename P>
PrecOps make_ops()
{
    PrecOps o;
    o.init_tables = &init_tables<P>;
    o.split = &split_op<P>;
    o.split_energy = &split_energy_impl<P>;
    o.round_trip_axis = &round_trip_axis<P>;
    o.energy = &energy_impl<P>;
    o.fft = &fft_impl<P>;
    o.norm2 = &norm2_impl<P>;
    o.adjoint = &adjoint_impl<P>;
    o.autotune = &autotune_impl<P>;
    o.set_variant = &set_variant<P>;
    o.n_variants = &axis_variants<P>;
    o.build_phase_into = &launch_build_phase<P>;
    o.table_entry_bytes = sizeof(cx<typename P::R>);
    o.batch_level_setup = &batch_level_setup<P>;
    o.batch_level_variant = &batch_level_variant<P>;
    o.batch_level_bind = &batch_level_bind<P>;
    o.batch_level_steps = &batch_level_steps<P>;
    o.batch_escalate = &batch_escalate<P>;
    o.batch_phase0 = &batch_phase0<P>;
    o.batch_finish = &batch_finish<P>;
    o.batch_autotune = &batch_autotune<P>;
    return o;
}

}  // namespace sfd
