/*
 * sfd_oracle.c -- FP64 CPU restatement of SuperfluidDynamics.jl's
 * Gross-Pitaevskii split-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the test suite, the
 * smoke check and the CPU-baseline leg of the benchmark call it to produce
 * expected values.  The product path (the HIP library) never links, loads or
 * calls it, and nothing here is a fallback for the GPU code.
 *
 * What it restates (reference file:line):
 *   - NumModelSplit2 step: K(dt/2) N(dt) K(dt/2)      splitting.jl:149-154
 *   - NumModelSplit1 step: K(dt) N(dt)                  splitting.jl:78-82
 *   - K = solveLapRot!: x, then y, then z round trips   splitting.jl:156-203
 *       x: exp(i (cD xi_x^2 - Om y xi_x) tau)           splitting.jl:190
 *       y: exp(i (cD xi_y^2 + Om x xi_y) tau)           splitting.jl:195
 *       z: exp(i  cD xi_z^2 tau)                        splitting.jl:200
 *     forward FFT unnormalised, inverse FFT then *1/n   Plans.jl:1132-1250
 *   - N = solveNL!: exp(-i (V + beta |psi|^2) dt) psi   splitting.jl:205-209
 *   - wavenumbers fftfreq(n, 2pi/dx)                    Plans.jl:829-830,967-969
 *   - coordinates LinRange(min,max,n+1)[1:n]            Grids.jl:78-82,119-125
 *   - norm sqrt(sum|psi|^2) sqrt(prod dx)               Fields.jl:172-180
 *   - energy (EOm, EDelta, Ebeta, E, mu) through the spectral derivative
 *     fields dx, dy, dz, rx = y dx, ry = -x dy          GrossPitaevskii.jl:61-136,
 *                                                       derivatives.jl:90-126
 *   - imaginary-time variant: the same sweeps with the factor i removed from
 *     every exponent, followed by normalize! (the renormalised gradient flow
 *     whose fixed point is the ground state of backward-euler.jl:72-90; the
 *     reference has no split-step imaginary-time scheme, this one is the
 *     project's own and is specified here).
 *
 * The FFT is the project's own (iterative radix-2 for powers of two, direct
 * DFT with an exact twiddle table otherwise); twiddles and phases are formed
 * in long double and rounded once.  OpenMP parallelises over lines.
 *
 * Layout: complex double interleaved (re, im), x fastest, (nx, ny[, nz]).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cplx;

/* ---------------------------------------------------------------- FFT --- */

typedef struct {
    int n;
    int pow2;
    cplx *w;      /* w[k] = exp(-2 pi i k / n), k in [0, n) */
    int *rev;     /* bit reversal permutation (pow2 only)   */
} fftplan;

static void plan_init(fftplan *p, int n)
{
    p->n = n;
    p->pow2 = (n & (n - 1)) == 0;
    p->w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        p->w[k].re = (double)cosl(a);
        p->w[k].im = (double)sinl(a);
    }
    p->rev = NULL;
    if (p->pow2) {
        int lg = 0;
        while ((1 << lg) < n) ++lg;
        p->rev = (int *)malloc(sizeof(int) * (size_t)n);
        for (int i = 0; i < n; ++i) {
            int r = 0;
            for (int b = 0; b < lg; ++b)
                if (i & (1 << b)) r |= 1 << (lg - 1 - b);
            p->rev[i] = r;
        }
    }
}

static void plan_free(fftplan *p)
{
    free(p->w);
    free(p->rev);
}

/* In-place unnormalised DFT of a contiguous line; sign = -1 forward
 * (exp(-2 pi i jk/n)), +1 backward.  tmp must hold n values. */
static void fft_line(const fftplan *p, cplx *a, int sign, cplx *tmp)
{
    const int n = p->n;
    if (n == 1) return;
    if (p->pow2) {
        for (int i = 0; i < n; ++i) {
            int r = p->rev[i];
            if (r > i) { cplx t = a[i]; a[i] = a[r]; a[r] = t; }
        }
        for (int len = 2; len <= n; len <<= 1) {
            const int half = len >> 1, step = n / len;
            for (int b = 0; b < n; b += len) {
                for (int j = 0; j < half; ++j) {
                    cplx w = p->w[j * step];
                    if (sign > 0) w.im = -w.im;
                    cplx u = a[b + j], v = a[b + j + half];
                    cplx t = { v.re * w.re - v.im * w.im, v.re * w.im + v.im * w.re };
                    a[b + j].re = u.re + t.re;        a[b + j].im = u.im + t.im;
                    a[b + j + half].re = u.re - t.re; a[b + j + half].im = u.im - t.im;
                }
            }
        }
    } else {
        for (int k = 0; k < n; ++k) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < n; ++j) {
                cplx w = p->w[(int)(((int64_t)j * k) % n)];
                if (sign > 0) w.im = -w.im;
                sr += a[j].re * w.re - a[j].im * w.im;
                si += a[j].re * w.im + a[j].im * w.re;
            }
            tmp[k].re = sr; tmp[k].im = si;
        }
        memcpy(a, tmp, sizeof(cplx) * (size_t)n);
    }
}

/* ------------------------------------------------------------- geometry --- */

/* Julia LinRange(a, b, n+1)[i+1] = (1 - t) a + t b, t = i/n (Grids.jl:78). */
static double lin_coord(double a, double b, int n, int i)
{
    double t = (double)i / (double)n;
    return (1.0 - t) * a + t * b;
}

/* fftfreq(n, 2pi/dx)[k+1] = kk * ((2pi/dx)/n)  (Plans.jl:829). */
static double fft_freq(int n, double dx, int k)
{
    const double twopi = 6.283185307179586; /* Julia's Float64(2pi) */
    double mult = (twopi / dx) / (double)n;
    int kk = (k < (n + 1) / 2) ? k : k - n;
    return (double)kk * mult;
}

typedef struct {
    int ndim;
    int n[3];
    double lo[3], hi[3], d[3];
    int64_t nsite;
    double *coord[3];
    double *xi[3];
    fftplan plan[3];
} geom;

static void geom_init(geom *g, int ndim, const int *n, const double *lo, const double *hi)
{
    g->ndim = ndim;
    g->nsite = 1;
    for (int a = 0; a < 3; ++a) {
        g->n[a] = a < ndim ? n[a] : 1;
        g->lo[a] = a < ndim ? lo[a] : 0.0;
        g->hi[a] = a < ndim ? hi[a] : 1.0;
        g->d[a] = (g->hi[a] - g->lo[a]) / (double)g->n[a];
        g->nsite *= g->n[a];
        g->coord[a] = (double *)malloc(sizeof(double) * (size_t)g->n[a]);
        g->xi[a] = (double *)malloc(sizeof(double) * (size_t)g->n[a]);
        for (int i = 0; i < g->n[a]; ++i) {
            g->coord[a][i] = lin_coord(g->lo[a], g->hi[a], g->n[a], i);
            g->xi[a][i] = fft_freq(g->n[a], g->d[a], i);
        }
        plan_init(&g->plan[a], g->n[a]);
    }
}

static void geom_free(geom *g)
{
    for (int a = 0; a < 3; ++a) {
        free(g->coord[a]);
        free(g->xi[a]);
        plan_free(&g->plan[a]);
    }
}

static double dvol(const geom *g)
{
    double v = 1.0;
    for (int a = 0; a < g->ndim; ++a) v *= g->d[a];
    return v;
}

/* Line bookkeeping: along axis `ax`, a line is fixed by the two other
 * indices; element j sits at base + j*stride. */
static void line_geom(const geom *g, int ax, int64_t line, int64_t *base, int64_t *stride, int *o1, int *o2)
{
    const int nx = g->n[0], ny = g->n[1];
    if (ax == 0) {                      /* line = y + ny*z */
        *o1 = (int)(line % ny); *o2 = (int)(line / ny);
        *base = line * nx; *stride = 1;
    } else if (ax == 1) {               /* line = x + nx*z */
        *o1 = (int)(line % nx); *o2 = (int)(line / nx);
        *base = (int64_t)(*o1) + (int64_t)(*o2) * nx * ny; *stride = nx;
    } else {                            /* line = x + nx*y */
        *o1 = (int)(line % nx); *o2 = (int)(line / nx);
        *base = (int64_t)(*o1) + (int64_t)(*o2) * nx; *stride = (int64_t)nx * ny;
    }
}

/* ------------------------------------------------------ linear sweeps --- */

/* One round trip along `ax` with multiplier exp(s (cD xi^2 + r) tau), where
 * s = i in real time and s = 1 in imaginary time; r = -Om y xi (x axis),
 * +Om x xi (y axis), 0 (z axis).  splitting.jl:188-201. */
static void sweep_axis(const geom *g, cplx *psi, int ax, double cD, double Om, double tau, int imag)
{
    const int n = g->n[ax];
    const int64_t nlines = g->nsite / n;
    const double inv_n = 1.0 / (double)n;
#pragma omp parallel
    {
        cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)n);
        cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)n);
#pragma omp for schedule(static)
        for (int64_t line = 0; line < nlines; ++line) {
            int64_t base, stride; int o1, o2;
            line_geom(g, ax, line, &base, &stride, &o1, &o2);
            for (int j = 0; j < n; ++j) buf[j] = psi[base + j * stride];
            fft_line(&g->plan[ax], buf, -1, tmp);
            for (int k = 0; k < n; ++k) {
                const double xi = g->xi[ax][k];
                double arg;
                if (ax == 0)      arg = (cD * (xi * xi) - (Om * g->coord[1][o1]) * xi) * tau;
                else if (ax == 1) arg = (cD * (xi * xi) + (Om * g->coord[0][o1]) * xi) * tau;
                else              arg = (cD * (xi * xi)) * tau;
                cplx m;
                if (imag) { m.re = exp(arg); m.im = 0.0; }
                else      { m.re = cos(arg); m.im = sin(arg); }
                cplx v = buf[k];
                buf[k].re = m.re * v.re - m.im * v.im;
                buf[k].im = m.re * v.im + m.im * v.re;
            }
            fft_line(&g->plan[ax], buf, +1, tmp);
            for (int j = 0; j < n; ++j) {
                cplx v = buf[j];
                v.re *= inv_n; v.im *= inv_n;
                psi[base + j * stride] = v;
            }
        }
        free(buf);
        free(tmp);
    }
}

/* solveLapRot!: x, y[, z] in that order (splitting.jl:156-203). */
static void lap_rot(const geom *g, cplx *psi, double cD, double Om, double tau, int imag)
{
    for (int ax = 0; ax < g->ndim; ++ax) sweep_axis(g, psi, ax, cD, Om, tau, imag);
}

/* solveNL!: psi <- exp(-i (V + beta |psi|^2) dt) psi (splitting.jl:205-209);
 * imaginary time: exp(-(V + beta |psi|^2) dt). */
static void solve_nl(const geom *g, cplx *psi, const double *V, double beta, double dt, int imag)
{
    const int64_t ns = g->nsite;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ns; ++i) {
        cplx v = psi[i];
        double a2 = v.re * v.re + v.im * v.im;
        double a = (V ? V[i] : 0.0) + beta * a2;
        double th = a * dt;
        cplx m;
        if (imag) { m.re = exp(-th); m.im = 0.0; }
        else      { m.re = cos(th); m.im = -sin(th); }
        psi[i].re = m.re * v.re - m.im * v.im;
        psi[i].im = m.re * v.im + m.im * v.re;
    }
}

static double sum_abs2(const geom *g, const cplx *psi)
{
    double s = 0.0;
    const int64_t ns = g->nsite;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < ns; ++i) s += psi[i].re * psi[i].re + psi[i].im * psi[i].im;
    return s;
}

/* norm: sqrt(sum |psi|^2) * sqrt(prod dx) (Fields.jl:172-180). */
static double field_norm(const geom *g, const cplx *psi)
{
    return sqrt(sum_abs2(g, psi)) * sqrt(dvol(g));
}

static void scale_field(const geom *g, cplx *psi, double s)
{
    const int64_t ns = g->nsite;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ns; ++i) { psi[i].re /= s; psi[i].im /= s; }
}

/* ------------------------------------------------------------ public --- */

/* Advance psi by nsteps split steps.
 * order 2: K(dt/2) N(dt) K(dt/2) (splitting.jl:149-154); order 1: K(dt) N(dt)
 * (splitting.jl:78-82).  imag != 0: imaginary time, normalize! after each
 * step (Fields.jl:196-200).  V may be NULL (PotentialZero). */
int orc_split_step(int ndim, const int *n, const double *lo, const double *hi,
                   double *psi_ri, const double *V,
                   double cD, double beta, double Om, double dt,
                   int64_t nsteps, int order, int imag)
{
    if (ndim < 1 || ndim > 3 || (order != 1 && order != 2)) return -3;
    geom g;
    geom_init(&g, ndim, n, lo, hi);
    cplx *psi = (cplx *)psi_ri;
    for (int64_t s = 0; s < nsteps; ++s) {
        if (order == 2) {
            lap_rot(&g, psi, cD, Om, dt * 0.5, imag);
            solve_nl(&g, psi, V, beta, dt, imag);
            lap_rot(&g, psi, cD, Om, dt * 0.5, imag);
        } else {
            lap_rot(&g, psi, cD, Om, dt, imag);
            solve_nl(&g, psi, V, beta, dt, imag);
        }
        if (imag) scale_field(&g, psi, field_norm(&g, psi));
    }
    geom_free(&g);
    return 0;
}

/* One linear round trip along one axis (the mul_d! / multiplier / ldiv_d!
 * triple of solveLapRot!), exposed for unit checks. */
int orc_sweep_axis(int ndim, const int *n, const double *lo, const double *hi,
                   double *psi_ri, int ax, double cD, double Om, double tau, int imag)
{
    if (ax < 0 || ax >= ndim) return -3;
    geom g;
    geom_init(&g, ndim, n, lo, hi);
    sweep_axis(&g, (cplx *)psi_ri, ax, cD, Om, tau, imag);
    geom_free(&g);
    return 0;
}

/* Unnormalised forward (sign -1) or normalised inverse (sign +1) FFT along
 * one axis, in place (mul_d! / ldiv_d!, Plans.jl:1132-1250). */
int orc_fft_axis(int ndim, const int *n, double *psi_ri, int ax, int sign)
{
    if (ax < 0 || ax >= ndim) return -3;
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    geom g;
    geom_init(&g, ndim, n, lo, hi);
    cplx *psi = (cplx *)psi_ri;
    const int nn = g.n[ax];
    const int64_t nlines = g.nsite / nn;
#pragma omp parallel
    {
        cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)nn);
        cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)nn);
#pragma omp for schedule(static)
        for (int64_t line = 0; line < nlines; ++line) {
            int64_t base, stride; int o1, o2;
            line_geom(&g, ax, line, &base, &stride, &o1, &o2);
            for (int j = 0; j < nn; ++j) buf[j] = psi[base + j * stride];
            fft_line(&g.plan[ax], buf, sign, tmp);
            for (int j = 0; j < nn; ++j) {
                cplx v = buf[j];
                if (sign > 0) { v.re *= 1.0 / nn; v.im *= 1.0 / nn; }
                psi[base + j * stride] = v;
            }
        }
        free(buf);
        free(tmp);
    }
    geom_free(&g);
    return 0;
}

double orc_norm(int ndim, const int *n, const double *lo, const double *hi, const double *psi_ri)
{
    geom g;
    geom_init(&g, ndim, n, lo, hi);
    double r = field_norm(&g, (const cplx *)psi_ri);
    geom_free(&g);
    return r;
}

/* Fill coordinates / wavenumbers of one axis (Grids.jl:78-82, Plans.jl:829). */
int orc_axis_tables(int n, double lo, double hi, double *coord, double *xi)
{
    double d = (hi - lo) / (double)n;
    for (int i = 0; i < n; ++i) {
        coord[i] = lin_coord(lo, hi, n, i);
        xi[i] = fft_freq(n, d, i);
    }
    return 0;
}

/* Spectral derivative along ax of psi: IFFT_ax(mult(k) * FFT_ax psi), where
 * mult = i xi (first derivative) or i c xi (c = coordinate of the line's
 * other axis, the rotation fields rx/ry).  derivatives.jl:90-126. */
static void deriv_axis(const geom *g, const cplx *psi, cplx *out, int ax, int kind)
{
    const int n = g->n[ax];
    const int64_t nlines = g->nsite / n;
#pragma omp parallel
    {
        cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)n);
        cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)n);
#pragma omp for schedule(static)
        for (int64_t line = 0; line < nlines; ++line) {
            int64_t base, stride; int o1, o2;
            line_geom(g, ax, line, &base, &stride, &o1, &o2);
            for (int j = 0; j < n; ++j) buf[j] = psi[base + j * stride];
            fft_line(&g->plan[ax], buf, -1, tmp);
            for (int k = 0; k < n; ++k) {
                double xi = g->xi[ax][k];
                /* kind 0: i xi;  kind 1 (ax 0): i y xi;  kind 1 (ax 1): -i x xi */
                double f = xi;
                if (kind == 1) f = (ax == 0) ? g->coord[1][o1] * xi : -(g->coord[0][o1] * xi);
                cplx v = buf[k];
                buf[k].re = -f * v.im;
                buf[k].im = f * v.re;
            }
            fft_line(&g->plan[ax], buf, +1, tmp);
            for (int j = 0; j < n; ++j) {
                cplx v = buf[j];
                v.re *= 1.0 / n; v.im *= 1.0 / n;
                out[base + j * stride] = v;
            }
        }
        free(buf);
        free(tmp);
    }
}

/* energy: out = {EOm, EDelta, Ebeta, E, mu}  (GrossPitaevskii.jl:61-136).
 * EOm    = sum Re(i conj(psi) Om (rx + ry)) dV,  rx = y dx, ry = -x dy
 * EDelta = (-cD sum |grad psi|^2 + sum V |psi|^2) dV
 * Ebeta  = sum 0.5 beta |psi|^4 dV
 * E      = -EOm + EDelta + Ebeta,   mu = -EOm + EDelta + 2 Ebeta. */
int orc_energy(int ndim, const int *n, const double *lo, const double *hi,
               const double *psi_ri, const double *V, double cD, double beta, double Om,
               double *out)
{
    geom g;
    geom_init(&g, ndim, n, lo, hi);
    const cplx *psi = (const cplx *)psi_ri;
    const int64_t ns = g.nsite;
    cplx *d = (cplx *)malloc(sizeof(cplx) * (size_t)ns);
    cplx *r = (cplx *)malloc(sizeof(cplx) * (size_t)ns);
    cplx *rr = (cplx *)calloc((size_t)ns, sizeof(cplx));
    double grad2 = 0.0;
    for (int ax = 0; ax < ndim; ++ax) {
        deriv_axis(&g, psi, d, ax, 0);
        double s = 0.0;
#pragma omp parallel for reduction(+ : s)
        for (int64_t i = 0; i < ns; ++i) s += d[i].re * d[i].re + d[i].im * d[i].im;
        grad2 += s;
        if (ax < 2 && ndim >= 2) {
            deriv_axis(&g, psi, r, ax, 1);
#pragma omp parallel for
            for (int64_t i = 0; i < ns; ++i) { rr[i].re += r[i].re; rr[i].im += r[i].im; }
        }
    }
    double sOm = 0.0, sV = 0.0, sb = 0.0;
#pragma omp parallel for reduction(+ : sOm, sV, sb)
    for (int64_t i = 0; i < ns; ++i) {
        cplx p = psi[i];
        double a2 = p.re * p.re + p.im * p.im;
        /* Re(i * conj(p) * (Om * rr)) */
        double qr = Om * rr[i].re, qi = Om * rr[i].im;
        double cr = p.re * qr + p.im * qi;   /* conj(p) * q, real */
        double ci = p.re * qi - p.im * qr;   /* imag */
        sOm += -ci;
        sV += (V ? V[i] : 0.0) * a2;
        sb += 0.5 * beta * (a2 * a2);
        (void)cr;
    }
    const double dv = dvol(&g);
    double EOm = sOm * dv;
    double ED = (-cD * grad2 + sV) * dv;
    double Eb = sb * dv;
    out[0] = EOm;
    out[1] = ED;
    out[2] = Eb;
    out[3] = -EOm + ED + Eb;
    out[4] = -EOm + ED + Eb * 2.0;
    free(d); free(r); free(rr);
    geom_free(&g);
    return 0;
}

/* ------------------------------------------------------------- adjoint ---
 * Reverse-mode gradient of L = 1 - |<target|psi_N>|^2, <a|b> = sum conj(a) b dV,
 * through nsteps real-time Strang steps psi1 = K psi_n, psi2 = exp(-i th) psi1
 * with th = (V + beta |psi1|^2) dt, psi_{n+1} = K psi2 (splitting.jl:149-209).
 * Gradient convention: G = dL/du + i dL/dv for psi = u + i v, so that
 * dL = Re(conj(G) dpsi).  Backward step:
 *   G <- K^dagger G   (sweeps over -tau in reverse axis order)
 *   w = exp(i th) G, q = Im(conj(w) psi1),
 *   G <- w + 2 beta dt q psi1,  dL/dV(x) += q dt,  dL/dbeta += q dt |psi1|^2
 *   G <- K^dagger G.
 * Terminal: G_N = -2 c dV target, c = <target|psi_N>.
 * All pre-nonlinear states are stored (small test problems only). */
static void lap_rot_adj(const geom *g, cplx *G, double cD, double Om, double tau)
{
    for (int ax = g->ndim - 1; ax >= 0; --ax) sweep_axis(g, G, ax, cD, Om, -tau, 0);
}

int orc_adjoint(int ndim, const int *n, const double *lo, const double *hi,
                const double *psi0_ri, const double *V, const double *target_ri,
                double cD, double beta, double Om, double dt, int64_t nsteps,
                double *gradV, double *grad_beta, double *loss)
{
    if (ndim < 1 || ndim > 3 || nsteps < 0) return -3;
    geom g;
    geom_init(&g, ndim, n, lo, hi);
    const int64_t ns = g.nsite;
    const double dv = dvol(&g);
    const double tau = 0.5 * dt;
    cplx *psi = (cplx *)malloc(sizeof(cplx) * (size_t)ns);
    cplx *G = (cplx *)malloc(sizeof(cplx) * (size_t)ns);
    cplx *store = (cplx *)malloc(sizeof(cplx) * (size_t)ns * (size_t)(nsteps > 0 ? nsteps : 1));
    const cplx *tgt = (const cplx *)target_ri;
    memcpy(psi, psi0_ri, sizeof(cplx) * (size_t)ns);
    for (int64_t s = 0; s < nsteps; ++s) {
        lap_rot(&g, psi, cD, Om, tau, 0);
        memcpy(store + (size_t)s * ns, psi, sizeof(cplx) * (size_t)ns);
        solve_nl(&g, psi, V, beta, dt, 0);
        lap_rot(&g, psi, cD, Om, tau, 0);
    }
    double cr = 0.0, ci = 0.0;
    for (int64_t i = 0; i < ns; ++i) {
        cr += tgt[i].re * psi[i].re + tgt[i].im * psi[i].im;
        ci += tgt[i].re * psi[i].im - tgt[i].im * psi[i].re;
    }
    cr *= dv;
    ci *= dv;
    *loss = 1.0 - (cr * cr + ci * ci);
    for (int64_t i = 0; i < ns; ++i) {
        /* -2 c dV target */
        G[i].re = -2.0 * dv * (cr * tgt[i].re - ci * tgt[i].im);
        G[i].im = -2.0 * dv * (cr * tgt[i].im + ci * tgt[i].re);
        gradV[i] = 0.0;
    }
    double gb = 0.0;
    for (int64_t s = nsteps - 1; s >= 0; --s) {
        const cplx *p1 = store + (size_t)s * ns;
        lap_rot_adj(&g, G, cD, Om, tau);
        for (int64_t i = 0; i < ns; ++i) {
            const cplx p = p1[i];
            const double a2 = p.re * p.re + p.im * p.im;
            const double th = ((V ? V[i] : 0.0) + beta * a2) * dt;
            const double c = cos(th), sn = sin(th);
            const cplx w = {c * G[i].re - sn * G[i].im, c * G[i].im + sn * G[i].re};
            const double q = w.re * p.im - w.im * p.re;
            G[i].re = w.re + 2.0 * beta * dt * q * p.re;
            G[i].im = w.im + 2.0 * beta * dt * q * p.im;
            gradV[i] += q * dt;
            gb += q * dt * a2;
        }
        lap_rot_adj(&g, G, cD, Om, tau);
    }
    *grad_beta = gb;
    free(psi);
    free(G);
    free(store);
    geom_free(&g);
    return 0;
}

/* L = 1 - |<target|psi_N>|^2 after nsteps real-time Strang steps. */
double orc_overlap_loss(int ndim, const int *n, const double *lo, const double *hi,
                        const double *psi0_ri, const double *V, const double *target_ri,
                        double cD, double beta, double Om, double dt, int64_t nsteps)
{
    geom g;
    geom_init(&g, ndim, n, lo, hi);
    const int64_t ns = g.nsite;
    const double dv = dvol(&g);
    cplx *psi = (cplx *)malloc(sizeof(cplx) * (size_t)ns);
    memcpy(psi, psi0_ri, sizeof(cplx) * (size_t)ns);
    for (int64_t s = 0; s < nsteps; ++s) {
        lap_rot(&g, psi, cD, Om, 0.5 * dt, 0);
        solve_nl(&g, psi, V, beta, dt, 0);
        lap_rot(&g, psi, cD, Om, 0.5 * dt, 0);
    }
    const cplx *tgt = (const cplx *)target_ri;
    double cr = 0.0, ci = 0.0;
    for (int64_t i = 0; i < ns; ++i) {
        cr += tgt[i].re * psi[i].re + tgt[i].im * psi[i].im;
        ci += tgt[i].re * psi[i].im - tgt[i].im * psi[i].re;
    }
    cr *= dv;
    ci *= dv;
    free(psi);
    geom_free(&g);
    return 1.0 - (cr * cr + ci * ci);
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
