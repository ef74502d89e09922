/*
 * sfd.h -- C ABI of the MI355X (gfx950) Gross-Pitaevskii split-step library
 * (libsfdhip.so).  Plain pointers and sizes only; every device pointer is a
 * HIP device allocation owned by the caller (a Julia HIP array, a torch
 * tensor, or memory from sfd_alloc).
 *
 * Data layout is the reference's pen_x layout: complex values interleaved
 * (re, im), x fastest, extents (nx, ny[, nz]) (Fields.jl:104,147).  The
 * potential V is a real array of the same extents and precision.
 *
 * Every entry point returns SFD_OK (0) or a negative SFD_ERR_* code; the
 * message of the last failure on the calling host thread is returned by
 * sfd_last_error().  Work is enqueued asynchronously on the context's
 * stream, except where a function returns a host scalar (documented below),
 * which synchronises that stream.  One context per stream; contexts share no
 * mutable state.
 */
#ifndef SFD_H
#define SFD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFD_OK 0
#define SFD_ERR_NULL (-1)  /* a required pointer is NULL                   */
#define SFD_ERR_SHAPE (-2) /* unsupported extents (see sfd_ctx_create)     */
#define SFD_ERR_ARG (-3)   /* invalid scalar argument or option            */
#define SFD_ERR_OOM (-4)   /* device allocation failed                     */
#define SFD_ERR_HIP (-5)   /* a HIP runtime call or kernel launch failed   */

/* SFD_F32 applies every rotation (DFT constants, twiddles, kinetic and
 * nonlinear phases) as FP32 lifting steps, whose forward/inverse pairs carry
 * no systematic rounding error. */
#define SFD_F32 1  /* complex FP32 field and arithmetic, FP32 V, FP64 reductions */
#define SFD_F64 2  /* complex FP64 field and arithmetic, FP64 V                  */
#define SFD_F32X 3 /* complex FP32 field and FP32 V in memory, FP64 arithmetic   */

#define SFD_REAL_TIME 0 /* exp(i ...) flows: NumModelSplit1/2                */
#define SFD_IMAG_TIME 1 /* exp(...) flows + renormalisation after every step */

#define SFD_FORWARD (-1) /* unnormalised exp(-2 pi i jk/n)                   */
#define SFD_BACKWARD 1   /* exp(+2 pi i jk/n) scaled by 1/n                  */

typedef struct sfd_ctx sfd_ctx;

/* Create a context for one grid: Grid(size, bounds; FT) (Grids.jl:64-129)
 * plus the FFT plan state of Plan(f; t=FFTPlan()) (Plans.jl:821-999):
 * wavenumbers fftfreq(n, 2pi/dx), twiddles, and the workspace of the
 * reductions.  ndim is 2 or 3; n[d] is a power of two in [2, 2048] or any
 * other length in [3, 1024] (Grids.jl:64-82 takes any n): lines of other
 * lengths are transformed by Bluestein's chirp method on the power of two
 * m >= 2n - 1, at the cost of four m-point FFTs per round trip instead of
 * two n-point ones (HBM traffic unchanged).  box_min/box_max are the bounds
 * ((xmin,xmax),...).  stream is a
 * hipStream_t (NULL = the default stream).  Geometry (coordinates, dx,
 * wavenumbers) is kept in FP64 for both precisions. */
int sfd_ctx_create(sfd_ctx** out, int ndim, const int64_t* n, const double* box_min,
                   const double* box_max, int precision, void* stream);
int sfd_ctx_destroy(sfd_ctx* ctx);
int sfd_ctx_set_stream(sfd_ctx* ctx, void* stream);

/* Advance psi in place by nsteps split steps.  Replaces timeStep! of
 * NumModelSplit2 (order = 2, splitting.jl:149-154) and NumModelSplit1
 * (order = 1, splitting.jl:78-82), i.e. solveLapRot! (splitting.jl:156-203)
 * and solveNL! (splitting.jl:205-209), for nsteps consecutive calls.
 * V may be NULL (PotentialZero).  mode SFD_IMAG_TIME runs the imaginary-time
 * variant: every exponent loses its factor i and the field is renormalised
 * to norm 1 (normalize!, Fields.jl:196-200) after each step.
 * Sweep order of a Strang step: the innermost axis's two half-step flows
 * commute with the other axes' and meet around the nonlinear phase in one
 * pass, x, y, [z N z], x, y (2D without rotation: x, [y N y], x).  With
 * Omega = 0 every flow commutes, and a call of several steps shares flows
 * between steps: in real time the half-step flows that meet between two
 * nonlinear phases run as one flow of dt per axis, y(dt/2), z(dt/2),
 * x(dt/2) + N, y(dt), z(dt), x(dt) + N, ..., y(dt/2), z(dt/2), x(dt/2),
 * with the nonlinear phase fused into the x round trip: D passes over the
 * field per step.  In imaginary time (the norm of the state between
 * steps is needed) each step runs as x, y, [z N z], y, x and the x flows of
 * consecutive steps share one pass, [x x]: 2D - 2 passes per step.  Only the
 * grouping of commuting linear flows changes, so the results equal the
 * reference order's up to rounding. */
int sfd_split_step(sfd_ctx* ctx, void* psi, const void* V, double coeffD, double beta,
                   double Omega, double dt, int64_t nsteps, int order, int mode);

/* sfd_split_step without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments,
 * kernels and results as sfd_split_step; the phase tables of the call, the
 * renormalisation sums and partials live in a static 16 MiB device array of
 * the library (one per GPU, allocated with the code object, held until the
 * process ends) instead of the context's lazily allocated cache and buffers,
 * and are rebuilt on every call.  Grids whose tables and partials exceed
 * 16 MiB (e.g. rotating FP64 grids with nx*ny of 2^19 or more) return SFD_ERR_ARG.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (two streams, two host threads) race and must be ordered by the caller. */
int sfd_split_step_noalloc(sfd_ctx* ctx, void* psi, const void* V, double coeffD, double beta,
                           double Omega, double dt, int64_t nsteps, int order, int mode);

/* solveLapRot!(n, tau) (splitting.jl:156-203): the linear flow over tau,
 * one round trip per axis in the order x, y[, z].  mode SFD_IMAG_TIME
 * replaces exp(i ...) by exp(...) (no renormalisation here). */
int sfd_lap_rot(sfd_ctx* ctx, void* psi, double coeffD, double Omega, double tau, int mode);

/* One axis of solveLapRot!: mul_d!, multiplier, ldiv_d! for axis 0/1/2
 * (splitting.jl:188-191, 193-196, 198-201). */
int sfd_lap_rot_axis(sfd_ctx* ctx, void* psi, int axis, double coeffD, double Omega, double tau,
                     int mode);

/* solveNL!(n, dt) (splitting.jl:205-209): psi *= exp(-i (V + beta|psi|^2) dt)
 * pointwise; SFD_IMAG_TIME: psi *= exp(-(V + beta|psi|^2) dt).  V may be NULL. */
int sfd_solve_nl(sfd_ctx* ctx, void* psi, const void* V, double beta, double dt, int mode);

/* sum |psi|^2 accumulated in FP64 (the sum inside norm, Fields.jl:172-180).
 * Synchronises the stream. */
int sfd_norm2(sfd_ctx* ctx, const void* psi, double* out_sum);

/* normalize! (Fields.jl:196-200): psi /= norm(psi), norm = sqrt(sum|psi|^2)
 * sqrt(prod dx).  Writes the norm before scaling to *out_norm when non-NULL.
 * Synchronises the stream. */
int sfd_normalize(sfd_ctx* ctx, void* psi, double* out_norm);

/* energy (GrossPitaevskii.jl:61-136): out[0..4] = EOm, EDelta, Ebeta, E, mu
 * with mu = -EOm + EDelta + 2 Ebeta.  Kinetic and rotation sums are taken in
 * spectral space by Parseval (one forward sweep per axis) instead of the
 * reference's gradient fields.  V may be NULL.  Synchronises the stream. */
int sfd_energy(sfd_ctx* ctx, const void* psi, const void* V, double coeffD, double beta,
               double Omega, double* out5);

/* The loop of solve! (NumModels.jl:32-37) without its writers: nsteps
 * iterations of (energy(n), timeStep!(n)).  energies (host, 5*nsteps
 * doubles) receives at [5*s .. 5*s+4] the (EOmega, EDelta, Ebeta, E, mu) of
 * sfd_energy for the state before step s; the steps are those of
 * sfd_split_step.  In real time with Omega = 0 the sums come out of each
 * step's first-half sweeps, so the energy costs no extra pass over psi (the
 * y and z sums are taken after the x flow, which leaves them unchanged to
 * round-off); with rotation or in imaginary time (where the flows damp)
 * each step is preceded by sfd_energy.
 * Synchronises the stream (every 4096 steps and at return). */
int sfd_split_step_energy(sfd_ctx* ctx, void* psi, const void* V, double coeffD, double beta,
                          double Omega, double dt, int64_t nsteps, int order, int mode,
                          double* energies);

/* One directional transform (mul_x!/mul_y!/mul_z! with direction
 * SFD_FORWARD, ldiv_x!/ldiv_y!/ldiv_z! with SFD_BACKWARD; Plans.jl:1132-1250).
 * axis 0 = x, 1 = y, 2 = z.  Input and output are both in pen_x layout
 * (frequency k of a line stored where sample k sits); in == out is allowed. */
int sfd_fft_axis(sfd_ctx* ctx, const void* in, void* out, int axis, int direction);

/* Autotuner (no counterpart in the reference, whose FFT plans come from
 * FFTW/CUFFT planning, Plans.jl:847-848,991-993): for each axis, time every
 * compiled launch variant of the round-trip sweep (line split N = n1*n2,
 * workgroup shape) on a scratch field and keep the fastest for this context;
 * on the x axis and the innermost axis the double round trips of a step
 * ([x x], [d N d]) are timed too and get their own fastest variant.
 * scratch: a device buffer of nsites complex values (overwritten with
 * unit-modulus pseudo-random data, since sweep timings depend on the data),
 * or NULL to allocate one temporarily.  times_ms (optional, 24 floats)
 * receives the per-launch time
 * of variant v of axis a at [a*8 + v] (-1 where none).  Synchronises. */
int sfd_autotune(sfd_ctx* ctx, void* scratch, float* times_ms);

/* Current launch variant of an axis's single round trips: its index, the
 * number of variants compiled for this line length and precision, and
 * (optional) its shape {n1, n2, ts, g, pipelined, minimum waves per SIMD,
 * halves}: a line of halves*n1*n2 points, halves = 2 splitting it into
 * interleaved halves joined by a radix-2 stage. */
int sfd_ctx_get_variant(sfd_ctx* ctx, int axis, int* index, int* count, int* shape7);
/* Force a launch variant (index < count) on an axis's single and double
 * round trips; results are identical up to floating-point rounding. */
int sfd_ctx_set_variant(sfd_ctx* ctx, int axis, int index);

/* Sweep timer (no counterpart in the reference, whose benchmark scripts time
 * whole steps, benchmarks/scalability_gp.jl:36-51): with enable != 0 every
 * sweep launch of the context is bracketed by two HIP events on its stream
 * and its duration summed per slot: 0, 1, 2 = one-axis round trips of x, y,
 * z; 3 = the double round trip [d N d] of a Strang step; 4 = forward,
 * inverse and energy sweeps; 5 = the double round trip [x x] of chained
 * Strang steps.  The events cost a few microseconds between launches, so
 * time whole steps with the timer off. */
int sfd_ctx_time_sweeps(sfd_ctx* ctx, int enable);
/* ms_sum[6] and count[6] (either may be NULL) receive the summed durations
 * and launch counts per slot since the last reset; reset != 0 clears them.
 * Synchronises on the recorded events. */
int sfd_ctx_sweep_times(sfd_ctx* ctx, double* ms_sum, int64_t* count, int reset);

/* Batched imaginary-time ground states (no single counterpart in the
 * reference, which solves one field at a time with NumModelBackwardEuler,
 * backward-euler.jl:72-90): ninst independent problems of possibly
 * different extents.  sfd_batch_create takes 2D instances, n[2*i],
 * n[2*i+1] on boxes box[4*i..4*i+3] = xmin, xmax, ymin, ymax;
 * sfd_batch_create_nd takes ndim[i] = 2 or 3, n[3*i..3*i+2] and
 * box[6*i..6*i+5] (unused entries ignored).  Extents are powers of two in
 * [2, 1024]; at most 32 distinct (ndim, extents) shapes. */
typedef struct sfd_batch sfd_batch;
int sfd_batch_create(sfd_batch** out, int ninst, const int64_t* n, const double* box, int precision,
                     void* stream);
int sfd_batch_create_nd(sfd_batch** out, int ninst, const int* ndim, const int64_t* n, const double* box,
                        int precision, void* stream);
int sfd_batch_destroy(sfd_batch* b);

#define SFD_BATCH_GRAPH 1    /* replay chunks of check_every steps as hipGraphs */
#define SFD_BATCH_ESCALATE 2 /* FP32 storage: finish each instance in FP64 (below) */

/* Advance every field psi[i] (V[i] may be NULL; V itself may be NULL) with
 * the imaginary-time Strang split of sfd_split_step(mode = SFD_IMAG_TIME):
 * sweeps exp((coeffD xi^2 + rotation terms of Omega[i]) dt/2) per axis, then
 * exp(-(V + beta[i]|psi|^2) dt), renormalised to norm 1 after each step in
 * FP64, until ||psi_{n+1} - psi_n|| < tol (checked every check_every steps)
 * or max_steps.  Omega may be NULL (no rotation).  Fields of equal shape
 * share each launch; with SFD_BATCH_GRAPH a chunk of check_every steps is
 * captured once per set of unfinished shape groups and replayed.
 * With SFD_BATCH_ESCALATE (FP32 or F32X batches): the rounding of FP32
 * states keeps ||Delta psi|| near 1e-7, so each instance runs in its FP32
 * storage until its increment is below 1e-6, or below 1e-4 while shrinking by
 * less than 0.8 per check, and then continues from that state in FP64 (FP64
 * field, FP64 snapshot, FP64 ||Delta psi||) until tol; with tol >= 1e-6 an
 * instance may also finish in FP32.  Without rotation (a real Hamiltonian,
 * whose ground state is real up to a global phase) the state taken to FP64
 * is put back on the dominant phase phi0 of its initial state,
 * psi <- e^{i phi0} Re(e^{-i phi} psi), phi = arg(sum psi^2)/2 of the state:
 * the exact flow keeps that phase, and the FP32 rounding noise out of it
 * would decay only at the slow rate of phase gradients.  On return every field is normalised
 * (an escalated one rounded once from its FP64 state); psi64_out[i]
 * (optional, complex FP64 device buffers of the instance's size, entries
 * may be NULL) receives the final state in FP64; steps_out[i], dpsi_out[i]
 * and level_out[i] (optional) receive the steps taken, the last
 * ||Delta psi|| and 1 if the instance finished in FP64.  Synchronises. */
int sfd_batch_imag_solve(sfd_batch* b, void* const* psi, const void* const* V, const double* beta,
                         const double* Omega, double coeffD, double dt, int64_t max_steps, double tol,
                         int check_every, int flags, int64_t* steps_out, double* dpsi_out,
                         void* const* psi64_out, int* level_out);

/* sfd_batch_imag_solve with Omega = NULL, flags = use_graph ? SFD_BATCH_GRAPH : 0
 * and no FP64 outputs (the interface of the first ABI version). */
int sfd_batch_imag_run(sfd_batch* b, void* const* psi, const void* const* V, const double* beta,
                       double coeffD, double dt, int64_t max_steps, double tol, int check_every,
                       int use_graph, int64_t* steps_out, double* dpsi_out);

/* Autotuner of a batch: for every shape group and axis, times each compiled
 * launch variant of the batched sweep (identity round trips on the batch's
 * own scratch, the fields are not touched; the double round trip on x and on
 * the folded axis, as a solve runs them) and keeps the fastest, like
 * sfd_autotune for one context.  An FP32 or F32X batch also tunes the FP64
 * level its instances finish on (SFD_BATCH_ESCALATE), whose buffers it then
 * allocates.  times_ms (may be NULL) receives at least 24*ninst floats of
 * the batch's own precision: [24 g + 8 axis + variant] = ms per launch, -1
 * where there is none, g in the batch's internal group order.  No reference
 * counterpart (FFTW plan flags, Plans.jl:847-848, play this role there). */
int sfd_batch_autotune(sfd_batch* b, float* times_ms);

/* Adjoint of the real-time Strang split (no counterpart in the reference):
 * loss L = 1 - |<target|psi_N>|^2, <a|b> = sum conj(a) b dV, after nsteps
 * NumModelSplit2 steps from psi0 (psi0 is not modified).  Returns *loss and
 * dL/dbeta on the host and writes dL/dV(x) (FP64, nsites values) to the
 * device buffer gradV when non-NULL; psi_final (optional) receives psi_N.
 * The reverse sweep recomputes forward states from checkpoints taken every
 * checkpoint_every steps (<= 0: ceil(sqrt(nsteps))), on S side streams of
 * its own (S = 2 for fields of at most 16 MiB, else 1); device memory:
 * nsteps/checkpoint_every + (S + 1)*checkpoint_every + S + 2 fields.
 * Synchronises. */
int sfd_adjoint(sfd_ctx* ctx, const void* psi0, const void* V, const void* target, double coeffD,
                double beta, double Omega, double dt, int64_t nsteps, int64_t checkpoint_every,
                double* gradV, double* grad_beta, double* loss, void* psi_final);

/* <a|b> = sum conj(a) b dV (FP64, out2 = {re, im}).  Synchronises. */
int sfd_inner(sfd_ctx* ctx, const void* a, const void* b, double* out2);

/* out[j] = sum_x field[x] * basis[j*nsites + x] (FP64 device arrays; the
 * chain rule from dL/dV to the coefficients of V = sum_j a_j basis_j).
 * Synchronises. */
int sfd_project(sfd_ctx* ctx, const double* field, const double* basis, int nbasis, double* out);

/* Convenience memory helpers on the context's device and stream. */
int sfd_alloc(sfd_ctx* ctx, size_t bytes, void** out);
int sfd_free(sfd_ctx* ctx, void* ptr);
int sfd_upload(sfd_ctx* ctx, void* dst_device, const void* src_host, size_t bytes);
int sfd_download(sfd_ctx* ctx, void* dst_host, const void* src_device, size_t bytes);
int sfd_synchronize(sfd_ctx* ctx);

/* Last error message of the calling host thread ("" when none). */
const char* sfd_last_error(void);
/* ABI version (major * 100 + minor). */
int sfd_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SFD_H */
