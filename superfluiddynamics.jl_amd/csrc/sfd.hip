// sfd.hip -- C ABI of libsfdhip.so (declared in include/sfd.h) and the
// precision-independent helpers of the library.
//
// The context replaces the reference's Grid + FFT Plan pair (Grids.jl:64-129,
// Plans.jl:958-999) for one field shape: it owns the FP64 geometry
// (coordinates, wavenumbers), the twiddle tables, a cache of phase-multiplier
// tables keyed by (axis, tau, coeffD, Omega, mode) and the reduction
// workspace.  Fields and potentials stay owned by the caller.  The
// precision-dependent work runs through the PrecOps table of the context's
// precision (sfd_f32.hip, sfd_f64.hip, sfd_f32x.hip).
#include "sfd_runtime.hpp"

#include <mutex>

// Device memory of sfd_split_step_noalloc: the pending scale and norm sums,
// the norm partials of the renormalising sweep and the phase tables of the
// call, in one statically sized array of every device the library runs on
// (held until the process ends).  Host code never uses its name as an
// address: each device's address comes from hipGetSymbolAddress once, at the
// first sfd_ctx_create on that device, and is kept in g_scratch_ptr.
static constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace sfd {

namespace {
thread_local std::string g_last_error;

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};
bool g_scratch_tried[kMaxDevices] = {};

// Looks the scratch array up on the current device (once per device).
void scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_tried[device]) return;
    g_scratch_tried[device] = true;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_ptr[device] = p;
    else (void)hipGetLastError();  // leave the device without scratch; the entry point reports it
}

void* scratch_ptr(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

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

// out[k] = sum_b partial[k*nb + b] for k < K, in a fixed order
// (deterministic: the order depends only on nb and the workgroup size).
// Each lane adds a strided subset with four independent accumulators, so its
// loads are in flight together; waves then combine through LDS.  Runs on
// one workgroup of kFinalizeThreads.

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

__global__ __launch_bounds__(kFinalizeThreads) void k_finalize_sets(const FinalizeSets f)
{
    int i = 0, k = blockIdx.x;
    while (k >= f.K[i]) k -= f.K[i++];
    finalize_rows(f.partial[i] + k * f.nb[i], f.nb[i], 1, f.out[i] + k);
}

void launch_finalize(hipStream_t s, const double* partial, long long nb, int K, double* out, double dv,
                     double* scale_out)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinalizeThreads), 0, s, partial, nb, K, out, dv, scale_out);
}

void launch_finalize_sets(hipStream_t s, const FinalizeSets& f)
{
    hipLaunchKernelGGL(k_finalize_sets, dim3(f.rows()), dim3(kFinalizeThreads), 0, s, f);
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

int upload_sync(void* dst, const void* src, size_t bytes)
{
    SFD_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SFD_OK;
}

// Rows of the phase table of an axis: n[other axis] with a rotation term
// (x and y axes, Omega != 0), else 1.
int phase_rows(const sfd_ctx* c, int axis, double Om)
{
    const bool rot = (Om != 0.0) && axis < 2 && c->ndim >= 2;
    return rot ? c->n[axis == 0 ? 1 : 0] : 1;
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

// Launch geometry of one sweep under the axis's current variant (of its
// double round trips with `twice`): returns grid/block and fills the line
// fields of the argument block.
Launch sweep_geometry(const sfd_ctx* c, int axis, SweepArgs& a, bool persistent, bool twice)
{
    const VariantInfo& v = twice ? c->vi2[axis] : c->vi[axis];
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

// Workgroups of a non-persistent sweep: an upper bound on the partial sums
// any sweep of the axis writes.
long long sweep_nblocks(const sfd_ctx* c, int axis)
{
    SweepArgs a{};
    return std::max(sweep_geometry(c, axis, a, false).nblocks, sweep_geometry(c, axis, a, false, true).nblocks);
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

// Side streams of the adjoint's recompute: 2 for fields of at most 16 MiB
// (their sweeps leave most CUs idle), else 1.  The environment variable
// SFD_ADJOINT_SIDE_STREAMS = 1 or 2, read at each call, forces the count, so
// that tests can run both orderings; the results do not depend on it.
int adjoint_side_streams(size_t field_bytes)
{
    if (const char* e = std::getenv("SFD_ADJOINT_SIDE_STREAMS")) {
        if (std::strcmp(e, "1") == 0) return 1;
        if (std::strcmp(e, "2") == 0) return 2;
    }
    return field_bytes <= ((size_t)16 << 20) ? 2 : 1;
}

void release(sfd_ctx* c)
{
    for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
    c->tev.clear();
    for (int ax = 0; ax < 3; ++ax) {
        if (c->tw[ax]) (void)hipFree(c->tw[ax]);
        if (c->tw2[ax]) (void)hipFree(c->tw2[ax]);
        if (c->xi[ax]) (void)hipFree(c->xi[ax]);
        if (c->coord[ax]) (void)hipFree(c->coord[ax]);
    }
    for (auto& p : c->phases) (void)hipFree(p.dev);
    if (c->partial) (void)hipFree(c->partial);
    if (c->epartial) (void)hipFree(c->epartial);
    if (c->ehist) (void)hipFree(c->ehist);
    if (c->scal) (void)hipFree(c->scal);
}

// Line geometry of instance `in` along `ax` under its group's variant at
// `level` (2D and 3D; the strided axes take x-columns in tiles of NL).
void batch_geom(SweepArgs& a, const sfd_batch::Inst& in, const sfd_batch::Group& G, int level, int ax)
{
    const long long nx = in.n[0], ny = in.n[1], nz = in.ndim == 3 ? in.n[2] : 1;
    if (ax == 0) {
        a.S = 1;
        a.nlines = ny * nz;
    } else {
        a.ncols = (int)nx;
        if (ax == 1) { a.S = nx; a.OS = nx * ny; }
        else         { a.S = nx * ny; a.OS = nx; }
        a.ntx = G.ntx[level][ax];
    }
    a.ntiles = G.ntiles[level][ax];
}

void batch_drop_graphs(sfd_batch* b)
{
    for (auto& cg : b->graphs) {
        if (cg.exec) (void)hipGraphExecDestroy(cg.exec);
        if (cg.graph) (void)hipGraphDestroy(cg.graph);
    }
    b->graphs.clear();
}

void batch_release(sfd_batch* b)
{
    batch_drop_graphs(b);
    for (auto& G : b->groups)
        for (int L = 0; L < sfd_batch::kLevels; ++L)
            for (int ax = 0; ax < 3; ++ax)
                if (G.tw[L][ax]) (void)hipFree(G.tw[L][ax]);
    for (auto& lv : b->lv) {
        for (void* p : lv.phase)
            if (p) (void)hipFree(p);
        void* ptrs[] = {lv.d_inst, lv.d_desc, lv.d_done, lv.d_partial, lv.d_snap, lv.d_field, lv.d_V};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
    }
    for (auto& in : b->inst)
        for (double* g : in.geo)
            if (g) (void)hipFree(g);
    void* ptrs[] = {b->d_scale, b->d_snap_scale, b->d_dpsi, b->d_dprev, b->d_esc,   b->d_steps,
                    b->d_dpart, b->d_out64,     b->d_real, b->d_sq,   b->d_phi0};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (b->ev) (void)hipEventDestroy(b->ev);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (hipEvent_t e : b->gdone) (void)hipEventDestroy(e);
    for (hipStream_t st : b->gstream) (void)hipStreamDestroy(st);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

namespace {

// Groups of equal (ndim, extents) in group-major internal order, the FP64
// geometry of every instance, the shared per-instance arrays, one stream per
// group, and level 0.
int batch_setup(sfd_batch* b)
{
    std::vector<sfd_batch::Inst> sorted;
    for (auto& in : b->inst) {
        int gidx = -1;
        for (size_t g = 0; g < b->groups.size(); ++g) {
            const auto& G = b->groups[g];
            if (G.ndim == in.ndim && G.n[0] == in.n[0] && G.n[1] == in.n[1] && G.n[2] == in.n[2]) gidx = (int)g;
        }
        if (gidx < 0) {
            sfd_batch::Group G;
            G.ndim = in.ndim;
            for (int a = 0; a < 3; ++a) G.n[a] = in.n[a];
            b->groups.push_back(G);
            gidx = (int)b->groups.size() - 1;
        }
        in.group = gidx;
    }
    if (b->groups.size() > 32) return fail(SFD_ERR_ARG, "batch: at most 32 distinct grid shapes (got %zu)",
                                           b->groups.size());
    for (size_t g = 0; g < b->groups.size(); ++g) {
        b->groups[g].first = (int)sorted.size();
        for (auto& in : b->inst)
            if (in.group == (int)g) sorted.push_back(in);
        b->groups[g].count = (int)sorted.size() - b->groups[g].first;
    }
    b->inst = sorted;
    b->sites = 0;
    for (auto& in : b->inst) {
        in.soff = b->sites;
        b->sites += in.nsite;
        for (int ax = 0; ax < in.ndim; ++ax) {
            const int n = in.n[ax];
            std::vector<double> t(2 * (size_t)n);
            for (int k = 0; k < n; ++k) {
                t[k] = fft_freq(n, in.d[ax], k);
                t[n + k] = lin_coord(in.lo[ax], in.hi[ax], n, k);
            }
            SFD_HIP(hipMalloc(&in.geo[ax], t.size() * sizeof(double)));
            SFD_TRY(upload_sync(in.geo[ax], t.data(), t.size() * sizeof(double)));
        }
    }
    const int ni = b->ninst;
    SFD_HIP(hipMalloc(&b->d_scale, ni * sizeof(double)));
    SFD_HIP(hipMalloc(&b->d_snap_scale, ni * sizeof(double)));
    SFD_HIP(hipMalloc(&b->d_dpsi, ni * sizeof(double)));
    SFD_HIP(hipMalloc(&b->d_dprev, ni * sizeof(double)));
    SFD_HIP(hipMalloc(&b->d_esc, ni * sizeof(int)));
    SFD_HIP(hipMalloc(&b->d_real, ni * sizeof(int)));
    SFD_HIP(hipMalloc(&b->d_sq, (size_t)ni * 2 * sfd_batch::kDiffBlocks * sizeof(double)));
    SFD_HIP(hipMalloc(&b->d_phi0, ni * sizeof(double)));
    SFD_HIP(hipMalloc(&b->d_steps, ni * sizeof(long long)));
    SFD_HIP(hipMalloc(&b->d_dpart, (size_t)ni * sfd_batch::kDiffBlocks * sizeof(double)));
    SFD_HIP(hipMalloc(&b->d_out64, ni * sizeof(void*)));
    SFD_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    for (size_t g = 0; g < b->groups.size(); ++g) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        SFD_HIP(hipStreamCreate(&st));
        b->gstream.push_back(st);
        SFD_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        b->gdone.push_back(ev);
    }
    b->lv[0].prec = b->prec;
    return ops_for(b->prec).batch_level_setup(b, 0);
}

// Enqueue k imaginary-time Strang steps of every (group, level) set in
// `mask` (bit 2 g + level), chained where the group chains; the last step
// is a convergence check, after which level-0 instances at their FP32
// floor move to FP64 when `escalate`.  Each size group runs on its own stream (forked from and
// joined back to the batch stream), so the groups' small launches overlap.
int batch_enqueue(sfd_batch* b, unsigned long long mask, int k, double tol, bool escalate)
{
    SFD_HIP(hipEventRecord(b->fork, b->stream));
    for (size_t g = 0; g < b->groups.size(); ++g) {
        const bool act[2] = {((mask >> (2 * g)) & 1ull) != 0, ((mask >> (2 * g + 1)) & 1ull) != 0};
        if (!act[0] && !act[1]) continue;
        hipStream_t st = b->gstream[g];
        SFD_HIP(hipStreamWaitEvent(st, b->fork, 0));
        // the two levels advance disjoint sets of instances (done flags), so
        // each runs its k steps in turn; escalation follows both checks
        for (int L = 0; L < sfd_batch::kLevels; ++L)
            if (act[L]) SFD_TRY(ops_for(b->lv[L].prec).batch_level_steps(b, L, (int)g, st, k, tol, L == 0 && escalate));
        if (escalate && act[0]) SFD_TRY(ops_for(b->prec).batch_escalate(b, (int)g, st));
        SFD_HIP(hipEventRecord(b->gdone[g], st));
        SFD_HIP(hipStreamWaitEvent(b->stream, b->gdone[g], 0));
    }
    return SFD_OK;
}

// The step of each group for this solve, and its chained form.  The first
// sweep of a step applies the pending renormalisation, the last one sums
// the norm, and a renormalisation (k_batch_finalize) ends it.
// - With rotation (K_x and K_y do not commute): x, y + N, x, y in 2D;
//   x, y, [z N z], x, y in 3D.  Not chained.
// - Without: x, [y N y], x in 2D (the folded Strang step); x, y, [z N z],
//   y, x in 3D.  Chained steps share their x flows: x, [y N y], [x x],
//   [y N y], ..., x in 2D and x, y, [z N z], y, [x x], y, ..., y, x in 3D,
//   where [x x] sums the norm of the state between its round trips (the one
//   that ends a step) and the sweep after it applies the renormalisation.
void batch_kinds(sfd_batch* b, const double* Omega)
{
    using K = sfd_batch::Kind;
    constexpr int F = sfd_batch::kFinalize;
    for (auto& G : b->groups) {
        bool rot = false;
        for (int i = G.first; i < G.first + G.count; ++i)
            rot = rot || (Omega && Omega[b->inst[i].user] != 0.0);
        G.chain = !rot;
        G.open.clear();
        G.body.clear();
        G.close.clear();
        if (G.ndim == 3 && rot) {
            G.kinds = {K{0, kFlagScale, false}, K{1, 0, false}, K{2, kFlagNL, true}, K{0, 0, false},
                       K{1, kFlagNorm, false}};
            G.step = {0, 1, 2, 3, 4, F};
        } else if (G.ndim == 3) {
            G.kinds = {K{0, kFlagScale, false}, K{1, 0, false},          K{2, kFlagNL, true},
                       K{1, 0, false},          K{0, kFlagNorm, false},  K{0, kFlagNormMid, true},
                       K{1, kFlagScale, false}};
            G.step = {0, 1, 2, 3, 4, F};
            G.open = {0, 1, 2, 3};
            G.body = {5, F, 6, 2, 3};
            G.close = {4, F};
        } else if (rot) {
            G.kinds = {K{0, kFlagScale, false}, K{1, kFlagNL, false}, K{0, 0, false}, K{1, kFlagNorm, false}};
            G.step = {0, 1, 2, 3, F};
        } else {
            G.kinds = {K{0, kFlagScale, false}, K{1, kFlagNL, true}, K{0, kFlagNorm, false},
                       K{0, kFlagNormMid, true}, K{1, kFlagNL | kFlagScale, true}};
            G.step = {0, 1, 2, F};
            G.open = {0, 1};
            G.body = {3, F, 4};
            G.close = {2, F};
        }
    }
}

int batch_solve(sfd_batch* b, void* const* psi, const void* const* V, const double* beta, const double* Omega,
                double cD, double dt, int64_t max_steps, double tol, int check_every, int flags, int64_t* steps_out,
                double* dpsi_out, void* const* psi64_out, int* level_out)
{
    const int ni = b->ninst;
    const bool escalate = (flags & SFD_BATCH_ESCALATE) && b->prec != SFD_F64;
    const bool use_graph = (flags & SFD_BATCH_GRAPH) != 0;
    for (int i = 0; i < ni; ++i)
        if (!psi[i]) return fail(SFD_ERR_NULL, "batch: field %d is null", i);
    if (escalate && !b->lv[1].prec) {  // marked set up only once it is
        SFD_TRY(ops_for(SFD_F64).batch_level_setup(b, 1));
        b->lv[1].prec = SFD_F64;
    }
    batch_kinds(b, Omega);
    batch_drop_graphs(b);
    // order after the caller's pending work
    SFD_HIP(hipEventRecord(b->ev, b->user_stream));
    SFD_HIP(hipStreamWaitEvent(b->stream, b->ev, 0));
    SFD_TRY(ops_for(b->prec).batch_level_bind(b, 0, psi, V, beta, Omega, cD, dt));
    if (escalate) SFD_TRY(ops_for(SFD_F64).batch_level_bind(b, 1, psi, V, beta, Omega, cD, dt));
    std::vector<void*> out64(ni, nullptr);
    std::vector<int> real(ni, 1);
    for (int i = 0; i < ni; ++i) {
        if (psi64_out) out64[i] = psi64_out[b->inst[i].user];
        if (Omega && Omega[b->inst[i].user] != 0.0) real[i] = 0;
    }
    SFD_HIP(hipMemcpyAsync(b->d_out64, out64.data(), ni * sizeof(void*), hipMemcpyHostToDevice, b->stream));
    SFD_HIP(hipMemcpyAsync(b->d_real, real.data(), ni * sizeof(int), hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(k_batch_reset, dim3((ni + 255) / 256), dim3(256), 0, b->stream, b->d_scale, b->lv[0].d_done,
                       escalate ? b->lv[1].d_done : (int*)nullptr, b->d_steps, b->d_dpsi, b->d_dprev, b->d_esc, ni);
    SFD_HIP(hipGetLastError());
    if (escalate) SFD_TRY(ops_for(b->prec).batch_phase0(b));
    std::vector<int> done0(ni, 0), done1(ni, 1);
    unsigned long long mask = 0;
    for (size_t g = 0; g < b->groups.size(); ++g) mask |= 1ull << (2 * g);
    int64_t left = max_steps;
    while (left > 0 && mask) {
        const int k = (int)(left < check_every ? left : check_every);
        if (use_graph && k == check_every) {
            sfd_batch::CachedGraph* cg = nullptr;
            for (auto& c : b->graphs)
                if (c.mask == mask && c.k == k) cg = &c;
            if (!cg) {
                if (b->graphs.size() >= 8) batch_drop_graphs(b);
                SFD_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                int rc = batch_enqueue(b, mask, k, tol, escalate);
                hipGraph_t graph = nullptr;
                hipError_t e = hipStreamEndCapture(b->stream, &graph);
                if (rc != SFD_OK) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return rc;
                }
                if (e != hipSuccess) return fail(SFD_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
                hipGraphExec_t exec = nullptr;
                e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                if (e != hipSuccess) {
                    (void)hipGraphDestroy(graph);
                    return fail(SFD_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
                }
                b->graphs.push_back({mask, k, graph, exec});
                cg = &b->graphs.back();
            }
            SFD_HIP(hipGraphLaunch(cg->exec, b->stream));
        } else {
            SFD_TRY(batch_enqueue(b, mask, k, tol, escalate));
        }
        left -= k;
        SFD_HIP(hipMemcpyAsync(done0.data(), b->lv[0].d_done, ni * sizeof(int), hipMemcpyDeviceToHost, b->stream));
        if (escalate)
            SFD_HIP(hipMemcpyAsync(done1.data(), b->lv[1].d_done, ni * sizeof(int), hipMemcpyDeviceToHost, b->stream));
        SFD_HIP(hipStreamSynchronize(b->stream));
        mask = 0;
        for (size_t g = 0; g < b->groups.size(); ++g) {
            const auto& G = b->groups[g];
            for (int i = G.first; i < G.first + G.count; ++i) {
                if (!done0[i]) mask |= 1ull << (2 * g);
                if (escalate && !done1[i]) mask |= 1ull << (2 * g + 1);
            }
        }
    }
    // leave every field normalised, and in FP64 where asked
    SFD_TRY(ops_for(b->prec).batch_finish(b));
    std::vector<long long> st(ni);
    std::vector<double> dp(ni);
    std::vector<int> esc(ni);
    SFD_HIP(hipMemcpyAsync(st.data(), b->d_steps, ni * sizeof(long long), hipMemcpyDeviceToHost, b->stream));
    SFD_HIP(hipMemcpyAsync(dp.data(), b->d_dpsi, ni * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    SFD_HIP(hipMemcpyAsync(esc.data(), b->d_esc, ni * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    SFD_HIP(hipStreamSynchronize(b->stream));
    for (int i = 0; i < ni; ++i) {
        const int u = b->inst[i].user;
        if (steps_out) steps_out[u] = st[i];
        if (dpsi_out) dpsi_out[u] = dp[i];
        if (level_out) level_out[u] = esc[i] == 2 ? 1 : 0;
    }
    return SFD_OK;
}

// Instance records of sfd_batch_create_nd (bounds checked, extents checked).
int batch_create(sfd_batch** out, int ninst, const int* ndim, const int64_t* n, const double* box, int precision,
                 void* stream)
{
    g_last_error.clear();
    if (!out || !n || !box) return fail(SFD_ERR_NULL, "sfd_batch_create: null argument");
    *out = nullptr;
    if (ninst < 1) return fail(SFD_ERR_ARG, "sfd_batch_create: ninst must be >= 1");
    if (precision != SFD_F32 && precision != SFD_F64 && precision != SFD_F32X)
        return fail(SFD_ERR_ARG, "precision must be SFD_F32, SFD_F64 or SFD_F32X (got %d)", precision);
    std::vector<sfd_batch::Inst> inst;
    for (int i = 0; i < ninst; ++i) {
        sfd_batch::Inst in{};
        in.ndim = ndim ? ndim[i] : 2;
        if (in.ndim != 2 && in.ndim != 3) return fail(SFD_ERR_SHAPE, "instance %d: ndim must be 2 or 3", i);
        in.nsite = 1;
        for (int a = 0; a < 3; ++a) {
            in.n[a] = 1;
            in.lo[a] = 0.0;
            in.hi[a] = 1.0;
            in.d[a] = 1.0;
            in.geo[a] = nullptr;
        }
        for (int a = 0; a < in.ndim; ++a) {
            const int64_t m = ndim ? n[3 * i + a] : n[2 * i + a];
            if (m < 2 || m > 1024 || (m & (m - 1)))
                return fail(SFD_ERR_SHAPE, "instance %d axis %d: n = %lld is not a power of two in [2, 1024]", i, a,
                            (long long)m);
            in.n[a] = (int)m;
            in.lo[a] = ndim ? box[6 * i + 2 * a] : box[4 * i + 2 * a];
            in.hi[a] = ndim ? box[6 * i + 2 * a + 1] : box[4 * i + 2 * a + 1];
            if (!(in.hi[a] > in.lo[a]) || !std::isfinite(in.lo[a]) || !std::isfinite(in.hi[a]))
                return fail(SFD_ERR_ARG, "instance %d axis %d: bounds must satisfy min < max", i, a);
            in.d[a] = (in.hi[a] - in.lo[a]) / (double)m;
            in.nsite *= m;
        }
        in.group = -1;
        in.user = i;
        inst.push_back(in);
    }
    auto* b = new sfd_batch();
    b->ninst = ninst;
    b->prec = precision;
    b->user_stream = static_cast<hipStream_t>(stream);
    b->inst = inst;
    hipError_t e = hipGetDevice(&b->device);
    if (e == hipSuccess) e = hipStreamCreate(&b->stream);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        batch_release(b);
        delete b;
        return fail(SFD_ERR_HIP, "sfd_batch_create: %s", hipGetErrorString(e));
    }
    const int rc = batch_setup(b);
    if (rc != SFD_OK) {
        batch_release(b);
        delete b;
        return rc;
    }
    *out = b;
    return SFD_OK;
}

}  // namespace

// Sweep timer: pair k uses events 2k (before the launch) and 2k + 1 (after
// it), both on the context's stream; a full pool is summed before reuse.
int timer_begin(sfd_ctx* c, int slot)
{
    if (c->tused == sfd_ctx::kTimerPairs) SFD_TRY(timer_drain(c));
    c->tslot[c->tused] = slot;
    SFD_HIP(hipEventRecord(c->tev[2 * c->tused], c->stream));
    return SFD_OK;
}

int timer_end(sfd_ctx* c)
{
    SFD_HIP(hipEventRecord(c->tev[2 * c->tused + 1], c->stream));
    ++c->tused;
    return SFD_OK;
}

int timer_drain(sfd_ctx* c)
{
    for (int k = 0; k < c->tused; ++k) {
        SFD_HIP(hipEventSynchronize(c->tev[2 * k + 1]));
        float ms = 0.f;
        SFD_HIP(hipEventElapsedTime(&ms, c->tev[2 * k], c->tev[2 * k + 1]));
        c->tsum[c->tslot[k]] += ms;
        c->tcount[c->tslot[k]] += 1;
    }
    c->tused = 0;
    return SFD_OK;
}

}  // namespace sfd

using namespace sfd;

// ===================================================================== C ABI
extern "C" {

const char* sfd_last_error(void) { return g_last_error.c_str(); }

int sfd_version(void) { return 101; }

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
        const bool pow2 = m >= 2 && (m & (m - 1)) == 0;
        if (m < 2 || (pow2 && m > 2048) || (!pow2 && m > 1024))
            return fail(SFD_ERR_SHAPE,
                        "n[%d] = %lld: line lengths must be powers of two in [2, 2048] or other lengths in [3, 1024]",
                        a, (long long)m);
        if (!(box_max[a] > box_min[a]) || !std::isfinite(box_min[a]) || !std::isfinite(box_max[a]))
            return fail(SFD_ERR_ARG, "axis %d: bounds must satisfy min < max (Grids.jl:69-76)", a);
    }
    auto* c = new sfd_ctx();
    c->ndim = ndim;
    c->prec = precision;
    c->stream = static_cast<hipStream_t>(stream);
    for (int a = 0; a < ndim; ++a) {
        c->n[a] = (int)n[a];
        c->blu[a] = (n[a] & (n[a] - 1)) != 0;
        c->m[a] = c->n[a];
        if (c->blu[a])
            for (c->m[a] = 8; c->m[a] < 2 * c->n[a] - 1;) c->m[a] *= 2;
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
    scratch_setup(c->device);
    int rc = ops_for(c->prec).init_tables(c);
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
    // Tables, partials and workspaces are reused in stream order: work still
    // queued on the old stream must end before the new stream touches them.
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (s != ctx->stream) SFD_HIP(hipStreamSynchronize(ctx->stream));
    ctx->stream = s;
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
    return ops_for(ctx->prec).split(ctx, psi, V, coeffD, beta, Omega, dt, nsteps, order, mode, nullptr, kFoldChain, nullptr);
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
    return ops_for(ctx->prec).split_energy(ctx, psi, V, coeffD, beta, Omega, dt, nsteps, order, mode,
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
    return ops_for(ctx->prec).round_trip_axis(ctx, psi, axis, tau, coeffD, Omega, imag);
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
    int rc = ops_for(ctx->prec).norm2(ctx, psi, ctx->scal + 1);
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
    return ops_for(ctx->prec).energy(ctx, psi, V, coeffD, beta, Omega, out5);
}

int sfd_fft_axis(sfd_ctx* ctx, const void* in, void* out, int axis, int direction)
{
    SFD_TRY(check_ctx(ctx));
    if (!in || !out) return fail(SFD_ERR_NULL, "sfd_fft_axis: null argument");
    if (axis < 0 || axis >= ctx->ndim) return fail(SFD_ERR_ARG, "axis %d out of range", axis);
    if (direction != SFD_FORWARD && direction != SFD_BACKWARD) return fail(SFD_ERR_ARG, "direction must be -1 or 1");
    return ops_for(ctx->prec).fft(ctx, in, out, axis, direction);
}

int sfd_batch_create(sfd_batch** out, int ninst, const int64_t* n, const double* box, int precision, void* stream)
{
    return batch_create(out, ninst, nullptr, n, box, precision, stream);
}

int sfd_batch_create_nd(sfd_batch** out, int ninst, const int* ndim, const int64_t* n, const double* box,
                        int precision, void* stream)
{
    if (!ndim) return fail(SFD_ERR_NULL, "sfd_batch_create_nd: ndim is null");
    return batch_create(out, ninst, ndim, n, box, precision, stream);
}

int sfd_batch_autotune(sfd_batch* b, float* times_ms)
{
    g_last_error.clear();
    if (!b) return fail(SFD_ERR_NULL, "sfd_batch_autotune: null batch");
    SFD_HIP(hipSetDevice(b->device));
    SFD_TRY(ops_for(b->prec).batch_autotune(b, 0, times_ms));
    if (b->prec == SFD_F64) return SFD_OK;
    // the FP64 level that escalated instances finish on has its own variants
    if (!b->lv[1].prec) {
        SFD_TRY(ops_for(SFD_F64).batch_level_setup(b, 1));
        b->lv[1].prec = SFD_F64;
    }
    return ops_for(SFD_F64).batch_autotune(b, 1, nullptr);
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

int sfd_batch_imag_solve(sfd_batch* b, void* const* psi, const void* const* V, const double* beta,
                         const double* Omega, double coeffD, double dt, int64_t max_steps, double tol,
                         int check_every, int flags, int64_t* steps_out, double* dpsi_out, void* const* psi64_out,
                         int* level_out)
{
    g_last_error.clear();
    if (!b) return fail(SFD_ERR_NULL, "null batch");
    if (!psi || !beta) return fail(SFD_ERR_NULL, "sfd_batch_imag_solve: null argument");
    if (max_steps < 0 || check_every < 1) return fail(SFD_ERR_ARG, "max_steps >= 0 and check_every >= 1 required");
    if (flags & ~(SFD_BATCH_GRAPH | SFD_BATCH_ESCALATE)) return fail(SFD_ERR_ARG, "unknown flags 0x%x", flags);
    if (!std::isfinite(coeffD) || !std::isfinite(dt) || !std::isfinite(tol))
        return fail(SFD_ERR_ARG, "non-finite scalar argument");
    for (int i = 0; i < b->ninst; ++i)
        if (!std::isfinite(beta[i]) || (Omega && !std::isfinite(Omega[i])))
            return fail(SFD_ERR_ARG, "instance %d: non-finite beta or Omega", i);
    SFD_HIP(hipSetDevice(b->device));
    return batch_solve(b, psi, V, beta, Omega, coeffD, dt, max_steps, tol, check_every, flags, steps_out, dpsi_out,
                       psi64_out, level_out);
}

int sfd_batch_imag_run(sfd_batch* b, void* const* psi, const void* const* V, const double* beta, double coeffD,
                       double dt, int64_t max_steps, double tol, int check_every, int use_graph, int64_t* steps_out,
                       double* dpsi_out)
{
    return sfd_batch_imag_solve(b, psi, V, beta, nullptr, coeffD, dt, max_steps, tol, check_every,
                                use_graph ? SFD_BATCH_GRAPH : 0, steps_out, dpsi_out, nullptr, nullptr);
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
    return ops_for(ctx->prec).adjoint(ctx, psi0, V, target, coeffD, beta, Omega, dt, nsteps, checkpoint_every,
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
    return ops_for(ctx->prec).autotune(ctx, scratch, times_ms);
}

int sfd_ctx_get_variant(sfd_ctx* ctx, int axis, int* index, int* count, int* shape7)
{
    if (!ctx || !index || !count) return fail(SFD_ERR_NULL, "sfd_ctx_get_variant: null argument");
    if (axis < 0 || axis >= ctx->ndim) return fail(SFD_ERR_ARG, "axis %d out of range", axis);
    *index = ctx->var[axis];
    *count = ops_for(ctx->prec).n_variants(ctx, axis);
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
    return ops_for(ctx->prec).set_variant(ctx, axis, index, 3);
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

int sfd_split_step_noalloc(sfd_ctx* ctx, void* psi, const void* V, double coeffD, double beta, double Omega,
                           double dt, int64_t nsteps, int order, int mode)
{
    SFD_TRY(check_ctx(ctx));
    if (!psi) return fail(SFD_ERR_NULL, "sfd_split_step_noalloc: psi is null");
    if (order != 1 && order != 2) return fail(SFD_ERR_ARG, "order must be 1 or 2 (got %d)", order);
    if (mode != SFD_REAL_TIME && mode != SFD_IMAG_TIME) return fail(SFD_ERR_ARG, "mode must be 0 or 1");
    if (nsteps < 0) return fail(SFD_ERR_ARG, "nsteps must be >= 0");
    if (!std::isfinite(coeffD) || !std::isfinite(beta) || !std::isfinite(Omega) || !std::isfinite(dt))
        return fail(SFD_ERR_ARG, "non-finite scalar argument");
    unsigned char* base = static_cast<unsigned char*>(scratch_ptr(ctx->device));
    if (!base) return fail(SFD_ERR_HIP, "sfd_split_step_noalloc: no scratch array on device %d", ctx->device);
    const PrecOps& ops = ops_for(ctx->prec);
    // layout: [scale and sums: 256 B][norm partials][phase table of each
    // axis for tau][for dt, when the call chains real-time steps]
    size_t need = 256;
    long long nparts = 0;
    for (int ax = 0; ax < ctx->ndim; ++ax) nparts = std::max(nparts, sweep_nblocks(ctx, ax));
    const size_t part_off = need;
    if (mode == SFD_IMAG_TIME) need += align256((size_t)nparts * sizeof(double));
    const bool whole = mode == SFD_REAL_TIME && order == 2 && Omega == 0.0 && nsteps >= 2 && ctx->ndim >= 2;
    Workspace ws{};
    size_t tab_off[3] = {0, 0, 0}, dt_off[3] = {0, 0, 0};
    for (int ax = 0; ax < ctx->ndim; ++ax) {
        ws.rows[ax] = phase_rows(ctx, ax, Omega);
        const size_t tb = align256((size_t)ws.rows[ax] * ctx->n[ax] * ops.table_entry_bytes);
        tab_off[ax] = need;
        need += tb;
        if (whole) {
            dt_off[ax] = need;
            need += tb;
        }
    }
    if (need > kScratchBytes)
        return fail(SFD_ERR_ARG,
                    "sfd_split_step_noalloc: this grid needs %zu bytes of scratch (phase tables and partial sums), "
                    "more than the %zu of the library's static array; use sfd_split_step",
                    need, kScratchBytes);
    ws.scal = reinterpret_cast<double*>(base);
    ws.partial = reinterpret_cast<double*>(base + part_off);
    const double tau = order == 2 ? 0.5 * dt : dt;
    const int imag = mode == SFD_IMAG_TIME;
    ws.tau = tau;
    for (int ax = 0; ax < ctx->ndim; ++ax) {
        ws.phase[ax] = base + tab_off[ax];
        SFD_TRY(ops.build_phase_into(ctx, ctx->stream, ws.phase[ax], ax, tau, coeffD, Omega, imag, ws.rows[ax]));
        if (whole) {
            ws.phase_dt[ax] = base + dt_off[ax];
            SFD_TRY(ops.build_phase_into(ctx, ctx->stream, ws.phase_dt[ax], ax, dt, coeffD, Omega, imag,
                                         ws.rows[ax]));
        }
    }
    return ops.split(ctx, psi, V, coeffD, beta, Omega, dt, nsteps, order, mode, nullptr, kFoldChain, &ws);
}

int sfd_ctx_time_sweeps(sfd_ctx* ctx, int enable)
{
    SFD_TRY(check_ctx(ctx));
    if (enable && ctx->tev.empty()) {
        ctx->tev.resize(2 * sfd_ctx::kTimerPairs, nullptr);
        for (auto& e : ctx->tev) SFD_HIP(hipEventCreate(&e));
    }
    if (!enable && ctx->tused > 0) SFD_TRY(timer_drain(ctx));
    ctx->time_sweeps = enable != 0;
    return SFD_OK;
}

int sfd_ctx_sweep_times(sfd_ctx* ctx, double* ms_sum, int64_t* count, int reset)
{
    SFD_TRY(check_ctx(ctx));
    if (ctx->tused > 0) SFD_TRY(timer_drain(ctx));
    for (int k = 0; k < sfd_ctx::kTimerSlots; ++k) {
        if (ms_sum) ms_sum[k] = ctx->tsum[k];
        if (count) count[k] = ctx->tcount[k];
        if (reset) {
            ctx->tsum[k] = 0.0;
            ctx->tcount[k] = 0;
        }
    }
    return SFD_OK;
}

}  // extern "C"

