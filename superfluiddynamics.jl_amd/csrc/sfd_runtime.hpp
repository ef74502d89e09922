// sfd_runtime.hpp -- state and helpers shared by the translation units of
// libsfdhip.so: the context and batch structures, error reporting, the
// reduction kernels' launchers and the per-precision operation tables.
//
// The library is compiled as four translation units: sfd.hip (the C ABI and
// the precision-independent helpers) and sfd_f32.hip / sfd_f64.hip /
// sfd_f32x.hip, each instantiating every precision-dependent routine of
// sfd_impl.hpp for one precision and exporting them as a PrecOps table.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sfd.h"
#include "sfd_variants.hpp"
#include "sfd_batch.hpp"
#include "sfd_adjoint.hpp"

namespace sfd {

// Records the message of a failure for sfd_last_error() and returns code.
int fail(int code, const char* fmt, ...);

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

// Reductions of per-workgroup partial sums (kernels in sfd.hip).
constexpr int kFinalizeThreads = 1024;

// Up to three independent fixed-order sums in one launch (the per-axis
// energy sums of a step): partial[i] holds K[i] rows of nb[i] partials,
// summed into out[i][0 .. K[i]).
struct FinalizeSets {
    const double* partial[3];
    long long nb[3];
    int K[3];
    double* out[3];
    int rows() const { return K[0] + K[1] + K[2]; }
};

// out[k] = sum_b partial[k*nb + b] for k < K (one workgroup, fixed order);
// with scale_out, also *scale_out = 1 / sqrt(out[0] * dv).
void launch_finalize(hipStream_t s, const double* partial, long long nb, int K, double* out, double dv = 0.0,
                     double* scale_out = nullptr);
void launch_finalize_sets(hipStream_t s, const FinalizeSets& f);

struct PhaseEntry {
    int axis;
    double tau, cD, Om;
    int imag;
    void* dev;
    int rows;
    size_t bytes;  // capacity of dev
};

// Launch geometry of one sweep.
struct Launch {
    dim3 grid, block;
    long long nblocks;
};

// Device memory of one call handed in by its caller instead of the
// context's lazily grown reduction buffers and phase-table cache (the
// allocation-free entry point): the pending scale and norm sums, the norm
// partials of the renormalising sweep, and the phase table of each axis for
// the call's single tau.
struct Workspace {
    double* scal;       // [0] pending scale, [1] sum |psi|^2
    double* partial;    // per-workgroup sums of the renormalising sweep
    void* phase[3];     // phase tables of the half step tau
    void* phase_dt[3];  // phase tables of the whole step dt (chained real-time steps), or null
    int rows[3];
    double tau;
};

// Sweep orders of the split step (split_impl):
//   kFoldNone   x, y, z, N, x, y, z: the order of splitting.jl:149-154;
//   kFoldInner  x, y, [z N z], x, y: the innermost axis's two flows meet
//               around the nonlinear phase in one double round trip;
//   kFoldChain  several steps at Omega = 0 share their flows: in real
//               time y(tau), z(tau), x(tau) + N, y(dt), z(dt), x(dt) + N,
//               ..., y(tau), z(tau), x(tau); where the state between two
//               steps is needed, x, y, [z N z], y, x per step, so that the
//               last x flow of a step and the first of the next meet in one
//               double round trip [x x].
enum FoldLevel : int { kFoldNone = 0, kFoldInner = 1, kFoldChain = 2 };

}  // namespace sfd

struct sfd_ctx {
    int ndim = 0;
    int n[3] = {1, 1, 1};
    // Bluestein axes (n not a power of two) transform on m = 2^ceil(log2(2n - 1))
    // points; elsewhere m = n
    int m[3] = {1, 1, 1};
    bool blu[3] = {false, false, false};
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1}, d[3] = {1, 1, 1};
    long long nsite = 1;
    int prec = SFD_F64;
    int device = 0;
    hipStream_t stream = nullptr;
    // per axis: launch variant of k_sweep and its twiddle table
    int var[3] = {0, 0, 0};
    sfd::VariantInfo vi[3] = {};
    void* tw[3] = {nullptr, nullptr, nullptr};
    // the variant, shape and twiddles of the axis's double round trips
    // ([d N d], [x x]), chosen apart from its single round trips
    int var2[3] = {0, 0, 0};
    sfd::VariantInfo vi2[3] = {};
    void* tw2[3] = {nullptr, nullptr, nullptr};
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
    std::vector<sfd::PhaseEntry> phases;
    bool pin_phases = false;    // no eviction: two streams read the cache
    // Sweep timer (sfd_ctx_time_sweeps): a HIP event before and after every
    // sweep launch, summed per slot: 0..2 one-axis round trips of x, y, z,
    // 3 the double round trip [d N d], 4 forward/inverse/energy sweeps,
    // 5 the double round trip [x x] of chained Strang steps.
    static constexpr int kTimerSlots = 6;
    static constexpr int kTimerPairs = 128;
    bool time_sweeps = false;
    std::vector<hipEvent_t> tev;  // 2 * kTimerPairs events, created on enable
    int tslot[kTimerPairs] = {};
    int tused = 0;                // pairs recorded and not yet summed
    double tsum[kTimerSlots] = {};
    long long tcount[kTimerSlots] = {};

    size_t csize() const { return prec == SFD_F64 ? 16 : 8; }
    double dvol() const
    {
        double v = 1.0;
        for (int a = 0; a < ndim; ++a) v *= d[a];
        return v;
    }
};

// ------------------------------------------------------------ batch solver
struct sfd_batch {
    static constexpr int kDiffBlocks = 64;
    static constexpr int kMaxKinds = 7;  // sweep descriptors of a group: a step and the chained sweeps
    static constexpr int kLevels = 2;    // 0: the batch's precision, 1: FP64 after escalation
    struct Inst {
        int ndim;
        int n[3];
        double lo[3], hi[3], d[3];
        long long nsite;
        long long soff;  // offset of the instance's sites in the batch's own buffers
        int group;
        int user;        // index in the caller's ordering
        double* geo[3];  // device FP64 wavenumbers [0, n) and coordinates [n, 2n) per axis
    };
    // One sweep of a step: axis, fused flags (kFlagScale / kFlagNL /
    // kFlagNorm / kFlagNormMid) and whether it is a double round trip
    // ([d N d], or [x x] of chained steps).
    struct Kind {
        int axis, flags;
        bool twice;
    };
    static constexpr int kFinalize = -1;  // in a sweep list: the renormalisation that ends a step
    struct Group {
        int ndim;
        int n[3];
        int first = 0, count = 0;
        // per level and axis: launch variant, its twiddles, tiles and x-tiles
        sfd::VariantInfo vi[kLevels][3] = {};
        int var[kLevels][3] = {};
        void* tw[kLevels][3] = {};
        long long ntiles[kLevels][3] = {};
        int ntx[kLevels][3] = {};
        // The sweeps of the current solve (kinds[k] has descriptor set k) and
        // sequences of them: `step` is one Strang step; chained steps run
        // `open`, then `body` once per step after the first, then `close`
        // (chain = false: the group does not chain its steps).
        std::vector<Kind> kinds;
        std::vector<int> step, open, body, close;
        bool chain = false;
    };
    // Device state of one precision level.
    struct Level {
        int prec = 0;                          // SFD_F32 / SFD_F32X / SFD_F64; 0 = not set up
        sfd::BatchInstDev* d_inst = nullptr;   // per instance: field, snapshot, partials
        sfd::SweepArgs* d_desc = nullptr;      // [kMaxKinds][ninst]
        int* d_done = nullptr;
        double* d_partial = nullptr;
        long long parts_per_inst_max = 0;      // partials reserved per instance (see batch setup)
        std::vector<long long> poff;           // per instance: offset of its partials
        void* d_snap = nullptr;                // snapshots, in the level's storage precision
        void* d_field = nullptr;               // level 1: the FP64 fields of escalated instances
        void* d_V = nullptr;                   // level 1: FP64 copies of the potentials
        std::vector<void*> phase;              // [ninst * 3] phase tables of the current solve
        std::vector<size_t> phase_bytes;
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
    long long sites = 0;
    Level lv[kLevels];
    double* d_scale = nullptr;
    double* d_snap_scale = nullptr;
    double* d_dpsi = nullptr;
    double* d_dprev = nullptr;
    int* d_esc = nullptr;
    int* d_real = nullptr;         // per instance: 1 when the Hamiltonian is real (no rotation)
    double* d_sq = nullptr;        // escalation: per-workgroup sums of psi^2, [ninst][2][kDiffBlocks]
    double* d_phi0 = nullptr;      // per instance: dominant phase of the initial state
    long long* d_steps = nullptr;
    double* d_dpart = nullptr;
    void** d_out64 = nullptr;      // per solve: the caller's FP64 outputs (internal order)
    // hipGraphs of one chunk, keyed by the active (group, level) mask
    struct CachedGraph {
        unsigned long long mask;
        int k;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<CachedGraph> graphs;
};

namespace sfd {

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

// ------------------------------------------- precision-independent helpers
double lin_coord(double a, double b, int n, int i);
double fft_freq(int n, double dx, int k);
int upload_sync(void* dst, const void* src, size_t bytes);
int phase_rows(const sfd_ctx* c, int axis, double Om);
int ensure_buf(sfd_ctx* c, double** buf, long long* cap, long long need);
int ensure_partial(sfd_ctx* c, long long need);
Launch sweep_geometry(const sfd_ctx* c, int axis, SweepArgs& a, bool persistent, bool twice = false);
long long sweep_nblocks(const sfd_ctx* c, int axis);
int check_ctx(sfd_ctx* c);
long long grid_blocks(long long n);
void combine_energy(const sfd_ctx* c, const double* kin, const double* rot, const double* pw, double cD,
                    double beta, double Om, double* out5);
int adjoint_side_streams(size_t field_bytes);
void release(sfd_ctx* c);
int timer_begin(sfd_ctx* c, int slot);
int timer_end(sfd_ctx* c);
int timer_drain(sfd_ctx* c);
void batch_geom(SweepArgs& a, const sfd_batch::Inst& in, const sfd_batch::Group& G, int level, int ax);
void batch_drop_graphs(sfd_batch* b);
void batch_release(sfd_batch* b);

// ------------------------------------------------ per-precision routines
// One table per precision (sfd_f32.hip, sfd_f64.hip, sfd_f32x.hip).
struct PrecOps {
    int (*init_tables)(sfd_ctx* c);
    int (*split)(sfd_ctx* c, void* psi, const void* V, double cD, double beta, double Om, double dt, int64_t nsteps,
                 int order, int mode, double* ehist, int fold, const Workspace* ws);
    int (*split_energy)(sfd_ctx* c, void* psi, const void* V, double cD, double beta, double Om, double dt,
                        int64_t nsteps, int order, int mode, double* energies);
    int (*round_trip_axis)(sfd_ctx* c, void* psi, int axis, double tau, double cD, double Om, int imag);
    int (*energy)(sfd_ctx* c, const void* psi, const void* V, double cD, double beta, double Om, double* out5);
    int (*fft)(sfd_ctx* c, const void* in, void* out, int axis, int dir);
    int (*norm2)(sfd_ctx* c, const void* psi, double* dev_out);
    int (*adjoint)(sfd_ctx* c, const void* psi0, const void* V, const void* target, double cD, double beta,
                   double Om, double dt, int64_t N, int64_t every, double* gradV, double* grad_beta, double* loss,
                   void* psi_final);
    int (*autotune)(sfd_ctx* c, void* scratch, float* times_out);
    int (*set_variant)(sfd_ctx* c, int axis, int idx, int which);  // which: 1 single, 2 double, 3 both
    int (*n_variants)(const sfd_ctx* c, int axis);
    int (*build_phase_into)(const sfd_ctx* c, hipStream_t stream, void* dev, int axis, double tau, double cD,
                            double Om, int imag, int rows);
    size_t table_entry_bytes;  // bytes of one phase-table entry
    // batch level routines (level 0: the batch's precision; 1: FP64)
    int (*batch_level_setup)(sfd_batch* b, int level);
    int (*batch_level_variant)(sfd_batch* b, int level, int group, int axis, int idx);
    int (*batch_level_bind)(sfd_batch* b, int level, void* const* psi, const void* const* V, const double* beta,
                            const double* Omega, double cD, double dt);
    int (*batch_level_steps)(sfd_batch* b, int level, int group, hipStream_t st, int k, double tol,
                             bool escalate);
    int (*batch_escalate)(sfd_batch* b, int group, hipStream_t st);
    int (*batch_phase0)(sfd_batch* b);
    int (*batch_finish)(sfd_batch* b);
    int (*batch_autotune)(sfd_batch* b, int level, float* times);
};

extern const PrecOps kOpsF32, kOpsF64, kOpsF32X;

inline const PrecOps& ops_for(int prec)
{
    return prec == SFD_F32 ? kOpsF32 : prec == SFD_F64 ? kOpsF64 : kOpsF32X;
}

}  // namespace sfd
