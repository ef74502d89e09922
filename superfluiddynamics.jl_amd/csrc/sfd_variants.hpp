// sfd_variants.hpp -- the compiled launch variants of k_sweep and their
// dispatch.  A variant fixes the line split N = N1*N2 and the workgroup
// shape: TS x-columns per lane group (1 = contiguous x axis) and G groups,
// i.e. NL = TS*G lines and NL*N2 threads per workgroup.  The autotuner in
// sfd.hip times every variant of an axis once per context and keeps the
// fastest; index 0 is the default.
#pragma once
#include "sfd_sweep.hpp"

namespace sfd {

struct VariantInfo {
    int n1, n2, ts, g, pipe, waves, h;
};

// pipe = 1: the round trip runs as the persistent, software-pipelined
// k_sweep_pipe (the other modes use k_sweep with the same shape).
// W: minimum wavefronts per SIMD the register allocation must allow.
// H = 2: lines of 2*N1*N2 points as two interleaved halves of N1*N2 points
// joined by a radix-2 stage across lane pairs (H*N2 lanes per line), which
// keeps N1 = 16 points per lane for 512-point lines.
// T = 1: the round trips read their tables from LDS (k_sweep TL).
template <int N1_, int N2_, int TS_, int G_, int P_ = 0, int W_ = 1, int H_ = 1, int T_ = 0>
struct V {
    static constexpr int N1 = N1_, N2 = N2_, TS = TS_, G = G_, P = P_, W = W_, H = H_;
    static constexpr bool T = T_ != 0;
    static_assert(P_ == 0 || H_ == 1, "the pipelined kernel takes whole lines");
};
template <typename... Vs>
struct VL {
    static constexpr int size = sizeof...(Vs);
};

// Precision of a context: storage scalar S (HBM) and compute scalar R.
struct PF32 {
    using S = float;
    using R = float;
};
struct PF64 {
    using S = double;
    using R = double;
};
struct PF32X {  // complex FP32 fields, FP64 arithmetic
    using S = float;
    using R = double;
};

// Per (precision, N): contiguous (x axis) and strided (y/z axes) variants.
// Every variant keeps its exchange buffer within 64 KiB of LDS and at least
// one full wavefront per workgroup.
template <typename P, int N>
struct Variants;

// the lists are passed parenthesised, so that their commas survive the macro
template <typename T>
struct Unparen;
template <typename T>
struct Unparen<void(T)> {
    using type = T;
};

#define SFD_VARIANTS(R_, N_, CONTIG, STRIDED)                  \
    template <>                                                \
    struct Variants<R_, N_> {                                  \
        using contig = typename Unparen<void CONTIG>::type;    \
        using strided = typename Unparen<void STRIDED>::type;  \
    };

// ---------------------------------------------------------------- FP32
SFD_VARIANTS(PF32, 2, (VL<V<2, 1, 1, 256>>), (VL<V<2, 1, 16, 4>>))
SFD_VARIANTS(PF32, 4, (VL<V<2, 2, 1, 128>>), (VL<V<2, 2, 16, 2>>))
SFD_VARIANTS(PF32, 8, (VL<V<4, 2, 1, 128>>), (VL<V<4, 2, 16, 2>>))
SFD_VARIANTS(PF32, 16, (VL<V<4, 4, 1, 64>>), (VL<V<4, 4, 16, 1>>))
SFD_VARIANTS(PF32, 32, (VL<V<8, 4, 1, 64>>), (VL<V<8, 4, 16, 1>>))
SFD_VARIANTS(PF32, 64, (VL<V<8, 8, 1, 32>, V<16, 4, 1, 64>>), (VL<V<8, 8, 16, 1>, V<8, 8, 32, 1>>))
SFD_VARIANTS(PF32, 128, (VL<V<16, 8, 1, 32>, V<16, 8, 1, 16>, V<32, 4, 1, 32>, V<16, 8, 1, 16, 1>>),
             (VL<V<16, 8, 16, 1>, V<16, 8, 32, 1>, V<32, 4, 32, 1>, V<16, 8, 16, 1, 1>>))
SFD_VARIANTS(PF32, 256, (VL<V<16, 16, 1, 16>, V<16, 16, 1, 8>, V<32, 8, 1, 16>, V<16, 16, 1, 16, 1>, V<16, 16, 1, 8, 1>,
                            V<16, 16, 1, 32>, V<16, 16, 1, 32, 1>>),
             (VL<V<16, 16, 16, 1>, V<16, 16, 32, 1>, V<32, 8, 16, 1>, V<16, 16, 16, 1, 1>, V<16, 16, 8, 1, 1>>))
SFD_VARIANTS(PF32, 512,
             (VL<V<32, 16, 1, 8>, V<32, 16, 1, 4>, V<16, 16, 1, 8, 0, 1, 2>, V<16, 16, 1, 4, 0, 1, 2>,
                 V<32, 16, 1, 8, 1>, V<32, 16, 1, 4, 1>, V<32, 16, 1, 8, 0, 1, 1, 1>>),
             (VL<V<32, 16, 16, 1>, V<32, 16, 8, 1>, V<16, 16, 16, 1, 0, 1, 2>, V<16, 16, 8, 1, 0, 1, 2>,
                 V<32, 16, 16, 1, 1>, V<32, 16, 32, 1, 1>>))
SFD_VARIANTS(PF32, 1024, (VL<V<32, 32, 1, 4>, V<32, 32, 1, 2>>), (VL<V<32, 32, 8, 1>, V<32, 32, 4, 1>>))
SFD_VARIANTS(PF32, 2048, (VL<V<32, 32, 1, 2, 0, 1, 2>>), (VL<V<32, 32, 4, 1, 0, 1, 2>>))

// ---------------------------------------------------------------- FP64
SFD_VARIANTS(PF64, 2, (VL<V<2, 1, 1, 256>>), (VL<V<2, 1, 16, 4>>))
SFD_VARIANTS(PF64, 4, (VL<V<2, 2, 1, 128>>), (VL<V<2, 2, 16, 2>>))
SFD_VARIANTS(PF64, 8, (VL<V<4, 2, 1, 128>>), (VL<V<4, 2, 16, 2>>))
SFD_VARIANTS(PF64, 16, (VL<V<4, 4, 1, 64>>), (VL<V<4, 4, 16, 1>>))
SFD_VARIANTS(PF64, 32, (VL<V<8, 4, 1, 64>>), (VL<V<8, 4, 16, 1>>))
SFD_VARIANTS(PF64, 64, (VL<V<8, 8, 1, 32>, V<16, 4, 1, 32>>), (VL<V<8, 8, 16, 1>, V<8, 8, 32, 1>>))
SFD_VARIANTS(PF64, 128, (VL<V<16, 8, 1, 16>, V<16, 8, 1, 8>, V<32, 4, 1, 16>>),
             (VL<V<16, 8, 16, 1>, V<16, 8, 32, 1>, V<32, 4, 16, 1>>))
SFD_VARIANTS(PF64, 256, (VL<V<16, 16, 1, 8>, V<16, 16, 1, 4>, V<16, 16, 1, 8, 1>>),
             (VL<V<16, 16, 16, 1>, V<16, 16, 8, 1>, V<16, 16, 8, 1, 1>>))
SFD_VARIANTS(PF64, 512,
             (VL<V<16, 16, 1, 4, 0, 1, 2>, V<32, 16, 1, 4>, V<16, 16, 1, 8, 0, 1, 2>, V<16, 16, 1, 2, 0, 1, 2, 1>,
                 V<16, 16, 1, 4, 0, 1, 2, 1>>),
             (VL<V<16, 16, 8, 1, 0, 1, 2>, V<32, 16, 8, 1>, V<32, 16, 4, 1>, V<16, 16, 16, 1, 0, 1, 2>,
                 V<16, 16, 4, 1, 0, 1, 2, 1>, V<16, 16, 8, 1, 0, 1, 2, 1>>))
SFD_VARIANTS(PF64, 1024, (VL<V<32, 32, 1, 2>>), (VL<V<32, 32, 4, 1>, V<32, 32, 2, 1>>))
SFD_VARIANTS(PF64, 2048, (VL<V<32, 32, 1, 1, 0, 1, 2>>), (VL<V<32, 32, 2, 1, 0, 1, 2>>))

#undef SFD_VARIANTS

// FP32 storage with FP64 arithmetic has the registers and LDS of FP64.
template <int N>
struct Variants<PF32X, N> : Variants<PF64, N> {
};

template <typename P, int MODE, typename L>
struct Dispatch;

template <typename P, int MODE>
struct Dispatch<P, MODE, VL<>> {
    static bool info(int, VariantInfo*) { return false; }
    static bool launch(int, dim3, dim3, hipStream_t, const SweepArgs&, bool) { return false; }
    static bool launch_batch(int, dim3, dim3, hipStream_t, const SweepArgs*, const int*, bool) { return false; }
    static int occupancy(int) { return 0; }
};

template <typename P, int MODE, typename V0, typename... Vs>
struct Dispatch<P, MODE, VL<V0, Vs...>> {
    using S = typename P::S;
    using R = typename P::R;
    static bool info(int idx, VariantInfo* out)
    {
        if (idx == 0) {
            *out = {V0::N1, V0::N2, V0::TS, V0::G, V0::P, V0::W, V0::H};
            return true;
        }
        return Dispatch<P, MODE, VL<Vs...>>::info(idx - 1, out);
    }
    // persistent: run the pipelined kernel of a pipelined variant (the
    // caller sized the grid for it); otherwise its plain k_sweep shape
    static bool launch(int idx, dim3 grid, dim3 block, hipStream_t s, const SweepArgs& a, bool persistent)
    {
        if (idx == 0) {
            // energy round trips run k_sweep: with the sums' registers the
            // pipelined kernel drops to two wavefronts per SIMD and is slower
            if constexpr (V0::P && (MODE == kRoundTrip || MODE == kRoundTripImag)) {
                if (persistent) {
                    hipLaunchKernelGGL((k_sweep_pipe<S, R, V0::N1, V0::N2, V0::TS, V0::G, rt_imag(MODE)>), grid,
                                       block, 0, s, a);
                    return true;
                }
            }
            if constexpr (rt_twice(MODE) && V0::TS == 1) {
                // the x axis runs only the double round trip [x x] of chained
                // Strang steps: real time, no nonlinear phase, no potential
                if constexpr (MODE == kRoundTripTwice || MODE == kRoundTripTwiceImag) {
                    if (a.V || (a.flags & (kFlagNL | kFlagScale | kFlagNorm))) return false;
                    hipLaunchKernelGGL(
                        (k_sweep<S, R, V0::N1, V0::N2, V0::TS, V0::G, MODE, V0::W, V0::H, V0::T, true>), grid,
                        block, 0, s, a);
                    return true;
                } else if constexpr (MODE == kRoundTripTwiceEnergyX) {
                    // with the energy sums of the state between them (V for the
                    // pointwise sums)
                    if (a.flags & (kFlagNL | kFlagScale | kFlagNorm)) return false;
                    if (a.V)
                        hipLaunchKernelGGL(
                            (k_sweep<S, R, V0::N1, V0::N2, V0::TS, V0::G, MODE, V0::W, V0::H, V0::T, false>),
                            grid, block, 0, s, a);
                    else
                        hipLaunchKernelGGL(
                            (k_sweep<S, R, V0::N1, V0::N2, V0::TS, V0::G, MODE, V0::W, V0::H, V0::T, true>), grid,
                            block, 0, s, a);
                    return true;
                } else {
                    return false;
                }
            } else if constexpr (MODE == kRoundTripTwiceEnergyX) {
                return false;  // x axis only
            } else if constexpr (rt_twice(MODE) && !rt_energy(MODE)) {
                if (!a.V)
                    hipLaunchKernelGGL(
                        (k_sweep<S, R, V0::N1, V0::N2, V0::TS, V0::G, MODE, V0::W, V0::H, V0::T, true>), grid,
                        block, 0, s, a);
                else
                    hipLaunchKernelGGL((k_sweep<S, R, V0::N1, V0::N2, V0::TS, V0::G, MODE, V0::W, V0::H, V0::T>),
                                       grid, block, 0, s, a);
                return true;
            } else {
                hipLaunchKernelGGL((k_sweep<S, R, V0::N1, V0::N2, V0::TS, V0::G, MODE, V0::W, V0::H, V0::T>), grid,
                                   block, 0, s, a);
                return true;
            }
        }
        return Dispatch<P, MODE, VL<Vs...>>::launch(idx - 1, grid, block, s, a, persistent);
    }
    // twice: two round trips per line in one pass ([y N y] / [z N z] on
    // strided variants, [x x] of chained steps on contiguous ones)
    static bool launch_batch(int idx, dim3 grid, dim3 block, hipStream_t s, const SweepArgs* descs,
                             const int* done, bool twice)
    {
        if (idx == 0) {
            if (twice) {
                hipLaunchKernelGGL((k_batch_sweep<S, R, V0::N1, V0::N2, V0::TS, V0::G, true, V0::H, V0::T>), grid,
                                   block, 0, s, descs, done);
                return true;
            }
            hipLaunchKernelGGL((k_batch_sweep<S, R, V0::N1, V0::N2, V0::TS, V0::G, false, V0::H, V0::T>), grid,
                               block, 0, s, descs, done);
            return true;
        }
        return Dispatch<P, MODE, VL<Vs...>>::launch_batch(idx - 1, grid, block, s, descs, done, twice);
    }
    // resident workgroups per CU of the persistent round-trip kernel
    static int occupancy(int idx)
    {
        if (idx == 0) {
            if constexpr (V0::P) {
                int nb = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &nb,
                        reinterpret_cast<const void*>(&k_sweep_pipe<S, R, V0::N1, V0::N2, V0::TS, V0::G, false>),
                        V0::TS * V0::G * V0::N2, 0) != hipSuccess)
                    return 1;
                return nb > 0 ? nb : 1;
            }
            return 0;
        }
        return Dispatch<P, MODE, VL<Vs...>>::occupancy(idx - 1);
    }
};

// Runtime lookup by line length.
template <typename P, int N>
int variant_count(bool contig)
{
    return contig ? Variants<P, N>::contig::size : Variants<P, N>::strided::size;
}

template <typename P, int MODE, int N>
bool launch_variant_n(bool contig, int idx, dim3 grid, dim3 block, hipStream_t s, const SweepArgs& a,
                      bool persistent)
{
    if (contig) return Dispatch<P, MODE, typename Variants<P, N>::contig>::launch(idx, grid, block, s, a, persistent);
    return Dispatch<P, MODE, typename Variants<P, N>::strided>::launch(idx, grid, block, s, a, persistent);
}

template <typename P, int N>
bool info_variant_n(bool contig, int idx, VariantInfo* out)
{
    if (contig) return Dispatch<P, 0, typename Variants<P, N>::contig>::info(idx, out);
    return Dispatch<P, 0, typename Variants<P, N>::strided>::info(idx, out);
}

#define SFD_FOR_EACH_N(M) M(2) M(4) M(8) M(16) M(32) M(64) M(128) M(256) M(512) M(1024) M(2048)

template <typename P>
int n_variants(int n, bool contig)
{
#define SFD_CASE(N_) \
    if (n == N_) return variant_count<P, N_>(contig);
    SFD_FOR_EACH_N(SFD_CASE)
#undef SFD_CASE
    return 0;
}

template <typename P, int N>
int occupancy_variant_n(bool contig, int idx)
{
    if (contig) return Dispatch<P, 0, typename Variants<P, N>::contig>::occupancy(idx);
    return Dispatch<P, 0, typename Variants<P, N>::strided>::occupancy(idx);
}

template <typename P>
int variant_occupancy(int n, bool contig, int idx)
{
#define SFD_CASE(N_) \
    if (n == N_) return occupancy_variant_n<P, N_>(contig, idx);
    SFD_FOR_EACH_N(SFD_CASE)
#undef SFD_CASE
    return 0;
}

template <typename P>
bool variant_info(int n, bool contig, int idx, VariantInfo* out)
{
#define SFD_CASE(N_) \
    if (n == N_) return info_variant_n<P, N_>(contig, idx, out);
    SFD_FOR_EACH_N(SFD_CASE)
#undef SFD_CASE
    return false;
}

template <typename P, int N>
bool launch_batch_n(bool contig, int idx, dim3 grid, dim3 block, hipStream_t s, const SweepArgs* d, const int* done,
                    bool twice)
{
    if (contig)
        return Dispatch<P, 0, typename Variants<P, N>::contig>::launch_batch(idx, grid, block, s, d, done, twice);
    return Dispatch<P, 0, typename Variants<P, N>::strided>::launch_batch(idx, grid, block, s, d, done, twice);
}

template <typename P>
bool launch_batch(int n, bool contig, int idx, dim3 grid, dim3 block, hipStream_t s, const SweepArgs* d,
                  const int* done, bool twice = false)
{
#define SFD_CASE(N_) \
    if (n == N_) return launch_batch_n<P, N_>(contig, idx, grid, block, s, d, done, twice);
    SFD_FOR_EACH_N(SFD_CASE)
#undef SFD_CASE
    return false;
}

template <typename P, int MODE>
bool launch_variant(int n, bool contig, int idx, dim3 grid, dim3 block, hipStream_t s, const SweepArgs& a,
                    bool persistent)
{
#define SFD_CASE(N_) \
    if (n == N_) return launch_variant_n<P, MODE, N_>(contig, idx, grid, block, s, a, persistent);
    SFD_FOR_EACH_N(SFD_CASE)
#undef SFD_CASE
    return false;
}

// Bluestein sweeps (lines of n points, n not a power of two) run on the
// first variant of the M-point lists, M = 2^ceil(log2(2n - 1)) in [8, 2048].
template <typename L>
struct FirstOf;
template <typename V0, typename... Vs>
struct FirstOf<VL<V0, Vs...>> {
    using type = V0;
};

template <typename P, int MODE, int N>
bool launch_blu_n(bool contig, dim3 grid, dim3 block, hipStream_t s, const SweepArgs& a)
{
    using S = typename P::S;
    using R = typename P::R;
    using VC = typename FirstOf<typename Variants<P, N>::contig>::type;
    using VS = typename FirstOf<typename Variants<P, N>::strided>::type;
    if (contig) {
        if constexpr (rt_twice(MODE)) return false;  // the double round trip runs on a strided axis only
        else hipLaunchKernelGGL((k_sweep_blu<S, R, VC::N1, VC::N2, VC::TS, VC::G, MODE, VC::H>), grid, block, 0, s, a);
    } else {
        hipLaunchKernelGGL((k_sweep_blu<S, R, VS::N1, VS::N2, VS::TS, VS::G, MODE, VS::H>), grid, block, 0, s, a);
    }
    return true;
}

#define SFD_FOR_EACH_M(M) M(8) M(16) M(32) M(64) M(128) M(256) M(512) M(1024) M(2048)

template <typename P, int MODE>
bool launch_blu(int m, bool contig, dim3 grid, dim3 block, hipStream_t s, const SweepArgs& a)
{
#define SFD_CASE(N_) \
    if (m == N_) return launch_blu_n<P, MODE, N_>(contig, grid, block, s, a);
    SFD_FOR_EACH_M(SFD_CASE)
#undef SFD_CASE
    return false;
}

}  // namespace sfd
