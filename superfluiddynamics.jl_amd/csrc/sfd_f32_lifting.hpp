// Generated by gen_f32_lifting.py -- FP32 lifting coefficients of the rotation
// by theta = -2 pi m / 64, m = -8..8 (entry m + 8): t = tan(theta/2), s = sin(theta).
#pragma once
namespace sfd {
__device__ constexpr float kLiftT[17] = {
    0x1.a8279a0000000p-2f, 0x1.6e64a00000000p-2f, 0x1.36a0840000000p-2f, 0x1.007fa80000000p-2f, 0x1.975f5e0000000p-3f, 0x1.2fcac80000000p-3f, 0x1.936bb80000000p-4f, 0x1.9272780000000p-5f, -0x0.0p+0f, -0x1.9272780000000p-5f, -0x1.936bb80000000p-4f, -0x1.2fcac80000000p-3f, -0x1.975f5e0000000p-3f, -0x1.007fa80000000p-2f, -0x1.36a0840000000p-2f, -0x1.6e64a00000000p-2f, -0x1.a8279a0000000p-2f};
__device__ constexpr float kLiftS[17] = {
    0x1.6a09e60000000p-1f, 0x1.44cf320000000p-1f, 0x1.1c73b40000000p-1f, 0x1.e2b5d40000000p-2f, 0x1.87de2a0000000p-2f, 0x1.2940620000000p-2f, 0x1.8f8b840000000p-3f, 0x1.917a6c0000000p-4f, -0x0.0p+0f, -0x1.917a6c0000000p-4f, -0x1.8f8b840000000p-3f, -0x1.2940620000000p-2f, -0x1.87de2a0000000p-2f, -0x1.e2b5d40000000p-2f, -0x1.1c73b40000000p-1f, -0x1.44cf320000000p-1f, -0x1.6a09e60000000p-1f};
}  // namespace sfd
