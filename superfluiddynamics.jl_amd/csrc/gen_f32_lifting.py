"""Generate sfd_f32_lifting.hpp: FP32 lifting coefficients of the in-register
DFT rotations.

A rotation by theta is applied as three shears (lifting steps),
    x -= t y;  y += s x;  x -= t y,   t = tan(theta/2), s = sin(theta),
and undone by the same steps with -t and -s.  With t and s rounded to FP32
the forward and inverse steps are still exact inverses of each other in
exact arithmetic, so a forward-inverse round trip has no systematic error:
what remains is the unbiased rounding of each fused multiply-add (the
lifting scheme of integer FFTs; S. Oraintara, Y.-J. Chen and T. Q. Nguyen,
"Integer fast Fourier transform", IEEE Trans. Signal Process. 50(3), 2002).
A plain complex multiply by rounded (cos, sin) instead changes the norm by
the fixed amount |w|^2 - 1 (-5.7e-8 for pi/8) on every pass.

Entry m + 8 (m = -8..8) is the rotation by theta = -2 pi m / 64.

Run:  python3 gen_f32_lifting.py > sfd_f32_lifting.hpp
"""
import math

import numpy as np


def main():
    print("// Generated by gen_f32_lifting.py -- FP32 lifting coefficients of the rotation")
    print("// by theta = -2 pi m / 64, m = -8..8 (entry m + 8): t = tan(theta/2), s = sin(theta).")
    print("#pragma once")
    print("namespace sfd {")
    ts, ss = [], []
    for m in range(-8, 9):
        th = -2.0 * math.pi * m / 64.0
        ts.append(float(np.float32(math.tan(th / 2))))
        ss.append(float(np.float32(math.sin(th))))
    print("__device__ constexpr float kLiftT[17] = {")
    print("    " + ", ".join(f"{v.hex()}f" for v in ts) + "};")
    print("__device__ constexpr float kLiftS[17] = {")
    print("    " + ", ".join(f"{v.hex()}f" for v in ss) + "};")
    print("}  // namespace sfd")


if __name__ == "__main__":
    main()
