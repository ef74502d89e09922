"""Writes the golden fixtures of the split-step path (run from the repo root:
``python tests/golden/make_golden.py``).

Inputs are seeded states; expected outputs come from the FP64 CPU oracle
(oracle/sfd_oracle.c), whose pins are in tests/test_oracle.py.  The
reference itself (Julia + FFTW) cannot run in this container, so no vector
here was produced by it.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "superfluiddynamics.jl_amd")]

from helpers import mesh, random_potential, smooth_state  # noqa: E402
from oracle import oracle  # noqa: E402

CASES = [
    # name, shape (x-first), bounds, V?, params
    ("split2_2d_32x16_rot", (32, 16), ((-4, 4), (-2, 2)), True,
     dict(coeffD=-0.5, beta=5.0, Omega=0.6, dt=4e-3, nsteps=12, order=2, imag=False)),
    ("split1_2d_16x64", (16, 64), ((-2, 2), (-8, 8)), True,
     dict(coeffD=-0.5, beta=2.0, Omega=0.0, dt=5e-3, nsteps=9, order=1, imag=False)),
    ("split2_3d_8x16x8_rot", (8, 16, 8), ((-2, 2), (-3, 3), (-2, 2)), True,
     dict(coeffD=-0.4, beta=3.0, Omega=0.3, dt=3e-3, nsteps=7, order=2, imag=False)),
    ("split2_3d_16cube", (16, 16, 16), ((0, 6.283185307179586),) * 3, False,
     dict(coeffD=-0.5, beta=10.0, Omega=0.0, dt=2e-3, nsteps=10, order=2, imag=False)),
    ("imag2_2d_32x32", (32, 32), ((-5, 5), (-5, 5)), True,
     dict(coeffD=-0.5, beta=30.0, Omega=0.0, dt=1e-2, nsteps=20, order=2, imag=True)),
    ("imag1_3d_8cube_rot", (8, 8, 8), ((-3, 3),) * 3, True,
     dict(coeffD=-0.5, beta=8.0, Omega=0.2, dt=1e-2, nsteps=15, order=1, imag=True)),
]


def main():
    arrays = {}
    meta = {}
    for name, shape, bounds, withV, prm in CASES:
        psi = smooth_state(shape, bounds, seed=len(name))
        V = random_potential(shape, bounds, amp=4.0) if withV else None
        out = oracle.split_step(psi.copy(), bounds, V=V, **prm)
        arrays[name + "/psi0"] = psi
        arrays[name + "/out"] = out
        if V is not None:
            arrays[name + "/V"] = V
        meta[name] = dict(shape=shape, bounds=bounds, params=prm)
    np.savez_compressed(os.path.join(HERE, "split_cases.npz"), **arrays)
    with open(os.path.join(HERE, "split_cases.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    print("wrote", len(CASES), "cases")


if __name__ == "__main__":
    main()
