"""Pins of the FP32 lifting rotations (csrc/gen_f32_lifting.py,
csrc/sfd_f32_lifting.hpp): the committed table is the generator's output,
each rotation is undone exactly by the same steps with negated coefficients
(exact rational arithmetic, FP32-rounded coefficients), and each rotation is
the intended one to FP32 accuracy.  A complex multiply by FP32-rounded
(cos, sin) instead changes the norm by a fixed |w|^2 - 1 per use, which is
what the lifting form removes."""
import math
import os
import re
import subprocess
import sys
from fractions import Fraction

import numpy as np

CSRC = os.path.join(os.path.dirname(__file__), "..", "superfluiddynamics.jl_amd", "csrc")


def _table():
    text = open(os.path.join(CSRC, "sfd_f32_lifting.hpp")).read()
    vals = {}
    for name in ("kLiftT", "kLiftS"):
        body = re.search(name + r"\[17\] = \{(.*?)\};", text, re.S).group(1)
        vals[name] = [float.fromhex(v.strip().rstrip("f")) for v in body.split(",") if v.strip()]
    return vals["kLiftT"], vals["kLiftS"]


def _lift(x, y, t, s):
    x = x - t * y
    y = y + s * x
    x = x - t * y
    return x, y


def test_committed_table_is_the_generators_output():
    out = subprocess.run([sys.executable, os.path.join(CSRC, "gen_f32_lifting.py")], check=True,
                         capture_output=True, text=True).stdout
    assert out == open(os.path.join(CSRC, "sfd_f32_lifting.hpp")).read()


def test_lifting_pairs_are_exact_inverses():
    T, S = _table()
    rng = np.random.default_rng(0)
    for t, s in zip(T, S):
        ft, fs = Fraction(t), Fraction(s)
        assert float(np.float32(t)) == t and float(np.float32(s)) == s  # FP32 values
        for _ in range(4):
            x0, y0 = (Fraction(float(np.float32(v))) for v in rng.standard_normal(2))
            x, y = _lift(x0, y0, ft, fs)
            x, y = _lift(x, y, -ft, -fs)
            assert (x, y) == (x0, y0)


def test_lifting_rotates_by_the_table_angle():
    T, S = _table()
    for k, (t, s) in enumerate(zip(T, S)):
        th = -2.0 * math.pi * (k - 8) / 64.0
        for x0, y0 in ((1.0, 0.0), (0.0, 1.0), (0.6, -0.8)):
            x, y = _lift(x0, y0, t, s)
            ex = (x0 * math.cos(th) - y0 * math.sin(th), x0 * math.sin(th) + y0 * math.cos(th))
            assert math.hypot(x - ex[0], y - ex[1]) < 1.5e-7


def test_rounded_multiplier_has_a_fixed_modulus_error():
    """The reason for lifting: FP32 (cos, sin) of pi/8 has |w|^2 - 1 = -5.7e-8,
    a loss of norm on every forward/inverse pass through it."""
    c, s = np.float32(math.cos(math.pi / 8)), np.float32(math.sin(math.pi / 8))
    m2 = Fraction(float(c)) ** 2 + Fraction(float(s)) ** 2 - 1
    assert -6e-8 < float(m2) < -5e-8
