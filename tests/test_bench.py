"""The benchmark's accounting, on the CPU: the algorithmic bytes per
site-update that price the roofline, the committed traffic summary it
reports, and the single-GPU contract of its command line."""
import json
import os

import pytest

import bench


@pytest.mark.parametrize("prec, ndim, with_v, passes, want", [
    ("f32", 3, False, None, 96),   # 3D Split2, V = 0: 2D = 6 round trips x (8 B in + 8 B out)
    ("f64", 3, False, None, 192),
    ("f64", 2, True, None, 136),   # 2D Split2 with V: 4 x 32 B + one 8 B read of V
    ("f32", 3, False, 5, 80),      # the folded step: x, y, [z N z], x, y
    ("f64", 3, False, 5, 160),
    ("f32", 2, True, 3, 52),       # 2D folded: x, [y N y], x, plus V in FP32
])
def test_bytes_per_site_update(prec, ndim, with_v, passes, want):
    assert bench.bytes_per_site_update(prec, ndim, with_v, passes) == want


@pytest.mark.parametrize("ndim, steps, chained, want", [
    (3, 1, True, 5.0),         # one step alone: x, y, [z N z], y, x
    (3, 2, True, 4.5),         # y, z, x + N, y, z, x + N, y, z, x: 9 sweeps
    (3, 200, True, 3.015),
    (3, 200, False, 5.0),
    (2, 4, True, 2.5),         # y, x + N, y, x + N, ..., y, x: 10 sweeps
])
def test_passes_per_step(ndim, steps, chained, want):
    assert bench.passes_per_step(ndim, steps, chained) == pytest.approx(want)


@pytest.mark.parametrize("n, slot", [(n, s) for n in (256, 512) for s in (
    "x round trip", "y round trip", "z round trip", "double round trip [d N d]")] +
    [(256, "double round trip [x x]"), (512, "double round trip [x x]")])
def test_committed_traffic_is_close_to_the_algorithmic_bytes(slot, n):
    """The calibrated counter bytes of every kernel of the step sit within 1 %
    above one read and one write of the n³ complex FP32 field (no re-reads),
    but for the 512³ y and z round trips' 32-column tiles (3 % re-fetched)."""
    t = bench.load_traffic(slot, n)
    assert t is not None
    alg = 2 * 8 * n ** 3
    # the 512³ y and z round trips' 32-column tiles re-fetch 3 % (one workgroup per CU)
    wide = n == 512 and slot in ("y round trip", "z round trip")
    assert alg <= t < (1.04 if wide else 1.01) * alg


def test_traffic_summary_is_plain_json():
    path = os.path.join(bench.ROOT, "profiles", "traffic_256cube_f32.json")
    with open(path) as fh:
        d = json.load(fh)
    assert d["algorithmic_bytes_per_launch"] == 2 * 8 * 256 ** 3
    assert bench.load_traffic("other sweeps") is None
    assert bench.load_traffic("x round trip", 128) is None


def test_kernel_names_of_the_sweep_slots():
    v = [(16, 16, 1, 8, 1, 1, 1), (16, 16, 16, 1, 1, 1, 1), (16, 16, 16, 1, 0, 1, 1)]
    assert bench.kernel_name(v, 0) == "k_sweep_pipe<float, float, 16, 16, 1, 8>"
    assert bench.kernel_name(v, 2) == "k_sweep<float, float, 16, 16, 16, 1>"
    assert bench.kernel_name(v, 3) == "k_sweep<float, float, 16, 16, 16, 1, 7>"


def test_dominant_kernel_sums_the_slots_of_one_kernel():
    """The y and z round trips of a chained step share a kernel when they
    share a variant: their times add up against the x round trip's."""
    v = [(16, 16, 1, 8, 1, 1, 1), (16, 16, 16, 1, 1, 1, 1), (16, 16, 16, 1, 1, 1, 1)]
    sw = {"x round trip": {"ms_per_step": 0.050}, "y round trip": {"ms_per_step": 0.045},
          "z round trip": {"ms_per_step": 0.046}, "other sweeps": {"ms_per_step": 0.2}}
    assert bench.dominant_kernel(v, sw) == ("k_sweep_pipe<float, float, 16, 16, 16, 1>",
                                            ["y round trip", "z round trip"])
    v2 = [v[0], v[1], (16, 16, 8, 1, 1, 1, 1)]
    assert bench.dominant_kernel(v2, sw) == ("k_sweep_pipe<float, float, 16, 16, 1, 8>", ["x round trip"])


def test_more_than_one_gpu_is_refused(capsys):
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err
