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


def test_committed_traffic_is_close_to_the_algorithmic_bytes():
    """The calibrated counter bytes of the dominant sweep sit within 1 % above
    one read and one write of the 256³ complex FP32 field (no re-reads)."""
    t = bench.load_traffic("traffic_sweep_strided_f32_256.json")
    assert t is not None
    alg = 2 * 8 * 256 ** 3
    assert alg <= t < 1.01 * alg


def test_traffic_summary_is_plain_json():
    path = os.path.join(bench.ROOT, "profiles", "traffic_sweep_strided_f32_256.json")
    with open(path) as fh:
        d = json.load(fh)
    assert "traffic_bytes_per_launch" in d


def test_more_than_one_gpu_is_refused(capsys):
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err
