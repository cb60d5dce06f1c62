"""Host-side pieces of bench.py (no GPU): the one-GPU guard, the algorithmic
byte counts behind the roofline, and the field-error reduction that the
accuracy report is built from."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_algorithmic_bytes_follow_the_state_layout():
    # one RK state is phi, pi (2 values) and E, Z (4 values) per cell, f32
    state = (2 + 4) * 4
    stage1 = state + state              # U_in read, U_out written (U^n is U_in)
    stage23 = 2 * state + state         # U_in and U^n read, U_out written
    assert bench.BYTES_PER_SITE_UPDATE == stage1 + 2 * stage23
    assert bench.ALG_READ_SHARE == pytest.approx((state + 2 * 2 * state) / bench.BYTES_PER_SITE_UPDATE)


def test_block_errors_match_dense_norms():
    rng = np.random.default_rng(5)
    ref = rng.standard_normal((6, 5, 7, 4))
    ref[..., 3] *= 1e-3  # a small component, as the momentum is
    got = (ref + 1e-6 * rng.standard_normal(ref.shape)).astype(np.float32)
    blk, comps = bench.block_errors(got, ref, ("E", "Zx", "Zy", "Zz"))
    d = got.astype(np.float64) - ref
    assert blk["rel_l2"] == pytest.approx(np.linalg.norm(d) / np.linalg.norm(ref), rel=1e-12)
    assert blk["rel_linf"] == pytest.approx(np.abs(d).max() / np.abs(ref).max(), rel=1e-12)
    for i, nm in enumerate(("E", "Zx", "Zy", "Zz")):
        assert comps[nm]["rel_l2"] == pytest.approx(np.linalg.norm(d[..., i]) / np.linalg.norm(ref[..., i]),
                                                    rel=1e-12)
        assert comps[nm]["rel_linf"] == pytest.approx(np.abs(d[..., i]).max() / np.abs(ref[..., i]).max(),
                                                      rel=1e-12)
    # the small component carries the larger relative error
    assert comps["Zz"]["rel_l2"] > 100 * comps["E"]["rel_l2"]


def test_block_errors_of_identical_states_are_zero():
    ref = np.arange(2 * 3 * 4 * 2, dtype=np.float64).reshape(2, 3, 4, 2)
    blk, comps = bench.block_errors(ref.astype(np.float32), ref, ("phi", "pi"))
    assert blk == {"rel_l2": 0.0, "rel_linf": 0.0}
    assert comps["phi"]["rel_l2"] == 0.0 and comps["pi"]["rel_linf"] == 0.0
