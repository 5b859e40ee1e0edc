"""SIRT checkpoint/resume: the file round trip on the CPU, and on the GPU
n + m iterations equal n iterations, a checkpoint, and m more from it."""
import json

import numpy as np
import pytest

import ctr
from ctr.checkpoint import load_sirt_state, save_sirt_state


def test_state_round_trip(tmp_path):
    x = np.random.default_rng(5).random(4 * 5 * 6).astype(np.float32)
    stem = str(tmp_path / "state")
    save_sirt_state(stem, x, 7, (4, 5, 6), {"lambda": 1.0, "seed": 5})
    y, k, h = load_sirt_state(stem)
    assert np.array_equal(x, y) and k == 7 and h["meta"]["seed"] == 5 and h["shape"] == [4, 5, 6]


def test_state_rejects_bad_input(tmp_path):
    stem = str(tmp_path / "s")
    with pytest.raises(ValueError):
        save_sirt_state(stem, np.zeros(10, np.float32), 1, (2, 2, 2))
    save_sirt_state(stem, np.zeros(8, np.float32), 1, (2, 2, 2))
    with open(stem + ".json") as f:
        h = json.load(f)
    h["shape"] = [2, 2, 3]
    with open(stem + ".json", "w") as f:
        json.dump(h, f)
    with pytest.raises(ValueError):
        load_sirt_state(stem)


@pytest.mark.gpu
def test_resume_equals_uninterrupted(tmp_path):
    g = ctr.geometry.cone_circular(24, 1.0, 20, 48, 40, 1.0, 120.0, 240.0)
    b = np.random.default_rng(6).random(g.nray).astype(np.float32)
    ctx = ctr.Context(0)
    try:
        full, r_full, _ = ctx.sirt3d(g, b, 6)
        half, r_half, _ = ctx.sirt3d(g, b, 3)
        stem = str(tmp_path / "sirt")
        save_sirt_state(stem, half, 3, g.vol_shape, {"lambda": 1.0})
        x, k, _ = load_sirt_state(stem)
        rest, r_rest, _ = ctx.sirt3d(g, b, 6 - k, x0=x)
    finally:
        ctx.close()
    assert np.array_equal(full, rest)
    assert np.allclose(r_full[3:], r_rest, rtol=0, atol=0)
