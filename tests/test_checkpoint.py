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


def test_crash_between_volume_and_header_keeps_old_state(tmp_path):
    """A save interrupted after the new volume is written and before the
    header is replaced resumes from the old state, never the new volume
    under the old count."""
    from ctr import checkpoint
    rng = np.random.default_rng(8)
    x3, x6 = rng.random(60).astype(np.float32), rng.random(60).astype(np.float32)
    stem = str(tmp_path / "run")
    save_sirt_state(stem, x3, 3, (3, 4, 5))
    checkpoint.write_volume(stem, x6, 6)            # the crash: no header replace after it
    y, k, _ = load_sirt_state(stem)
    assert k == 3 and np.array_equal(y, x3)
    save_sirt_state(stem, x6, 6, (3, 4, 5))         # a completed save commits the new state
    y, k, h = load_sirt_state(stem)
    assert k == 6 and np.array_equal(y, x6)
    assert not (tmp_path / "run.k3.raw").exists()   # the superseded volume is removed


def test_corrupt_volume_is_rejected(tmp_path):
    stem = str(tmp_path / "c")
    save_sirt_state(stem, np.arange(8, dtype=np.float32), 2, (2, 2, 2))
    raw = tmp_path / "c.k2.raw"
    data = bytearray(raw.read_bytes())
    data[5] ^= 0x40
    raw.write_bytes(bytes(data))
    with pytest.raises(ValueError, match="CRC"):
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
