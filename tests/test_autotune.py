"""Autotuner: the variants it chooses between compute the same operators,
so results after tuning equal the default-tuned results (bit-exact for the
backprojection and forward projection, to summation order for the adjoint)."""
import numpy as np
import pytest

import ctr
from ctr import geometry
from conftest import report


def _geom():
    return geometry.cone_circular(40, 1.0, 30, 64, 48, 1.0, 100.0, 200.0)


@pytest.mark.gpu
def test_autotune_choices_and_results():
    g = _geom()
    rng = np.random.default_rng(4)
    vol = rng.uniform(0, 1, g.nvox).astype(np.float32)
    ref_ctx = ctr.Context(0)
    p0 = ref_ctx.cone_fp(g, vol)
    f0 = ref_ctx.fdk(g, p0)
    a0 = ref_ctx.cone_bp_adjoint(g, p0)
    ref_ctx.close()

    ctx = ctr.Context(0)
    try:
        r = ctx.autotune_cone(g)
        report("autotune", **{k: str(v) for k, v in r.items()})
        assert r["bp_pitch"] in (64, 72, 80) and r["adj_cm"] in (3, 4, 6) and r["fp_transpose"] in (0, 1)
        assert all(t > 0 for t in r["bp_ms"] + r["adj_ms"] + r["fp_ms"])
        p1 = ctx.cone_fp(g, vol)
        f1 = ctx.fdk(g, p1)
        a1 = ctx.cone_bp_adjoint(g, p1)
        # the flag path: tuned on first use, then reused
        p2 = ctx.cone_fp(g, vol, flags=ctr.AUTOTUNE)
    finally:
        ctx.close()
    assert np.array_equal(p0, p1) and np.array_equal(p1, p2)
    assert np.array_equal(f0, f1)
    assert np.linalg.norm(a1 - a0) <= 1e-6 * np.linalg.norm(a0)
