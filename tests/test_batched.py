"""Batched SIRT3D: uneven problems advanced in one graph-replayed iteration
(per-problem kernels on forked streams, or job-table kernels shared by all
problems) must give exactly the iterates of one ctr_sirt3d call per problem
(the same per-ray and per-tile arithmetic)."""
import numpy as np
import pytest

import ctr
from ctr import geometry
from conftest import report


def _problems():
    rng = np.random.default_rng(3)
    specs = [(24, 20, 40, 1.0), (32, 36, 56, 1.2), (16, 12, 32, 0.9)]   # (N, views, detector, pitch)
    out = []
    for n, views, det, pitch in specs:
        g = geometry.cone_circular(n, 1.0, views, det, det, pitch, 90.0, 180.0)
        vol = rng.uniform(0, 1, n ** 3).astype(np.float32)
        out.append((g, vol))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("device", [False, True])
def test_batched_sirt_matches_single(ctx, device):
    probs = _problems()
    bs = [ctx.cone_fp(g, v) for g, v in probs]
    n_iter = 7
    if device:
        import torch
        bd = [torch.from_numpy(b).cuda() for b in bs]
        xs, res = ctx.sirt3d_batched([(g, b) for (g, _), b in zip(probs, bd)], n_iter)
        xs = [x.cpu().numpy() for x in xs]
    else:
        xs, res = ctx.sirt3d_batched([(g, b) for (g, _), b in zip(probs, bs)], n_iter)
    for i, ((g, _), b) in enumerate(zip(probs, bs)):
        x1, r1, _ = ctx.sirt3d(g, b, n_iter)
        dx = float(np.abs(xs[i] - x1).max())
        dr = float(np.abs(res[i] - r1).max())
        report(f"sirt3d_batched[{i}, device={device}]", max_dx=dx, max_dres=dr, res_last=float(res[i][-1]))
        assert dx <= 1e-6 * max(float(np.abs(x1).max()), 1.0)
        assert dr <= 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("streams,graph", [("0", "1"), ("4", "0"), ("1", "1"), ("2", "1")])
def test_batched_sirt_launch_forms_match_single(ctx, monkeypatch, streams, graph):
    """Every launch form of the batch (job-table kernels, per-problem kernels
    on 1, 2 or 4 streams, with or without graph capture) gives the single-call
    iterates."""
    monkeypatch.setenv("CTR_BATCH_STREAMS", streams)
    monkeypatch.setenv("CTR_BATCH_GRAPH", graph)
    probs = _problems()
    bs = [ctx.cone_fp(g, v) for g, v in probs]
    n_iter = 5
    xs, res = ctx.sirt3d_batched([(g, b) for (g, _), b in zip(probs, bs)], n_iter)
    for i, ((g, _), b) in enumerate(zip(probs, bs)):
        x1, r1, _ = ctx.sirt3d(g, b, n_iter)
        dx = float(np.abs(xs[i] - x1).max())
        dr = float(np.abs(res[i] - r1).max())
        report(f"sirt3d_batched[{i}, streams={streams}, graph={graph}]", max_dx=dx, max_dres=dr)
        assert dx <= 1e-6 * max(float(np.abs(x1).max()), 1.0)
        assert dr <= 1e-9


@pytest.mark.gpu
def test_batched_sirt_rejects_empty_batch(ctx):
    st = ctr._lib.lib().ctr_sirt3d_batched(ctx.h, None, 0, None, 0)
    assert st == ctr._lib.EINVAL
