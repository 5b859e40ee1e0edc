"""Learned SIRT update: training data, training and the plain-vs-learned
comparison, driven from Python over the library's CNN kernels.

    python -m ctr.learned train --out weights.npz [--phantoms 6 --steps 1500]
    python -m ctr.learned eval --weights weights.npz [--seed 2000]

Problem (sparse-view cone beam): 256^3 random-ellipsoid phantom (1 mm
voxels), 90 views over 360 degrees, 512^2 flat detector at 1 mm, SOD 600,
SDD 1200, Poisson noise with I0 = 1e5 photons per ray.  The line integrals
are scaled so that the largest is 4 before the noise is drawn (the phantom's
0-1 amplitudes carry no physical unit) and scaled back afterwards.

Network input at iterate x_k: (x_k, s_k / rms(s_k)) with s_k the plain SIRT
step (lambda = 1, clamp at 0); training target: (x_{k+m} - x_k - s_k) / rms(s_k)
with x_{k+m} the plain SIRT iterate m = 16 iterations ahead, on random patches
of the recorded iterates: one learned step is trained to stand for m plain
ones.  Inference: ctr_learned_sirt3d, judged against the plain step (a
learned step is kept only when its data residual is below the plain step's).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import _lib, geometry, phantoms

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "learned_sirt_256_90.npz")
CORRECTION_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "emulated_correction_256_90.npz")

# The emulated SIRT of the sparse-view problem: 8 plain iterations, the rest
# of 215 emulated by a degree-28 Chebyshev polynomial on the iterate's
# support; the two-pass variant starts at 6, refines the support with a
# degree-14 pass and polishes with 2 plain iterations.
EMULATED = dict(n_start=8, n_target=215, degree=28, n_polish=0)
EMULATED_TWO_PASS = dict(n_start=6, n_target=215, degree=28, n_polish=2, support_degree=14)
STARTS = (0, 1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48)
AHEAD = 16
SNAPSHOTS = tuple(sorted(set(STARTS) | {k + AHEAD for k in STARTS}))


def make_problem(ctx, seed, n=256, views=90, i0=1e5):
    """(geometry, true volume, noisy projections) as CUDA tensors."""
    import torch
    g = geometry.cone_circular(n, 256.0 / n, views, 512, 512, 1.0, 600.0, 1200.0)
    vol = torch.from_numpy(phantoms.random_phantom(n, seed).ravel()).cuda()
    b = ctx.cone_fp(g, vol)
    torch.cuda.synchronize()
    scale = 4.0 / max(float(b.max()), 1e-12)
    bn = phantoms.poisson_noise(b.cpu().numpy() * scale, i0, seed) / scale
    return g, vol, torch.from_numpy(bn).cuda()


def plain_trajectory(ctx, g, b, n_iter=200, snapshots=SNAPSHOTS):
    """Plain SIRT from zero: {k: x_k} at the snapshot iterations, x_n, residuals."""
    import torch
    x = torch.zeros(g.nvox, dtype=torch.float32, device=b.device)
    snaps, res = {}, [1.0]
    done = 0
    for k in sorted(set(snapshots) | {n_iter}):
        if k > n_iter:
            break
        if k > done:
            x, r2, _ = ctx.sirt3d(g, b, k - done, x0=x)
            res += list(r2[1:])
            done = k
        if k in snapshots:
            snaps[k] = x.clone()
    return snaps, x, np.array(res)


def collect(ctx, seeds, n=256, views=90, ahead=AHEAD, log=print):
    """Training records [(x_k, s_k, rms_k, target_k)] from plain trajectories."""
    import torch
    recs = []
    for seed in seeds:
        t0 = time.perf_counter()
        g, vol, b = make_problem(ctx, seed, n, views)
        snaps, x200, res = plain_trajectory(ctx, g, b, snapshots=set(STARTS) | {k + ahead for k in STARTS})
        for k in STARTS:
            xk = snaps[k]
            s = ctx.sirt3d_step(g, b, xk)
            rms = float(torch.sqrt(torch.mean(s * s)))
            if not rms > 0:
                continue
            recs.append((xk, s, rms, (snaps[k + ahead] - xk - s) / rms))
        torch.cuda.synchronize()
        log(f"  phantom {seed}: residual(200) {res[-1]:.5f}, {time.perf_counter() - t0:.1f} s", flush=True)
    return recs


def collect_on_policy(ctx, net, seeds, n=256, views=90, steps=10, ahead=AHEAD, log=print):
    """Records at the iterates the learned loop visits (judged against the
    plain step), labelled by `ahead` plain SIRT iterations from each."""
    import torch
    recs = []
    for seed in seeds:
        t0 = time.perf_counter()
        g, vol, b = make_problem(ctx, seed, n, views)
        x = torch.zeros(g.nvox, dtype=torch.float32, device=b.device)
        for _ in range(steps):
            s = ctx.sirt3d_step(g, b, x)
            rms = float(torch.sqrt(torch.mean(s * s)))
            if rms > 0:
                ahead_x, _, _ = ctx.sirt3d(g, b, ahead, x0=x.clone(), residuals=False)
                recs.append((x.clone(), s, rms, (ahead_x - x - s) / rms))
            x, _, _ = ctx.learned_sirt3d(g, b, net, 1, x0=x, judge_vs_plain=True)
        torch.cuda.synchronize()
        log(f"  on-policy phantom {seed}: {time.perf_counter() - t0:.1f} s", flush=True)
    return recs


def train(net, recs, n, steps=1500, batch=4, patch=64, lr=1e-3, border=4, seed=0, log=print):
    import torch
    rng = np.random.default_rng(seed)
    shape = (patch, patch, patch)
    losses = []
    for it in range(steps):
        xs, gs, ts, sc = [], [], [], []
        for _ in range(batch):
            xk, s, rms, tgt = recs[rng.integers(len(recs))]
            z0, y0, x0 = (int(rng.integers(0, n - patch + 1)) for _ in range(3))
            sl = (slice(z0, z0 + patch), slice(y0, y0 + patch), slice(x0, x0 + patch))
            xs.append(xk.view(n, n, n)[sl].reshape(-1))
            gs.append(s.view(n, n, n)[sl].reshape(-1))
            ts.append(tgt.view(n, n, n)[sl].reshape(-1))
            sc.append(1.0 / rms)
        X, G, T = (torch.stack(v).contiguous() for v in (xs, gs, ts))
        step_lr = lr * (0.1 if it > 0.8 * steps else 1.0)
        losses.append(net.train_step(shape, X, G, np.array(sc, np.float32), T, border=border, lr=step_lr))
        if it % 100 == 0 or it == steps - 1:
            log(f"  step {it}: loss {np.mean(losses[-100:]):.4e}", flush=True)
    return losses


def evaluate(ctx, net, seed=2000, n=256, views=90, n_plain=200, n_learned=10, log=print):
    """Plain SIRT (n_plain iterations) against the learned loop (n_learned steps)."""
    import torch
    g, vol, b = make_problem(ctx, seed, n, views)
    # warm both paths (tables, footprints, workspaces)
    ctx.sirt3d(g, b, 1)
    ctx.learned_sirt3d(g, b, net, 1, judge_vs_plain=True)
    xc = torch.zeros(g.nvox, dtype=torch.float32, device=b.device)
    net.forward((n, n, n), xc, xc, 1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        net.forward((n, n, n), xc, xc, 1.0)
    torch.cuda.synchronize()
    t_cnn = (time.perf_counter() - t0) / 5
    t0 = time.perf_counter()
    xp, rp, _ = ctx.sirt3d(g, b, n_plain)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    xl, rl, acc = ctx.learned_sirt3d(g, b, net, n_learned, judge_vs_plain=True)
    torch.cuda.synchronize()
    t_learned = time.perf_counter() - t0
    target = float(rp[-1])
    reached = float(rl[-1]) <= target
    # plain iterations needed for the learned residual, and their share of the plain time
    k_eq = int(np.argmax(rp <= rl[-1])) if np.any(rp <= rl[-1]) else n_plain
    t_plain_eq = t_plain * k_eq / n_plain
    img = float(torch.linalg.norm(xl - xp) / torch.linalg.norm(xp))
    out = {
        "problem": f"{n}^3 random ellipsoids (seed {seed}), {views} views, 512^2 detector, Poisson I0=1e5",
        "plain_iters": n_plain, "plain_s": round(t_plain, 4), "plain_ms_per_iter": round(1e3 * t_plain / n_plain, 3),
        "target_residual": target,
        "learned_steps": n_learned, "learned_s": round(t_learned, 4), "learned_residual": float(rl[-1]),
        "accepted": acc.tolist(), "reached_target": bool(reached),
        "plain_iters_for_learned_residual": k_eq,
        "speedup_at_target": round(t_plain / t_learned, 3) if reached else None,
        "speedup_at_equal_residual": round(t_plain_eq / t_learned, 3),
        "image_rel_l2_vs_plain200": img,
        "cnn_forward_ms": round(1e3 * t_cnn, 3),
        "residuals_learned": [round(float(v), 6) for v in rl],
        "residuals_plain_at_steps": [round(float(rp[k]), 6) for k in range(0, n_learned + 1)],
    }
    log(json.dumps(out), flush=True)
    return out


def collect_correction(ctx, seeds, n=256, views=90, log=print):
    """Records (x_e, s_e, rms, target) for the correction network: x_e the
    emulated SIRT (EMULATED), s_e the plain step there, target
    (x_200 - x_e - s_e) / rms(s_e) with x_200 plain SIRT of 200 iterations."""
    import torch
    recs = []
    for seed in seeds:
        t0 = time.perf_counter()
        g, vol, b = make_problem(ctx, seed, n, views)
        xe, _, _ = ctx.sirt3d_emulated(g, b, None, **EMULATED)
        s = ctx.sirt3d_step(g, b, xe)
        rms = float(torch.sqrt(torch.mean(s * s)))
        x200, _, _ = ctx.sirt3d(g, b, 200, residuals=False)
        recs.append((xe, s, rms, (x200 - xe - s) / rms))
        torch.cuda.synchronize()
        log(f"  phantom {seed}: {time.perf_counter() - t0:.1f} s", flush=True)
    return recs


def m2_measure(ctx, seed=2000, n=256, views=90, n_plain=200):
    """The time-to-residual comparison on the held-out problem, as one dict.
    Plain SIRT's residual after n_plain iterations is the target; the plain
    time is the time plain SIRT takes to first reach it (per-iteration time
    times the iteration).  Two methods:
      * emulated + learned: EMULATED, then one step of the correction network
        judged against the plain step (the packaged correction weights);
      * two-pass emulated: EMULATED_TWO_PASS (no network).
    Each reports its wall time, residual, image rel-L2 against plain SIRT's
    200th iterate, the speed-up and the speed-up at equal residual (plain
    time to the method's residual over the method's time)."""
    import torch
    g, vol, b = make_problem(ctx, seed, n, views)
    ctx.sirt3d(g, b, 2)
    ctx.sirt3d_emulated(g, b, None, n_start=2, n_target=10, degree=2, n_polish=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xp, rp, _ = ctx.sirt3d(g, b, n_plain)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    target = float(rp[-1])
    k_first = int(np.argmax(rp <= target))
    t_iter = t_plain / n_plain
    nxp = float(torch.linalg.norm(xp))

    def summary(name, x, res, t):
        k_eq = int(np.argmax(rp <= res)) if np.any(rp <= res) else n_plain
        return {"method": name, "s": round(t, 4), "residual": res, "reached_target": bool(res <= target),
                "image_rel_l2_vs_plain200": float(torch.linalg.norm(x - xp)) / nxp,
                "speedup": round(t_iter * k_first / t, 3),
                "plain_iters_for_residual": k_eq, "speedup_at_equal_residual": round(t_iter * k_eq / t, 3)}

    out = {"problem": f"{n}^3 random ellipsoids (seed {seed}), {views} views, 512^2 detector, Poisson I0=1e5",
           "target_residual": target, "plain_iters_to_target": k_first, "plain_s_to_target": round(t_iter * k_first, 4),
           "plain_ms_per_iter": round(1e3 * t_iter, 3), "methods": []}
    net = None
    if os.path.exists(CORRECTION_WEIGHTS):
        net = _lib.Cnn(ctx, _lib.cnn_init_params(0, 0.0))
        with np.load(CORRECTION_WEIGHTS, allow_pickle=False) as f:
            net.set_params(f["params"])
        xe, _, _ = ctx.sirt3d_emulated(g, b, None, **EMULATED)     # warm the network path too
        ctx.learned_sirt3d(g, b, net, 1, x0=xe, judge_vs_plain=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xe, _, _ = ctx.sirt3d_emulated(g, b, None, **EMULATED)
        xl, rl, acc = ctx.learned_sirt3d(g, b, net, 1, x0=xe, judge_vs_plain=True)
        torch.cuda.synchronize()
        out["methods"].append(dict(summary("emulated + learned correction", xl, float(rl[-1]),
                                           time.perf_counter() - t0), accepted=int(acc[0]), **EMULATED))
        net.close()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x2, r2, fb = ctx.sirt3d_emulated(g, b, None, **EMULATED_TWO_PASS)
    torch.cuda.synchronize()
    out["methods"].append(dict(summary("two-pass emulated", x2, float(r2[2]), time.perf_counter() - t0),
                               fell_back=fb, **EMULATED_TWO_PASS))
    best = min(out["methods"], key=lambda m: m["image_rel_l2_vs_plain200"])
    out["best_image"] = best["method"]
    torch.cuda.empty_cache()
    return out


def m2_measure_cnn_step(ctx, seed=2000, n=256, views=90, n_plain=200):
    """The time-to-residual comparison on the held-out problem, as one dict:
    plain SIRT's 200-iteration residual is the target; the plain time is the
    time plain SIRT takes to first reach it; the learned time is the wall
    time of the learned loop with the packaged weights."""
    import torch
    net = _lib.Cnn(ctx, _lib.cnn_init_params(0, 0.0))
    try:
        with np.load(WEIGHTS, allow_pickle=False) as f:
            net.set_params(f["params"])
        r = evaluate(ctx, net, seed=seed, n=n, views=views, n_plain=n_plain, log=lambda *a, **k: None)
    finally:
        net.close()
    torch.cuda.empty_cache()
    return {"method": "cnn-step", "problem": r["problem"], "target_residual": r["target_residual"],
            "plain_s_to_target": r["plain_s"], "plain_ms_per_iter": r["plain_ms_per_iter"],
            "learned_s": r["learned_s"], "learned_residual": r["learned_residual"],
            "reached_target": r["reached_target"], "image_rel_l2_vs_plain200": r["image_rel_l2_vs_plain200"],
            "speedup": r["speedup_at_target"], "speedup_at_equal_residual": r["speedup_at_equal_residual"]}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=("train", "eval", "all", "train-correction", "m2"))
    ap.add_argument("--out", default="learned_sirt_weights.npz")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--phantoms", type=int, default=6)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--rounds", type=int, default=1, help="on-policy data rounds after the first training")
    ap.add_argument("--ahead", type=int, default=AHEAD, help="plain iterations one learned step stands for")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2000)
    args = ap.parse_args(argv)
    import torch
    ctx = _lib.Context(0)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    if args.cmd == "m2":
        print(json.dumps(m2_measure(ctx, seed=args.seed, n=args.n)), flush=True)
        ctx.close()
        return 0
    if args.cmd == "train-correction":
        net = _lib.Cnn(ctx, _lib.cnn_init_params(2, 0.01))
        print(f"collecting {args.phantoms} training phantoms (emulated SIRT and 200-iteration targets)", flush=True)
        recs = collect_correction(ctx, range(3000, 3000 + args.phantoms), n=args.n)
        print(f"training {args.steps} steps on {len(recs)} records", flush=True)
        losses = train(net, recs, args.n, steps=args.steps, lr=args.lr)
        np.savez(args.out, params=net.get_params(), losses=np.array(losses))
        print(f"saved {args.out}", flush=True)
        net.close()
        ctx.close()
        return 0
    net = _lib.Cnn(ctx, _lib.cnn_init_params(0, 0.0))
    if args.weights:
        with np.load(args.weights, allow_pickle=False) as f:
            net.set_params(f["params"])
    if args.cmd in ("train", "all"):
        if not args.weights:
            net.set_params(_lib.cnn_init_params(0, 0.01))
        print(f"collecting {args.phantoms} training phantoms", flush=True)
        recs = collect(ctx, range(1000, 1000 + args.phantoms), n=args.n, ahead=args.ahead)
        print(f"training {args.steps} steps on {len(recs)} iterates", flush=True)
        losses = train(net, recs, args.n, steps=args.steps, lr=args.lr)
        for rnd in range(args.rounds):
            # on-policy rounds: label the iterates the learned loop actually visits
            recs += collect_on_policy(ctx, net, range(1000, 1000 + args.phantoms), n=args.n, ahead=args.ahead)
            print(f"round {rnd + 1}: training {args.steps // 2} steps on {len(recs)} iterates", flush=True)
            losses += train(net, recs, args.n, steps=args.steps // 2, lr=args.lr * 0.5, seed=rnd + 1)
        np.savez(args.out, params=net.get_params(), losses=np.array(losses))
        print(f"saved {args.out}", flush=True)
        del recs
    if args.cmd in ("eval", "all"):
        evaluate(ctx, net, seed=args.seed, n=args.n)
    net.close()
    ctx.close()


if __name__ == "__main__":
    sys.exit(main())
