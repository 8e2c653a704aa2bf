"""Train the warm-start CNN on patches solved by this library's own SOR solver, then evaluate it.

    python -m gpu_testbed_poisson_gravity_amd.train_warm_start --patches 400000 --steps 120000 --batch 256 \
        --lr 1e-3 --decay-frac 0.5 --res-weight 1.0 --out stage1.npy
    python -m gpu_testbed_poisson_gravity_amd.train_warm_start --patches 400000 --steps 120000 --batch 256 \
        --lr 3e-4 --decay-frac 0.5 --res-weight 1.0 --seed 2 --init stage1.npy \
        --out gpu_testbed_poisson_gravity_amd/weights/warm_start_gs5.npy

(the shipped weights: a first run, then a second from its weights on fresh patches at a lower rate)

Data: ghost size 5 (18^3 solve cube), density fields half sums of 1..8 Plummer spheres and half lognormal
(sigma 1.5, correlation length 3 cells), generated on the host cores before the GPU is touched.  Targets:
the solver's converged potential (relative residual 1e-6) on the whole 16^3 solve interior, as the
correction (phi - interp) / s with s = max|interp| of the patch, in the network's depth-to-space order.  Training runs
entirely in the library's HIP kernels (pgamd_cnn_train_step: MFMA forward, backward and Adam); torch only
holds device buffers and draws the minibatch indices.

Evaluation on held-out seeded patches: iterations to relative residual 1e-5 (omega 1.8, min 10 / max
1000) from the zero guess, from the interpolated guess and from the network's guess (with the SOR kernel's
guess check at 1e-2), and the time of each batched solve.  The weights (.npy, 41568 FP32 values) and a
JSON record of the run are written side by side.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import inputs as I

GS = 5
POI = 4.0 * np.pi * I.NEWTON_G


def _gen_chunk(job):
    kind, n, seed = job
    if kind == "plummer":
        r, c, _, _ = I.plummer(n, GS, seed=seed)
    else:
        r, c, _, _ = I.lognormal(n, GS, seed=seed, chunk=128)
    return r, c


def generate(npatch: int, seed: int, lognormal_frac: float = 0.5, workers: int = 16, chunk: int = 500):
    """Seeded host-side generation of npatch GS5 patches (rho [n,16^3], coarse potential [n,12^3])."""
    import multiprocessing as mp

    nlog = int(round(npatch * lognormal_frac))
    jobs = []
    for kind, total, base in (("plummer", npatch - nlog, 0), ("lognormal", nlog, 1)):
        for k, b in enumerate(range(0, total, chunk)):
            jobs.append((kind, min(chunk, total - b), seed * 1_000_003 + base * 100_000 + k))
    if workers > 1:
        with mp.get_context("fork").Pool(workers) as pool:
            parts = pool.map(_gen_chunk, jobs, chunksize=1)
    else:
        parts = [_gen_chunk(j) for j in jobs]
    rho = np.concatenate([p[0] for p in parts])
    cpot = np.concatenate([p[1] for p in parts])
    return rho, cpot


def solver_args(pg, rel_iters=(10, 1000)):
    """Poisson only (no gravity update): omega 1.8, dh 1, Poi_Coeff 4 pi G."""
    return pg.solver_args(dt=1e-3, dh=1.0, sor_min_iter=rel_iters[0], sor_max_iter=rel_iters[1], sor_omega=1.8,
                          poi_coeff=POI, graacc=False)


def make_targets(pg, ctx, net, d_rho, d_cpot, tol=1e-6, max_iter=2000, chunk=32768):
    """Network inputs x0 [n,10^3*16] (BF16 bits), scales [n] and targets [n,512,8] on the device.

    The solve writes the whole 18^3 cube (output window = ghost size, Poisson only); the target is the
    interior correction (phi - interp) / s in the network's depth-to-space order."""
    import torch

    n = d_rho.shape[0]
    R, N = 16, 18
    x0 = torch.empty((n, pg.CNN.INPUT), dtype=torch.int16, device="cuda")
    scale = torch.empty(n, device="cuda")
    target = torch.empty((n, 512, 8), device="cuda")
    iters = torch.empty(n, dtype=torch.int32, device="cuda")
    status = torch.empty(n, dtype=torch.int32, device="cuda")
    args = solver_args(pg, (10, max_iter))
    ctx.set_options(pot_ghost_size=GS, gra_ghost_size=GS, init_mode=pg.INIT_INTERP, term_mode=pg.TERM_RELRES,
                    rel_tol=tol, fuse_gravity=0, check_every=4, guess_fallback_tol=0.0)
    for b in range(0, n, chunk):
        m = min(chunk, n - b)
        out = torch.empty((m, N, N, N), device="cuda")
        guess = torch.empty((m, R, R, R), device="cuda")
        ctx.solve_device(d_rho[b:b + m], d_cpot[b:b + m], out, None, m, args, iters=iters[b:b + m],
                         status=status[b:b + m])
        net.prepare(d_rho[b:b + m], d_cpot[b:b + m], m, GS, args, x0[b:b + m], scale[b:b + m])
        # an untrained network (zero output layer) returns the interpolated interior unchanged
        net.warm_start(d_rho[b:b + m], d_cpot[b:b + m], m, args, guess)
        corr = (out[:, 1:17, 1:17, 1:17] - guess) / scale[b:b + m, None, None, None]
        target[b:b + m] = corr.reshape(m, 8, 2, 8, 2, 8, 2).permute(0, 1, 3, 5, 2, 4, 6).reshape(m, 512, 8)
    ctx.set_options(gra_ghost_size=1)
    torch.cuda.synchronize()
    return x0, scale, target, iters.cpu().numpy(), status.cpu().numpy()


def evaluate(pg, ctx, net, rho, cpot, tol=1e-5, fallback_tol=1e-2, repeats=3):
    """Iterations and batched solve time to relative residual tol from the zero guess, the interpolated
    guess and the network's guess (with the guess check).  Every patch must converge."""
    import torch

    n = rho.shape[0]
    d_rho, d_cpot = torch.as_tensor(rho).cuda(), torch.as_tensor(cpot).cuda()
    out = torch.empty((n, 10, 10, 10), device="cuda")
    guess = torch.empty((n, 16, 16, 16), device="cuda")
    iters = torch.empty(n, dtype=torch.int32, device="cuda")
    status = torch.empty(n, dtype=torch.int32, device="cuda")
    args = solver_args(pg)
    res = {}
    for name, mode in (("zero", pg.INIT_ZERO), ("interp", pg.INIT_INTERP), ("cnn", pg.INIT_GUESS)):
        ctx.set_options(pot_ghost_size=GS, init_mode=mode, term_mode=pg.TERM_RELRES, rel_tol=tol, fuse_gravity=0,
                        check_every=4, guess_fallback_tol=fallback_tol if mode == pg.INIT_GUESS else 0.0)
        times = []
        for _ in range(repeats + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if mode == pg.INIT_GUESS:
                net.warm_start(d_rho, d_cpot, n, args, guess)
            ctx.solve_device(d_rho, d_cpot, out, None, n, args, guess=guess if mode == pg.INIT_GUESS else None,
                             iters=iters, status=status)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        it, st = iters.cpu().numpy(), status.cpu().numpy()
        t = float(np.median(times[1:]))
        res[name] = {"mean_iters": float(it.mean()), "p50_iters": float(np.percentile(it, 50)),
                     "p99_iters": float(np.percentile(it, 99)), "max_iters": int(it.max()),
                     "converged": int(((st & pg.STATUS_CONVERGED) != 0).sum()),
                     "fallbacks": int(((st & pg.STATUS_FALLBACK) != 0).sum()), "ms": t * 1e3,
                     "patch_solves_per_s": n / t}
    res["speedup_vs_zero"] = res["zero"]["ms"] / res["cnn"]["ms"]
    res["iteration_ratio_vs_zero"] = res["cnn"]["mean_iters"] / res["zero"]["mean_iters"]
    return res


def learning_rate(step: int, steps: int, lr: float, decay_frac: float) -> float:
    """Constant rate lr, then a linear decay to lr/10 over the last decay_frac of the steps (for steps > 1
    the last step runs at lr/10)."""
    frac = step / max(steps - 1, 1)
    d0 = 1.0 - decay_frac
    return lr if frac < d0 else lr * (1.0 - 0.9 * (frac - d0) / decay_frac)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--patches", type=int, default=200000)
    ap.add_argument("--lognormal-frac", type=float, default=0.5)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--decay-frac", type=float, default=0.25,
                    help="fraction of the steps, at the end, over which the rate decays linearly to lr/10")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--eval-patches", type=int, default=16384)
    ap.add_argument("--time-limit", type=float, default=1e9, help="stop training after this many seconds")
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--res-weight", type=float, default=0.0,
                    help="weight of the residual term of the loss (pgamd_cnn_set_loss; 0: mean squared error only)")
    ap.add_argument("--init", default=None, help="start from these weights (.npy)")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "weights", "warm_start_gs5.npy"))
    a = ap.parse_args(argv)
    if not 0.0 < a.decay_frac <= 1.0:
        ap.error("--decay-frac must lie in (0, 1]")
    t_start = time.time()

    # host-side generation first, before anything initialises the GPU
    rho, cpot = generate(a.patches, a.seed, a.lognormal_frac, a.workers)
    ev_rho, ev_cpot = generate(a.eval_patches, 7919 + a.seed, 0.5, a.workers)   # held out: other seeds
    nep = a.eval_patches - int(round(a.eval_patches * 0.5))                      # Plummer first, then lognormal
    print(f"[train] generated {rho.shape[0]} training + {a.eval_patches} held-out patches in "
          f"{time.time() - t_start:.0f} s", flush=True)

    import torch

    from . import CNN, Context, load

    load()
    ctx = Context(0)
    net = CNN(ctx, seed=a.seed)
    net.set_loss(a.res_weight)
    ctx.set_options(pot_ghost_size=GS)
    d_rho, d_cpot = torch.as_tensor(rho).cuda(), torch.as_tensor(cpot).cuda()
    del rho, cpot
    pg = sys.modules[__package__]
    x0, scale, target, it_t, st_t = make_targets(pg, ctx, net, d_rho, d_cpot)
    del d_rho, d_cpot
    print(f"[train] targets: solver iterations mean {it_t.mean():.1f} max {it_t.max()}, converged "
          f"{int((st_t & 1).sum())}/{it_t.size}, target rms {float(target.square().mean().sqrt()):.3e}",
          flush=True)
    # train on targets normalised by their RMS; the output layer is linear, so the scale folds back into it
    tscale = float(target.square().mean().sqrt())
    target /= tscale
    o6 = 5 * (16 * 432 + 16)              # output layer W6, b6
    if a.init:
        p = np.load(a.init, allow_pickle=False).astype(np.float32)
        p[o6:] /= tscale
        net.set_params(p)

    n = x0.shape[0]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(a.seed)
    curve, t0, done = [], time.time(), 0
    ema = None
    for step in range(a.steps):
        idx = torch.randint(0, n, (a.batch,), device="cuda", generator=gen)
        xb, tb = x0.index_select(0, idx), target.index_select(0, idx)
        want = step % 100 == 0 or step == a.steps - 1
        lr = learning_rate(step, a.steps, a.lr, a.decay_frac)
        loss = net.train_step(xb, tb, a.batch, lr, want_loss=want)
        done = step + 1
        if want:
            ema = loss if ema is None else 0.9 * ema + 0.1 * loss
            if step % 500 == 0 or step == a.steps - 1:
                curve.append([step, loss])
                print(f"[train] step {step} loss {loss:.4e} (normalised; ema {ema:.4e}) lr {lr:.1e} "
                      f"{time.time() - t0:.0f} s", flush=True)
        if time.time() - t0 > a.time_limit:
            print(f"[train] time limit reached after {done} steps", flush=True)
            break
    torch.cuda.synchronize()
    train_s = time.time() - t0
    p = net.get_params()
    p[o6:] *= tscale                      # undo the target normalisation
    net.set_params(p)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    net.save(a.out)
    del x0, target

    ev = {}
    for name, sl in (("plummer", slice(0, nep)), ("lognormal", slice(nep, None))):
        ev[name] = evaluate(pg, ctx, net, ev_rho[sl], ev_cpot[sl])
        z, k = ev[name]["zero"], ev[name]["cnn"]
        print(f"[eval] {name}: iterations zero {z['mean_iters']:.1f} interp {ev[name]['interp']['mean_iters']:.1f} "
              f"cnn {k['mean_iters']:.1f} (fallbacks {k['fallbacks']}); time zero {z['ms']:.2f} ms cnn "
              f"{k['ms']:.2f} ms", flush=True)
    record = {"patches": n, "lognormal_frac": a.lognormal_frac, "steps": done, "batch": a.batch, "lr": a.lr,
              "res_weight": a.res_weight,
              "lr_schedule": f"constant, then linear decay to lr/10 over the last {a.decay_frac:g} of the steps",
              "init": a.init, "target_rms": tscale,
              "seed": a.seed, "train_seconds": train_s, "ms_per_step": 1e3 * train_s / max(done, 1),
              "loss_curve": curve, "target_solver": {"rel_tol": 1e-6, "mean_iters": float(it_t.mean())},
              "eval_patches": a.eval_patches, "eval": ev, "device": ctx.device_info()}
    with open(os.path.splitext(a.out)[0] + ".json", "w") as fh:
        json.dump(record, fh, indent=1)
    net.close()
    ctx.close()
    print(json.dumps({k: record[k] for k in ("steps", "ms_per_step")}), flush=True)


if __name__ == "__main__":
    main()
