"""Benchmark: patch-solves/s at relative residual <= 1e-5 (FP32) on one MI355X.

Workload (the configuration the metric is quoted on): 16384 patches, POT_GHOST_SIZE = 5 (18^3 solve cube),
density = 1..8 seeded Plummer spheres per patch, omega = 1.8, relative-residual stopping at 1e-5 with
min 10 / max 1000 iterations, zero initial guess (plain) and learned warm start (when trained weights are
present).  A step is one batched solve of all 16384 patches with inputs resident in HBM.

Prints one JSON line (rank 0).  `--gpus N` with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "patch-solves/s at rel. residual ≤1e-5: warm-start vs plain SOR, 1 MI355X"
HBM_PEAK = 8.0e12          # B/s, MI355X published
LDS_PEAK = 256 * 128 * 2.4e9  # B/s: 256 CUs x 128 B/clk x 2.4 GHz
LDS_BYTES_PER_UPDATE = 36  # 7 phi reads + 1 rho read + 1 write, FP32


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--npatch", type=int, default=16384)
    ap.add_argument("--gs", type=int, default=5)
    ap.add_argument("--check-every", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=512, help="patches timed on the CPU oracle (0: skip)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--pad", type=int, default=None, help="SOR split-row padding (default: library choice)")
    ap.add_argument("--ppw", type=int, default=None, help="patches per wavefront (default: library choice)")
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import torch

    import gpu_testbed_poisson_gravity_amd as pg
    from gpu_testbed_poisson_gravity_amd import inputs as I

    dev = torch.device("cuda", 0)
    t0 = time.time()
    rho, cpot, flu, _ = I.plummer(a.npatch, a.gs, seed=a.seed, energy=1.0)
    log(f"[bench] inputs generated in {time.time() - t0:.1f}s")
    g = pg.geometry(a.gs)
    d_rho = torch.from_numpy(rho).to(dev)
    d_cpot = torch.from_numpy(cpot).to(dev)
    d_flu0 = torch.from_numpy(flu).to(dev)
    d_flu = d_flu0.clone()
    d_out = torch.empty((a.npatch, 10, 10, 10), device=dev)
    d_it = torch.zeros(a.npatch, dtype=torch.int32, device=dev)
    d_st = torch.zeros(a.npatch, dtype=torch.int32, device=dev)
    poi = 4.0 * np.pi * 1e-5
    args = pg.solver_args(dt=1e-3, sor_min_iter=10, sor_max_iter=1000, sor_omega=1.8, poi_coeff=poi)
    ctx = pg.Context(0)
    ctx.set_options(pot_ghost_size=a.gs, gra_ghost_size=1, pot_scheme=pg.SOR, init_mode=pg.INIT_ZERO,
                    term_mode=pg.TERM_RELRES, check_every=a.check_every, rel_tol=1e-5, fuse_gravity=1)
    if a.pad is not None or a.ppw is not None:
        ppw0, pad0 = ctx.get_layout(a.gs)
        ctx.set_layout(a.gs, a.ppw if a.ppw is not None else ppw0, a.pad if a.pad is not None else pad0)
    stream = torch.cuda.current_stream(dev)

    def step():
        d_flu.copy_(d_flu0)  # restore the fluid so every step does identical work (tiny D2D copy)
        ctx.solve_device(d_rho, d_cpot, d_out, d_flu, a.npatch, args, iters=d_it, status=d_st,
                         stream=stream.cuda_stream)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize(dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * a.steps)]
    torch.cuda.synchronize(dev)
    t_wall = time.perf_counter()
    for s in range(a.steps):
        d_flu.copy_(d_flu0)
        ev[2 * s].record(stream)
        ctx.solve_device(d_rho, d_cpot, d_out, d_flu, a.npatch, args, iters=d_it, status=d_st,
                         stream=stream.cuda_stream)
        ev[2 * s + 1].record(stream)
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t_wall
    kern_ms = [ev[2 * s].elapsed_time(ev[2 * s + 1]) for s in range(a.steps)]
    ms = float(np.mean(kern_ms))
    iters = d_it.cpu().numpy()
    status = d_st.cpu().numpy()
    value = a.npatch / (ms * 1e-3)
    R = g["R"]
    upd = float(iters.astype(np.float64).sum()) * R ** 3
    lds_alg = upd * LDS_BYTES_PER_UPDATE
    hbm_alg = a.npatch * 4.0 * (R ** 3 + g["C"] ** 3 + 10 ** 3 + 4 * 512 + 5 * 512)
    roof = {"bound": "lds", "achieved": lds_alg / (ms * 1e-3), "peak": LDS_PEAK, "unit": "B/s",
            "frac": lds_alg / (ms * 1e-3) / LDS_PEAK, "traffic": None,
            "hbm_achieved": hbm_alg / (ms * 1e-3), "hbm_peak": HBM_PEAK}

    cpu = None
    if a.cpu_sample > 0:
        import oracle as O

        n = min(a.cpu_sample, a.npatch)
        nt = len(os.sched_getaffinity(0))
        p = O.make_params(gs=a.gs, init_mode=O.INIT_ZERO, term_mode=O.TERM_RELRES, check_every=a.check_every,
                          rel_tol=1e-5, sor_min_iter=10, sor_max_iter=1000, sor_omega=1.8, poi_coeff=poi, dt=1e-3)
        t = time.perf_counter()
        _, _, it_cpu, _ = O.solve(rho[:n], cpot[:n], flu[:n], p, nthreads=nt)
        dt = time.perf_counter() - t
        cpu = {"value": n / dt, "unit": "patch-solves/s", "cores": nt, "kind": "port",
               "sample": f"first {n} of the {a.npatch} patches, FP32 oracle, OpenMP over patches",
               "mean_iters": float(it_cpu.mean())}

    out = {
        "metric": METRIC, "value": value, "unit": "patch-solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{a.npatch} patches gs={a.gs} (18^3) Plummer 1-8 spheres, SOR omega=1.8, "
                               f"rel.res 1e-5, min10/max1000, zero guess, fused gravity",
                   "npatch": a.npatch, "gs": a.gs, "check_every": a.check_every},
        "plain": {"value": value, "ms_per_step": ms, "mean_iters": float(iters.mean()),
                  "max_iters": int(iters.max()), "converged": int((status == 1).sum())},
        "wall_ms_per_step": wall / a.steps * 1e3,
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
