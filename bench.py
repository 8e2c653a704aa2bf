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
LDS_PEAK = 256 * 128 * 2.4e9  # B/s: 256 CUs x 128 B/clk x 2.4 GHz, used when the measurement is unavailable
MFMA_BF16_PEAK = 2.5e15    # FLOP/s, dense BF16, MI355X published
LDS_BYTES_PER_UPDATE = 36  # 7 phi reads + 1 rho read + 1 write, FP32
# Arithmetic of one SOR cell update in the reference's operand order, no contraction: five adds of the
# neighbours, 6*phi, two subtractions, omega/6 times the residual, and the add into phi.
VALU_OPS_PER_UPDATE = 10
# Vector issue rate without packed or fused operations: 256 CUs x 64 lanes per clock x 2.4 GHz (FP32 and FP64).
VALU_PEAK = 256 * 64 * 2.4e9
# Warm-start network FLOPs per patch that the model needs: five 3^3 convolutions 16 -> 16 channels and the
# 16 -> 8 output layer on the 8^3 grid, K = 27 taps x 16 channels = 432:
#   5 x 512 x 16 x 432 x 2 + 512 x 8 x 432 x 2 = 38,928,384.
# The roofline counts these only; the matrix cores also multiply the zero-weight K pad (480 of 432) and the
# output layer's 8 unused rows, reported beside it as the issued figure.
CNN_USEFUL_FLOP = 5 * 512 * 16 * 432 * 2 + 512 * 8 * 432 * 2
WEIGHTS = os.path.join(ROOT, "gpu_testbed_poisson_gravity_amd", "weights", "warm_start_gs5.npy")
# HBM bytes per launch of each kernel on the default workload, measured with rocprofv3 FETCH_SIZE and WRITE_SIZE
# (each scaled by what the counter reads for a 1 GiB device copy); lines "traffic <kernel> <leg> <bytes>"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_plummer16384_gs5.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(a):
    """{(kernel, leg): HBM bytes per launch} for the default workload (16384 Plummer patches, ghost size 5),
    empty for other sizes: the counters were collected on that workload only."""
    out = {}
    if a.npatch != 16384 or a.gs != 5 or not os.path.exists(TRAFFIC_FILE):
        return out
    with open(TRAFFIC_FILE) as f:
        for line in f:
            w = line.split()
            if len(w) == 4 and w[0] == "traffic":
                out[(w[1], w[2])] = float(w[3])
    return out


def measure_peaks(ctx):
    """LDS read and HBM copy bandwidth measured on the device (GB/s), or None if the entry fails."""
    try:
        lds, hbm = ctx.measure_peaks()
    except Exception as e:  # the peaks only feed the roofline report
        log(f"[bench] peak measurement failed: {e}")
        return None
    log(f"[bench] measured peaks: LDS read {lds / 1e12:.1f} TB/s, HBM copy {hbm / 1e12:.2f} TB/s")
    return {"lds_read_GBps": lds / 1e9, "hbm_copy_GBps": hbm / 1e9}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--npatch", type=int, default=16384)
    ap.add_argument("--gs", type=int, default=5)
    ap.add_argument("--check-every", type=int, default=2,
                    help="relative-residual test every k iterations after Min_Iter (same rule on every leg)")
    ap.add_argument("--cpu-sample", type=int, default=16384,
                    help="patches timed on the host cores by the reference CPU solver and by the oracle (0: skip; "
                         "default: the whole 16384-patch workload, about a second each on 16 host cores)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--pad", type=int, default=None, help="SOR split-row padding (default: library choice)")
    ap.add_argument("--ppw", type=int, default=None, help="patches per wavefront (default: library choice)")
    ap.add_argument("--weights", default=WEIGHTS, help="trained warm-start network (.npy)")
    ap.add_argument("--fallback-tol", type=float, default=1e-2,
                    help="guess check: a network guess with relative residual above this restarts from zero")
    ap.add_argument("--mode", choices=["both", "plain", "warm"], default="both")
    ap.add_argument("--guess", choices=["cube", "interior"], default="cube",
                    help="warm leg: the network writes the whole solve cube and the solve reads its Dirichlet layer "
                         "from it (cube, init GUESS_CUBE), or the interior only (init GUESS: the solve "
                         "interpolates the Dirichlet layer itself)")
    ap.add_argument("--chunks", type=int, default=0,
                    help="also time the pipelined warm start (pgamd_cnn_solve_device) over this many slices; "
                         "0: serial only (measured faster on MI355X: the two kernels do not share CUs well)")
    ap.add_argument("--workload", choices=["config2", "config3", "config5"], default="config2",
                    help="config2 (default): the metric's workload; config3: 32768 mixed-GS lognormal patches, "
                         "half SOR half MG, one captured graph; config5: 65536 mixed-GS "
                         "patches (power law / Plummer / lognormal), warm start with guess check, SOR+MG, gravity")
    ap.add_argument("--asyn-streams", default="1,2,4,8",
                    help="stream counts of the host-array pipeline leg (pgamd_asyn_poisson_gravity_solver with pinned "
                         "buffers, PCIe included; empty: skip)")
    ap.add_argument("--fp64", type=int, default=1, help="1: also time the FP64 device path on the workload")
    ap.add_argument("--level-box", type=int, default=12,
                    help="refined box of M^3 coarse patches (8 M^3 fine patches) for the prepare/solve/scatter "
                         "cycle of two AMR levels; 0 skips it")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (a known-size reference for "
                         "rocprof FETCH_SIZE / WRITE_SIZE calibration)")
    return ap.parse_args()


OMEGA_BY_GS = [1.49, 1.57, 1.62, 1.65, 1.69]


def run_mixed(a):
    """Mixed ghost sizes and schemes, one captured graph of per-ghost-size launches (configurations 3 and 5)."""
    import torch

    import gpu_testbed_poisson_gravity_amd as pg
    from gpu_testbed_poisson_gravity_amd import inputs as I

    c5 = a.workload == "config5"
    n = a.npatch if a.npatch != 16384 else (65536 if c5 else 32768)
    kinds = ("powerlaw", "plummer", "lognormal") if c5 else ("lognormal",)
    t0 = time.time()
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    gs, kind, rho, cpot, flu = I.mixed_batch(n, a.seed, kinds=kinds, workers=workers)
    sch = np.where(np.random.default_rng(a.seed + 1).permutation(n) % 2 == 0, pg.SOR, pg.MG).astype(np.int32)
    log(f"[bench] {a.workload}: {n} patches generated in {time.time() - t0:.1f}s")
    dev = torch.device("cuda", 0)
    poi = 4.0 * np.pi * 1e-5
    args = pg.solver_args(dt=1e-3, sor_min_iter=10, sor_max_iter=1000, sor_omega=1.69, mg_max_iter=100, mg_npre=2,
                          mg_npost=2, mg_tol=1e-5, poi_coeff=poi)
    omega = np.asarray(OMEGA_BY_GS, np.float32)
    ctx = pg.Context(0)
    peaks = measure_peaks(ctx)
    lds_peak = peaks["lds_read_GBps"] * 1e9 if peaks else LDS_PEAK
    net = None
    init = pg.INIT_INTERP
    if c5 and os.path.exists(a.weights):
        net = pg.CNN(ctx)
        net.load_params(a.weights)
        init = pg.INIT_GUESS
    ctx.set_options(pot_ghost_size=5, gra_ghost_size=1, init_mode=init, term_mode=pg.TERM_RELRES, rel_tol=1e-5,
                    check_every=a.check_every, fuse_gravity=1, guess_fallback_tol=a.fallback_tol if net else 0.0)
    # autotune every ghost size on its own SOR patches (the mixed launches use the row padding per ghost size
    # and the patches per wavefront of ghost size 1)
    tuned = {}
    for g in range(1, 6):
        idx = np.nonzero((gs == g) & (sch == pg.SOR))[0][:4096]
        if idx.size == 0:
            continue
        d_r = torch.as_tensor(np.stack([rho[i] for i in idx])).to(dev)
        d_c = torch.as_tensor(np.stack([cpot[i] for i in idx])).to(dev)
        d_o = torch.empty((idx.size, 10, 10, 10), device=dev)
        ctx.set_options(pot_ghost_size=g, init_mode=pg.INIT_INTERP)
        ga = pg.solver_args(dt=1e-3, sor_min_iter=10, sor_max_iter=1000, sor_omega=float(omega[g - 1]), poi_coeff=poi,
                            graacc=False)
        tuned[g] = ctx.autotune(g, d_r, d_c, d_o, idx.size, ga, repeats=2)[:2]
    ctx.set_options(pot_ghost_size=5, init_mode=init)
    plan = ctx.plan(gs, sch)
    d_rho = torch.as_tensor(pg.pack(rho, plan.rho_off, plan.rho_floats)).to(dev)
    d_cpot = torch.as_tensor(pg.pack(cpot, plan.pot_in_off, plan.pot_in_floats)).to(dev)
    d_flu0 = torch.as_tensor(flu).to(dev)
    d_flu = d_flu0.clone()
    d_out = torch.empty((n, 10, 10, 10), device=dev)
    d_guess = torch.zeros(plan.rho_floats, device=dev) if net is not None else None
    d_it = torch.zeros(n, dtype=torch.int32, device=dev)
    d_st = torch.zeros(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    R = 6 + 2 * gs.astype(np.float64)

    def leg(init_mode):
        """One start (zero, interpolation or the network's guess): a captured graph of the per-ghost-size
        launches, timed with events around the network and around the graph."""
        use_net = init_mode == pg.INIT_GUESS
        ctx.set_options(init_mode=init_mode, guess_fallback_tol=a.fallback_tol if use_net else 0.0)
        graph = ctx.graph_mixed(plan, d_rho, d_cpot, d_out, d_flu, args, guess=d_guess if use_net else None,
                                omega_by_ghost=omega, iters=d_it, status=d_st)

        def step(e=None):
            d_flu.copy_(d_flu0)
            if e is not None:
                e[0].record(stream)
            if use_net:
                net.warm_start_mixed(plan, d_rho, d_cpot, args, d_guess, stream=stream.cuda_stream)
            if e is not None:
                e[1].record(stream)
            graph.launch(stream.cuda_stream)
            if e is not None:
                e[2].record(stream)

        for _ in range(a.warmup):
            step()
        torch.cuda.synchronize(dev)
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(a.steps)]
        for s_ in range(a.steps):
            step(evs[s_])
        torch.cuda.synchronize(dev)
        graph.close()
        pre = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
        sol = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
        it, st = d_it.cpu().numpy(), d_st.cpu().numpy()
        r = {"value": n / ((pre + sol) * 1e-3), "ms_per_step": pre + sol, "network_ms": pre, "solve_ms": sol,
             "sor": {"mean_iters": float(it[sch == pg.SOR].mean()), "max_iters": int(it[sch == pg.SOR].max())},
             "mg": {"mean_vcycles": float(it[sch == pg.MG].mean()), "max_vcycles": int(it[sch == pg.MG].max())},
             "converged": int(((st & pg.STATUS_CONVERGED) != 0).sum()),
             "fallbacks": int(((st & pg.STATUS_FALLBACK) != 0).sum()),
             "fallbacks_sor": int(((st & pg.STATUS_FALLBACK) != 0)[sch == pg.SOR].sum()),
             "fallbacks_mg": int(((st & pg.STATUS_FALLBACK) != 0)[sch == pg.MG].sum()),
             "sor_lds_bytes": float((it * R ** 3 * LDS_BYTES_PER_UPDATE)[sch == pg.SOR].sum())}
        log(f"[bench] {a.workload} start {init_mode}: {r['value']:.4g} patch-solves/s, network {pre:.3f} ms, "
            f"solve {sol:.3f} ms, SOR {r['sor']['mean_iters']:.1f} iterations, MG {r['mg']['mean_vcycles']:.2f} "
            f"V-cycles, fallbacks {r['fallbacks']}")
        return r

    legs = {}
    if net is not None:            # config 5: the network against both starts without it
        legs["zero"] = leg(pg.INIT_ZERO)
        legs["interp"] = leg(pg.INIT_INTERP)
        legs["warm"] = leg(pg.INIT_GUESS)
        head = legs["warm"]
    else:
        head = legs["interp"] = leg(pg.INIT_INTERP)
    sol = head["solve_ms"]
    out = {
        "metric": f"patch-solves/s at rel. residual <=1e-5, mixed ghost sizes ({a.workload}), 1 MI355X",
        "value": head["value"], "unit": "patch-solves/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{n} patches, GS uniform in 1..5, densities {'/'.join(kinds)}, half SOR (omega by "
                               f"GS {OMEGA_BY_GS}) half MG (2+2 smoothing), rel. residual 1e-5, fused gravity, one "
                               f"captured graph (one launch per ghost size and scheme), autotuned layouts {tuned}; start: "
                               + ("learned warm start (guess check 1e-2)" if net is not None else "interpolation"),
                   "npatch": n},
        "network_ms": head["network_ms"], "solve_ms": sol, "sor": head["sor"], "mg": head["mg"],
        "converged": head["converged"], "fallbacks": head["fallbacks"],
        "legs": {k: {q: v for q, v in r.items() if q != "sor_lds_bytes"} for k, r in legs.items()},
        "warm_vs_zero": legs["warm"]["value"] / legs["zero"]["value"] if "warm" in legs else None,
        "warm_vs_interp": legs["warm"]["value"] / legs["interp"]["value"] if "warm" in legs else None,
        "roofline": {"bound": "lds", "achieved": head["sor_lds_bytes"] / (sol * 1e-3) / 1e9, "peak": lds_peak / 1e9,
                     "unit": "GB/s", "frac": head["sor_lds_bytes"] / (sol * 1e-3) / lds_peak, "traffic": None,
                     "peak_basis": "measured LDS read bandwidth (pgamd_measure_peaks)" if peaks else "placeholder",
                     "note": "SOR patches' algorithmic LDS bytes over the whole graph time (MG time included)"},
        "cpu_baseline": None, "peaks_measured": peaks,
    }
    print(json.dumps(out))
    plan.close()
    if net is not None:
        net.close()
    ctx.close()


def asyn_pipeline(a, pg, torch, ctx, rho, cpot, flu, args, kernel_ms):
    """The reference call shape with host arrays (CUAPI_Asyn_PoissonGravitySolver.cu:143-265): pinned buffers from
    pgamd_mem_allocate_poisson_gravity, the batch split over GPU_NStream streams, each stream copying its slice in,
    solving it and copying it out.  Reported: PCIe-inclusive patch-solves/s per stream count (host wall clock
    around the call and pgamd_synchronize, median of the timed calls), and beside them the three phases alone on
    the same bytes: host-to-device copies, the solve (the plain leg's kernel time) and device-to-host copies."""
    n = a.npatch - a.npatch % 8
    ctx.set_options(pot_ghost_size=a.gs, gra_ghost_size=1, pot_scheme=pg.SOR, init_mode=pg.INIT_ZERO,
                    term_mode=pg.TERM_RELRES, check_every=a.check_every, rel_tol=1e-5, fuse_gravity=1,
                    guess_fallback_tol=0.0)
    h_rho, h_cpot, h_out, h_flu = ctx.mem_allocate(n // 8)
    h_rho[:] = rho[:n]
    h_cpot[:] = cpot[:n]
    legs = {}
    for ns in [int(v) for v in a.asyn_streams.split(",") if v]:
        times = []
        for k in range(a.warmup + a.steps):
            h_flu[:] = flu[:n]                                   # identical work every call (outside the clock)
            t0 = time.perf_counter()
            ctx.asyn_poisson_gravity_solver(h_rho, h_cpot, h_out, h_flu, n // 8, args.dt, args.dh, args.sor_min_iter,
                                            args.sor_max_iter, args.sor_omega, args.mg_max_iter, args.mg_npre_smooth,
                                            args.mg_npost_smooth, args.mg_tolerated_error, args.poi_coeff,
                                            args.int_scheme, False, True, True, gpu_nstream=ns)
            ctx.synchronize()
            if k >= a.warmup:
                times.append(time.perf_counter() - t0)
        ms = float(np.median(times)) * 1e3
        legs[str(ns)] = {"ms": ms, "patch_solves_per_s": n / (ms * 1e-3)}
        log(f"[bench] host-array pipeline, {ns} stream(s): {ms:.2f} ms per call, {n / (ms * 1e-3):.4g} patch-solves/s")
    it, _ = ctx.patch_stats(n)
    ctx.mem_free()
    # the phases alone on the same bytes (pinned torch buffers, events on one stream)
    dev = torch.device("cuda", 0)
    h_in = [torch.from_numpy(x[:n]).pin_memory() for x in (rho, cpot, flu)]
    d_in = [torch.empty_like(x, device=dev) for x in h_in]
    d_out = [torch.empty((n, 10, 10, 10), device=dev), torch.empty((n, 5, 8, 8, 8), device=dev)]
    h_o = [torch.empty(x.shape, pin_memory=True) for x in d_out]

    def timed(fn):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        fn()
        torch.cuda.synchronize(dev)
        ts = []
        for _ in range(max(3, a.steps // 4)):
            ev[0].record()
            fn()
            ev[1].record()
            torch.cuda.synchronize(dev)
            ts.append(ev[0].elapsed_time(ev[1]))
        return float(np.median(ts))

    h2d = timed(lambda: [d.copy_(h, non_blocking=True) for d, h in zip(d_in, h_in)])
    d2h = timed(lambda: [h.copy_(d, non_blocking=True) for h, d in zip(h_o, d_out)])
    h2d_bytes = float(sum(x.numel() * 4 for x in h_in))
    d2h_bytes = float(sum(x.numel() * 4 for x in d_out))
    best = min(legs.values(), key=lambda r: r["ms"])
    return {"npatch": n, "streams": legs, "h2d_ms": h2d, "kernel_ms": kernel_ms, "d2h_ms": d2h,
            "h2d_GBps": h2d_bytes / (h2d * 1e-3) / 1e9, "d2h_GBps": d2h_bytes / (d2h * 1e-3) / 1e9,
            "h2d_bytes": h2d_bytes, "d2h_bytes": d2h_bytes, "serial_sum_ms": h2d + kernel_ms + d2h,
            "best_ms": best["ms"], "overlap_saving": 1.0 - best["ms"] / (h2d + kernel_ms + d2h),
            "mean_iters": float(it.mean()),
            "note": "PCIe-inclusive rate of the host-array entry: never the headline value"}


def level_cycle(a, pg, ctx, args):
    """Two AMR levels through the host-side cycle a GAMER time step runs: pgamd_prepare_patch_groups gathers
    every patch group of the refined box (density with ghost cells from the fine siblings, coarse potential from
    the father's neighbourhood, fluid) into the pinned buffers, pgamd_asyn_poisson_gravity_solver solves them
    (PCIe included), pgamd_scatter_patch_groups writes potential and fluid back.  A Plummer sphere (M = 1,
    a = 6 cells) at the centre of the box; each stage timed on the host clock (median of the timed calls)."""
    from gpu_testbed_poisson_gravity_amd import inputs as I

    m, gs = a.level_box, a.gs
    ncp = m + 2
    F, L = 16 * m, 8 * ncp
    c = 8.0 * ncp
    f = 16.0 + np.arange(F) + 0.5
    r2 = (f[:, None, None] - c) ** 2 + (f[None, :, None] - c) ** 2 + (f[None, None, :] - c) ** 2
    fine = np.zeros((5, F, F, F), np.float32)
    fine[0] = 3.0 / (4.0 * np.pi * 6.0 ** 3) * (1.0 + r2 / 36.0) ** -2.5
    fine[4] = 1.0
    del r2
    X = 2.0 * np.arange(L) + 1.0
    coarse = (-I.NEWTON_G / np.sqrt((X[:, None, None] - c) ** 2 + (X[None, :, None] - c) ** 2 +
                                    (X[None, None, :] - c) ** 2 + 36.0)).astype(np.float32)
    d = I.two_level_box(fine, coarse, (1, 1, 1), seed=a.seed)
    del fine
    ng = d["group_first"].size
    n = 8 * ng
    ctx.set_options(pot_ghost_size=gs, gra_ghost_size=1, pot_scheme=pg.SOR, init_mode=pg.INIT_INTERP,
                    term_mode=pg.TERM_RELRES, check_every=a.check_every, rel_tol=1e-5, fuse_gravity=1,
                    guess_fallback_tol=0.0)
    h_rho, h_cpot, h_out, h_flu = ctx.mem_allocate(ng)
    fine_pot = np.zeros((n, 8, 8, 8), np.float32)
    fine_flu = d["fine_flu"].copy()
    nt = min(16, len(os.sched_getaffinity(0)))
    ts = {"prepare": [], "solve": [], "scatter": []}
    for k in range(a.warmup + max(3, a.steps // 4)):
        t0 = time.perf_counter()
        pg.prepare_patch_groups(gs, d["group_first"], d["group_father"], d["fine_flu"], d["fine_sib"],
                                d["coarse_pot"], d["coarse_sib"], h_rho=h_rho, h_pot_in=h_cpot, h_flu=h_flu,
                                nthreads=nt)
        t1 = time.perf_counter()
        ctx.asyn_poisson_gravity_solver(h_rho, h_cpot, h_out, h_flu, ng, args.dt, args.dh, args.sor_min_iter,
                                        args.sor_max_iter, args.sor_omega, args.mg_max_iter, args.mg_npre_smooth,
                                        args.mg_npost_smooth, args.mg_tolerated_error, args.poi_coeff,
                                        args.int_scheme, False, True, True, gpu_nstream=4)
        ctx.synchronize()
        t2 = time.perf_counter()
        pg.scatter_patch_groups(1, d["group_first"], h_out, h_flu, fine_pot=fine_pot, fine_flu=fine_flu, nthreads=nt)
        t3 = time.perf_counter()
        if k >= a.warmup:
            ts["prepare"].append(t1 - t0)
            ts["solve"].append(t2 - t1)
            ts["scatter"].append(t3 - t2)
    it, st = ctx.patch_stats(n)
    ctx.mem_free()
    dev_leg = level_cycle_device(a, pg, ctx, args, d, gs)
    if a.level_box == 12 and gs == 5:
        dev_leg["gather_roofline"]["traffic"] = measured_traffic(a).get(("pack_groups_kernel", "device"))
    ms = {k: float(np.median(v)) * 1e3 for k, v in ts.items()}
    tot = sum(ms.values())
    g = pg.geometry(gs)
    moved = n * 4.0 * (g["R"] ** 3 + g["C"] ** 3 + 2 * 5 * 512)      # bytes written into the pinned buffers
    out = {"fine_patches": n, "patch_groups": ng, "coarse_patches": int(ncp ** 3), "host_threads": nt,
           "prepare_ms": ms["prepare"], "solve_ms": ms["solve"], "scatter_ms": ms["scatter"], "total_ms": tot,
           "patch_solves_per_s": n / (tot * 1e-3), "prepare_GBps": moved / (ms["prepare"] * 1e-3) / 1e9,
           "mean_iters": float(it.mean()), "converged": int(np.sum((st & pg.STATUS_CONVERGED) != 0)),
           "note": "host-side AMR cycle (gather, PCIe-inclusive solve, scatter): never the headline value",
           "device": dev_leg}
    log(f"[bench] level cycle ({n} fine patches, {nt} host threads): prepare {ms['prepare']:.2f} ms, solve "
        f"{ms['solve']:.2f} ms, scatter {ms['scatter']:.2f} ms -> {out['patch_solves_per_s']:.4g} patch-solves/s")
    return out


def level_cycle_device(a, pg, ctx, args, d, gs):
    """The same cycle with both levels resident in HBM: pack_groups_kernel gathers the batch
    (pgamd_prepare_patch_groups_device), the solve runs on it (interpolated start, the context's rule),
    scatter_groups_kernel writes potential and fluid back.  Each phase timed with events on one stream.
    Gather roofline (HBM): per patch it writes (R^3 + C^3 + 5*8^3) floats and must read at least the patch's
    fluid (5*8^3) and its share of the coarse level (8^3/8), each once."""
    import torch

    dev = torch.device("cuda", 0)
    g = pg.geometry(gs)
    ng = d["group_first"].size
    n = 8 * ng
    t = {k: torch.as_tensor(np.ascontiguousarray(v)).to(dev) for k, v in d.items()
         if k not in ("fine_coords", "coarse_coords")}
    nfine, ncoarse = t["fine_flu"].shape[0], t["coarse_pot"].shape[0]
    rho = torch.empty((n, g["R"], g["R"], g["R"]), device=dev)
    pin = torch.empty((n, g["C"], g["C"], g["C"]), device=dev)
    flu = torch.empty((n, 5, 8, 8, 8), device=dev)
    pst = torch.empty(n, dtype=torch.int32, device=dev)
    out = torch.empty((n, 10, 10, 10), device=dev)
    it = torch.empty(n, dtype=torch.int32, device=dev)
    st = torch.empty(n, dtype=torch.int32, device=dev)
    fine_pot = torch.zeros((nfine, 8, 8, 8), device=dev)
    fine_flu = t["fine_flu"].clone()
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream

    def step(e=None):
        if e is not None:
            e[0].record(stream)
        pg.prepare_patch_groups_device(gs, ng, t["group_first"], t["group_father"], nfine, t["fine_flu"],
                                       t["fine_sib"], ncoarse, t["coarse_pot"], t["coarse_sib"], rho, pin, flu, pst,
                                       stream=sp)
        if e is not None:
            e[1].record(stream)
        ctx.solve_device(rho, pin, out, flu, n, args, iters=it, status=st, stream=sp)
        if e is not None:
            e[2].record(stream)
        pg.scatter_patch_groups_device(1, ng, t["group_first"], out, flu, nfine, fine_pot, fine_flu, stream=sp)
        if e is not None:
            e[3].record(stream)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize(dev)
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(a.steps)]
    for k in range(a.steps):
        step(evs[k])
    torch.cuda.synchronize(dev)
    ph = [float(np.mean([e[i].elapsed_time(e[i + 1]) for e in evs])) for i in range(3)]
    written = n * 4.0 * (g["R"] ** 3 + g["C"] ** 3 + 5 * 512)
    minread = n * 4.0 * (5 * 512 + 512 / 8)
    gather_alg = written + minread
    tot = sum(ph)
    res = {"prepare_ms": ph[0], "solve_ms": ph[1], "scatter_ms": ph[2], "total_ms": tot,
           "patch_solves_per_s": n / (tot * 1e-3), "mean_iters": float(it.float().mean().item()),
           "pack_flags": int((pst != 0).sum().item()),
           "gather_roofline": {"kernel": "pack_groups_kernel", "bound": "hbm", "algorithmic_bytes": gather_alg,
                               "achieved": gather_alg / (ph[0] * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                               "unit": "GB/s", "frac": gather_alg / (ph[0] * 1e-3) / HBM_PEAK,
                               "bytes_basis": "written (R^3 + C^3 + 5*8^3 floats) + minimum read (5*8^3 + 8^3/8) "
                                              "per patch"},
           "note": "levels resident in HBM: gather kernel, solve, scatter kernel; nothing crosses PCIe"}
    log(f"[bench] level cycle on the device: gather {ph[0]:.3f} ms, solve {ph[1]:.3f} ms, scatter {ph[2]:.3f} ms "
        f"-> {res['patch_solves_per_s']:.4g} patch-solves/s; gather {res['gather_roofline']['achieved']:.0f} GB/s")
    return res


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if a.workload != "config2":
        run_mixed(a)
        return
    import torch

    import gpu_testbed_poisson_gravity_amd as pg
    from gpu_testbed_poisson_gravity_amd import inputs as I

    dev = torch.device("cuda", 0)
    t0 = time.time()
    rho, cpot, flu, _ = I.plummer(a.npatch, a.gs, seed=a.seed, energy=1.0)
    log(f"[bench] inputs generated in {time.time() - t0:.1f}s")
    g = pg.geometry(a.gs)
    R = g["R"]
    d_rho = torch.from_numpy(rho).to(dev)
    d_cpot = torch.from_numpy(cpot).to(dev)
    d_flu0 = torch.from_numpy(flu).to(dev)
    d_flu = d_flu0.clone()
    d_out = torch.empty((a.npatch, 10, 10, 10), device=dev)
    N = g["N"]
    d_guess = torch.empty((a.npatch, N, N, N) if a.guess == "cube" else (a.npatch, R, R, R), device=dev)
    d_it = torch.zeros(a.npatch, dtype=torch.int32, device=dev)
    d_st = torch.zeros(a.npatch, dtype=torch.int32, device=dev)
    poi = 4.0 * np.pi * 1e-5
    args = pg.solver_args(dt=1e-3, sor_min_iter=10, sor_max_iter=1000, sor_omega=1.8, poi_coeff=poi)
    ctx = pg.Context(0)
    peaks = measure_peaks(ctx)
    lds_peak = peaks["lds_read_GBps"] * 1e9 if peaks else LDS_PEAK
    if a.pad is not None or a.ppw is not None:
        ppw0, pad0 = ctx.get_layout(a.gs)
        ctx.set_layout(a.gs, a.ppw if a.ppw is not None else ppw0, a.pad if a.pad is not None else pad0)
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream
    if a.copy_mib > 0:
        src = torch.ones(a.copy_mib * 2 ** 18, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize(dev)
        del src, dst

    def run(init_mode, net, pipelined=False):
        """warmup + timed steps; events bracket the network phase and the solve separately (serial legs)
        or the whole pipelined warm-start solve."""
        ctx.set_options(pot_ghost_size=a.gs, gra_ghost_size=1, pot_scheme=pg.SOR, init_mode=init_mode,
                        term_mode=pg.TERM_RELRES, check_every=a.check_every, rel_tol=1e-5, fuse_gravity=1,
                        guess_fallback_tol=a.fallback_tol if net is not None else 0.0)

        def step(e=None):
            d_flu.copy_(d_flu0)  # restore the fluid so every step does identical work (D2D, outside events)
            if e is not None:
                e[0].record(stream)
            if pipelined:
                if e is not None:
                    e[1].record(stream)
                net.solve_device(d_rho, d_cpot, d_guess, d_out, d_flu, a.npatch, args, iters=d_it, status=d_st,
                                 nchunk=a.chunks, stream=sp)
            else:
                if net is not None and a.guess == "cube":
                    net.warm_start_cube(d_rho, d_cpot, a.npatch, args, d_guess, stream=sp)
                elif net is not None:
                    net.warm_start(d_rho, d_cpot, a.npatch, args, d_guess, stream=sp)
                if e is not None:
                    e[1].record(stream)
                ctx.solve_device(d_rho, d_cpot, d_out, d_flu, a.npatch, args,
                                 guess=d_guess if net is not None else None, iters=d_it, status=d_st, stream=sp)
            if e is not None:
                e[2].record(stream)

        for _ in range(a.warmup):
            step()
        torch.cuda.synchronize(dev)
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(a.steps)]
        t_wall = time.perf_counter()
        for s in range(a.steps):
            step(evs[s])
        torch.cuda.synchronize(dev)
        wall = time.perf_counter() - t_wall
        pre = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
        sol = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
        it, st = d_it.cpu().numpy(), d_st.cpu().numpy()
        return {"value": a.npatch / ((pre + sol) * 1e-3), "ms_per_step": pre + sol, "network_ms": pre,
                "solve_ms": sol, "wall_ms_per_step": wall / a.steps * 1e3, "mean_iters": float(it.mean()),
                "p99_iters": float(np.percentile(it, 99)), "max_iters": int(it.max()),
                "converged": int(((st & pg.STATUS_CONVERGED) != 0).sum()),
                "fallbacks": int(((st & pg.STATUS_FALLBACK) != 0).sum()), "iters": it}

    traffic = measured_traffic(a)
    occ = ctx.layout_occupancy(a.gs)
    log(f"[bench] SOR residency at ghost size {a.gs}: {occ[0]} workgroups, {occ[1]} patches per CU")

    def lds_roof(r, leg):
        alg = float(r["iters"].astype(np.float64).sum()) * R ** 3 * LDS_BYTES_PER_UPDATE
        ach = alg / (r["solve_ms"] * 1e-3)
        # density, the start (coarse cube; or the guess: interior plus the coarse cube, or the whole cube),
        # the output window and the fluid (5 fields read, 4 written)
        start = {"plain": g["C"] ** 3, "interp": g["C"] ** 3}.get(
            leg, N ** 3 if a.guess == "cube" else R ** 3 + g["C"] ** 3)
        hbm = a.npatch * 4.0 * (R ** 3 + start + 10 ** 3 + 4 * 512 + 5 * 512)
        ops = float(r["iters"].astype(np.float64).sum()) * R ** 3 * VALU_OPS_PER_UPDATE
        return {"kernel": "sor_uniform_kernel", "bound": "lds", "achieved": ach / 1e9, "peak": lds_peak / 1e9,
                "valu": {"ops_per_update": VALU_OPS_PER_UPDATE, "achieved_Tops": ops / (r["solve_ms"] * 1e-3) / 1e12,
                         "peak_Tops": VALU_PEAK / 1e12, "frac": ops / (r["solve_ms"] * 1e-3) / VALU_PEAK,
                         "basis": "FP32 arithmetic of the updates alone; the kernel's real limit (no packed FP32)"},
                "resident_patches_per_cu": occ[1],
                "unit": "GB/s", "frac": ach / lds_peak, "traffic": traffic.get(("sor_uniform_kernel", leg)),
                "peak_basis": "measured LDS read bandwidth (pgamd_measure_peaks)" if peaks else "placeholder",
                "frac_of_placeholder": ach / LDS_PEAK,
                "algorithmic_bytes_per_launch": alg, "launch_ms": r["solve_ms"],
                "hbm_algorithmic_bytes_per_launch": hbm,
                "hbm_achieved": hbm / (r["solve_ms"] * 1e-3) / 1e9, "hbm_peak": HBM_PEAK / 1e9}

    plain = interp = warm = warm_serial = None
    if a.mode in ("both", "plain"):
        plain = run(pg.INIT_ZERO, None)
        log(f"[bench] plain: {plain['value']:.4g} patch-solves/s, {plain['mean_iters']:.1f} iterations")
        # the reference's own start (the interpolated coarse potential, CPU_PoissonSolver_SOR.cpp:65): the
        # method without the network
        interp = run(pg.INIT_INTERP, None)
        log(f"[bench] interp: {interp['value']:.4g} patch-solves/s, {interp['mean_iters']:.1f} iterations")
    fp64 = None
    if a.fp64 and a.mode in ("both", "plain"):
        # the FP64 device path (the reference's FLOAT8 build) on the same workload, plain leg's settings
        ctx.set_options(pot_ghost_size=a.gs, gra_ghost_size=1, pot_scheme=pg.SOR, init_mode=pg.INIT_ZERO,
                        term_mode=pg.TERM_RELRES, check_every=a.check_every, rel_tol=1e-5, fuse_gravity=1,
                        guess_fallback_tol=0.0)
        a64 = pg.solver_args_f64(dt=1e-3, sor_min_iter=10, sor_max_iter=1000, sor_omega=1.8, poi_coeff=poi)
        r64, c64, f64 = (x.double() for x in (d_rho, d_cpot, d_flu0))
        fl64 = f64.clone()
        o64 = torch.empty((a.npatch, 10, 10, 10), dtype=torch.float64, device=dev)
        steps64 = max(2, a.steps // 4)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        for k in range(1 + steps64):
            fl64.copy_(f64)
            if k == 1:
                ev[0].record(stream)
            ctx.solve_device_f64(r64, c64, o64, fl64, a.npatch, a64, iters=d_it, status=d_st, stream=sp)
        ev[1].record(stream)
        torch.cuda.synchronize(dev)
        ms64 = ev[0].elapsed_time(ev[1]) / steps64      # includes the D2D fluid restore (0.05 ms at 2.7 TB/s)
        it64 = d_it.cpu().numpy()
        fp64 = {"value": a.npatch / (ms64 * 1e-3), "ms_per_step": ms64, "mean_iters": float(it64.mean()),
                "iters_sum": float(it64.astype(np.float64).sum()),
                "converged": int(((d_st.cpu().numpy() & pg.STATUS_CONVERGED) != 0).sum()), "steps": steps64,
                "note": "sor64_kernel, zero guess to relative residual 1e-5 in FP64; the plain leg in FP64"}
        log(f"[bench] fp64 plain: {fp64['value']:.4g} patch-solves/s, {fp64['mean_iters']:.1f} iterations")
        del r64, c64, f64, fl64, o64
    net = None
    if a.mode in ("both", "warm"):
        if os.path.exists(a.weights):
            net = pg.CNN(ctx)
            net.load_params(a.weights)
            warm_serial = run(pg.INIT_GUESS_CUBE if a.guess == "cube" else pg.INIT_GUESS, net)
            log(f"[bench] warm (serial): {warm_serial['value']:.4g} patch-solves/s, "
                f"{warm_serial['mean_iters']:.1f} iterations, network {warm_serial['network_ms']:.3f} ms + "
                f"solve {warm_serial['solve_ms']:.3f} ms")
            warm = warm_serial
            if a.chunks > 0:
                warm_pipe = run(pg.INIT_GUESS_CUBE if a.guess == "cube" else pg.INIT_GUESS, net, pipelined=True)
                log(f"[bench] warm (pipelined, {a.chunks} slices): {warm_pipe['value']:.4g} patch-solves/s, "
                    f"{warm_pipe['mean_iters']:.1f} iterations")
                if warm_pipe["value"] > warm["value"]:
                    warm = warm_pipe
        else:
            log(f"[bench] no trained weights at {a.weights}: warm-start leg skipped")
    head_r = warm if warm is not None else plain
    # per-kernel rooflines come from the serial legs (each kernel timed alone on the stream)
    roof_sor = lds_roof(warm_serial, "warm") if warm_serial is not None else lds_roof(plain, "plain")
    roofs = {"sor": roof_sor}
    if warm_serial is not None and plain is not None:
        roofs["sor_plain"] = lds_roof(plain, "plain")
    if interp is not None:
        roofs["sor_interp"] = lds_roof(interp, "interp")
    if fp64 is not None:
        ops64 = fp64["iters_sum"] * R ** 3 * VALU_OPS_PER_UPDATE
        roofs["sor64"] = {"kernel": "sor64_kernel", "bound": "valu", "achieved": ops64 / (fp64["ms_per_step"] * 1e-3) / 1e12,
                          "peak": VALU_PEAK / 1e12, "unit": "Tops/s (FP64)",
                          "frac": ops64 / (fp64["ms_per_step"] * 1e-3) / VALU_PEAK,
                          "traffic": traffic.get(("sor64_kernel", "fp64")), "launch_ms": fp64["ms_per_step"]}
    roof = roof_sor
    if warm_serial is not None:
        flop = CNN_USEFUL_FLOP * a.npatch
        ach = flop / (warm_serial["network_ms"] * 1e-3)
        roofs["cnn"] = {"kernel": "fused_guess_kernel (interpolation, packing, 6 MFMA convolutions in LDS, guess)",
                        "bound": "mfma", "achieved": ach / 1e12, "peak": MFMA_BF16_PEAK / 1e12, "unit": "TFLOP/s",
                        "frac": ach / MFMA_BF16_PEAK, "traffic": traffic.get(("fused_guess_kernel", "warm")),
                        "algorithmic_flop_per_launch": flop, "flop_basis": "useful: 38.93 MFLOP per patch",
                        "launch_ms": warm_serial["network_ms"],
                        "hbm_algorithmic_bytes_per_launch":
                            a.npatch * 4.0 * (R ** 3 + g["C"] ** 3 + (N ** 3 if a.guess == "cube" else R ** 3))}
        if warm_serial["network_ms"] > warm_serial["solve_ms"]:
            roof = roofs["cnn"]

    cpu = cpu_port = None
    if a.cpu_sample > 0:
        import oracle as O

        n = min(a.cpu_sample, a.npatch)
        n -= n % 8                                              # whole patch groups for the reference entry
        nt = len(os.sched_getaffinity(0))
        nt = min(nt, int(os.environ.get("OMP_NUM_THREADS") or nt))   # the job's CPU share
        sample = (f"all {n} patches of the workload" if n == a.npatch else f"first {n} of the {a.npatch} patches")
        # the reference's own CPU solver (CPU_PoissonGravitySolver, compiled from its sources by oracle/Makefile
        # at -O3, OpenMP over patches): interpolated start and the stall rule only, so it runs the plain leg's
        # mean iteration count as a fixed count (Min_Iter = Max_Iter), then the gravity update
        k = int(round(plain["mean_iters"])) if plain is not None else 52
        if O.ref_available("SOR", a.gs, 1, "f32"):
            ref = O.RefSolver("SOR", a.gs, 1, "f32")
            t = time.perf_counter()
            ref.solve(rho[:n], cpot[:n], flu[:n], n // 8, 1e-3, 1.0, k, k, 1.8, 1, 1, 1, 1e-6, poi, pg.INT_CQUAD,
                      nthreads=nt)
            dt = time.perf_counter() - t
            cpu = {"value": n / dt, "unit": "patch-solves/s", "cores": nt, "kind": "reference",
                   "sample": sample + f", the reference CPU_PoissonGravitySolver (g++ -O3, OpenMP over patches, "
                                      f"{nt} threads), SOR omega 1.8 at {k} iterations (the plain leg's mean) from "
                                      "its interpolated start, then the gravity update",
                   "iterations": k, "seconds": dt}
        p = O.make_params(gs=a.gs, init_mode=O.INIT_ZERO, term_mode=O.TERM_RELRES, check_every=a.check_every,
                          rel_tol=1e-5, sor_min_iter=10, sor_max_iter=1000, sor_omega=1.8, poi_coeff=poi, dt=1e-3)
        t = time.perf_counter()
        _, _, it_cpu, _ = O.solve(rho[:n], cpot[:n], flu[:n], p, nthreads=nt)
        dt = time.perf_counter() - t
        cpu_port = {"value": n / dt, "unit": "patch-solves/s", "cores": nt, "kind": "port",
                    "sample": sample + ", zero guess to relative residual 1e-5, the FP32 oracle (C++ restatement, "
                                       "g++ -O3, OpenMP over patches)",
                    "mean_iters": float(it_cpu.mean()), "seconds": dt}
        if cpu is None:
            cpu = cpu_port

    asyn = None
    if a.asyn_streams and plain is not None:
        asyn = asyn_pipeline(a, pg, torch, ctx, rho, cpot, flu, args, plain["solve_ms"])
    level = level_cycle(a, pg, ctx, args) if a.level_box > 0 else None

    def strip(r):
        return None if r is None else {k: v for k, v in r.items() if k != "iters"}

    out = {
        "metric": METRIC, "value": head_r["value"], "unit": "patch-solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": head_r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{a.npatch} patches gs={a.gs} (18^3 cube), Plummer 1-8 spheres, SOR omega=1.8, "
                               f"rel. residual 1e-5, min 10 / max 1000 iterations, fused gravity; value = "
                               + ("learned warm start (network, then SOR from its guess with the guess check at 1e-2)"
                                  if warm is not None else "zero guess (no trained weights)"),
                   "npatch": a.npatch, "gs": a.gs, "check_every": a.check_every},
        "plain": strip(plain), "interp": strip(interp), "warm": strip(warm),
        "warm_vs_plain": (warm["value"] / plain["value"]) if (warm is not None and plain is not None) else None,
        "warm_vs_interp": (warm["value"] / interp["value"]) if (warm is not None and interp is not None) else None,
        "wall_ms_per_step": head_r["wall_ms_per_step"],
        "roofline": roof, "roofline_kernels": roofs, "cpu_baseline": cpu, "cpu_port": cpu_port,
        "peaks_measured": peaks, "asyn_pipeline": asyn, "level_cycle": level, "fp64": fp64,
    }
    print(json.dumps(out))
    if net is not None:
        net.close()
    ctx.close()


if __name__ == "__main__":
    main()
