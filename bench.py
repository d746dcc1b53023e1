"""MPDATA step benchmark on one MI355X.

Workload (default): ring3d plain MPDATA -- 128^3 triply periodic FP64 grid, solid-body
rotation about the axis (1,1,1)/sqrt(3) through the centre scaled to a maximum face
Courant number of 0.4, a Gaussian blob (sigma 8 cells) 30 cells off the axis, iord=2
with the nonoscillatory limiter, 128 steps.  A step is one MPDATA time step over the
whole grid; the fields are resident in HBM before timing starts.

Prints ONE JSON line:
  value        cell-updates/s over the K timed steps (HIP events on the launch stream)
  roofline     algorithmic bytes of a step (40 B per cell-update: read c^n and the three
               face Courant numbers, write c^{n+1}) over the summed average durations of
               the step's kernels, against MI355X's 8 TB/s HBM peak; per-kernel figures
               in roofline.kernels
  parity       max|gpu - oracle| / max|oracle| after 1 step and after the full run
  cpu_baseline the oracle (FP64 C restatement of the reference step, 1 thread) timed on
               the host cores over the same run, which is also the parity reference

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-steps S] [--no-cpu]
  python bench.py --calibrate-copy N     (plain device copy of N doubles, for counters)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published peak
BYTES_PER_CELL = 40    # c^n + U + V + W read, c^{n+1} written (FP64)


def face_courant_from_cell_winds(u, v, w, dt=1.0):
    """Face Courant numbers of a periodic grid with unit spacing from steady cell-centred
    winds (u0 == u1): 0.25*dt*(2u_i + 2u_{i+1}) (the formula of uu.c:132-134)."""
    U = 0.25 * dt * ((u + np.roll(u, -1, axis=2)) * 2.0)
    V = 0.25 * dt * ((v + np.roll(v, -1, axis=1)) * 2.0)
    W = 0.25 * dt * ((w + np.roll(w, -1, axis=0)) * 2.0)
    return U, V, W


def tuned_config(g):
    """The launch shapes a grid steps with: the antidiffusive kernel's tile and z-chunk,
    and the limiter kernel's (its own when the tuner chose one)."""
    (tx, ty, zc), (lx, ly, lz) = g.tiles()
    out = {"tile": [tx, ty], "zc": zc}
    if lx:
        out["limiter_tile"] = [lx, ly]
        out["limiter_zc"] = lz
    return out


def rotation_case(n=128, cmax=0.4, sigma=8.0, offset=30.0):
    """Solid-body rotation about (1,1,1)/sqrt(3) through the grid centre, scaled to a
    maximum face Courant number cmax; a Gaussian blob of width sigma centred `offset`
    cells from the axis along (1,-1,0)/sqrt(2).  Arrays are (z, y, x) raveled."""
    ax = np.arange(n, dtype=np.float64) - n / 2
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    k = np.ones(3) / np.sqrt(3.0)
    u = k[1] * Z - k[2] * Y  # omega * (k x r)
    v = k[2] * X - k[0] * Z
    w = k[0] * Y - k[1] * X
    U, V, W = face_courant_from_cell_winds(u, v, w)
    s = cmax / max(np.abs(U).max(), np.abs(V).max(), np.abs(W).max())
    U, V, W = U * s, V * s, W * s
    c0 = offset * np.array([1.0, -1.0, 0.0]) / np.sqrt(2.0)
    r2 = (X - c0[0]) ** 2 + (Y - c0[1]) ** 2 + (Z - c0[2]) ** 2
    c = np.exp(-0.5 * r2 / sigma ** 2)
    return {"U": U.ravel().copy(), "V": V.ravel().copy(), "W": W.ravel().copy(),
            "c": c.ravel().copy(), "shape": (n, n, n)}


def host_cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def relerr(got, want):
    return float(np.abs(got - want).max() / np.abs(want).max())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=128)
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--iord", type=int, default=2)
    ap.add_argument("--path", type=int, default=0, help="0 auto, 1 staged, 2 fused")
    ap.add_argument("--arith", type=int, default=1, help="0 exact, 1 fast")
    ap.add_argument("--cpu-steps", type=int, default=128,
                    help="oracle steps: the CPU baseline sample and the parity reference")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="N")
    ap.add_argument("--ensemble", type=int, default=192,
                    help="instances of the batched-ensemble comparison (0: skip)")
    ap.add_argument("--large", type=int, default=256,
                    help="edge of the large single-grid roofline run (0: skip)")
    ap.add_argument("--adjoint", type=int, default=256,
                    help="edge of the source-inversion adjoint leg (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=512,
                    help="time steps of the adjoint leg's loss window")
    ap.add_argument("--dropin", type=int, default=128,
                    help="edge of the drop-in (FRAME, smolar semantics) leg (0: skip)")
    ap.add_argument("--ring2d", type=int, default=128,
                    help="edge of the periodic 2-D ring leg, 1024 steps (0: skip)")
    ap.add_argument("--deform", type=int, default=1,
                    help="run the 512x512x256 iord-3 deformational leg (0: skip)")
    ap.add_argument("--deform-oracle-steps", type=int, default=1,
                    help="steps of the deformational leg checked against the oracle")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU (every box has one); the "
                          "step shards by z-slabs with halo exchange only by design -- run "
                          "with --gpus 1"}))
        sys.exit(2)

    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import mpdata_amd as M

    if args.calibrate_copy:
        ms = M.calibration_copy(args.calibrate_copy, 10)
        nbytes = 16 * args.calibrate_copy
        print(json.dumps({"calibration_copy_doubles": args.calibrate_copy, "ms": ms,
                          "GB/s": nbytes / ms / 1e6}))
        return

    n = args.n
    case = rotation_case(n)
    cells = n ** 3

    def make_grid():
        g = M.Grid(n, n, n, M.PERIODIC, iord=args.iord, nonosc=True)
        g.set_path(args.path)
        g.set_arith(args.arith)
        g.set_courant(case["U"], case["V"], case["W"])
        g.set_field(case["c"])
        return g

    g = make_grid()
    tuned = None
    if args.path != 1 and args.iord >= 2:
        g.autotune(2)  # each kernel's tile shape and z-chunk, chosen on this grid (state kept)
        tuned = tuned_config(g)
    g.step(args.warmup)
    g.synchronize()
    t0 = time.perf_counter()
    ev_ms = g.step_timed(args.steps)
    wall = time.perf_counter() - t0
    kms = g.kernel_times(8)
    g.close()
    ms_per_step = ev_ms / args.steps
    value = cells * args.steps / (ev_ms * 1e-3)

    # kernels of one step, in launch order, with their own input/output bytes per cell
    if args.path == 1:
        names = ["staged"] * len(kms)
        own = [None] * len(kms)
    else:
        names, own = [], []
        for p in range(1, args.iord):
            names += [f"k_f1 (pass {p})", f"k_f2 (pass {p})"]
            last = p == args.iord - 1
            own += [64 if p == 1 else 56, 48 if last else 72]
    kern_ms = sum(kms)
    achieved = cells * BYTES_PER_CELL / (kern_ms * 1e-3) / 1e9
    kernels = [{"name": nm, "avg_ms": t, "io_bytes_per_cell": b,
                "io_GB/s": (cells * b / (t * 1e-3) / 1e9) if b else None}
               for nm, t, b in zip(names, kms, own)]
    dominant = max(kernels, key=lambda k: k["avg_ms"])

    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_rotation{n}_iord{args.iord}.json")
    if os.path.exists(tfile) and args.path == 0 and args.arith == 1:
        t = json.load(open(tfile))
        # counters measured with the same tile shape and z-chunk length only
        if t.get("config") == tuned:
            traffic = t.get("hbm_bytes_per_step")

    line = {
        "metric": "cell-updates/s at <=1e-12 rel err (FP64 MPDATA step, 1 GPU)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic rotation wind and Gaussian blob, no RNG)",
        "config": {"workload": f"ring3d: periodic {n}^3 solid-body rotation about (1,1,1), "
                               f"max Courant 0.4, Gaussian blob sigma 8, iord={args.iord} "
                               "nonoscillatory", "grid": [n, n, n], "iord": args.iord,
                   "nonoscillatory": True, "boundary": "periodic",
                   "path": ["auto(fused)", "staged", "fused"][args.path],
                   "arith": ["exact", "fast"][args.arith], "autotuned": tuned},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_step": cells * BYTES_PER_CELL,
                     "step_kernel_ms": kern_ms, "dominant_kernel": dominant["name"],
                     "kernels": kernels},
        "wall_s_timed": wall,
    }

    import oracle_lib as O

    g = make_grid()
    g.step(1)
    got1 = g.field()
    want1 = O.per_step(n, n, n, case["U"], case["V"], case["W"], args.iord, True, case["c"])
    par = {"steps_1": {"rel_err": relerr(got1, want1)}}
    if not args.no_cpu:
        ks = max(1, args.cpu_steps)
        g.step(ks - 1)
        gotk = g.field()
        mass, cmin, cmax_ = g.stats()
        threads = O.get_threads()
        O.set_threads(1)
        t0 = time.perf_counter()
        wantk = O.per_run(n, n, n, case["U"], case["V"], case["W"], args.iord, True,
                          case["c"], ks)
        t_cpu = time.perf_counter() - t0
        # the same restatement over z planes on all the threads OpenMP is given
        nthr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or threads
        O.set_threads(nthr)
        t0 = time.perf_counter()
        wantp = O.per_run(n, n, n, case["U"], case["V"], case["W"], args.iord, True,
                          case["c"], ks)
        t_par = time.perf_counter() - t0
        O.set_threads(threads)
        par[f"steps_{ks}"] = {"rel_err": relerr(gotk, wantk),
                              "mass_rel_drift": abs(mass - case["c"].sum()) / case["c"].sum(),
                              "min": cmin}
        line["cpu_baseline"] = {
            "value": cells * ks / t_cpu, "unit": "cell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle/cpu_ref.c (FP64 restatement of the reference step with the "
                      f"same periodic + limiter options, 1 thread, {host_cpu_model()}) over "
                      f"{ks} steps of the same {n}^3 iord={args.iord} workload",
        }
        line["cpu_parallel"] = {
            "value": cells * ks / t_par, "unit": "cell-updates/s", "cores": nthr,
            "kind": "port",
            "sample": f"the same oracle loops with OpenMP over z planes ({nthr} threads, this "
                      f"build's parallel CPU, not the reference's), same {ks} steps",
            "bitwise_equal_to_1_thread": bool(np.array_equal(wantp, wantk)),
        }
    g.close()
    line["parity"] = par
    if args.ensemble:
        line["ensemble"] = ensemble_leg(M, args.ensemble)
    if args.large:
        line["large"] = large_leg(M, args.large, args.iord, args.arith)
    if args.ring2d:
        line["ring2d"] = ring2d_leg(M, args.ring2d)
    if args.dropin:
        line["dropin"] = dropin_leg(M, args.dropin)
    if args.deform:
        line["deform"] = deform_leg(M, oracle_steps=args.deform_oracle_steps)
    if args.adjoint:
        line["adjoint"] = adjoint_leg(M, args.adjoint, nsteps=args.adjoint_steps)
    print(json.dumps(line))


def ring2d_leg(M, n, nsteps=1024):
    """Config 1: periodic n x n 2-D grid, rotation scaled to max Courant 0.5, cone of
    height 1 and radius 15 cells at (64, 96), iord 2 nonoscillatory, 1024 steps; timed,
    then the final field against the oracle's trajectory."""
    import mpdata_amd.configs as K
    import oracle_lib as O

    U, V, W, c = K.ring2d(n)
    g = M.Grid(n, n, 1, M.PERIODIC, iord=2, nonosc=True)
    g.set_courant(U, V, W)
    g.set_field(c)
    ms = g.step_timed(nsteps) / nsteps
    got = g.field()
    mass, cmin, _ = g.stats()
    g.close()
    want = O.per_run(n, n, 1, U, V, W, 2, True, c, nsteps)
    return {"workload": f"periodic {n}x{n} rotation (max Courant 0.5), cone r=15 at (64,96), "
                        f"iord 2 nonoscillatory, {nsteps} steps",
            "ms_per_step": ms, "cell_updates_per_s": n * n / (ms * 1e-3),
            "rel_err_vs_oracle": relerr(got, want),
            "mass_rel_drift": abs(mass - c.sum()) / c.sum(), "min": cmin}


def dropin_leg(M, n, nsteps=64, order=2):
    """The drop-in path: smolar() semantics (FRAME mode: the reference's face layouts,
    frame held, its reduced patterns; no limiter) on a device-resident grid, which is
    what msmolar() steps.  n^3 with the rotation case's velocities as cell-centred winds
    (dx = dt = 1, max Courant ~0.4), order 2.  One step against the oracle's smolar()
    restatement (bitwise expected), then nsteps timed steps; cell-updates count the
    (n-2)^3 interior cells the reference updates."""
    import oracle_lib as O

    case = rotation_case(n)
    u, v, w, c = case["U"], case["V"], case["W"], case["c"]
    one = np.ones(n)
    g = M.Grid(n, n, n, M.FRAME, iord=order, nonosc=False)
    g.set_winds(one, one, one, 1.0, u, u, v, v, w, w)
    g.set_field(c)
    g.step(1)
    got = g.field()
    _, _, want = O.call_smolar(O.oracle().orc_smolar, n, n, n, one, one, one, 1.0, u, u, v, v,
                               w, w, c.copy(), np.zeros_like(c), order)
    g.step(3)
    ms = g.step_timed(nsteps) / nsteps
    g.close()
    return {"workload": f"FRAME (smolar semantics) {n}^3, rotation winds as cell-centred "
                        f"velocities, order {order}, device-resident (the msmolar path)",
            "ms_per_step": ms, "cell_updates_per_s": (n - 2) ** 3 / (ms * 1e-3),
            "rel_err_1_step_vs_oracle": relerr(got, want)}


def deform_leg(M, nsteps=200, oracle_steps=1):
    """Config 3: 512 x 512 x 256 periodic, deformational wind from 4 seeded Fourier modes
    per potential component (max Courant 0.5), 32 seeded Gaussians, iord 3
    nonoscillatory, 200 timed steps; oracle_steps steps against the oracle at full size
    first (one by default: the oracle runs ~1 s per step on 16 cores), and the whole
    trajectory against the exact-arithmetic kernels afterwards."""
    import mpdata_amd.configs as K
    import oracle_lib as O

    shape = (512, 512, 256)
    cells = shape[0] * shape[1] * shape[2]
    g = M.Grid(*shape, M.PERIODIC, iord=3, nonosc=True)
    K.deform(g, shape, seed=3)
    U, V, W = g.get_courant()
    c0 = g.field()
    mass0 = c0.sum()
    g.step(oracle_steps)
    got = g.field()
    t0 = time.perf_counter()
    want = c0
    for done in range(0, oracle_steps, 25):  # in chunks, a progress line after each
        want = O.per_run(*shape, U, V, W, 3, True, want, min(25, oracle_steps - done))
        if oracle_steps > 1:
            print(f"deform: oracle step {min(done + 25, oracle_steps)} of {oracle_steps}",
                  file=sys.stderr, flush=True)
    t_cpu = (time.perf_counter() - t0) / oracle_steps
    err = relerr(got, want)
    del U, V, W, got, want
    g.autotune(1)
    g.step(2)
    ms = g.step_timed(nsteps) / nsteps
    kms = g.kernel_times(2)
    mass, cmin, _ = g.stats()
    tuned = tuned_config(g)
    # the whole trajectory (oracle_steps + 2 + nsteps + 2 steps; the autotuner restores
    # the state it found) again in exact arithmetic, which equals the oracle's bit for bit
    taken = oracle_steps + 2 + nsteps + 2
    fast = g.field()
    g.set_arith(M.ARITH_EXACT)
    g.set_field(c0)
    g.step(taken)
    err_run = relerr(fast, g.field())
    del fast, c0
    g.close()
    ach = cells * BYTES_PER_CELL / (sum(kms) * 1e-3) / 1e9
    return {"workload": "periodic 512x512x256 deformational wind (seed 3, 4 modes per "
                        "component, max Courant 0.5), 32 Gaussians, iord 3 nonoscillatory",
            "autotuned": tuned, "steps": nsteps, "ms_per_step": ms,
            "cell_updates_per_s": cells / (ms * 1e-3), "kernel_ms": kms,
            "achieved_GB/s": ach, "frac_of_hbm_peak": ach / HBM_PEAK_GBS,
            f"rel_err_{oracle_steps}_step{'s' if oracle_steps > 1 else ''}_vs_oracle": err,
            f"rel_err_{taken}_steps_vs_exact": err_run,
            "oracle_cell_updates_per_s": cells / t_cpu,
            "oracle_threads": O.get_threads(),
            "mass_rel_drift": abs(mass - mass0) / mass0, "min": cmin}


def ensemble_leg(M, n):
    """Config 4: n instances of uneven sizes, iord 2 nonoscillatory, 100 steps as one
    autotuned hipGraph replay against per-instance launch sequences; afterwards one
    instance of each size against a single grid (bitwise) and the smallest against the
    oracle."""
    import mpdata_amd.configs as K

    import oracle_lib as O

    t0 = time.perf_counter()
    b = K.ensemble(n, seed=4)
    cells = b.cells()
    setup = time.perf_counter() - t0
    # the first instance of each size, checked after the run against single grids
    dims = K.ensemble_dims(n, seed=4)
    firsts = sorted({d: i for i, d in reversed(list(enumerate(dims)))}.values())
    start = {i: (b.instances[i].get_courant(), b.instances[i].field()) for i in firsts}
    tx, ty, zc, pers = b.autotune(2)
    lim = b.tiles()[1]
    b.step(2, M.BATCH_GRAPH)
    ms_graph = b.step_timed(100, M.BATCH_GRAPH) / 100
    ms_fused = b.step_timed(10, M.BATCH_FUSED) / 10
    b.step(1, M.BATCH_PER_INSTANCE)
    ms_per = b.step_timed(10, M.BATCH_PER_INSTANCE) / 10
    taken = 2 + 100 + 10 + 1 + 10
    mass_ok = all(np.isfinite(g.stats()[0]) for g in b.instances[:4])
    # every mode gives the bits of a single grid (default launch shape), and the smallest
    # instance stays within the tolerance of the oracle over the whole run
    same, err = True, None
    for i in firsts:
        (U, V, W), c = start[i]
        got = b.instances[i].field()
        g = M.Grid(*dims[i], M.PERIODIC, iord=2, nonosc=True)
        g.set_courant(U, V, W)
        g.set_field(c)
        g.step(taken)
        same = same and bool(np.array_equal(got, g.field()))
        g.close()
        if dims[i] == min(dims, key=lambda d: d[0] * d[1] * d[2]):
            err = relerr(got, O.per_run(*dims[i], U, V, W, 2, True, c, taken))
    b.close()
    return {"workload": f"{n} periodic instances from {{64^3, 96^3, 128^3, 160x160x96, "
                        f"192x192x64}} (seed 4), rotation + 10% seeded modes, iord 2 "
                        "nonoscillatory", "cells": cells, "setup_s": setup,
            "autotuned": dict({"tile": [tx, ty], "zc": zc, "persistent": pers},
                              **({"limiter_tile": list(lim[:2]), "limiter_zc": lim[2]}
                                 if lim[0] else {})),
            "graph_ms_per_step": ms_graph, "fused_launch_ms_per_step": ms_fused,
            "per_instance_ms_per_step": ms_per,
            "graph_cell_updates_per_s": cells / (ms_graph * 1e-3),
            "per_instance_cell_updates_per_s": cells / (ms_per * 1e-3),
            "speedup_graph_vs_per_instance": ms_per / ms_graph, "finite": mass_ok,
            f"bitwise_vs_single_grids_after_{taken}_steps":
                {"instances": firsts, "equal": same},
            f"rel_err_smallest_vs_oracle_after_{taken}_steps": err}


def large_leg(M, n, iord, arith):
    """Roofline of the single-grid step at n^3 (the size the 70%-of-HBM target names); one
    step against the oracle, and the whole run against the exact-arithmetic kernels."""
    import mpdata_amd.configs as K
    import oracle_lib as O

    g = M.Grid(n, n, n, M.PERIODIC, iord=iord, nonosc=True)
    g.set_arith(arith)
    K.ring3d(g, n)
    U, V, W = g.get_courant()
    c0 = g.field()
    g.autotune(2)
    g.step(1)
    err1 = relerr(g.field(), O.per_step(n, n, n, U, V, W, iord, True, c0))
    del U, V, W
    g.step(1)
    ms = g.step_timed(20) / 20
    kms = g.kernel_times(4)
    tuned = tuned_config(g)
    taken = 2 + 20 + 4
    fast = g.field()
    g.set_arith(M.ARITH_EXACT)
    g.set_field(c0)
    g.step(taken)
    err_run = relerr(fast, g.field())
    del fast, c0
    g.close()
    cells = n ** 3
    kern = sum(kms)
    ach = cells * BYTES_PER_CELL / (kern * 1e-3) / 1e9
    return {"workload": f"periodic {n}^3 rotation, iord {iord} nonoscillatory",
            "autotuned": tuned, "ms_per_step": ms, "cell_updates_per_s": cells / (ms * 1e-3),
            "kernel_ms": kms, "achieved_GB/s": ach, "frac_of_hbm_peak": ach / HBM_PEAK_GBS,
            "rel_err_1_step_vs_oracle": err1, f"rel_err_{taken}_steps_vs_exact": err_run}


def adjoint_leg(M, n, nsteps=512, every=16, K=32, nrec=64):
    """Config 5: n^3 periodic, deformational wind (seed 5), 32 Gaussian blobs, emission of
    8 seeded positive Gaussians added every step; 64 seeded receptors sampled every 16
    steps against a twin run whose emission amplitudes are scaled by (1 + 0.2 xi);
    gradient of J = 1/2 sum (c - obs)^2 w.r.t. c^0 and the emission, checkpoints every 32
    steps, iord 2 with the limiter.  The gradient run keeps the records of as many final
    segments as the device memory holds and recomputes the others; it is repeated with
    only the last segment kept (pure checkpointing, same gradient bit for bit).  The
    adjoint's cost is the gradient run's time beyond a plain exact forward of the same
    steps, in forward steps: (forward sweep + backward - plain forward) / plain forward.
    Also reported: central
    differences along a direction scaled by c^0 (limiter on, and the same with the limiter
    off, where the step is differentiable), and the gradient against the CPU oracle's at
    32^3 x 64 steps."""
    import mpdata_amd.configs as K_
    import oracle_lib as O

    rng = K_.SplitMix64(5005)
    rec = np.array(sorted({rng.randint(n ** 3) for _ in range(nrec * 2)})[:nrec], np.int32)
    xi = [rng.uniform(-1.0, 1.0) for _ in range(8)]

    def fresh(nonosc=True, scale=None, size=n):
        g = M.Grid(size, size, size, M.PERIODIC, iord=2, nonosc=nonosc)
        cen, sig, amp = K_.deform(g, (size, size, size), seed=5, emission_blobs=8)
        if scale is not None:
            K_.generate_blobs(g, cen, sig, [a * f for a, f in zip(amp, scale)],
                              periodic=True, target=K_.TARGET_EMISSION)
        return g

    twin = fresh(scale=[1.0 + 0.2 * x for x in xi])
    obs = twin.sample_receptors(nsteps, rec, every)
    twin.close()
    g = fresh()
    c0 = g.field()
    # the gradient run's steps are exact-arithmetic fused steps: their launch shapes are
    # tuned in that arithmetic (the plain forward it is compared with uses the same)
    g.set_arith(M.ARITH_EXACT)
    g.autotune(2)
    tuned = tuned_config(g)
    # records of as many final segments as the device memory holds are kept from the
    # forward sweep (the rest are recomputed from their checkpoints)
    g.set_adjoint_tape(-1)
    J, g0, gq, (tf, tb, tr) = g.receptor_gradient(nsteps, K, rec, every, obs)
    kept = g.adjoint_tape()[1]
    # the same gradient with only the last segment's record kept (pure checkpointing)
    g.set_field(c0)
    g.set_adjoint_tape(0)
    J_c, g0_c, gq_c, (tf_c, tb_c, tr_c) = g.receptor_gradient(nsteps, K, rec, every, obs)
    same = bool(J_c == J and np.array_equal(g0_c, g0) and np.array_equal(gq_c, gq))
    # a plain forward of the same steps in the driver's (exact) arithmetic
    g.set_field(c0)
    g.set_arith(M.ARITH_EXACT)
    tp = g.step_timed(nsteps)
    g.close()
    d = np.random.default_rng(5).standard_normal(c0.size) * c0

    def fd_check(nonosc, grad, epss):
        res = {}
        for e in epss:
            vals = []
            for sgn in (1, -1):
                gx = fresh(nonosc)
                gx.set_field(c0 + sgn * e * d)
                vals.append(gx.receptor_gradient(nsteps, K, rec, every, obs, grads=False)[0])
                gx.close()
            fd, an = (vals[0] - vals[1]) / (2 * e), float(grad @ d)
            res[f"{e:g}"] = {"fd": fd, "adjoint": an,
                             "rel_diff": abs(fd - an) / max(abs(an), 1e-300)}
        return res

    fd_lim = fd_check(True, g0, (1e-6, 1e-8))
    g = fresh(nonosc=False)
    _, g0u, _, _ = g.receptor_gradient(nsteps, K, rec, every, obs)
    g.close()
    fd_unlim = fd_check(False, g0u, (1e-6,))

    # CPU reference gradient at 32^3 x 64 steps (same generators, limiter on)
    m, ms, mrec = 32, 64, np.array(sorted({rng.randint(32 ** 3) for _ in range(40)})[:16], np.int32)
    gs = fresh(size=m)
    U, V, W = gs.get_courant()
    cs = gs.field()
    # the emission increment is read back through a zero-wind step: c + q - c
    gz = M.Grid(m, m, m, M.PERIODIC, iord=2, nonosc=True)
    z = np.zeros(m ** 3)
    gz.set_courant(z, z, z)
    gz.set_field(z)
    K_.deform(gz, (m, m, m), seed=5, emission_blobs=8)
    gz.set_courant(z, z, z)
    gz.set_field(z)
    gz.step(1)
    qs = gz.field()
    gz.close()
    obs_s = np.zeros((ms // 16, mrec.size))
    Jg, g0g, gqg, _ = gs.receptor_gradient(ms, 16, mrec, 16, obs_s)
    gs.close()
    Jc, g0c, gqc = O.receptor_loss_and_gradient(m, m, m, U, V, W, True, cs, qs, ms, mrec, 16,
                                                obs_s)
    cpu = {"grid": [m, m, m], "steps": ms, "J_rel_diff": abs(Jg - Jc) / Jc,
           "grad_c0_rel_diff": float(np.abs(g0g - g0c).max() / np.abs(g0c).max()),
           "grad_q_rel_diff": float(np.abs(gqg - gqc).max() / np.abs(gqc).max())}
    return {"workload": f"periodic {n}^3 deformational wind (seed 5), 8-source emission, "
                        f"{nsteps} steps, {nrec} receptors every {every} steps, checkpoints "
                        f"every {K}, iord 2 nonoscillatory",
            "autotuned": tuned, "J": J, "forward_ms_per_step": tp / nsteps,
            "forward_sweep_ms_per_step": tf / nsteps, "backward_ms_per_step": tb / nsteps,
            "adjoint_over_forward": (tf + tb - tp) / tp,
            "kept_segments": kept, "segments": -(-nsteps // K),
            "recompute_ms_per_step": tr / nsteps,
            "adjoint_steps_only_over_forward": (tb - tr) / tp,
            "checkpointing_only": {
                "forward_sweep_ms_per_step": tf_c / nsteps,
                "backward_ms_per_step": tb_c / nsteps,
                "recompute_ms_per_step": tr_c / nsteps,
                "adjoint_over_forward": (tf_c + tb_c - tp) / tp,
                "same_gradient_bitwise": same},
            "fd_limiter_on": fd_lim, "fd_limiter_off": fd_unlim, "cpu_reference": cpu,
            "grad_norms": {"c0": float(np.linalg.norm(g0)), "q": float(np.linalg.norm(gq))}}


if __name__ == "__main__":
    main()
