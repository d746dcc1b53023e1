"""MPDATA step benchmark on one MI355X.

Workload (the headline line): the large single-instance case -- a 512 x 512 x 256
triply periodic FP64 grid, divergence-free deformational wind curl(A) of 4 seeded
Fourier modes per potential component scaled to a maximum face Courant number of 0.5,
the field a seeded sum of 32 positive Gaussians (sigma 4..16 cells), iord=3 with the
nonoscillatory limiter.  A step is one MPDATA time step over the whole grid; the fields
are resident in HBM before timing starts.  The largest single-GPU configuration, and the
one whose working set (~7 GB) does not fit the 256 MB Infinity Cache.

Prints ONE JSON line:
  value        cell-updates/s over the K timed steps (HIP events on the launch stream)
  roofline     algorithmic bytes of a step (40 B per cell-update: read c^n and the three
               face Courant numbers, write c^{n+1}) over the summed average durations of
               the step's kernels (each launched between HIP events of its own), against
               MI355X's 8 TB/s HBM peak; traffic = rocprof HBM bytes of a step (FETCH_SIZE
               + WRITE_SIZE, calibrated against a plain copy) when measured for the same
               launch shapes; per-kernel figures in roofline.kernels; roofline.valu, the
               vector-ALU side from the rocprof instruction counts of the same launch
               shapes: VALU issue rate, FP64 TFLOP/s and the HBM fraction the step would
               reach at full issue (its instruction-throughput ceiling)
  parity       max|gpu - oracle| / max|oracle| after the CPU-baseline steps, and the whole
               timed trajectory against the exact-arithmetic kernels (which equal the
               oracle bit for bit)
  cpu_baseline the oracle (FP64 C restatement of the reference step with the periodic and
               limiter options, 1 thread) timed on the host over the first steps of the
               same workload, which are also the parity reference; cpu_calibration times
               the reference's own smolar() (compiled from its sources, when present)
               against the restatement's on the same frame-mode input
Legs (objects on the same line): ring3d 128^3 (the cache-resident rotation case), the
256^3 rotation, the 192-instance ensemble, ring2d, the drop-in frame path and the
source-inversion adjoint.

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
# the vector ALU side of MI355X: 256 CUs, each issuing at most one wave64 VALU instruction
# per cycle (FP64 included), at the 2.4 GHz peak engine clock; FP64 vector peak 78.6 TFLOP/s
CUS, CLOCK_GHZ, FP64_PEAK_TFLOPS = 256, 2.4, 78.6


def valu_roofline(counts, kern_ms, hbm_frac):
    """The step's vector-ALU side from its rocprof instruction counts (wave64 VALU
    instructions and FP64 FLOP per step, measured with the same launch shapes) over the
    live kernel time: the VALU issue rate as a fraction of one instruction per CU and
    cycle, the FP64 rate against its peak, and the HBM fraction the same step would reach
    at full issue (the instruction-throughput ceiling in the HBM roofline's terms)."""
    if not counts:
        return None
    issue = counts["valu_wave_instr_per_step"] / (kern_ms * 1e-3 * CUS * CLOCK_GHZ * 1e9)
    tflops = counts["fp64_flop_per_step"] / (kern_ms * 1e-3) / 1e12
    return {"valu_wave_instr_per_step": counts["valu_wave_instr_per_step"],
            "issue_frac": issue, "fp64_TFLOP/s": tflops,
            "fp64_frac": tflops / FP64_PEAK_TFLOPS,
            "hbm_frac_at_full_issue": hbm_frac / issue,
            "source": counts.get("valu_source")}
DEFORM_SHAPE = (512, 512, 256)


def face_courant_from_cell_winds(u, v, w, dt=1.0):
    """Face Courant numbers of a periodic grid with unit spacing from steady cell-centred
    winds (u0 == u1): 0.25*dt*(2u_i + 2u_{i+1}) (the formula of uu.c:132-134)."""
    U = 0.25 * dt * ((u + np.roll(u, -1, axis=2)) * 2.0)
    V = 0.25 * dt * ((v + np.roll(v, -1, axis=1)) * 2.0)
    W = 0.25 * dt * ((w + np.roll(w, -1, axis=0)) * 2.0)
    return U, V, W


def tuned_config(g):
    """The launch shapes a grid steps with: one kernel per corrective pass (32x16 tiles,
    its z-chunk), or the antidiffusive kernel's tile and z-chunk and the limiter kernel's
    (its own when the tuner chose one)."""
    if g.path() == 3:
        return {"kernel": "one per corrective pass", "tile": [32, 16], "zc": g.pass_zchunk(),
                "waves": "8 tile + 4 ring"}
    if g.path() == 4:
        return {"kernel": "one per corrective pass, single-plane form", "tile": [16, 16]}
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


def host_cores():
    """Logical CPUs of the host and those this process may run on."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"logical_cpus": os.cpu_count(), "available_to_process": avail,
            "model": host_cpu_model()}


def relerr(got, want):
    return float(np.abs(got - want).max() / np.abs(want).max())


def step_kernels(M, iord, kms, path=0):
    """Names and own input/output bytes per cell of the kernels of one step, in launch
    order, with their measured average durations."""
    if path == 1:
        return [{"name": "staged", "avg_ms": t, "io_bytes_per_cell": None, "io_GB/s": None}
                for t in kms]
    names, own = [], []
    if len(kms) == iord - 1:  # one kernel per corrective pass
        for p in range(1, iord):
            last = p == iord - 1
            names.append(f"k_f12 (pass {p})")
            # reads c^n, the pass input (later passes), the three faces; writes the
            # result (and the limited faces when another pass follows)
            own.append((32 if p == 1 else 40) + (8 if last else 32))
        return [{"name": nm, "avg_ms": t, "io_bytes_per_cell": b}
                for nm, t, b in zip(names, kms, own)]
    for p in range(1, iord):
        names += [f"k_f1 (pass {p})", f"k_f2 (pass {p})"]
        last = p == iord - 1
        own += [64 if p == 1 else 56, 48 if last else 72]
    return [{"name": nm, "avg_ms": t, "io_bytes_per_cell": b}
            for nm, t, b in zip(names, kms, own)]


def roofline(cells, kms, kernels, tfile, tuned):
    """Step-level roofline: 40 algorithmic bytes per cell over the summed kernel time."""
    kern_ms = sum(kms)
    achieved = cells * BYTES_PER_CELL / (kern_ms * 1e-3) / 1e9
    for k in kernels:
        b = k["io_bytes_per_cell"]
        k["io_GB/s"] = (cells * b / (k["avg_ms"] * 1e-3) / 1e9) if b else None
    dominant = max(kernels, key=lambda k: k["avg_ms"])
    traffic, counts = None, None
    if tfile and os.path.exists(tfile):
        t = json.load(open(tfile))
        # counters measured with the same launch shapes only
        if t.get("config") == tuned:
            traffic = t.get("hbm_bytes_per_step")
            if "valu_wave_instr_per_step" in t:
                counts = t
    frac = achieved / HBM_PEAK_GBS
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": frac, "traffic": traffic,
            "algorithmic_bytes_per_step": cells * BYTES_PER_CELL, "step_kernel_ms": kern_ms,
            "dominant_kernel": dominant["name"], "kernels": kernels,
            "valu": valu_roofline(counts, kern_ms, frac)}


def cpu_calibration(O, n=128, order=2, reps=2):
    """The restatement's frame-mode step (orc_smolar) against the reference's own
    smolar() compiled from its sources (oracle/_ref), same n^3 input, best of reps."""
    if not O.have_ref():
        return None
    case = rotation_case(n)
    u, v, w, c = case["U"], case["V"], case["W"], case["c"]
    one = np.ones(n)
    out, best = {}, {}
    for name, fn in (("reference", O.ref().smolar), ("restatement", O.oracle().orc_smolar)):
        best[name] = 1e30
        for _ in range(reps):
            t0 = time.perf_counter()
            _, _, res = O.call_smolar(fn, n, n, n, one, one, one, 1.0, u, u, v, v, w, w,
                                      c.copy(), np.zeros_like(c), order)
            best[name] = min(best[name], time.perf_counter() - t0)
        out[name] = res
    cells = (n - 2) ** 3
    return {"input": f"frame mode {n}^3 order {order}, rotation winds (the reference's "
                     "semantics, no limiter), 1 thread",
            "reference_cell_updates_per_s": cells / best["reference"],
            "restatement_cell_updates_per_s": cells / best["restatement"],
            "restatement_over_reference_time": best["restatement"] / best["reference"],
            "bitwise_equal": bool(np.array_equal(out["reference"], out["restatement"]))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--arith", type=int, default=1, help="0 exact, 1 fast")
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="oracle steps of the headline workload: the CPU baseline sample "
                         "and the parity reference")
    ap.add_argument("--repeats", type=int, default=5,
                    help="further timed runs of K headline steps after the timed one, "
                         "reported with their median beside the value (0: none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="N")
    ap.add_argument("--ring3d", type=int, default=128,
                    help="edge of the cache-resident rotation leg (0: skip)")
    ap.add_argument("--ensemble", type=int, default=192,
                    help="instances of the batched-ensemble comparison (0: skip)")
    ap.add_argument("--large", type=int, default=256,
                    help="edge of the large rotation roofline leg (0: skip)")
    ap.add_argument("--adjoint", type=int, default=256,
                    help="edge of the source-inversion adjoint leg (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=512,
                    help="time steps of the adjoint leg's loss window")
    ap.add_argument("--adjoint-checks", type=int, default=1,
                    help="full-size dot-product test, FD directions and eps sweeps (0: skip)")
    ap.add_argument("--dropin", type=int, default=128,
                    help="edge of the drop-in (FRAME, smolar semantics) leg (0: skip)")
    ap.add_argument("--ring2d", type=int, default=128,
                    help="edge of the periodic 2-D ring leg, 1024 steps (0: skip)")
    ap.add_argument("--legs", type=int, default=1, help="0: the headline only")
    ap.add_argument("--path", type=int, default=0, choices=(0, 2, 3),
                    help="step path of the headline, ring3d and large grids: 0 automatic, 2 "
                         "two kernels per corrective pass, 3 one kernel per pass")
    ap.add_argument("--only", default="",
                    help="run one workload alone and print its object (for profiling): "
                         "headline, ring3d, large, ensemble, ring2d, dropin, adjoint")
    ap.add_argument("--profile", action="store_true",
                    help="skip the exact-arithmetic re-runs and CPU checks (the exact and "
                         "fast kernels share names, which would mix in a profiler's "
                         "statistics)")
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

    if args.profile:
        args.no_cpu = True
    if args.only and args.only != "headline":
        legs = {"ring3d": lambda: ring3d_leg(M, args.ring3d, args.arith, check=not args.profile,
                                             path=args.path),
                "large": lambda: large_leg(M, args.large, 2, args.arith, check=not args.profile,
                                           path=args.path),
                "ensemble": lambda: ensemble_leg(M, args.ensemble),
                "ring2d": lambda: ring2d_leg(M, args.ring2d),
                "dropin": lambda: dropin_leg(M, args.dropin),
                "adjoint": lambda: adjoint_leg(M, args.adjoint, nsteps=args.adjoint_steps,
                                               checks=bool(args.adjoint_checks))}
        print(json.dumps({args.only: legs[args.only]()}))
        return
    line = headline(M, args)
    if args.legs and not args.only:
        if args.ring3d:
            line["ring3d"] = ring3d_leg(M, args.ring3d, args.arith, path=args.path)
        if args.ensemble:
            line["ensemble"] = ensemble_leg(M, args.ensemble)
        if args.large:
            line["large"] = large_leg(M, args.large, 2, args.arith, path=args.path)
        if args.ring2d:
            line["ring2d"] = ring2d_leg(M, args.ring2d)
        if args.dropin:
            line["dropin"] = dropin_leg(M, args.dropin)
        if args.adjoint:
            line["adjoint"] = adjoint_leg(M, args.adjoint, nsteps=args.adjoint_steps,
                                          checks=bool(args.adjoint_checks))
    print(json.dumps(line))


def headline(M, args):
    """The 512 x 512 x 256 iord-3 deformational case: autotuned launch shapes, W warm-up
    steps, K timed steps; then the oracle (1 thread, the CPU baseline) over the first S
    steps of the same workload against a GPU run of S steps, and the whole timed
    trajectory against the exact-arithmetic kernels."""
    import mpdata_amd.configs as K
    import oracle_lib as O

    shape = DEFORM_SHAPE
    nx, ny, nz = shape
    cells = nx * ny * nz
    g = M.Grid(*shape, M.PERIODIC, iord=3, nonosc=True)
    g.set_arith(args.arith)
    g.set_path(args.path)
    K.deform(g, shape, seed=3)
    c0 = g.field()
    mass0 = c0.sum()
    g.autotune(1)  # each kernel's tile shape and z-chunk on this grid (state kept)
    tuned = tuned_config(g)
    print("bench: headline tuned", tuned, file=sys.stderr, flush=True)
    g.step(args.warmup)
    g.synchronize()
    t0 = time.perf_counter()
    ev_ms = g.step_timed(args.steps)
    wall = time.perf_counter() - t0
    kms = g.kernel_times(2)
    mass, cmin, _ = g.stats()
    # the state after the warm-up, timed and kernel-timing steps (compared below with the
    # same steps in exact arithmetic)
    fast = None if args.profile else g.field()
    # the spread of the figure: R more runs of K steps, each timed the same way
    reps = [] if args.profile else [g.step_timed(args.steps) / args.steps
                                    for _ in range(args.repeats)]
    ms_per_step = ev_ms / args.steps
    value = cells * args.steps / (ev_ms * 1e-3)
    kernels = step_kernels(M, 3, kms)
    tfile = os.path.join(ROOT, "profiles", "traffic_deform512_iord3.json")
    roof = roofline(cells, kms, kernels, tfile if args.arith == 1 else None, tuned)
    # those steps (the autotuner restores the state it found) again in exact arithmetic,
    # which equals the oracle's bit for bit
    taken = args.warmup + args.steps + 2
    err_run = None
    if not args.profile:
        g.set_arith(M.ARITH_EXACT)
        g.set_field(c0)
        g.step(taken)
        err_run = relerr(fast, g.field())
        del fast
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
        "data": "synthetic (seeded splitmix64 deformational wind and Gaussians, generated "
                "on the device)",
        "config": {"workload": "large single instance: periodic 512x512x256 FP64, "
                               "divergence-free deformational wind (4 Fourier modes per "
                               "potential component, seed 3, max Courant 0.5), 32 seeded "
                               "Gaussians sigma 4..16, iord=3 nonoscillatory",
                   "grid": list(shape), "iord": 3, "nonoscillatory": True,
                   "boundary": "periodic", "arith": ["exact", "fast"][args.arith],
                   "autotuned": tuned},
        "roofline": roof,
        "wall_s_timed": wall,
        "mass_rel_drift": abs(mass - mass0) / mass0, "min": cmin,
    }
    if reps:
        line["repeats"] = {"runs": len(reps), "steps_each": args.steps,
                           "ms_per_step": reps, "median_ms_per_step": float(np.median(reps)),
                           "median_cell_updates_per_s": cells / (float(np.median(reps)) * 1e-3)}
    par = {f"steps_{taken}_vs_exact_kernels": {"rel_err": err_run}} if err_run is not None else {}
    if not args.no_cpu and args.cpu_steps > 0:
        ks = args.cpu_steps
        g.set_arith(args.arith)
        g.set_field(c0)
        g.step(ks)
        got = g.field()
        U, V, W = g.get_courant()
        g.close()
        threads = O.get_threads()
        O.set_threads(1)
        t0 = time.perf_counter()
        want = O.per_run(*shape, U, V, W, 3, True, c0, ks)
        t_cpu = time.perf_counter() - t0
        print(f"bench: oracle 1 thread {t_cpu:.1f} s for {ks} steps", file=sys.stderr,
              flush=True)
        nthr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or threads
        O.set_threads(nthr)
        t0 = time.perf_counter()
        wantp = O.per_run(*shape, U, V, W, 3, True, c0, ks)
        t_par = time.perf_counter() - t0
        O.set_threads(threads)
        del U, V, W
        par[f"steps_{ks}_vs_oracle"] = {"rel_err": relerr(got, want)}
        line["cpu_baseline"] = {
            "value": cells * ks / t_cpu, "unit": "cell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle/cpu_ref.c (FP64 restatement of the reference step with the "
                      f"same periodic + limiter options, 1 thread of {host_cpu_model()}) "
                      f"over the first {ks} steps of the same 512x512x256 iord=3 workload",
            "host": host_cores(),
        }
        line["cpu_parallel"] = {
            "value": cells * ks / t_par, "unit": "cell-updates/s", "cores": nthr,
            "kind": "port",
            "sample": f"the same oracle loops with OpenMP over z planes ({nthr} threads, this "
                      f"build's parallel CPU, not the reference's), same {ks} steps",
            "bitwise_equal_to_1_thread": bool(np.array_equal(wantp, want)),
        }
        del want, wantp, got
        cal = cpu_calibration(O)
        if cal:
            line["cpu_calibration"] = cal
    else:
        g.close()
    line["parity"] = par
    return line


def ring3d_leg(M, n, arith, steps=128, check=True, path=0):
    """The cache-resident rotation case: n^3 periodic, solid-body rotation about (1,1,1)
    at max Courant 0.4, Gaussian blob sigma 8, iord 2 nonoscillatory; autotuned, 64
    warm-up and `steps` timed steps, the step's roofline, and the run against the
    oracle's (16 threads) after 1 and `steps` steps."""
    import oracle_lib as O

    case = rotation_case(n)
    cells = n ** 3
    g = M.Grid(n, n, n, M.PERIODIC, iord=2, nonosc=True)
    g.set_arith(arith)
    g.set_path(path)
    g.set_courant(case["U"], case["V"], case["W"])
    g.set_field(case["c"])
    g.autotune(2)
    tuned = tuned_config(g)
    g.step(64)
    ev_ms = g.step_timed(steps)
    kms = g.kernel_times(8)
    ms = ev_ms / steps
    tfile = os.path.join(ROOT, "profiles", f"traffic_rotation{n}_iord2.json")
    roof = roofline(cells, kms, step_kernels(M, 2, kms), tfile if arith == 1 else None, tuned)
    if not check:
        g.close()
        return {"ms_per_step": ms, "cell_updates_per_s": cells / (ms * 1e-3), "autotuned": tuned,
                "roofline": roof}
    # parity on the timed grid (its tuned launch shapes), from the initial field
    g.set_field(case["c"])
    g.step(1)
    got1 = g.field()
    g.step(steps - 1)
    gotk = g.field()
    mass, cmin, _ = g.stats()
    g.close()
    want1 = O.per_step(n, n, n, case["U"], case["V"], case["W"], 2, True, case["c"])
    wantk = O.per_run(n, n, n, case["U"], case["V"], case["W"], 2, True, case["c"], steps)
    return {"workload": f"ring3d: periodic {n}^3 solid-body rotation about (1,1,1), max "
                        "Courant 0.4, Gaussian blob sigma 8, iord=2 nonoscillatory "
                        "(fits the Infinity Cache)",
            "autotuned": tuned, "ms_per_step": ms, "cell_updates_per_s": cells / (ms * 1e-3),
            "roofline": roof, "rel_err_1_step_vs_oracle": relerr(got1, want1),
            f"rel_err_{steps}_steps_vs_oracle": relerr(gotk, wantk),
            "mass_rel_drift": abs(mass - case["c"].sum()) / case["c"].sum(), "min": cmin}


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
    # against a single grid of each size stepped alone (one kernel per pass too, in the
    # z-chunks of its own size: the same arithmetic per cell), and the smallest instance
    # against the oracle over the whole run
    worst, err = 0.0, None
    path = b.path()
    for i in firsts:
        (U, V, W), c = start[i]
        got = b.instances[i].field()
        g = M.Grid(*dims[i], M.PERIODIC, iord=2, nonosc=True)
        g.set_courant(U, V, W)
        g.set_field(c)
        g.step(taken)
        worst = max(worst, relerr(got, g.field()))
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
            "batched_path": "one kernel per corrective pass" if path == 3 else "two kernels per pass",
            f"rel_err_vs_single_grids_after_{taken}_steps":
                {"instances": firsts, "max": worst},
            f"rel_err_smallest_vs_oracle_after_{taken}_steps": err}


def large_leg(M, n, iord, arith, check=True, path=0):
    """Roofline of the single-grid step at n^3 (the size the 70%-of-HBM target names); one
    step against the oracle, and the whole run against the exact-arithmetic kernels."""
    import mpdata_amd.configs as K
    import oracle_lib as O

    g = M.Grid(n, n, n, M.PERIODIC, iord=iord, nonosc=True)
    g.set_arith(arith)
    g.set_path(path)
    K.ring3d(g, n)
    U, V, W = g.get_courant()
    c0 = g.field()
    g.autotune(2)
    g.step(1)
    err1 = relerr(g.field(), O.per_step(n, n, n, U, V, W, iord, True, c0)) if check else None
    del U, V, W
    g.step(40)  # warm-up after the CPU check (clocks)
    ms = g.step_timed(40) / 40
    kms = g.kernel_times(4)
    tuned = tuned_config(g)
    taken = 1 + 40 + 40 + 4
    if not check:
        g.close()
        return {"autotuned": tuned, "ms_per_step": ms, "cell_updates_per_s": n ** 3 / (ms * 1e-3),
                "kernel_ms": kms}
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
    tfile = os.path.join(ROOT, "profiles", f"traffic_rotation{n}_iord{iord}.json")
    counts = None
    if arith == 1 and os.path.exists(tfile):
        t = json.load(open(tfile))
        if t.get("config") == tuned and "valu_wave_instr_per_step" in t:
            counts = t
    return {"workload": f"periodic {n}^3 rotation, iord {iord} nonoscillatory",
            "autotuned": tuned, "ms_per_step": ms, "cell_updates_per_s": cells / (ms * 1e-3),
            "kernel_ms": kms, "achieved_GB/s": ach, "frac_of_hbm_peak": ach / HBM_PEAK_GBS,
            "valu": valu_roofline(counts, kern, ach / HBM_PEAK_GBS),
            "rel_err_1_step_vs_oracle": err1, f"rel_err_{taken}_steps_vs_exact": err_run}


def adjoint_leg(M, n, nsteps=512, every=16, K=32, nrec=64, checks=True):
    """Source inversion: n^3 periodic, deformational wind and 32 Gaussians (seed 5),
    emission of 8 seeded positive Gaussians added every step; 64 seeded receptors sampled
    every 16 steps against a twin run whose emission amplitudes are scaled by
    (1 + 0.2 xi); gradient of J = 1/2 sum (c - obs)^2 w.r.t. c^0 and the emission,
    checkpoints every 32 steps, iord 2 with the limiter.  The gradient run keeps the
    records of as many final segments as the device memory holds and recomputes the
    others; it is repeated with only the last segment kept (pure checkpointing, same
    gradient bit for bit).  The adjoint's cost is the gradient run's time beyond a plain
    exact forward of the same steps, in forward steps:
    (forward sweep + backward - plain forward) / plain forward.

    With `checks`: at full size, the dot-product test of one GPU adjoint step (a state of
    the run) against the oracle's tangent-linear model in both controls; central
    differences of J along 4 seeded c^0 directions and one emission direction at eps
    1e-6; and eps sweeps 1e-4 .. 1e-9 along one c^0 direction with the limiter off and
    on.  J is piecewise smooth in either case: the limiter's max/min selections, and even
    without them the donor fluxes' upwinding on the sign of the c-dependent antidiffusive
    velocity, put kinks in it, which a difference quotient straddles when eps is too
    large; the sweeps show where the quotient meets the adjoint.  Also the gradient
    against the CPU oracle's at 32^3 x 64 steps."""
    import mpdata_amd.configs as K_
    import oracle_lib as O

    rec, factors = K_.inversion_receptors(n ** 3, nrec=nrec)
    twin = K_.inversion_grid(n, factors=factors)
    obs = twin.sample_receptors(nsteps, rec, every)
    twin.close()
    g = K_.inversion_grid(n)
    c0 = g.field()
    q0 = g.emission()
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
    tp = g.step_timed(nsteps)
    print(f"bench: adjoint {(tf + tb - tp) / tp:.3f}x, checkpointing only "
          f"{(tf_c + tb_c - tp) / tp:.3f}x", file=sys.stderr, flush=True)
    out = {"workload": f"periodic {n}^3 deformational wind (seed 5), 8-source emission, "
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
           "grad_norms": {"c0": float(np.linalg.norm(g0)), "q": float(np.linalg.norm(gq))}}
    if checks:
        out.update(adjoint_checks(M, O, g, n, nsteps, every, K, rec, obs, c0, q0, g0, gq))
    g.close()

    # CPU reference gradient at 32^3 x 64 steps (same generators, limiter on)
    m, ms = 32, 64
    mrec, mfac = K_.inversion_receptors(m ** 3, nrec=16)
    gs = K_.inversion_grid(m)
    U, V, W = gs.get_courant()
    cs, qs = gs.field(), gs.emission()
    tw = K_.inversion_grid(m, factors=mfac)
    obs_s = tw.sample_receptors(ms, mrec, 16)
    tw.close()
    Jg, g0g, gqg, _ = gs.receptor_gradient(ms, 16, mrec, 16, obs_s)
    gs.close()
    Jc, g0c, gqc = O.receptor_loss_and_gradient(m, m, m, U, V, W, True, cs, qs, ms, mrec, 16,
                                                obs_s)
    out["cpu_reference"] = {
        "grid": [m, m, m], "steps": ms, "J_rel_diff": abs(Jg - Jc) / Jc,
        "grad_c0_rel_diff": float(np.abs(g0g - g0c).max() / np.abs(g0c).max()),
        "grad_q_rel_diff": float(np.abs(gqg - gqc).max() / np.abs(gqc).max())}
    return out


def adjoint_checks(M, O, g, n, nsteps, every, K, rec, obs, c0, q0, g0, gq):
    """Full-size checks of the source-inversion gradient (see adjoint_leg)."""
    res = {}
    # dot-product test of one adjoint step at a state of the run (8 steps in), both
    # controls: <TLM(dc, dq), lam> against <dc, ADJ lam> + <dq, lq>
    g.set_field(c0)
    g.step(8)
    c = g.field()
    U, V, W = g.get_courant()
    rng = np.random.default_rng(2024)
    lam = rng.standard_normal(c.size)
    lq = np.zeros(c.size)
    lin = g.adjoint_step(c, lam, lq)
    dc = rng.standard_normal(c.size) * c
    dq = rng.standard_normal(c.size) * q0
    t = O.tlm_step(n, n, n, U, V, W, True, c, dc, dq)
    lhs, rhs = float(t @ lam), float(dc @ lin) + float(dq @ lq)
    res["dot_product_test"] = {"state": "c^8 of the run", "lhs_tlm": lhs, "rhs_adjoint": rhs,
                               "rel_diff": abs(lhs - rhs) / abs(lhs)}
    del t, lin, lam, lq, U, V, W, c
    print(f"bench: adjoint dot-product test {abs(lhs - rhs) / abs(lhs):.2e}", file=sys.stderr,
          flush=True)

    def loss(grid, cc, qq):
        grid.set_field(cc)
        if qq is not None:
            grid.set_emission(qq)
        return grid.receptor_gradient(nsteps, K, rec, every, obs, grads=False)[0]

    def central(grid, d_c, d_q, eps, grad_c, grad_q):
        if d_c is not None:
            jp, jm = loss(grid, c0 + eps * d_c, None), loss(grid, c0 - eps * d_c, None)
            an = float(grad_c @ d_c)
        else:
            jp, jm = loss(grid, c0, q0 + eps * d_q), loss(grid, c0, q0 - eps * d_q)
            grid.set_emission(q0)
            an = float(grad_q @ d_q)
        fd = (jp - jm) / (2 * eps)
        return {"fd": fd, "adjoint": an, "rel_diff": abs(fd - an) / max(abs(an), 1e-300)}

    dirs = np.random.default_rng(5)
    # 4 seeded c^0 directions and one emission direction, scaled by the fields themselves
    # (relative perturbations: the fields span many orders of magnitude), eps 1e-6
    res["fd_eps_1e-6"] = {
        "c0_directions": [central(g, dirs.standard_normal(c0.size) * c0, None, 1e-6, g0, gq)
                          for _ in range(4)],
        "q_direction": central(g, None, dirs.standard_normal(c0.size) * q0, 1e-6, g0, gq)}
    d = dirs.standard_normal(c0.size) * c0
    epss = (1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9)
    res["eps_sweep_limiter_on"] = {f"{e:g}": central(g, d, None, e, g0, gq) for e in epss}
    # the same with the limiter off: its own gradient
    gu = M.Grid(n, n, n, M.PERIODIC, iord=2, nonosc=False)
    gu.set_arith(M.ARITH_EXACT)
    gu.set_courant(*g.get_courant())
    gu.set_emission(q0)
    gu.set_field(c0)
    _, g0u, gqu, _ = gu.receptor_gradient(nsteps, K, rec, every, obs)
    res["eps_sweep_limiter_off"] = {f"{e:g}": central(gu, d, None, e, g0u, gqu) for e in epss}
    gu.close()
    return res


if __name__ == "__main__":
    main()
