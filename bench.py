"""MPDATA step benchmark on one MI355X.

Workload (default): ring3d plain MPDATA -- 128^3 triply periodic FP64 grid, solid-body
rotation about the axis (1,1,1)/sqrt(3) through the centre scaled to a maximum face
Courant number of 0.4, a Gaussian blob (sigma 8 cells) 30 cells off the axis, iord=2
with the nonoscillatory limiter.  A step is one MPDATA time step over the whole grid;
the fields are resident in HBM before timing starts.

Prints ONE JSON line: cell-updates/s over the timed steps (hipEvents on the launch
stream, host clock around synchronizations as a cross-check), the roofline of the
dominant kernel (algorithmic bytes: 40 B per cell-update = read c^n and the three face
Courant numbers, write c^{n+1}), the parity of the measured path against the CPU oracle,
and the oracle timed on a bounded sample of the same workload on the host cores.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-steps S] [--no-cpu]
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


def face_courant_from_cell_winds(u, v, w, n, dt=1.0):
    """Face Courant numbers of a periodic grid with unit spacing from steady cell-centred
    winds (u0 == u1): 0.25*dt*(2u_i + 2u_{i+1}), the reference formula (uu.c:132-134)."""
    U = 0.25 * dt * ((u + np.roll(u, -1, axis=2)) * 2.0)
    V = 0.25 * dt * ((v + np.roll(v, -1, axis=1)) * 2.0)
    W = 0.25 * dt * ((w + np.roll(w, -1, axis=0)) * 2.0)
    return U, V, W


def rotation_case(n=128, cmax=0.4, sigma=8.0, offset=30.0):
    """Solid-body rotation about (1,1,1)/sqrt(3) through the grid centre, scaled to a
    maximum face Courant number cmax; a Gaussian blob of width sigma centred `offset`
    cells from the axis along (1,-1,0)/sqrt(2).  Arrays are (z, y, x) raveled."""
    ax = np.arange(n, dtype=np.float64) - n / 2
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    k = np.ones(3) / np.sqrt(3.0)
    # omega * (k x r)
    u = k[1] * Z - k[2] * Y
    v = k[2] * X - k[0] * Z
    w = k[0] * Y - k[1] * X
    U, V, W = face_courant_from_cell_winds(u, v, w, n)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=128)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--iord", type=int, default=2)
    ap.add_argument("--path", type=int, default=0, help="0 auto, 1 staged, 2 fused")
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: every box has one GPU and "
                          "the path shards by slab decomposition only in design "
                          "(see DESIGN.md); run with --gpus 1"}))
        sys.exit(2)

    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import mpdata_amd as M

    n = args.n
    case = rotation_case(n)
    cells = n ** 3
    g = M.Grid(n, n, n, M.PERIODIC, iord=args.iord, nonosc=True)
    g.set_path(args.path)
    g.set_courant(case["U"], case["V"], case["W"])
    g.set_field(case["c"])
    g.step(args.warmup)
    g.synchronize()
    t0 = time.perf_counter()
    ev_ms = g.step_timed(args.steps)
    wall = time.perf_counter() - t0
    ms_per_step = ev_ms / args.steps
    value = cells * args.steps / (ev_ms * 1e-3)
    mass, cmin, cmax_ = g.stats()
    g.close()

    achieved = cells * BYTES_PER_CELL / (ms_per_step * 1e-3) / 1e9
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
        "data": "synthetic (seedless analytic rotation wind and Gaussian blob)",
        "config": {"workload": f"ring3d periodic {n}^3 rotation, iord={args.iord} "
                               "nonoscillatory", "grid": [n, n, n], "iord": args.iord,
                   "nonoscillatory": True, "boundary": "periodic", "max_courant": 0.4,
                   "path": ["auto", "staged", "fused"][args.path]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "wall_s_timed": wall,
        "field_after": {"mass": mass, "min": cmin, "max": cmax_},
    }

    # parity of the measured path with the oracle on the same inputs
    import oracle_lib as O

    kpar = 2
    gp = M.Grid(n, n, n, M.PERIODIC, iord=args.iord, nonosc=True)
    gp.set_path(args.path)
    gp.set_courant(case["U"], case["V"], case["W"])
    gp.set_field(case["c"])
    gp.step(kpar)
    got = gp.field()
    gp.close()
    t0 = time.perf_counter()
    want = O.per_run(n, n, n, case["U"], case["V"], case["W"], args.iord, True, case["c"],
                     kpar)
    t_par = time.perf_counter() - t0
    line["parity"] = {"steps": kpar,
                      "rel_err": float(np.abs(got - want).max() / np.abs(want).max())}

    if not args.no_cpu:
        ks = max(1, args.cpu_steps)
        t0 = time.perf_counter()
        O.per_run(n, n, n, case["U"], case["V"], case["W"], args.iord, True, case["c"], ks)
        t_cpu = time.perf_counter() - t0
        line["cpu_baseline"] = {
            "value": cells * ks / t_cpu, "unit": "cell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle (FP64 C restatement, 1 thread, {host_cpu_model()}) on {ks} "
                      f"steps of the same {n}^3 iord={args.iord} nonoscillatory workload",
        }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
