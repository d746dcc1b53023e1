"""Writes the golden fixtures under tests/golden/ from the reference's own step.

Runs in a container where the reference tree is mounted and oracle/_ref/libsmolar_ref.so
has been built from it (`make -C oracle ref`: the reference's core/ and generators/ C
sources compiled -DDOUBLE -ffp-contract=off, malloc zero-filled).  Every fixture holds its
inputs and the reference's outputs, so the tests can check the oracle (and, through it, the
GPU path) without the reference present.

Fixtures (numpy .npz, no pickled objects):
  frame_3d.npz      13x11x9 frame-mode grid, non-uniform spacing, u0 != u1, orders 1..4,
                    1 step and 12 steps (pins smolar.c, smolar_pass.c, uu/vv/ww.c incl.
                    the vv.c:81 skipped column through order 4's ping-pong).
  frame_2d.npz      17x15x3 (2-D embedding, w = 0), orders 1..3, 12 steps.
  periodic.npz      reference smolar() on a grid padded by H=8 periodic cells, central
                    region after 3 steps: 16^3 and 24x20x1, orders 1..3.
  ring2d_32.npz     the ring2d driver's configuration at n=32 (reference generators
                    genringf2d/gengauss2d, order 3, dt 0.0125) and c after 1, 10, 50 steps.
  ring2d_128_1024.npz
                    the reference's own run `ring2d 128 1024`: sha256 of the ring2d.bin
                    and of the standard output it writes, and the final c plane.

Usage: python tests/golden/make_golden.py [ring2d_128]   (ring2d_128: that fixture only)
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402


def ref_steps(nx, ny, ns, dx, dy, ds, dt, u0, u1, v0, v1, w0, w1, c, order, nsteps):
    """nsteps reference smolar() calls, swapping c0/c1 as the ring drivers do."""
    lib = O.ref()
    c0 = c.copy()
    c1 = np.zeros_like(c)
    for _ in range(nsteps):
        rc, a0, a1 = O.call_smolar(lib.smolar, nx, ny, ns, dx, dy, ds, dt, u0, u1, v0, v1,
                                   w0, w1, c0, c1, order)
        assert rc == 0
        c0, c1 = a1, a0  # ring2d/test.c:62-63: the result becomes the next input
    return c0.copy()


def frame_case(rng, nx, ny, ns, orders, steps, two_d):
    n = nx * ny * ns
    d = {"dims": np.array([nx, ny, ns]), "dt": np.array(0.9)}
    d["dx"] = rng.uniform(0.8, 1.25, nx)
    d["dy"] = rng.uniform(0.8, 1.25, ny)
    d["ds"] = rng.uniform(0.8, 1.25, ns)
    for name in ("u0", "u1", "v0", "v1"):
        d[name] = rng.uniform(-0.3, 0.3, n)
    for name in ("w0", "w1"):
        d[name] = np.zeros(n) if two_d else rng.uniform(-0.3, 0.3, n)
    c = rng.uniform(0.0, 1.0, n)
    if two_d:
        c = np.tile(c[: nx * ny], ns)
    d["c"] = c
    for order in orders:
        for s in steps:
            d[f"out_o{order}_s{s}"] = ref_steps(nx, ny, ns, d["dx"], d["dy"], d["ds"],
                                                0.9, d["u0"], d["u1"], d["v0"], d["v1"],
                                                d["w0"], d["w1"], c, order, s)
    return d


def padded_periodic(nx, ny, nz, dx, dy, dz, dt, u, v, w, c, order, H, nsteps):
    """Periodic step from the reference: pad by H wrapped cells, step, keep the centre."""
    def pad3(a):
        return np.ascontiguousarray(np.pad(a.reshape(nz, ny, nx), H, mode="wrap")).ravel()
    dxp, dyp, dzp = (np.ascontiguousarray(np.pad(a, H, mode="wrap")) for a in (dx, dy, dz))
    up, vp, wp = pad3(u), pad3(v), pad3(w)
    for _ in range(nsteps):
        out = ref_steps(nx + 2 * H, ny + 2 * H, nz + 2 * H, dxp, dyp, dzp, dt,
                        up, up, vp, vp, wp, wp, pad3(c), order, 1)
        c = out.reshape(nz + 2 * H, ny + 2 * H, nx + 2 * H)[H:H + nz, H:H + ny, H:H + nx]
        c = np.ascontiguousarray(c).ravel()
    return c


def periodic_cases(rng):
    d = {}
    for tag, (nx, ny, nz) in (("a", (16, 16, 16)), ("b", (24, 20, 1))):
        n = nx * ny * nz
        d[f"{tag}_dims"] = np.array([nx, ny, nz])
        d[f"{tag}_dx"] = rng.uniform(0.9, 1.1, nx)
        d[f"{tag}_dy"] = rng.uniform(0.9, 1.1, ny)
        d[f"{tag}_dz"] = rng.uniform(0.9, 1.1, nz)
        d[f"{tag}_u"] = rng.uniform(-0.3, 0.3, n)
        d[f"{tag}_v"] = rng.uniform(-0.3, 0.3, n)
        d[f"{tag}_w"] = rng.uniform(-0.3, 0.3, n) if nz > 1 else np.zeros(n)
        d[f"{tag}_c"] = rng.uniform(0.0, 1.0, n)
        d[f"{tag}_dt"] = np.array(0.8)
        for order in (1, 2, 3):
            outs = [padded_periodic(nx, ny, nz, d[f"{tag}_dx"], d[f"{tag}_dy"],
                                    d[f"{tag}_dz"], 0.8, d[f"{tag}_u"], d[f"{tag}_v"],
                                    d[f"{tag}_w"], d[f"{tag}_c"], order, H, 3)
                    for H in (8, 10)]
            assert np.array_equal(outs[0], outs[1]), "padded result depends on H"
            d[f"{tag}_out_o{order}"] = outs[0]
    return d


def ring2d_setup(n):
    """tests/init.c:15-103 for m=3: genringf2d wind, gengauss2d cone, replicated layers.
    Returns (dx, dy, ds, u, v, w, c, xof, yof)."""
    lib = O.ref()
    PD = O.PD
    lib.genringf2d.argtypes = [C.c_int] * 4 + [PD, PD, C.c_double, C.c_double, PD, PD,
                                               C.c_double, C.c_int]
    lib.gengauss2d.argtypes = [C.c_int] * 4 + [C.c_double] * 3 + [PD, C.c_double]
    m = 3
    n2 = n * n
    dx = np.ones(n)
    dy = np.ones(n)
    ds = np.ones(m)
    u = np.zeros(n2 * m)
    v = np.zeros(n2 * m)
    w = np.zeros(n2 * m)
    rmin, rmax = 1.0, float(np.float32(n / 2.5))
    lib.genringf2d(n // 2, n // 2, n, n, O.ptr(dx), O.ptr(dy), rmin, rmax, O.ptr(u),
                   O.ptr(v), 1.0, 0)
    for i in range(1, m):
        u[i * n2:(i + 1) * n2] = u[:n2]
        v[i * n2:(i + 1) * n2] = v[:n2]
    c = np.zeros(n2 * m)
    xof, yof = n // 4, 0
    radius = float(n // 8)
    lib.gengauss2d(n // 2 + xof, n // 2 + yof, n, n, radius, 6.0, 6.0, O.ptr(c), 1.0)
    for i in range(1, m):
        c[i * n2:(i + 1) * n2] = c[:n2]
    return dx, dy, ds, u, v, w, c, xof, yof


def ring2d_case(n=32, steps=(1, 10, 50)):
    m = 3
    dx, dy, ds, u, v, w, c, _, _ = ring2d_setup(n)
    d = {"dims": np.array([n, n, m]), "dt": np.array(0.0125), "u": u, "v": v, "w": w,
         "c": c, "order": np.array(3)}
    for s in steps:
        d[f"out_s{s}"] = ref_steps(n, n, m, dx, dy, ds, 0.0125, u, u, v, v, w, w, c, 3, s)
    return d


def ring2d_run(n=128, nt=1024):
    """The reference's own headline run, `ring2d 128 1024` (tests/ring2d/test.c:10-71),
    with the reference's generators and smolar(): the sha256 of the ring2d.bin it writes
    (per step the u, v and c planes of layer 1, after the initial state), the sha256 of
    its standard output ("step 0", then "step <it> <probe cell, %f>"), and the final c
    plane of layer 1."""
    import hashlib
    m = 3
    n2 = n * n
    dx, dy, ds, u, v, w, c, xof, yof = ring2d_setup(n)
    lib = O.ref()
    binh, outh = hashlib.sha256(), hashlib.sha256()
    u1, v1 = u[n2:2 * n2].tobytes(), v[n2:2 * n2].tobytes()
    binh.update(u1 + v1 + c[n2:2 * n2].tobytes())
    outh.update(b"step 0\n")
    c0, c1 = c.copy(), np.zeros_like(c)
    probe = n // 2 + xof + n * (n // 2 + yof) + n2
    for it in range(1, nt + 1):
        rc, a0, a1 = O.call_smolar(lib.smolar, n, n, m, dx, dy, ds, 0.0125, u, u, v, v, w, w,
                                   c0, c1, 3)
        assert rc == 0
        binh.update(u1 + v1 + a1[n2:2 * n2].tobytes())
        outh.update(("step %d %f\n" % (it, a1[probe])).encode())
        c0, c1 = a1, a0  # ring2d/test.c:62-63
    return {"dims": np.array([n, n, m]), "nt": np.array(nt),
            "bin_sha256": np.array(binh.hexdigest()),
            "stdout_sha256": np.array(outh.hexdigest()),
            "c_final_layer1": c0[n2:2 * n2].copy()}


def main():
    if not O.have_ref():
        sys.exit("oracle/_ref/libsmolar_ref.so not built (make -C oracle ref)")
    if sys.argv[1:] == ["ring2d_128"]:  # this fixture alone
        np.savez_compressed(os.path.join(HERE, "ring2d_128_1024.npz"), **ring2d_run())
        return
    rng = np.random.default_rng(20101)
    np.savez_compressed(os.path.join(HERE, "frame_3d.npz"),
                        **frame_case(rng, 13, 11, 9, (1, 2, 3, 4), (1, 12), False))
    np.savez_compressed(os.path.join(HERE, "frame_2d.npz"),
                        **frame_case(rng, 17, 15, 3, (1, 2, 3), (1, 12), True))
    np.savez_compressed(os.path.join(HERE, "periodic.npz"), **periodic_cases(rng))
    np.savez_compressed(os.path.join(HERE, "ring2d_32.npz"), **ring2d_case())
    np.savez_compressed(os.path.join(HERE, "ring2d_128_1024.npz"), **ring2d_run())
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
