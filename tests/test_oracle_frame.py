"""Pins the oracle's FRAME-mode restatement of the reference step smolar()
(core/src/smolar.c:33-63) against the reference's own outputs.

* golden fixtures written by tests/golden/make_golden.py from the reference compiled
  -DDOUBLE (bitwise equality);
* the compiled reference itself on fresh random cases where oracle/_ref is built;
* closed forms the scheme fixes: zero wind leaves c unchanged, Courant 1 along x with
  V = W = 0 shifts the interior by exactly one cell per step at every order.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def run_oracle(d, order, nsteps, nx, ny, ns):
    c0 = d["c"].copy()
    c1 = np.zeros_like(c0)
    for _ in range(nsteps):
        rc, a0, a1 = O.call_smolar(O.oracle().orc_smolar, nx, ny, ns, d["dx"], d["dy"],
                                   d["ds"], float(d["dt"]), d["u0"], d["u1"], d["v0"],
                                   d["v1"], d["w0"], d["w1"], c0, c1, order)
        assert rc == 0
        c0, c1 = a1, a0
    return c0


@pytest.mark.parametrize("fixture,orders", [("frame_3d.npz", (1, 2, 3, 4)),
                                            ("frame_2d.npz", (1, 2, 3))])
def test_frame_matches_reference_fixtures(fixture, orders):
    d = dict(np.load(os.path.join(GOLD, fixture)))
    nx, ny, ns = (int(x) for x in d["dims"])
    for order in orders:
        for s in (1, 12):
            got = run_oracle(d, order, s, nx, ny, ns)
            np.testing.assert_array_equal(got, d[f"out_o{order}_s{s}"],
                                          err_msg=f"{fixture} order {order} step {s}")


def test_ring2d_trajectory_matches_reference():
    d = dict(np.load(os.path.join(GOLD, "ring2d_32.npz")))
    n, _, m = (int(x) for x in d["dims"])
    one = np.ones(n)
    c0 = d["c"].copy()
    c1 = np.zeros_like(c0)
    done = 0
    for target in (1, 10, 50):
        while done < target:
            rc, a0, a1 = O.call_smolar(O.oracle().orc_smolar, n, n, m, one, one, np.ones(m),
                                       float(d["dt"]), d["u"], d["u"], d["v"], d["v"],
                                       d["w"], d["w"], c0, c1, 3)
            c0, c1 = a1, a0
            done += 1
        np.testing.assert_array_equal(c0, d[f"out_s{target}"])


def test_pointer_permutation_contract():
    """smolar.c:58: *c1 holds the result; for even order the two buffers swap."""
    nx = ny = ns = 5
    n = nx * ny * ns
    z = np.zeros(n)
    one = np.ones(nx)
    for order in (1, 2, 3):
        c0 = np.random.default_rng(order).uniform(0, 1, n)
        c1 = np.zeros(n)
        rc, a0, a1 = O.call_smolar(O.oracle().orc_smolar, nx, ny, ns, one, one, one, 1.0,
                                   z, z, z, z, z, z, c0, c1, order)
        assert (a1 is c1) == (order % 2 == 1)


@pytest.mark.skipif(not O.have_ref(), reason="reference not compiled here (oracle/_ref)")
@pytest.mark.parametrize("dims", [(13, 11, 9), (3, 3, 3), (4, 5, 6), (9, 3, 4), (3, 8, 5)])
def test_frame_matches_compiled_reference(dims):
    nx, ny, ns = dims
    n = nx * ny * ns
    rng = np.random.default_rng(nx * 100 + ny * 10 + ns)
    dx, dy, ds = rng.uniform(0.8, 1.2, nx), rng.uniform(0.8, 1.2, ny), rng.uniform(0.8, 1.2, ns)
    winds = [rng.uniform(-0.35, 0.35, n) for _ in range(6)]
    c = rng.uniform(0, 1, n)
    for order in (1, 2, 3, 4, 5):
        outs = []
        for fn in (O.ref().smolar, O.oracle().orc_smolar):
            c0, c1 = c.copy(), np.full(n, 3.0)
            rc, a0, a1 = O.call_smolar(fn, nx, ny, ns, dx, dy, ds, 0.7, *winds, c0, c1, order)
            outs.append((a1.copy(), a0.copy()))
        np.testing.assert_array_equal(outs[0][0], outs[1][0])
        np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_zero_wind_is_identity():
    nx, ny, ns = 9, 8, 7
    n = nx * ny * ns
    z = np.zeros(n)
    c = np.random.default_rng(5).uniform(0, 1, n)
    for order in (1, 2, 3):
        c0, c1 = c.copy(), np.zeros(n)
        rc, a0, a1 = O.call_smolar(O.oracle().orc_smolar, nx, ny, ns, np.ones(nx),
                                   np.ones(ny), np.ones(ns), 1.0, z, z, z, z, z, z, c0, c1,
                                   order)
        np.testing.assert_array_equal(a1, c)


def test_unit_courant_shift_is_exact():
    """U = 1 everywhere, V = W = 0: every order moves the interior one cell in +x.  Values
    lie in [1, 1.9] so that c - (c - c_left) is exact (Sterbenz)."""
    nx, ny, ns = 12, 6, 5
    n = nx * ny * ns
    z = np.zeros(n)
    u = np.ones(n)  # Courant = 0.25*dt*(4u)/dx = 1 with dt=1, dx=1
    c = np.random.default_rng(6).uniform(1.0, 1.9, n)
    for order in (1, 2, 3):
        c0, c1 = c.copy(), np.zeros(n)
        rc, a0, a1 = O.call_smolar(O.oracle().orc_smolar, nx, ny, ns, np.ones(nx),
                                   np.ones(ny), np.ones(ns), 1.0, u, u, z, z, z, z, c0, c1,
                                   order)
        got = a1.reshape(ns, ny, nx)
        want = c.reshape(ns, ny, nx)
        np.testing.assert_array_equal(got[1:-1, 1:-1, 1:-1], want[1:-1, 1:-1, :-2])


def test_planted_mistake_is_caught():
    """The fixtures have teeth: a copy of the fixture output with one ulp changed in one
    cell, or the 2-D output compared against order 2 instead of 3, must not match."""
    d = dict(np.load(os.path.join(GOLD, "frame_3d.npz")))
    nx, ny, ns = (int(x) for x in d["dims"])
    got = run_oracle(d, 2, 1, nx, ny, ns)
    assert not np.array_equal(got, d["out_o3_s1"])
    bad = d["out_o2_s1"].copy()
    k = np.argmax(bad)
    bad[k] = np.nextafter(bad[k], 2.0)
    assert not np.array_equal(got, bad)
