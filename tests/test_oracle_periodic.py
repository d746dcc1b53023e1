"""Pins the oracle's periodic step.

* Unlimited step: the reference smolar() run on a grid padded by H=8 wrapped cells
  (tests/golden/periodic.npz; H=10 gave the identical centre) -- bitwise.
* Nonoscillatory limiter (not in the reference; Smolarkiewicz & Grabowski 1990): no new
  extrema (every cell ends inside the max/min of c^n and the first-pass field over its
  7-point neighbourhood), positivity, mass conservation, and teeth -- the unlimited
  scheme does break the bound on the same step profile.
* Emission: zero wind gives c + q exactly; the step is affine in q.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("tag", ["a", "b"])
def test_periodic_matches_padded_reference(tag):
    d = dict(np.load(os.path.join(GOLD, "periodic.npz")))
    nx, ny, nz = (int(x) for x in d[f"{tag}_dims"])
    U, V, W = O.per_courant(nx, ny, nz, d[f"{tag}_dx"], d[f"{tag}_dy"], d[f"{tag}_dz"],
                            float(d[f"{tag}_dt"]), d[f"{tag}_u"], d[f"{tag}_u"],
                            d[f"{tag}_v"], d[f"{tag}_v"], d[f"{tag}_w"], d[f"{tag}_w"])
    for order in (1, 2, 3):
        got = O.per_run(nx, ny, nz, U, V, W, order, False, d[f"{tag}_c"], 3)
        np.testing.assert_array_equal(got, d[f"{tag}_out_o{order}"], err_msg=f"order {order}")


def rotation_case(n=24, nz=1, cmax=0.5, seed=0):
    """Solid-body rotation in x-y with a sharp-edged slotted block: a demanding profile
    for monotonicity."""
    x = np.arange(n) - n / 2 + 0.5
    X, Y = np.meshgrid(x, x, indexing="xy")
    om = cmax / (n / 2)
    u = np.broadcast_to(-om * Y, (nz, n, n)).ravel().copy()
    v = np.broadcast_to(om * X, (nz, n, n)).ravel().copy()
    w = np.zeros(n * n * nz)
    c2 = ((np.abs(X - 4) < 5) & (np.abs(Y) < 5)).astype(float)
    c = np.broadcast_to(c2, (nz, n, n)).ravel().copy()
    one = np.ones(n)
    U, V, W = O.per_courant(n, n, nz, one, one, np.ones(nz), 1.0, u, u, v, v, w, w)
    return n, nz, U, V, W, c


def local_bounds(nx, ny, nz, *fields):
    hi = -np.inf
    lo = np.inf
    for f in fields:
        a = f.reshape(nz, ny, nx)
        for ax in range(3):
            for s in (-1, 0, 1):
                r = np.roll(a, s, axis=ax)
                hi = np.maximum(hi, r)
                lo = np.minimum(lo, r)
    return hi.ravel(), lo.ravel()


@pytest.mark.parametrize("iord", [2, 3])
def test_limiter_creates_no_new_extrema(iord):
    """Each limited corrective pass keeps every cell inside the max/min of c^n and the
    field entering that pass over the cell's 7-point neighbourhood."""
    n, nz, U, V, W, c = rotation_case()
    cur = c
    tol = 1e-14
    for _ in range(10):
        first = np.empty_like(cur)
        O.oracle().orc_per_donor(n, n, nz, O.ptr(U), O.ptr(V), O.ptr(W), O.ptr(cur),
                                 O.ptr(first))
        prev = first
        for k in range(2, iord + 1):
            nxt = O.per_step(n, n, nz, U, V, W, k, True, cur)
            hi, lo = local_bounds(n, n, nz, cur, prev)
            assert np.all(nxt <= hi + tol) and np.all(nxt >= lo - tol), f"pass {k}"
            prev = nxt
        cur = prev
    # teeth: the unlimited scheme overshoots on the same profile
    first = np.empty_like(c)
    O.oracle().orc_per_donor(n, n, nz, O.ptr(U), O.ptr(V), O.ptr(W), O.ptr(c), O.ptr(first))
    hi, lo = local_bounds(n, n, nz, c, first)
    raw = O.per_step(n, n, nz, U, V, W, 2, False, c)
    assert np.any(raw > hi + 1e-6) or np.any(raw < lo - 1e-6)


@pytest.mark.parametrize("nonosc", [False, True])
def test_mass_conservation_and_positivity(nonosc):
    rng = np.random.default_rng(3)
    nx, ny, nz = 12, 10, 8
    n = nx * ny * nz
    U, V, W = (rng.uniform(-0.15, 0.15, n) for _ in range(3))
    c = rng.uniform(0, 1, n) ** 4
    m0 = c.sum()
    out = O.per_run(nx, ny, nz, U, V, W, 3, nonosc, c, 20)
    assert abs(out.sum() - m0) <= 1e-13 * m0
    assert out.min() >= 0.0


def test_constant_field_and_unit_shift_with_limiter():
    nx, ny, nz = 10, 6, 4
    n = nx * ny * nz
    rng = np.random.default_rng(9)
    U, V, W = (rng.uniform(-0.2, 0.2, n) for _ in range(3))
    # a constant stays constant under a uniform (divergence-free) velocity
    U[:] = 0.3
    V[:] = -0.2
    W[:] = 0.1
    c = np.full(n, 0.7)
    for nonosc in (False, True):
        out = O.per_run(nx, ny, nz, U, V, W, 3, nonosc, c, 5)
        np.testing.assert_allclose(out, 0.7, rtol=0, atol=1e-15)
    U[:] = 1.0
    V[:] = 0.0
    W[:] = 0.0
    c = rng.uniform(1.0, 1.9, n)  # within a factor 2: the shift is exact (Sterbenz)
    for nonosc in (False, True):
        out = O.per_step(nx, ny, nz, U, V, W, 2, nonosc, c)
        np.testing.assert_array_equal(out.reshape(nz, ny, nx),
                                      np.roll(c.reshape(nz, ny, nx), 1, axis=2))


def test_emission_closed_form_and_affinity():
    rng = np.random.default_rng(4)
    nx, ny, nz = 8, 7, 6
    n = nx * ny * nz
    z = np.zeros(n)
    c = rng.uniform(0, 1, n)
    q = rng.uniform(0, 0.01, n)
    out = O.per_step(nx, ny, nz, z, z, z, 2, True, c, q)
    np.testing.assert_array_equal(out, c + q)
    U, V, W = (rng.uniform(-0.2, 0.2, n) for _ in range(3))
    a = O.per_step(nx, ny, nz, U, V, W, 2, True, c, q)
    b = O.per_step(nx, ny, nz, U, V, W, 2, True, c)
    np.testing.assert_array_equal(a, b + q)


def test_thread_count_does_not_change_results():
    """The oracle's periodic loops run OpenMP over z planes; every cell is the same
    expression whatever the thread count, so 1, 3 and 4 threads agree bit for bit."""
    rng = np.random.default_rng(6)
    nx, ny, nz = 12, 9, 10
    n = nx * ny * nz
    U, V, W = (rng.uniform(-0.2, 0.2, n) for _ in range(3))
    c = rng.uniform(0, 1, n) ** 2
    q = rng.uniform(0, 0.01, n)
    saved = O.get_threads()
    try:
        out = []
        for t in (1, 3, 4):
            O.set_threads(t)
            assert O.get_threads() == t
            out.append(O.per_run(nx, ny, nz, U, V, W, 3, True, c, 4, q))
    finally:
        O.set_threads(saved)
    for o in out[1:]:
        np.testing.assert_array_equal(o, out[0])
