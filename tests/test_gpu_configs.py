"""The benchmark configurations at their full sizes.

* ring2d (periodic 128 x 128, rotation scaled to max Courant 0.5, cone of radius 15 at
  (64, 96), iord 2 nonoscillatory, 1024 steps): the GPU trajectory against the oracle's,
  step by step -- bitwise with exact arithmetic, <= 1e-12 normwise with the fast kernels;
  mass to rounding and positivity at every step.  Positivity holds to rounding of the
  field's scale: in cells of size 1e-88 the limited update c1 - (fluxes) cancels to a
  few ulps of those cells and can leave values like -1e-88, in the oracle exactly as on
  the GPU.
* deform (512 x 512 x 256, seeded deformational wind and 32 Gaussians, iord 3
  nonoscillatory): one step against the oracle at full size (<= 1e-12 normwise), then
  mass conservation and positivity over further steps; in exact arithmetic, three steps
  bit for bit against the oracle's.
"""
import numpy as np
import pytest

import mpdata_amd as M
import mpdata_amd.configs as K
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if M.device_count() < 1:
        pytest.skip("no GPU")


def relerr(got, want):
    return float(np.abs(got - want).max() / np.abs(want).max())


@pytest.mark.parametrize("arith", [M.ARITH_EXACT, M.ARITH_FAST])
def test_ring2d_periodic_step_by_step(arith):
    n, nsteps = 128, 1024
    U, V, W, c = K.ring2d(n)
    assert abs(max(np.abs(U).max(), np.abs(V).max()) - 0.5) < 1e-15
    g = M.Grid(n, n, 1, M.PERIODIC, iord=2, nonosc=True)
    g.set_arith(arith)
    g.set_courant(U, V, W)
    g.set_field(c)
    mass0 = c.sum()
    want = c
    worst = 0.0
    for s in range(nsteps):
        g.step(1)
        got = g.field()
        want = O.per_step(n, n, 1, U, V, W, 2, True, want)
        if arith == M.ARITH_EXACT:
            if not np.array_equal(got, want):
                raise AssertionError(f"step {s + 1}: exact trajectory differs "
                                     f"(rel {relerr(got, want):.3e})")
        else:
            worst = max(worst, relerr(got, want))
            assert worst <= 1e-12, f"step {s + 1}: rel err {worst:.3e}"
        assert got.min() >= -1e-15 * got.max()
        assert abs(got.sum() - mass0) <= 1e-12 * mass0
    print(f"ring2d arith={arith}: worst normwise rel err over {nsteps} steps {worst:.3e}",
          flush=True)
    g.close()


def test_deform_512x512x256_iord3():
    shape = (512, 512, 256)
    nx, ny, nz = shape
    g = M.Grid(nx, ny, nz, M.PERIODIC, iord=3, nonosc=True)
    K.deform(g, shape, seed=3)
    U, V, W = g.get_courant()
    assert abs(max(np.abs(U).max(), np.abs(V).max(), np.abs(W).max()) - 0.5) < 1e-12
    c0 = g.field()
    assert c0.min() >= 0.0
    mass0 = c0.sum()
    # the donor pass keeps c >= 0 where the outgoing Courant numbers of every cell sum to
    # at most 1; the limited corrections then keep it there
    u3, v3, w3 = (a.reshape(nz, ny, nx) for a in (U, V, W))
    out = (np.maximum(u3, 0) - np.minimum(np.roll(u3, 1, axis=2), 0)
           + np.maximum(v3, 0) - np.minimum(np.roll(v3, 1, axis=1), 0)
           + np.maximum(w3, 0) - np.minimum(np.roll(w3, 1, axis=0), 0))
    positive = bool(out.max() <= 1.0)
    del out, u3, v3, w3
    print(f"case deform 1 step vs oracle (positivity guaranteed: {positive})", flush=True)
    g.step(1)
    got = g.field()
    want = O.per_step(nx, ny, nz, U, V, W, 3, True, c0)
    del U, V, W
    assert relerr(got, want) <= 1e-12
    del want
    print("case deform 20 steps: mass and positivity", flush=True)
    g.step(19)
    mass, cmin, cmax = g.stats()
    g.close()
    assert abs(mass - mass0) <= 1e-12 * mass0
    if positive:
        assert cmin >= -1e-15 * cmax


def test_deform_512x512x256_iord3_exact_step_by_step():
    """The exact-arithmetic kernels on the full deform case follow the oracle bit for bit,
    step by step (the fast kernels' whole 203-step run is stated against these by
    bench.py)."""
    shape = (512, 512, 256)
    nx, ny, nz = shape
    g = M.Grid(nx, ny, nz, M.PERIODIC, iord=3, nonosc=True)
    K.deform(g, shape, seed=3)
    g.set_arith(M.ARITH_EXACT)
    U, V, W = g.get_courant()
    want = g.field()
    for s in range(3):
        print(f"case deform exact step {s + 1}", flush=True)
        g.step(1)
        want = O.per_step(nx, ny, nz, U, V, W, 3, True, want)
        got = g.field()
        if not np.array_equal(got, want):
            raise AssertionError(f"step {s + 1}: exact trajectory differs "
                                 f"(rel {relerr(got, want):.3e})")
        del got
    g.close()
