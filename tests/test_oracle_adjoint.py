"""The oracle's tangent-linear model and adjoint of one periodic MPDATA step (iord 2,
with and without the nonoscillatory limiter; oracle/cpu_adj.c), pinned by what the
mathematics fixes rather than by restating the formulas:

* dot-product test <TLM dc, lam> = <dc, ADJ lam> on random fields, on fields with exact
  zeros (where the limiter's bounds tie and the ratio guards dominate) and on degenerate
  extents (nx = 1, 2);
* central finite differences of the forward step (oracle/cpu_ref.c) along random
  directions: a dropped or sign-flipped term of the linearisation fails them;
* Euler's theorem: the step is positively homogeneous of degree 1 in c (the
  antidiffusive ratio (p - m)/(p + m) and the limiter ratios are scale invariant), so the
  TLM applied to dc = c reproduces the step itself;
* emission: c_out = A(c) + q, so d c_out / dq is the identity;
* a multi-step receptor loss: the full-sweep gradient matches finite differences.
"""
import numpy as np
import pytest

import oracle_lib as O


def case(nx, ny, nz, seed, cmax=0.22, zeros=False):
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    U, V, W = (rng.uniform(-cmax, cmax, n) for _ in range(3))
    if nz == 1:
        W[:] = 0.0
    c = 0.05 + rng.uniform(0, 1, n) ** 2
    if zeros:
        c[rng.uniform(0, 1, n) < 0.4] = 0.0
    return rng, U, V, W, c


DIMS = [(12, 10, 8), (16, 9, 1), (7, 6, 5), (2, 6, 4), (1, 8, 6)]


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("nonosc", [False, True])
@pytest.mark.parametrize("zeros", [False, True])
def test_dot_product(dims, nonosc, zeros):
    nx, ny, nz = dims
    rng, U, V, W, c = case(nx, ny, nz, nx * 100 + ny * 10 + nz, zeros=zeros)
    for _ in range(4):
        dc = rng.standard_normal(c.size)
        lam = rng.standard_normal(c.size)
        t = O.tlm_step(nx, ny, nz, U, V, W, nonosc, c, dc)
        a = O.adj_step(nx, ny, nz, U, V, W, nonosc, c, lam)
        lhs, rhs = float(t @ lam), float(dc @ a)
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0), (lhs, rhs)


@pytest.mark.parametrize("nonosc", [False, True])
def test_tlm_matches_central_differences(nonosc):
    nx, ny, nz = 10, 9, 8
    rng, U, V, W, c = case(nx, ny, nz, 7)
    eps = 1e-6
    for _ in range(4):
        dc = rng.standard_normal(c.size)
        t = O.tlm_step(nx, ny, nz, U, V, W, nonosc, c, dc)
        fp = O.per_step(nx, ny, nz, U, V, W, 2, nonosc, c + eps * dc)
        fm = O.per_step(nx, ny, nz, U, V, W, 2, nonosc, c - eps * dc)
        fd = (fp - fm) / (2 * eps)
        # random smooth backgrounds sit away from the limiter's kinks for this eps
        assert np.abs(fd - t).max() <= 1e-6 * np.abs(t).max()


def test_central_differences_catch_a_wrong_linearisation():
    """The finite-difference check has teeth: dropping the antidiffusive correction
    (the TLM of the donor pass alone) misses it by far more than the tolerance."""
    nx, ny, nz = 10, 9, 8
    rng, U, V, W, c = case(nx, ny, nz, 7)
    dc = rng.standard_normal(c.size)
    eps = 1e-6
    fd = (O.per_step(nx, ny, nz, U, V, W, 2, False, c + eps * dc)
          - O.per_step(nx, ny, nz, U, V, W, 2, False, c - eps * dc)) / (2 * eps)
    donor_only = O.per_step(nx, ny, nz, U, V, W, 1, False, dc)   # the donor pass is linear
    assert np.abs(fd - donor_only).max() > 1e-3 * np.abs(fd).max()


@pytest.mark.parametrize("nonosc", [False, True])
def test_euler_homogeneity(nonosc):
    nx, ny, nz = 9, 8, 7
    _, U, V, W, c = case(nx, ny, nz, 11)
    t = O.tlm_step(nx, ny, nz, U, V, W, nonosc, c, c)
    f = O.per_step(nx, ny, nz, U, V, W, 2, nonosc, c)
    assert np.abs(t - f).max() <= 1e-13 * np.abs(f).max()


@pytest.mark.parametrize("nonosc", [False, True])
def test_emission_is_identity_in_q(nonosc):
    nx, ny, nz = 6, 5, 4
    rng, U, V, W, c = case(nx, ny, nz, 5)
    dq = rng.standard_normal(c.size)
    t = O.tlm_step(nx, ny, nz, U, V, W, nonosc, c, np.zeros_like(c), dq)
    np.testing.assert_array_equal(t, dq)
    lam = rng.standard_normal(c.size)
    lq = np.zeros_like(c)
    O.adj_step(nx, ny, nz, U, V, W, nonosc, c, lam, lq)
    np.testing.assert_array_equal(lq, lam)


def test_receptor_gradient_matches_central_differences():
    nx, ny, nz = 8, 7, 6
    rng, U, V, W, c = case(nx, ny, nz, 3)
    q = rng.uniform(0.0, 2e-3, c.size)
    rec = rng.choice(c.size, size=6, replace=False)
    nsteps, every = 6, 2
    obs = rng.uniform(0, 1, (nsteps // every, rec.size))
    J, g0, gq = O.receptor_loss_and_gradient(nx, ny, nz, U, V, W, False, c, q, nsteps, rec,
                                             every, obs)
    eps = 1e-6
    for _ in range(3):
        d = rng.standard_normal(c.size)
        Jp = O.receptor_loss_and_gradient(nx, ny, nz, U, V, W, False, c + eps * d, q, nsteps,
                                          rec, every, obs)[0]
        Jm = O.receptor_loss_and_gradient(nx, ny, nz, U, V, W, False, c - eps * d, q, nsteps,
                                          rec, every, obs)[0]
        assert abs((Jp - Jm) / (2 * eps) - g0 @ d) <= 1e-6 * max(abs(g0 @ d), 1e-3)
        Jp = O.receptor_loss_and_gradient(nx, ny, nz, U, V, W, False, c, q + eps * d, nsteps,
                                          rec, every, obs)[0]
        Jm = O.receptor_loss_and_gradient(nx, ny, nz, U, V, W, False, c, q - eps * d, nsteps,
                                          rec, every, obs)[0]
        assert abs((Jp - Jm) / (2 * eps) - gq @ d) <= 1e-6 * max(abs(gq @ d), 1e-3)
