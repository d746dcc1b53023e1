"""GPU adjoint of the periodic MPDATA step (iord 2) and the receptor-loss gradient.

* one adjoint step against the oracle's adjoint (oracle/cpu_adj.c);
* dot-product test against the oracle's tangent-linear model:
  <TLM dc, lam> = <dc, ADJ lam>;
* the checkpointed multi-step gradient against the oracle's full-sweep gradient, for
  several checkpoint strides;
* central finite differences of the GPU loss (eps = 1e-6) along random directions:
  unit-normal directions for c^0 (c = O(1)) and directions scaled by q itself for the
  emission (q ~ 1e-3: a unit-normal step of 1e-6 would be a 1e-3 relative change, enough
  to cross the limiter's kinks and leave the differentiable region).
Tolerances are written per test: the GPU adjoint gathers contributions in a different
order from the oracle's scatter, so agreement is to rounding (<= 1e-12 relative).
"""
import numpy as np
import pytest

import mpdata_amd as M
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if M.device_count() < 1:
        pytest.skip("no GPU")


def case(nx, ny, nz, seed, cmax=0.22):
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    U, V, W = (rng.uniform(-cmax, cmax, n) for _ in range(3))
    if nz == 1:
        W[:] = 0.0
    c = 0.05 + rng.uniform(0, 1, n) ** 2
    return rng, U, V, W, c


def grid(nx, ny, nz, nonosc, U, V, W):
    g = M.Grid(nx, ny, nz, M.PERIODIC, iord=2, nonosc=nonosc)
    g.set_arith(M.ARITH_EXACT)
    g.set_courant(U, V, W)
    return g


DIMS = [(12, 10, 8), (16, 9, 1), (7, 6, 5), (2, 6, 4), (1, 8, 6)]


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("nonosc", [False, True])
def test_adjoint_step_matches_oracle_and_tlm(dims, nonosc):
    nx, ny, nz = dims
    print(f"case dims={dims} nonosc={nonosc}", flush=True)
    rng, U, V, W, c = case(nx, ny, nz, nx * 100 + ny * 10 + nz)
    n = c.size
    g = grid(nx, ny, nz, nonosc, U, V, W)
    lam = rng.standard_normal(n)
    lq_gpu = np.zeros(n)
    got = g.adjoint_step(c, lam, lq_gpu)
    lq_cpu = np.zeros(n)
    want = O.adj_step(nx, ny, nz, U, V, W, nonosc, c, lam, lq_cpu)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    np.testing.assert_array_equal(lq_gpu, lq_cpu)
    for _ in range(3):
        dc = rng.standard_normal(n)
        t = O.tlm_step(nx, ny, nz, U, V, W, nonosc, c, dc)
        lhs, rhs = float(t @ lam), float(dc @ got)
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    g.close()


def loss_setup(nx, ny, nz, seed, nonosc):
    rng, U, V, W, c = case(nx, ny, nz, seed)
    n = c.size
    q = rng.uniform(0.0, 2e-3, n)
    rec = rng.choice(n, size=12, replace=False).astype(np.int32)
    return rng, U, V, W, c, q, rec


@pytest.mark.parametrize("nonosc", [False, True])
def test_receptor_gradient_matches_oracle(nonosc):
    nx, ny, nz = 14, 12, 10
    nsteps, every = 24, 4
    rng, U, V, W, c, q, rec = loss_setup(nx, ny, nz, 21, nonosc)
    obs = rng.uniform(0, 1, (nsteps // every, rec.size))
    Jc, g0c, gqc = O.receptor_loss_and_gradient(nx, ny, nz, U, V, W, nonosc, c, q, nsteps,
                                                rec, every, obs)
    for K in (1, 5, 24):
        print(f"case nonosc={nonosc} checkpoint={K}", flush=True)
        g = grid(nx, ny, nz, nonosc, U, V, W)
        g.set_field(c)
        g.set_emission(q)
        J, g0, gq, _ = g.receptor_gradient(nsteps, K, rec, every, obs)
        g.close()
        assert abs(J - Jc) <= 1e-13 * Jc
        assert np.abs(g0 - g0c).max() <= 1e-11 * np.abs(g0c).max()
        assert np.abs(gq - gqc).max() <= 1e-11 * np.abs(gqc).max()


@pytest.mark.parametrize("nonosc", [False, True])
def test_receptor_gradient_matches_central_differences(nonosc):
    nx, ny, nz = 16, 14, 12
    nsteps, every, K = 32, 8, 8
    rng, U, V, W, c, q, rec = loss_setup(nx, ny, nz, 33, nonosc)
    obs = rng.uniform(0, 1, (nsteps // every, rec.size))

    def loss(c0, qq, grads=False):
        g = grid(nx, ny, nz, nonosc, U, V, W)
        g.set_field(c0)
        g.set_emission(qq)
        out = g.receptor_gradient(nsteps, K, rec, every, obs, grads=grads)
        g.close()
        return out

    J, g0, gq, _ = loss(c, q, grads=True)
    eps = 1e-6
    for trial in range(4):
        d = rng.standard_normal(c.size)
        Jp = loss(c + eps * d, q)[0]
        Jm = loss(c - eps * d, q)[0]
        fd, an = (Jp - Jm) / (2 * eps), float(g0 @ d)
        print(f"c0 direction {trial}: fd {fd:.12e} adjoint {an:.12e}", flush=True)
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3)
        dq = q * d
        Jp = loss(c, q + eps * dq)[0]
        Jm = loss(c, q - eps * dq)[0]
        fd, an = (Jp - Jm) / (2 * eps), float(gq @ dq)
        print(f"q direction {trial}: fd {fd:.12e} adjoint {an:.12e}", flush=True)
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3)


def test_atomic_scatter_matches_gather():
    nx, ny, nz = 14, 12, 10
    nsteps, every = 24, 4
    rng, U, V, W, c, q, rec = loss_setup(nx, ny, nz, 21, True)
    obs = rng.uniform(0, 1, (nsteps // every, rec.size))
    out = []
    for atomic in (False, True):
        g = grid(nx, ny, nz, True, U, V, W)
        g.set_adjoint_scatter(atomic)
        g.set_field(c)
        g.set_emission(q)
        out.append(g.receptor_gradient(nsteps, 6, rec, every, obs))
        g.close()
    assert out[0][0] == out[1][0]
    for a, b in ((out[0][1], out[1][1]), (out[0][2], out[1][2])):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()


# shapes whose tiles are partly or fully inside the grid in x and y, with several
# z-chunks: the tiled kernels' edge rows read the halo waves' results
@pytest.mark.parametrize("dims", [(12, 10, 8), (1, 8, 6), (2, 6, 4), (40, 9, 5), (14, 12, 10),
                                  (16, 12, 8), (33, 25, 9)])
def test_tiled_limiter_adjoint_matches_per_cell_passes(dims):
    nx, ny, nz = dims
    rng, U, V, W, c = case(nx, ny, nz, 77 + nx)
    c[rng.uniform(0, 1, c.size) < 0.2] = 0.0  # ties with the bounds
    lam = rng.standard_normal(c.size)
    out = []
    for tiled in (True, False):
        g = grid(nx, ny, nz, True, U, V, W)
        g.set_adjoint_tiled(tiled)
        lq = np.zeros(c.size)
        out.append((g.adjoint_step(c, lam, lq), lq))
        g.close()
    (a, qa), (b, qb) = out
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()
    np.testing.assert_array_equal(qa, qb)
    want = O.adj_step(nx, ny, nz, U, V, W, True, c, lam)
    assert np.abs(a - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("nonosc", [False, True])
def test_kept_segment_records_give_the_same_gradient(nonosc):
    """Segments whose forward records are kept from the forward sweep (instead of being
    recomputed from their checkpoints) give the same loss and gradients bit for bit:
    only the last segment kept (default), two more kept by an exact byte budget, and all
    kept (-1, the free memory).  Sampling every K steps puts receptor residuals on every
    segment boundary, where a segment's first state is read from a kept record."""
    nx, ny, nz = 14, 12, 10
    nsteps = 24
    rng, U, V, W, c, q, rec = loss_setup(nx, ny, nz, 45, nonosc)
    per_step = nx * ny * nz * (60 if nonosc else 40)
    for K, every in ((5, 5), (4, 3), (7, 4), (1, 2)):
        obs = rng.uniform(0, 1, (nsteps // every, rec.size))
        out = []
        for tape in (0, 2 * K * per_step, -1):
            print(f"case nonosc={nonosc} K={K} every={every} tape={tape}", flush=True)
            g = grid(nx, ny, nz, nonosc, U, V, W)
            g.set_adjoint_tape(tape)
            g.set_field(c)
            g.set_emission(q)
            J, g0, gq, t = g.receptor_gradient(nsteps, K, rec, every, obs)
            end = g.field()
            nck = -(-nsteps // K)
            assert g.adjoint_tape() == (tape, {0: 1, -1: nck}.get(tape, min(3, nck)))
            if tape == -1:
                assert t[2] == 0.0  # nothing recomputed
            g.close()
            out.append((J, g0, gq, end))
        for J, g0, gq, end in out[1:]:
            assert J == out[0][0]
            np.testing.assert_array_equal(g0, out[0][1])
            np.testing.assert_array_equal(gq, out[0][2])
            np.testing.assert_array_equal(end, out[0][3])
    Jc, g0c, gqc = O.receptor_loss_and_gradient(nx, ny, nz, U, V, W, nonosc, c, q, nsteps,
                                                rec, 2, obs)
    assert abs(out[0][0] - Jc) <= 1e-13 * Jc
    assert np.abs(out[0][1] - g0c).max() <= 1e-11 * np.abs(g0c).max()
    assert np.abs(out[0][2] - gqc).max() <= 1e-11 * np.abs(gqc).max()


def test_receptor_sampling_rejects_frame_grids():
    """mpd_sample_receptors steps with the periodic kernels; a FRAME grid (face arrays in
    the reference's shorter layouts) is refused before any kernel is launched."""
    nx, ny, nz = 9, 8, 7
    n = nx * ny * nz
    g = M.Grid(nx, ny, nz, M.FRAME, iord=2, nonosc=False)
    one = np.ones(nx), np.ones(ny), np.ones(nz)
    z = np.zeros(n)
    g.set_winds(*one, 1.0, z, z, z, z, z, z)
    c = np.linspace(0.5, 1.5, n)
    g.set_field(c)
    with pytest.raises(M.MpdError) as e:
        g.sample_receptors(4, np.array([1, 5], np.int32), 2)
    assert e.value.code == 4  # MPD_ERR_UNSUPPORTED
    np.testing.assert_array_equal(g.field(), c)  # nothing was stepped
    g.close()
