"""GPU: the one-kernel-per-pass path (mpd_set_path(g, 3)) of the periodic nonoscillatory
step against the oracle.

Each corrective pass is one kernel that forms the donor field, the antidiffusive
velocities, the limiter ratios and the limited donor pass on a 32 x 16 tile with a
three-cell ring.  Cases cover grids narrower than the ring (the region wraps onto itself),
extents that are not multiples of the tile, z-chunks of one plane up to the whole
column, iord 2..4 (passes after the first read the previous pass's field and limited
faces) and the emission increment.  Exact arithmetic: bit for bit equal to the oracle;
fast arithmetic: <= 1e-12 normwise per step.  Grids of one z-plane take the kernel's
single-plane form (path 4), tested the same way.
"""
import numpy as np
import pytest

import mpdata_amd as M
import oracle_lib as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if M.device_count() < 1:
        pytest.skip("no GPU")


def relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def case(nx, ny, nz, seed, cmax=0.2):
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    U, V, W = (rng.uniform(-cmax, cmax, n) for _ in range(3))
    c = rng.uniform(0.0, 1.0, n) ** 3
    c[rng.uniform(0, 1, n) < 0.1] = 0.0  # cells on their bounds
    q = rng.uniform(0.0, 1e-3, n)
    return U, V, W, c, q


SHAPES = [(40, 20, 10), (33, 17, 5), (7, 6, 4), (64, 32, 9), (1, 5, 3), (70, 35, 2)]


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("arith", [M.ARITH_EXACT, M.ARITH_FAST])
def test_one_kernel_per_pass_matches_oracle(dims, arith):
    nx, ny, nz = dims
    U, V, W, c, q = case(nx, ny, nz, nx * 13 + ny * 5 + nz)
    for iord in (2, 3, 4):
        for with_q in (False, True):
            for zc in sorted({1, 2, 3, 5, nz}):
                print(f"case dims={dims} arith={arith} iord={iord} q={with_q} zc={zc}",
                      flush=True)
                g = M.Grid(nx, ny, nz, M.PERIODIC, iord=iord, nonosc=True)
                g.set_path(3)
                g.set_arith(arith)
                g.set_pass_zchunk(zc)
                g.set_courant(U, V, W)
                g.set_field(c)
                if with_q:
                    g.set_emission(q)
                want = c
                for s in range(3):
                    g.step(1)
                    want = O.per_step(nx, ny, nz, U, V, W, iord, True, want,
                                      q if with_q else None)
                    got = g.field()
                    if arith == M.ARITH_EXACT:
                        if not np.array_equal(got, want):
                            raise AssertionError(f"step {s}: differs, rel {relerr(got, want):.3e}")
                    else:
                        e = relerr(got, want)
                        assert e <= TOL, f"step {s}: rel err {e:.3e}"
                g.close()


def test_automatic_path_selection_and_agreement():
    """The automatic path takes one kernel per pass for nonoscillatory grids of at least
    8 planes and two kernels for thinner or unlimited ones; the two paths agree bit for
    bit in exact arithmetic, and the fast one-kernel path with the chosen z-chunk stays
    within 1e-12 of them."""
    nx, ny, nz = 256, 128, 128  # 2^22 cells
    U, V, W, c, q = case(nx, ny, nz, 5)
    out = {}
    for path in (0, 2, 3):
        for arith in (M.ARITH_EXACT, M.ARITH_FAST):
            print(f"case auto-path check path={path} arith={arith}", flush=True)
            g = M.Grid(nx, ny, nz, M.PERIODIC, iord=3, nonosc=True)
            g.set_path(path)
            g.set_arith(arith)
            g.set_courant(U, V, W)
            g.set_field(c)
            g.set_emission(q)
            assert g.path() == (3 if path == 0 else path)
            if g.path() == 3:
                assert 16 <= g.pass_zchunk() <= nz
            g.step(2)
            out[(path, arith)] = g.field()
            g.close()
    np.testing.assert_array_equal(out[(0, M.ARITH_EXACT)], out[(2, M.ARITH_EXACT)])
    np.testing.assert_array_equal(out[(3, M.ARITH_EXACT)], out[(2, M.ARITH_EXACT)])
    assert relerr(out[(0, M.ARITH_FAST)], out[(2, M.ARITH_EXACT)]) <= 2 * TOL
    for dims, nonosc, want in (((128, 128, 128), True, 3), ((7, 6, 8), True, 3),
                               ((128, 128, 7), True, 2), ((256, 128, 128), False, 2),
                               ((128, 128, 1), True, 4), ((128, 128, 1), False, 2)):
        g = M.Grid(*dims, M.PERIODIC, iord=2, nonosc=nonosc)
        assert g.path() == want
        g.set_path(3)
        assert g.path() == (3 if nonosc else 2)
        g.set_path(4)
        assert g.path() == (4 if nonosc and dims[2] == 1 else 2)
        g.close()


# (planes with fewer 16 x 16 tiles than 256 take 8 x 8 tiles, the last two 16 x 16)
PLANES = [(128, 128), (33, 17), (7, 5), (1, 9), (16, 16), (40, 3), (17, 1), (256, 256),
          (260, 250)]


@pytest.mark.parametrize("dims", PLANES)
@pytest.mark.parametrize("arith", [M.ARITH_EXACT, M.ARITH_FAST])
def test_single_plane_pass_matches_oracle(dims, arith):
    """Grids of one z-plane take the single-plane form of the one-kernel pass (path 4,
    automatic there): five stages over the plane, every z-neighbour the cell itself.
    Tiles of 8 x 8 or 16 x 16 with a three-cell ring, so the cases cover planes narrower
    than the ring in either axis and extents that are not multiples of the tile.  Exact arithmetic
    bit for bit equal to the oracle over three steps, fast within 1e-12 per step; the
    explicit and the automatic path agree bit for bit.  Path 3 on the same grids runs the
    streaming kernel itself over a column of one plane (its z-neighbours wrapped onto the
    cell's own column), to the same bars against the oracle."""
    nx, ny = dims
    U, V, W, c, q = case(nx, ny, 1, 7 * nx + ny)
    for iord in (2, 3, 4):
        for with_q in (False, True):
            fields = []
            for path in (4, 0, 3):
                print(f"case plane dims={dims} arith={arith} iord={iord} q={with_q} "
                      f"path={path}", flush=True)
                g = M.Grid(nx, ny, 1, M.PERIODIC, iord=iord, nonosc=True)
                g.set_path(path)
                assert g.path() == (3 if path == 3 else 4)
                g.set_arith(arith)
                g.set_courant(U, V, W)
                g.set_field(c)
                if with_q:
                    g.set_emission(q)
                want = c
                for s in range(3):
                    g.step(1)
                    want = O.per_step(nx, ny, 1, U, V, W, iord, True, want,
                                      q if with_q else None)
                    got = g.field()
                    if arith == M.ARITH_EXACT:
                        if not np.array_equal(got, want):
                            raise AssertionError(f"step {s}: differs, rel {relerr(got, want):.3e}")
                    else:
                        e = relerr(got, want)
                        assert e <= TOL, f"step {s}: rel err {e:.3e}"
                fields.append(got)
                g.close()
            np.testing.assert_array_equal(fields[0], fields[1])


@pytest.mark.parametrize("dims", [(40, 20, 10), (33, 17, 5), (7, 6, 4)])
def test_one_kernel_records_give_the_same_adjoint(dims):
    """The adjoint's forward records (c1, the unlimited faces, the unclamped ratios, the
    bound codes) taken from the one-kernel storing pass give the adjoint step and the
    checkpointed receptor gradient of the two-kernel records bit for bit, with several
    z-chunks (records on chunk boundaries)."""
    nx, ny, nz = dims
    U, V, W, c, q = case(nx, ny, nz, 900 + nx)
    rng = np.random.default_rng(nx)
    lam = rng.standard_normal(c.size)
    rec = rng.choice(c.size, size=min(10, c.size), replace=False).astype(np.int32)
    nsteps, every = 12, 3
    obs = rng.uniform(0, 1, (nsteps // every, rec.size))
    out = []
    for path, zc in ((2, 0), (3, 1), (3, 3), (3, 0)):
        print(f"case records dims={dims} path={path} zc={zc}", flush=True)
        g = M.Grid(nx, ny, nz, M.PERIODIC, iord=2, nonosc=True)
        g.set_arith(M.ARITH_EXACT)
        g.set_path(path)
        g.set_pass_zchunk(zc)
        g.set_courant(U, V, W)
        g.set_emission(q)
        lq = np.zeros(c.size)
        lin = g.adjoint_step(c, lam, lq)
        g.set_field(c)
        J, g0, gq, _ = g.receptor_gradient(nsteps, 4, rec, every, obs)
        g.close()
        out.append((lin, lq, J, g0, gq))
    for o in out[1:]:
        np.testing.assert_array_equal(o[0], out[0][0])
        np.testing.assert_array_equal(o[1], out[0][1])
        assert o[2] == out[0][2]
        np.testing.assert_array_equal(o[3], out[0][3])
        np.testing.assert_array_equal(o[4], out[0][4])
    want = O.adj_step(nx, ny, nz, U, V, W, True, c, lam)
    assert relerr(out[0][0], want) <= 1e-12
