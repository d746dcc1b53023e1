"""GPU: the one-kernel-per-pass path (mpd_set_path(g, 3)) of the periodic nonoscillatory
step against the oracle.

Each corrective pass is one kernel that forms the donor field, the antidiffusive
velocities, the limiter ratios and the limited donor pass on a 32 x 16 tile with a
three-cell ring.  Cases cover grids narrower than the ring (the region wraps onto itself),
extents that are not multiples of the tile, z-chunks of one plane up to the whole
column, iord 2..4 (passes after the first read the previous pass's field and limited
faces) and the emission increment.  Exact arithmetic: bit for bit equal to the oracle;
fast arithmetic: <= 1e-12 normwise per step.
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
    2^22 cells and two kernels otherwise; at such a size the two paths agree bit for bit
    in exact arithmetic, and the fast one-kernel path with the chosen z-chunk stays
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
    for dims, nonosc, want in (((128, 128, 128), True, 2), ((256, 128, 128), False, 2)):
        g = M.Grid(*dims, M.PERIODIC, iord=2, nonosc=nonosc)
        assert g.path() == want
        g.set_path(3)
        assert g.path() == (3 if nonosc else 2)
        g.close()
