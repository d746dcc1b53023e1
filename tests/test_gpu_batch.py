"""GPU: on-device generators and the batched ensemble path.

* generators against numpy evaluations of the same formulas;
* every batch mode (per-instance launches, one batched launch per stage, hipGraph
  replay; persistent or one-workgroup-per-item grids; several tile shapes) gives the
  same fields, bit for bit, and equals the oracle (exact arithmetic: bitwise; fast:
  <= 1e-12 normwise);
* a captured graph replayed several times, with eager work and synchronisations in
  between, matches the eager result after every replay;
* the autotuner leaves the fields unchanged.
"""
import math

import numpy as np
import pytest

import bench
import mpdata_amd as M
import mpdata_amd.configs as K
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if M.device_count() < 1:
        pytest.skip("no GPU")


def relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_rotation_generator_matches_numpy():
    n = 64
    g = M.Grid(n, n, n, M.PERIODIC, iord=2, nonosc=True)
    K.generate_wind(g, K.ring3d_wind((n, n, n), cmax=0.4))
    U, V, W = g.get_courant()
    case = bench.rotation_case(n, cmax=0.4)
    for got, want in ((U, case["U"]), (V, case["V"]), (W, case["W"])):
        assert relerr(got, want) <= 1e-13
    c, s = K.ring3d_blob((n, n, n), offset=10.0)
    K.generate_blobs(g, [c], [s], [1.0], periodic=False)
    ax = np.arange(n, dtype=float)
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    want = np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * s * s))
    assert relerr(g.field(), want.ravel()) <= 1e-14
    g.close()


def test_curl_generator_is_divergence_free_and_scaled():
    shape = (40, 32, 24)
    g = M.Grid(*shape, M.PERIODIC, iord=2, nonosc=True)
    rng = K.SplitMix64(3)
    K.generate_wind(g, K.wind_spec(modes=K.random_modes(rng, 4, 1.0), cmax=0.5))
    U, V, W = (a.reshape(shape[::-1]) for a in g.get_courant())
    assert abs(max(np.abs(U).max(), np.abs(V).max(), np.abs(W).max()) - 0.5) < 1e-15
    # face Courant numbers are face averages of a divergence-free field: the discrete
    # divergence is a second-order truncation error, small against the Courant scale
    div = (U - np.roll(U, 1, 2)) + (V - np.roll(V, 1, 1)) + (W - np.roll(W, 1, 0))
    assert np.abs(div).max() < 0.05
    g.close()


DIMS = [(24, 20, 16), (16, 16, 16), (33, 17, 9), (40, 8, 12), (8, 8, 8)]


def make_batch(iord, arith, seed=7, dims=DIMS):
    b = M.Batch(dims, iord=iord, nonosc=True)
    b.set_arith(arith)
    rng = np.random.default_rng(seed)
    init = []
    for g, (nx, ny, nz) in zip(b.instances, dims):
        n = nx * ny * nz
        U, V, W = (rng.uniform(-0.2, 0.2, n) for _ in range(3))
        c = rng.uniform(0, 1, n) ** 3
        g.set_courant(U, V, W)
        g.set_field(c)
        init.append((U, V, W, c))
    return b, init


# (mode, tile shape / z-chunk / persistent of both kernels, limiter kernel's own shape)
MODES = [(M.BATCH_PER_INSTANCE, None, None), (M.BATCH_FUSED, (32, 8, 16, 1), None),
         (M.BATCH_FUSED, (16, 16, 8, 0), None), (M.BATCH_GRAPH, (32, 16, 8, 1), None),
         (M.BATCH_GRAPH, (16, 8, 32, 0), None),
         (M.BATCH_FUSED, (16, 16, 8, 0), (32, 8, 4)), (M.BATCH_FUSED, (64, 8, 8, 1), (16, 8, 3)),
         (M.BATCH_GRAPH, (32, 8, 16, 1), (64, 8, 16)),
         (M.BATCH_PER_INSTANCE, (32, 8, 16, 1), (16, 16, 8))]


@pytest.mark.parametrize("iord", [2, 3])
@pytest.mark.parametrize("arith", [M.ARITH_EXACT, M.ARITH_FAST])
def test_batch_modes_agree_and_match_oracle(iord, arith):
    results = []
    for mode, tile, lim in MODES:
        print(f"case iord={iord} arith={arith} mode={mode} tile={tile} limiter={lim}",
              flush=True)
        b, init = make_batch(iord, arith)
        if tile:
            b.set_tile(*tile)
        if lim:
            b.set_limiter_tile(*lim)
            assert b.tiles() == (tile, lim)
        b.step(3, mode)
        results.append([g.field() for g in b.instances])
        b.close()
    for r in results[1:]:
        for a, c in zip(r, results[0]):
            np.testing.assert_array_equal(a, c)
    _, init = make_batch(iord, arith)
    for (nx, ny, nz), (U, V, W, c), got in zip(DIMS, init, results[0]):
        want = O.per_run(nx, ny, nz, U, V, W, iord, True, c, 3)
        if arith == M.ARITH_EXACT:
            np.testing.assert_array_equal(got, want)
        else:
            assert relerr(got, want) <= 1e-12


# thin instances: one z-plane (stepped alone on the single-plane form, batched on the
# streaming kernel over a column of one plane) and two planes
THIN = [(24, 20, 16), (20, 12, 1), (9, 7, 2), (8, 8, 8)]


@pytest.mark.parametrize("dims", [DIMS, THIN], ids=["dims", "thin"])
@pytest.mark.parametrize("iord", [2, 3])
@pytest.mark.parametrize("arith", [M.ARITH_EXACT, M.ARITH_FAST])
def test_one_kernel_batched_launches(iord, arith, dims):
    """Batched launches with one kernel per corrective pass (forced on a small batch):
    fused launches and graph replays, several z-chunks and emission on some instances,
    against per-instance launches and the oracle -- bit for bit in exact arithmetic,
    1e-12 in fast arithmetic (the two paths contract differently)."""
    rng = np.random.default_rng(iord)
    qs = [rng.uniform(0, 1e-3, nx * ny * nz) if i % 2 else None
          for i, (nx, ny, nz) in enumerate(dims)]
    results = []
    for mode, zc in ((M.BATCH_PER_INSTANCE, 0), (M.BATCH_FUSED, 0), (M.BATCH_FUSED, 3),
                     (M.BATCH_GRAPH, 5), (M.BATCH_GRAPH, 0)):
        print(f"case one-kernel batch iord={iord} arith={arith} mode={mode} zc={zc}",
              flush=True)
        b, init = make_batch(iord, arith, dims=dims)
        for g, q in zip(b.instances, qs):
            if q is not None:
                g.set_emission(q)
        b.set_path(3)
        b.set_pass_zchunk(zc)
        assert b.path() == 3
        b.step(3, mode)
        if mode == M.BATCH_GRAPH:
            b.step(2, mode)  # a second replay from the other parity
            b.step(2, M.BATCH_FUSED)
        else:
            b.step(4, mode)
        results.append([g.field() for g in b.instances])
        b.close()
    for r in results[1:]:
        for a, c in zip(r, results[0]):
            if arith == M.ARITH_EXACT:
                np.testing.assert_array_equal(a, c)
            else:
                assert relerr(a, c) <= 1e-12
    _, init = make_batch(iord, arith, dims=dims)
    for (nx, ny, nz), (U, V, W, c), q, got in zip(dims, init, qs, results[1]):
        want = O.per_run(nx, ny, nz, U, V, W, iord, True, c, 7, q)
        if arith == M.ARITH_EXACT:
            np.testing.assert_array_equal(got, want)
        else:
            assert relerr(got, want) <= 1e-12


def test_graph_replays_match_eager():
    ref, _ = make_batch(2, M.ARITH_FAST)
    gb, _ = make_batch(2, M.ARITH_FAST)
    for rep in range(4):
        print(f"case replay {rep}", flush=True)
        ref.step(2, M.BATCH_PER_INSTANCE)
        ref.instances[0].synchronize()
        gb.step(2, M.BATCH_GRAPH)
        gb.instances[0].synchronize()
        # eager work between replays
        _ = [g.stats() for g in gb.instances]
        for a, b in zip(gb.instances, ref.instances):
            np.testing.assert_array_equal(a.field(), b.field())
    ref.close()
    gb.close()


def test_autotune_restores_fields():
    b, _ = make_batch(2, M.ARITH_FAST)
    before = [g.field() for g in b.instances]
    tx, ty, zc, pers = b.autotune(1)
    shapes = [(32, 8), (16, 16), (64, 8), (16, 8), (32, 16)]
    assert (tx, ty) in shapes
    assert zc in (8, 16, 32, 64) and pers in (0, 1)
    first, lim = b.tiles()
    assert first == (tx, ty, zc, pers)
    assert lim == (0, 0, 0) or (lim[:2] in shapes and lim[2] in (8, 16, 32, 64))
    for g, f in zip(b.instances, before):
        np.testing.assert_array_equal(g.field(), f)
    b.step(2, M.BATCH_GRAPH)
    assert all(np.isfinite(g.field()).all() for g in b.instances)
    b.close()


def test_one_kernel_autotune_times_the_zchunk():
    """On one kernel per pass the autotuner times the batched launches' z-chunk: it
    reports the kernel's tile, a candidate chunk no taller than the tallest instance (or
    the default when every instance is shorter than the smallest candidate), leaves the
    fields as they were, and the tuned launches equal per-instance launches bit for bit."""
    b, _ = make_batch(2, M.ARITH_EXACT)
    b.set_path(3)
    before = [g.field() for g in b.instances]
    tx, ty, zc, pers = b.autotune(1)
    tallest = max(nz for _, _, nz in DIMS)
    assert (tx, ty, pers) == (32, 16, 0)
    assert zc in [z for z in (16, 24, 32, 48, 64, 96, 128) if z <= tallest] or (
        zc == 32 and tallest < 16)
    for g, f in zip(b.instances, before):
        np.testing.assert_array_equal(g.field(), f)
    ref, _ = make_batch(2, M.ARITH_EXACT)
    b.step(3, M.BATCH_GRAPH)
    ref.step(3, M.BATCH_PER_INSTANCE)
    for a, c in zip(b.instances, ref.instances):
        np.testing.assert_array_equal(a.field(), c.field())
    b.close()
    ref.close()


def test_small_ensemble_config_runs():
    b = K.ensemble(6, seed=4)
    m0 = [g.stats()[0] for g in b.instances]
    b.step(5, M.BATCH_GRAPH)
    for g, m in zip(b.instances, m0):
        mass, lo, hi = g.stats()
        assert abs(mass - m) <= 1e-12 * m and lo >= 0.0
    b.close()
