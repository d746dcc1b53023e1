"""GPU parity: the HIP path through the C ABI against the oracle and the reference's
golden fixtures.

Frame mode (the drop-in smolar()/msmolar()) must equal the reference bit for bit.
Periodic mode (both kernel paths) must agree with the oracle to a normwise relative
error <= 1e-12 per step and after N steps (max|gpu - oracle| / max|oracle|); the staged
path is compiled unfused and is expected to be bit-identical.
"""
import os

import numpy as np
import pytest

import mpdata_amd as M
import oracle_lib as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-12


def relerr(got, want):
    return float(np.abs(got - want).max() / max(np.abs(want).max(), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if M.device_count() < 1:
        pytest.skip("no GPU")


# ------------------------------------------------------------------ frame mode ---

@pytest.mark.parametrize("fixture,orders", [("frame_3d.npz", (1, 2, 3, 4)),
                                            ("frame_2d.npz", (1, 2, 3))])
def test_smolar_matches_reference_fixtures(fixture, orders):
    d = dict(np.load(os.path.join(GOLD, fixture)))
    nx, ny, ns = (int(x) for x in d["dims"])
    print(f"case {fixture}", flush=True)
    for order in orders:
        c0 = d["c"].copy()
        c1 = np.zeros_like(c0)
        for s in range(1, 13):
            a0, a1 = M.smolar(nx, ny, ns, d["dx"], d["dy"], d["ds"], float(d["dt"]), d["u0"],
                              d["u1"], d["v0"], d["v1"], d["w0"], d["w1"], c0, c1, order=order)
            c0, c1 = a1, a0
            if s in (1, 12):
                np.testing.assert_array_equal(c0, d[f"out_o{order}_s{s}"],
                                              err_msg=f"order {order} step {s}")


def test_smolar_pointer_contract_and_previous_pass():
    d = dict(np.load(os.path.join(GOLD, "frame_3d.npz")))
    nx, ny, ns = (int(x) for x in d["dims"])
    args = (d["dx"], d["dy"], d["ds"], float(d["dt"]), d["u0"], d["u1"], d["v0"], d["v1"],
            d["w0"], d["w1"])
    for order in (1, 2, 3):
        g0, g1 = d["c"].copy(), np.full(d["c"].size, 5.0)
        r0, r1 = M.smolar(nx, ny, ns, *args, g0, g1, order=order)
        o0, o1 = d["c"].copy(), np.full(d["c"].size, 5.0)
        _, q0, q1 = O.call_smolar(O.oracle().orc_smolar, nx, ny, ns, *args, o0, o1, order)
        assert (r1 is g1) == (q1 is o1)
        np.testing.assert_array_equal(r1, q1)
        np.testing.assert_array_equal(r0, q0)  # previous pass's field, as the reference


def test_smolar_large_arrays_through_the_staged_copies():
    """smolar() with host arrays of 8 MiB and more moves them through the pinned staging
    buffers in 8 MiB chunks (a full chunk and a partial one here, both directions, every
    array): both returned fields and the pointer contract against the oracle, bit for
    bit, over two calls with the returned buffers passed back in."""
    nx, ny, ns = 112, 110, 96  # 1.18 M cells: 9.5 MB per field
    n = nx * ny * ns
    rng = np.random.default_rng(112)
    u, v, w = (rng.uniform(-0.25, 0.25, n) for _ in range(3))
    c = rng.uniform(0.0, 1.0, n)
    sp = [rng.uniform(0.9, 1.1, m) for m in (nx, ny, ns)]
    for order in (2, 3):
        g0, g1 = c.copy(), np.zeros(n)
        o0, o1 = c.copy(), np.zeros(n)
        for call in range(2):
            print(f"case staged smolar order={order} call={call}", flush=True)
            g0, g1 = M.smolar(nx, ny, ns, *sp, 0.8, u, u, v, v, w, w, g0, g1, order=order)
            _, o0, o1 = O.call_smolar(O.oracle().orc_smolar, nx, ny, ns, *sp, 0.8, u, u, v, v,
                                      w, w, o0, o1, order)
            np.testing.assert_array_equal(g1, o1)
            np.testing.assert_array_equal(g0, o0)
            g0, g1 = g1, g0
            o0, o1 = o1, o0


def _device_free_bytes():
    """Free device memory (hipMemGetInfo of the HIP runtime the library loaded)."""
    import ctypes as C
    with open("/proc/self/maps") as f:
        path = next(line.split()[-1] for line in f if "libamdhip64.so" in line)
    hip = C.CDLL(path)
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_smolar_from_concurrent_and_short_lived_threads():
    """smolar() is reentrant: three threads calling it at once, each with host arrays
    large enough for the staged copies, get the oracle's fields bit for bit; and threads
    that call once and end keep no device memory each (the device state and the staging
    buffers are leased from pools per device, which grow only with concurrent calls)."""
    import threading
    nx, ny, ns = 112, 110, 96
    n = nx * ny * ns
    rng = np.random.default_rng(5)
    u, v, w = (rng.uniform(-0.25, 0.25, n) for _ in range(3))
    sp = [rng.uniform(0.9, 1.1, m) for m in (nx, ny, ns)]
    cs = [rng.uniform(0.0, 1.0, n) for _ in range(3)]
    want = []
    for c in cs:
        _, o0, o1 = O.call_smolar(O.oracle().orc_smolar, nx, ny, ns, *sp, 0.8, u, u, v, v,
                                  w, w, c.copy(), np.zeros(n), 2)
        want.append((o0, o1))
    got, errors = [None] * 3, []

    def run(i):
        try:
            got[i] = M.smolar(nx, ny, ns, *sp, 0.8, u, u, v, v, w, w, cs[i].copy(),
                              np.zeros(n), order=2)
        except Exception as e:  # reported below, in the test's thread
            errors.append(e)

    print("case concurrent smolar: 3 threads", flush=True)
    threads = [threading.Thread(target=run, args=(i,)) for i in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for (g0, g1), (o0, o1) in zip(got, want):
        np.testing.assert_array_equal(g0, o0)
        np.testing.assert_array_equal(g1, o1)
    free0 = _device_free_bytes()
    for k in range(6):
        print(f"case short-lived smolar thread {k}", flush=True)
        t = threading.Thread(target=run, args=(k % 3,))
        t.start()
        t.join()
        assert not errors, errors
    # one state of this grid holds ~15 arrays of 9.5 MB; six threads that each kept one
    # would take ~850 MB
    assert free0 - _device_free_bytes() < 64 << 20
    for (g0, g1), (o0, o1) in zip(got, want):
        np.testing.assert_array_equal(g0, o0)
        np.testing.assert_array_equal(g1, o1)


def test_ring2d_trajectory_matches_reference():
    d = dict(np.load(os.path.join(GOLD, "ring2d_32.npz")))
    n, _, m = (int(x) for x in d["dims"])
    one = np.ones(n)
    c0, c1 = d["c"].copy(), np.zeros_like(d["c"])
    for s in range(1, 51):
        a0, a1 = M.smolar(n, n, m, one, one, np.ones(m), float(d["dt"]), d["u"], d["u"],
                          d["v"], d["v"], d["w"], d["w"], c0, c1, order=3)
        c0, c1 = a1, a0
        if s in (1, 10, 50):
            np.testing.assert_array_equal(c0, d[f"out_s{s}"])


@pytest.mark.parametrize("order", [1, 2, 3])
def test_msmolar_matches_reference_fixture(order):
    d = dict(np.load(os.path.join(GOLD, "frame_3d.npz")))
    nx, ny, ns = (int(x) for x in d["dims"])
    ms = M.MSmolar()
    c0, c1 = d["c"].copy(), np.zeros_like(d["c"])
    for _ in range(12):
        c0, c1 = ms.step(nx, ny, ns, d["dx"], d["dy"], d["ds"], float(d["dt"]), d["u0"],
                         d["u1"], d["v0"], d["v1"], d["w0"], d["w1"], c0, c1, order=order)
    ms.gather(c0, c1)
    ms.close()
    np.testing.assert_array_equal(c1, d[f"out_o{order}_s12"])


def test_frame_grid_handle_matches_oracle():
    rng = np.random.default_rng(11)
    nx, ny, ns = 20, 17, 12
    n = nx * ny * ns
    sp = [rng.uniform(0.9, 1.1, k) for k in (nx, ny, ns)]
    winds = [rng.uniform(-0.3, 0.3, n) for _ in range(6)]
    c = rng.uniform(0, 1, n)
    g = M.Grid(nx, ny, ns, M.FRAME, iord=3, nonosc=False)
    g.set_winds(*sp, 0.8, *winds)
    g.set_field(c)
    g.step(5)
    got = g.field()
    c0, c1 = c.copy(), np.zeros(n)
    for _ in range(5):
        _, a0, a1 = O.call_smolar(O.oracle().orc_smolar, nx, ny, ns, *sp, 0.8, *winds, c0, c1, 3)
        c0, c1 = a1, a0
    np.testing.assert_array_equal(got, c0)


FRAME_SHAPES = [(70, 37, 21), (33, 9, 40), (3, 3, 3), (35, 3, 5), (3, 20, 9), (66, 18, 3)]


@pytest.mark.parametrize("dims", FRAME_SHAPES)
def test_frame_tiled_kernel_matches_per_face_kernels_and_oracle(dims):
    # the tiled donor + antidiffusive kernel (default) against the per-face kernels
    # (mpd_set_path 1) and the oracle's smolar(), bitwise: shapes spanning several
    # 32x8 tiles and z-chunks, and 3-cell extents (every face next to the frame)
    nx, ny, ns = dims
    rng = np.random.default_rng(nx * 1000 + ny * 10 + ns)
    n = nx * ny * ns
    sp = [rng.uniform(0.9, 1.1, k) for k in (nx, ny, ns)]
    winds = [rng.uniform(-0.3, 0.3, n) for _ in range(6)]
    c = rng.uniform(0, 1, n) ** 2
    print(f"case frame {dims}", flush=True)
    for order in (1, 2, 3, 4):
        out = []
        for path in (0, 1):
            g = M.Grid(nx, ny, ns, M.FRAME, iord=order, nonosc=False)
            g.set_path(path)
            g.set_winds(*sp, 0.7, *winds)
            g.set_field(c)
            g.step(3)
            out.append(g.field())
            g.close()
        np.testing.assert_array_equal(out[0], out[1], err_msg=f"order {order}")
        c0, c1 = c.copy(), np.zeros(n)
        for _ in range(3):
            _, a0, a1 = O.call_smolar(O.oracle().orc_smolar, nx, ny, ns, *sp, 0.7, *winds, c0,
                                      c1, order)
            c0, c1 = a1, a0
        np.testing.assert_array_equal(out[0], c0, err_msg=f"order {order} vs oracle")


# --------------------------------------------------------------- periodic mode ---

def per_case(nx, ny, nz, seed, cmax=0.2):
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    U, V, W = (rng.uniform(-cmax, cmax, n) for _ in range(3))
    if nz == 1:
        W[:] = 0.0
    c = rng.uniform(0, 1, n) ** 3
    return U, V, W, c


PER_CASES = [(16, 16, 16), (24, 20, 1), (13, 11, 7), (64, 8, 5), (1, 9, 10), (5, 1, 1)]


# (path, arith): staged exact, fused exact, fused fast, automatic (one kernel per pass
# for the nonoscillatory grids of at least 8 planes, its z-chunk from the cost model, and
# its single-plane form for those of one plane)
PATHS = [(1, M.ARITH_EXACT), (2, M.ARITH_EXACT), (2, M.ARITH_FAST), (0, M.ARITH_EXACT),
         (0, M.ARITH_FAST)]


@pytest.mark.parametrize("path,arith", PATHS)
@pytest.mark.parametrize("dims", PER_CASES)
def test_periodic_matches_oracle(dims, path, arith):
    nx, ny, nz = dims
    U, V, W, c = per_case(nx, ny, nz, nx * 31 + ny * 7 + nz)
    for iord in (1, 2, 3):
        for nonosc in (False, True):
            print(f"case dims={dims} path={path} arith={arith} iord={iord} nonosc={nonosc}",
                  flush=True)
            g = M.Grid(nx, ny, nz, M.PERIODIC, iord=iord, nonosc=nonosc)
            g.set_path(path if iord > 1 else 1)
            g.set_arith(arith)
            g.set_courant(U, V, W)
            g.set_field(c)
            want = c
            for s in range(4):
                g.step(1)
                want = O.per_step(nx, ny, nz, U, V, W, iord, nonosc, want)
                got = g.field()
                e = relerr(got, want)
                assert e <= TOL, f"iord {iord} nonosc {nonosc} step {s}: {e:.3e}"
                if arith == M.ARITH_EXACT or iord == 1:
                    # unfused IEEE arithmetic in the oracle's order: bit-identical
                    np.testing.assert_array_equal(got, want)
            g.close()


TILES = [(32, 8), (16, 16), (64, 8), (16, 8), (32, 16)]


@pytest.mark.parametrize("tile", TILES)
def test_fused_tiles_match_oracle(tile):
    for dims in [(70, 19, 9), (33, 17, 3), (128, 8, 40)]:
        nx, ny, nz = dims
        U, V, W, c = per_case(nx, ny, nz, 17 + nx)
        for iord in (2, 3):
            for nonosc in (False, True):
                want = O.per_run(nx, ny, nz, U, V, W, iord, nonosc, c, 3)
                for arith in (M.ARITH_EXACT, M.ARITH_FAST):
                    print(f"case tile={tile} dims={dims} iord={iord} nonosc={nonosc} "
                          f"arith={arith}", flush=True)
                    g = M.Grid(nx, ny, nz, M.PERIODIC, iord=iord, nonosc=nonosc)
                    g.set_path(2)
                    g.set_arith(arith)
                    g.set_tile(*tile)
                    g.set_courant(U, V, W)
                    g.set_field(c)
                    g.step(3)
                    got = g.field()
                    g.close()
                    if arith == M.ARITH_EXACT:
                        np.testing.assert_array_equal(got, want)
                    else:
                        assert relerr(got, want) <= TOL


def test_periodic_matches_padded_reference_fixture():
    d = dict(np.load(os.path.join(GOLD, "periodic.npz")))
    for tag in ("a", "b"):
        nx, ny, nz = (int(x) for x in d[f"{tag}_dims"])
        for order in (1, 2, 3):
            g = M.Grid(nx, ny, nz, M.PERIODIC, iord=order, nonosc=False)
            g.set_winds(d[f"{tag}_dx"], d[f"{tag}_dy"], d[f"{tag}_dz"], float(d[f"{tag}_dt"]),
                        d[f"{tag}_u"], d[f"{tag}_u"], d[f"{tag}_v"], d[f"{tag}_v"],
                        d[f"{tag}_w"], d[f"{tag}_w"])
            g.set_field(d[f"{tag}_c"])
            g.step(3)
            want = d[f"{tag}_out_o{order}"]
            assert relerr(g.field(), want) <= TOL
            g.close()


@pytest.mark.parametrize("path", [1, 2, 0])
def test_periodic_emission_and_long_run(path):
    nx, ny, nz = 48, 40, 32
    U, V, W, c = per_case(nx, ny, nz, 5, cmax=0.25)
    q = np.random.default_rng(6).uniform(0, 1e-3, nx * ny * nz)
    g = M.Grid(nx, ny, nz, M.PERIODIC, iord=2, nonosc=True)
    g.set_path(path)
    g.set_courant(U, V, W)
    g.set_field(c)
    g.set_emission(q)
    g.step(20)
    want = O.per_run(nx, ny, nz, U, V, W, 2, True, c, 20, q)
    got = g.field()
    assert relerr(got, want) <= TOL
    mass, lo, hi = g.stats()
    assert abs(mass - want.sum()) <= 1e-12 * want.sum()
    assert lo >= 0.0
    g.set_emission(None)
    g.close()


def test_mass_and_positivity_at_128():
    """Size-independent invariants on the 128^3 rotation case the bench measures."""
    import bench

    case = bench.rotation_case(128, cmax=0.4)
    g = M.Grid(128, 128, 128, M.PERIODIC, iord=2, nonosc=True)
    g.set_courant(case["U"], case["V"], case["W"])
    g.set_field(case["c"])
    m0, _, hi0 = g.stats()
    g.step(16)
    m1, lo, hi = g.stats()
    assert abs(m1 - m0) <= 1e-12 * m0
    assert lo >= 0.0 and hi <= hi0 * (1 + 1e-14)
    g.close()


def test_autotune_keeps_state_and_parity():
    nx, ny, nz = 70, 19, 9
    U, V, W, c = per_case(nx, ny, nz, 5)
    want = O.per_run(nx, ny, nz, U, V, W, 2, True, c, 3)
    g = M.Grid(nx, ny, nz, M.PERIODIC, iord=2, nonosc=True)
    g.set_path(2)  # the two-kernel path is the one with launch shapes to tune
    g.set_arith(M.ARITH_EXACT)
    g.set_courant(U, V, W)
    g.set_field(c)
    tx, ty, zc = g.autotune(1)
    print(f"case autotune -> tile {tx}x{ty} zc {zc}", flush=True)
    assert (tx, ty) in TILES and 1 <= zc <= nz
    np.testing.assert_array_equal(g.field(), c)
    g.step(3)
    np.testing.assert_array_equal(g.field(), want)
    g.close()
    with pytest.raises(M.MpdError):
        f = M.Grid(nx, ny, nz, M.FRAME, iord=2)
        try:
            f.autotune(1)
        finally:
            f.close()


@pytest.mark.parametrize("iord", [2, 3])
@pytest.mark.parametrize("zc", [1, 2, 3, 4, 5, 6, 9, 40])
def test_zchunk_lengths_match_oracle(zc, iord):
    """Every z-chunk length, bitwise: a chunk of zc planes streams zc + 2 planes through
    loops unrolled by four, then two, then one, so 1..9 cover each remainder (nz = 9
    also leaves short last chunks); iord 3 adds the limited faces stored for the next
    pass."""
    nx, ny, nz = 33, 17, 9
    U, V, W, c = per_case(nx, ny, nz, 8)
    want = O.per_run(nx, ny, nz, U, V, W, iord, True, c, 2)
    g = M.Grid(nx, ny, nz, M.PERIODIC, iord=iord, nonosc=True)
    g.set_path(2)  # the two-kernel path's chunks (test_gpu_onepass.py covers the other's)
    g.set_arith(M.ARITH_EXACT)
    g.set_zchunk(zc)
    g.set_courant(U, V, W)
    g.set_field(c)
    g.step(2)
    np.testing.assert_array_equal(g.field(), want)
    g.close()


@pytest.mark.parametrize("dims", [(33, 17, 9), (33, 17, 1)])
@pytest.mark.parametrize("path", [0, 2])
@pytest.mark.parametrize("iord,nonosc", [(2, True), (2, False), (3, True)])
def test_graph_replays_match_eager_steps(iord, nonosc, path, dims):
    """mpd_step replays blocks of 8 steps as captured graphs; the same steps launched one
    by one give identical fields -- over several replays with eager work (field reads)
    between them, odd step counts, a tile change that forces a new capture, and an
    emission array; on the automatic path (one kernel per pass with the limiter, in its
    single-plane form on the plane) and on two kernels per pass."""
    nx, ny, nz = dims
    U, V, W, c = per_case(nx, ny, nz, 12)
    q = np.random.default_rng(3).uniform(0.0, 1e-3, c.size)
    for arith in (M.ARITH_EXACT, M.ARITH_FAST):
        gs = []
        for on in (True, False):
            g = M.Grid(nx, ny, nz, M.PERIODIC, iord=iord, nonosc=nonosc)
            g.set_path(path)
            g.set_arith(arith)
            g.set_graph(on)
            g.set_courant(U, V, W)
            g.set_field(c)
            g.set_emission(q)
            gs.append(g)
        total = 0
        for n in (8, 17, 1, 16, 3, 24):
            print(f"case iord={iord} nonosc={nonosc} arith={arith} steps={n}", flush=True)
            for g in gs:
                g.step(n)
            total += n
            np.testing.assert_array_equal(gs[0].field(), gs[1].field())
            if n == 1:
                for g in gs:
                    g.set_tile(16, 16)
        if arith == M.ARITH_EXACT:
            want = O.per_run(nx, ny, nz, U, V, W, iord, nonosc, c, total, q)
            np.testing.assert_array_equal(gs[0].field(), want)
        for g in gs:
            g.close()


@pytest.mark.parametrize("iord", [2, 3])
def test_limiter_kernel_in_its_own_launch_shape(iord):
    """The limiter kernel may run in a tile shape and z-chunk of its own: exact arithmetic
    gives the same fields bit for bit as one shared shape, fast arithmetic stays within
    the tolerance of the oracle, and graph replays with mixed shapes match eager steps."""
    for dims in [(70, 19, 9), (33, 17, 3), (40, 24, 13)]:
        nx, ny, nz = dims
        U, V, W, c = per_case(nx, ny, nz, 29 + nx)
        want = O.per_run(nx, ny, nz, U, V, W, iord, True, c, 10)  # graph: a block of 8 + 2
        for arith in (M.ARITH_EXACT, M.ARITH_FAST):
            outs = []
            for lim, graph in (((0, 0, 0), False), ((16, 16, 5), False), ((64, 8, 0), False),
                               ((16, 8, 2), True)):
                print(f"case dims={dims} iord={iord} arith={arith} limiter={lim} graph={graph}",
                      flush=True)
                g = M.Grid(nx, ny, nz, M.PERIODIC, iord=iord, nonosc=True)
                g.set_path(2)
                g.set_arith(arith)
                g.set_tile(32, 8)
                g.set_zchunk(4)
                g.set_limiter_tile(*lim)
                assert g.tiles() == ((32, 8, 4), lim if lim[0] else (0, 0, 0))
                g.set_graph(graph)
                g.set_courant(U, V, W)
                g.set_field(c)
                g.step(10)
                outs.append(g.field())
                g.close()
            for o in outs:
                if arith == M.ARITH_EXACT:
                    np.testing.assert_array_equal(o, want)
                else:
                    assert relerr(o, want) <= TOL


def test_autotune_reports_both_kernels_shapes():
    nx, ny, nz = 40, 24, 13
    U, V, W, c = per_case(nx, ny, nz, 31)
    g = M.Grid(nx, ny, nz, M.PERIODIC, iord=2, nonosc=True)
    g.set_path(2)
    g.set_courant(U, V, W)
    g.set_field(c)
    before = g.field()
    f1 = g.autotune(1)
    t1, t2 = g.tiles()
    assert t1 == f1
    shapes = {(32, 8), (16, 16), (64, 8), (16, 8), (32, 16)}
    assert t1[:2] in shapes and t2[:2] in shapes and 1 <= t1[2] <= nz and 1 <= t2[2] <= nz
    np.testing.assert_array_equal(g.field(), before)  # tuning leaves the state unchanged
    g.step(3)
    assert relerr(g.field(), O.per_run(nx, ny, nz, U, V, W, 2, True, c, 3)) <= TOL
    g.close()


def test_autotune_leaves_one_kernel_grids_alone():
    """A grid whose steps take one kernel per corrective pass (the automatic path of a
    nonoscillatory grid) has no launch shapes to tune: the tuner returns at once, the
    grid's two-kernel settings, path and state unchanged."""
    nx, ny, nz = 40, 24, 13
    U, V, W, c = per_case(nx, ny, nz, 32)
    g = M.Grid(nx, ny, nz, M.PERIODIC, iord=2, nonosc=True)
    g.set_courant(U, V, W)
    g.set_field(c)
    assert g.path() == 3
    before = g.tiles()
    assert g.autotune(1) == before[0]
    assert g.tiles() == before and g.path() == 3
    np.testing.assert_array_equal(g.field(), c)
    g.step(3)
    assert relerr(g.field(), O.per_run(nx, ny, nz, U, V, W, 2, True, c, 3)) <= TOL
    g.close()


def test_autotune_with_many_passes():
    """iord 11: twenty kernels per step, more than the autotuner's first guess of a
    per-step kernel count; it times them all and the tuned grid matches the oracle."""
    nx, ny, nz = 20, 12, 8
    U, V, W, c = per_case(nx, ny, nz, 41)
    g = M.Grid(nx, ny, nz, M.PERIODIC, iord=11, nonosc=True)
    g.set_path(2)
    g.set_courant(U, V, W)
    g.set_field(c)
    g.autotune(1)
    t1, t2 = g.tiles()
    assert t1[0] > 0 and t1[2] >= 1
    g.step(2)
    assert relerr(g.field(), O.per_run(nx, ny, nz, U, V, W, 11, True, c, 2)) <= TOL
    g.close()
