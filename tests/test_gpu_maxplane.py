"""GPU: a periodic plane of the largest size the fused kernels take (2^29 cells).

The one-kernel pass addresses a column within a plane by a 32-bit byte offset, which is
why planes are limited to 2^29 cells (larger ones take the staged path, indexed in 64
bits).  At the limit itself, 32768 x 16384 cells of one z-plane, the single-plane form
(path 4), the streaming kernel over a column of one plane (path 3) and the staged path
(path 1) must give the same fields bit for bit in exact arithmetic over two iord-3
steps with the limiter, and the field's sum must stay that of the initial field to
rounding (the step is in flux form); one row more, every path request falls back to the
staged path.  About 70 GB of device memory at the peak (one grid at a time) and 30 GB of
host memory.
"""
import ctypes as C

import numpy as np
import pytest

import mpdata_amd as M

pytestmark = pytest.mark.gpu

NX, NY = 32768, 16384  # 2^29 cells


def _device_memory():
    """(free, total) device bytes (hipMemGetInfo of the HIP runtime the library loaded)."""
    M.device_count()
    with open("/proc/self/maps") as f:
        path = next(line.split()[-1] for line in f if "libamdhip64.so" in line)
    hip = C.CDLL(path)
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value, total.value


def test_largest_fused_plane_matches_staged_path():
    if M.device_count() < 1:
        pytest.skip("no GPU")
    if _device_memory()[0] < (96 << 30):
        pytest.skip("needs ~70 GB of free device memory")
    n = NX * NY
    rng = np.random.default_rng(29)
    U = rng.uniform(-0.3, 0.3, n)
    V = rng.uniform(-0.3, 0.3, n)
    W = np.zeros(n)
    c = rng.uniform(0.0, 1.0, n)
    c[: n // 16] = 0.0  # a band of cells on their lower bound
    mass0 = float(np.sum(c))
    out = {}
    for path in (1, 4, 3):
        print(f"case plane {NX}x{NY} iord 3 exact path {path}", flush=True)
        g = M.Grid(NX, NY, 1, M.PERIODIC, iord=3, nonosc=True)
        g.set_path(path)
        assert g.path() == path
        g.set_arith(M.ARITH_EXACT)
        g.set_courant(U, V, W)
        g.set_field(c)
        g.step(2)
        out[path] = g.field()
        g.close()
    for path in (4, 3):
        if not np.array_equal(out[path], out[1]):
            d = np.flatnonzero(out[path] != out[1])
            raise AssertionError(f"path {path}: {d.size} cells differ from the staged path, "
                                 f"first at {d[0]}")
    assert abs(float(np.sum(out[1])) - mass0) <= 1e-12 * mass0
    # one row more: beyond the fused kernels' plane limit every path request takes the
    # staged path
    g = M.Grid(NX, NY + 1, 1, M.PERIODIC, iord=3, nonosc=True)
    for path in (0, 2, 3, 4):
        g.set_path(path)
        assert g.path() == 1
    g.close()
