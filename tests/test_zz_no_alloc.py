"""The allocation-free periodic step, mpd_step_periodic_noalloc, through the C ABI.

* GPU: two calls in a row on caller-owned device arrays (the second on a smaller grid
  and another iord), then a third without the limiter, each equal bit for bit to
  mpd_step on a grid with the same Courant numbers, options and emission (the same
  fused kernels in the same arithmetic).  Device arrays come from the HIP runtime.
* CPU: invalid arguments -- too many cells for the fixed scratch, NULL arrays, equal
  input and output, iord 1 -- are refused before any device work (pointer values that
  are never dereferenced).
"""
import ctypes as C

import numpy as np
import pytest

import mpdata_amd as M

MPD_ERR_ARG = 1


class Hip:
    """The few HIP runtime calls the test needs for device arrays."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        self.check(self.lib.hipMalloc(C.byref(p), a.nbytes), "hipMalloc")
        self.bufs.append(p)
        self.check(self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, n):
        out = np.empty(n)
        self.check(self.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self.check(self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def case(dims, seed, cmax=0.2):
    rng = np.random.default_rng(seed)
    n = dims[0] * dims[1] * dims[2]
    U, V, W = (rng.uniform(-cmax, cmax, n) for _ in range(3))
    c = rng.uniform(0.0, 1.0, n) ** 2
    q = rng.uniform(0.0, 1e-3, n)
    return U, V, W, c, q


def grid_step(dims, iord, nonosc, U, V, W, c, q):
    g = M.Grid(*dims, M.PERIODIC, iord=iord, nonosc=nonosc)
    g.set_courant(U, V, W)
    g.set_field(c)
    g.set_emission(q)
    g.step(1)
    out = g.field()
    g.close()
    return out


@pytest.mark.gpu
def test_noalloc_step_equals_grid_step_on_consecutive_calls():
    if M.device_count() < 1:
        pytest.skip("no GPU")
    hip = Hip()
    try:
        calls = [((48, 40, 36), 3, True, True), ((24, 20, 16), 2, True, False),
                 ((30, 18, 12), 2, False, True)]
        for k, (dims, iord, nonosc, with_q) in enumerate(calls):
            print(f"case call {k}: dims={dims} iord={iord} nonosc={nonosc} q={with_q}",
                  flush=True)
            U, V, W, c, q = case(dims, 100 + k)
            n = c.size
            d = [hip.upload(a) for a in (U, V, W, c)]
            dq = hip.upload(q) if with_q else None
            dout = hip.upload(np.full(n, np.nan))
            rc = M.lib().mpd_step_periodic_noalloc(*dims, iord, int(nonosc), d[0], d[1], d[2],
                                                   dq, d[3], dout, None)
            assert rc == 0, M.lib().mpd_error_string().decode()
            got = hip.download(dout, n)
            want = grid_step(dims, iord, nonosc, U, V, W, c, q if with_q else None)
            np.testing.assert_array_equal(got, want)
            hip.free()
    finally:
        hip.free()


def test_noalloc_refuses_invalid_arguments_before_device_work():
    f = M.lib().mpd_step_periodic_noalloc
    fake = [C.c_void_p(0x1000 * (i + 1)) for i in range(6)]
    U, V, W, q, cin, cout = fake
    # 8 Mi cells is the iord-2 limit, 4 Mi at iord >= 3
    assert f(256, 256, 129, 2, 1, U, V, W, q, cin, cout, None) == MPD_ERR_ARG
    assert f(256, 256, 65, 3, 1, U, V, W, q, cin, cout, None) == MPD_ERR_ARG
    assert "exceed" in M.lib().mpd_error_string().decode()
    assert f(8, 8, 8, 2, 1, None, V, W, q, cin, cout, None) == MPD_ERR_ARG
    assert f(8, 8, 8, 2, 1, U, V, W, q, cin, None, None) == MPD_ERR_ARG
    assert f(8, 8, 8, 2, 1, U, V, W, q, cin, cin, None) == MPD_ERR_ARG
    assert f(8, 8, 8, 1, 1, U, V, W, q, cin, cout, None) == MPD_ERR_ARG
    assert f(0, 8, 8, 2, 1, U, V, W, q, cin, cout, None) == MPD_ERR_ARG
