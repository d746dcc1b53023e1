"""Python host binding of libsmolar_hip.so (MPDATA on AMD Instinct MI355X).

Mirrors the reference's C interface (smolar.h): ``smolar()`` with the same argument
meaning and pointer contract, ``msmolar()``/``msmolar_gather()`` over a device-resident
descriptor, plus the ``Grid`` class over the mpd_* extension interface (periodic
boundaries, nonoscillatory limiter, emission, batched ensembles, adjoint).

The shared library is the only compute path: importing this package on a machine
without it raises, and every call goes through the C ABI.  There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsmolar_hip.so")

FRAME = 0
PERIODIC = 1
ARITH_EXACT = 0
ARITH_FAST = 1
SOLENOIDAL = 0
DIVERGENT = 1
MSMOLAR_SERIAL, MSMOLAR_PTHREADS, MSMOLAR_OPENMP = 0, 1, 2

PD = C.POINTER(C.c_double)
PPD = C.POINTER(PD)
PI = C.POINTER(C.c_int)

ERRORS = {1: "MPD_ERR_ARG", 2: "MPD_ERR_DEVICE", 3: "MPD_ERR_NOMEM",
          4: "MPD_ERR_UNSUPPORTED", 5: "MPD_ERR_NUMERIC"}


class MpdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


# every exported function of include/*.h: name -> (restype, argtypes)
_SMOLAR = [C.c_int, C.c_int, C.c_int, PD, PD, PD, C.c_double, PD, PD, PD, PD, PD, PD,
           PPD, PPD, C.c_int, C.c_int]
_MSMOLAR = [C.POINTER(C.c_void_p)] + _SMOLAR + [C.c_int] * 4
SIGNATURES = {
    # smolar.h
    "smolar": (C.c_int, _SMOLAR),
    "msmolar": (C.c_int, _MSMOLAR),
    "msmolar_gather": (C.c_int, _MSMOLAR),
    "msmolar_free": (C.c_int, [C.POINTER(C.c_void_p)]),
    # timing.h
    "smolar_get_timer_resolution": (None, [C.c_void_p]),
    "smolar_get_time": (None, [C.c_void_p]),
    "smolar_get_time_diff": (C.c_double, [C.c_void_p, C.c_void_p]),
    "smolar_print_time_diff": (None, [C.c_void_p, C.c_void_p]),
    "smolar_device_timer_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "smolar_device_timer_start": (C.c_int, [C.c_void_p]),
    "smolar_device_timer_stop": (C.c_int, [C.c_void_p]),
    "smolar_device_timer_elapsed_ms": (C.c_int, [C.c_void_p, PD]),
    "smolar_device_timer_destroy": (C.c_int, [C.c_void_p]),
    # mpdata.h
    "mpd_error_string": (C.c_char_p, []),
    "mpd_device_count": (C.c_int, [PI]),
    "mpd_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, C.c_int,
                             C.c_int, C.c_int]),
    "mpd_destroy": (C.c_int, [C.c_void_p]),
    "mpd_set_winds": (C.c_int, [C.c_void_p, PD, PD, PD, C.c_double, PD, PD, PD, PD, PD, PD]),
    "mpd_set_courant": (C.c_int, [C.c_void_p, PD, PD, PD]),
    "mpd_get_courant": (C.c_int, [C.c_void_p, PD, PD, PD]),
    "mpd_set_field": (C.c_int, [C.c_void_p, PD]),
    "mpd_get_field": (C.c_int, [C.c_void_p, PD]),
    "mpd_get_previous_field": (C.c_int, [C.c_void_p, PD]),
    "mpd_set_emission": (C.c_int, [C.c_void_p, PD]),
    "mpd_get_emission": (C.c_int, [C.c_void_p, PD]),
    "mpd_step": (C.c_int, [C.c_void_p, C.c_int]),
    "mpd_step_periodic_noalloc": (C.c_int, [C.c_int] * 5 + [C.c_void_p] * 7),
    "mpd_synchronize": (C.c_int, [C.c_void_p]),
    "mpd_field_device_ptr": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "mpd_step_timed": (C.c_int, [C.c_void_p, C.c_int, PD]),
    "mpd_field_stats": (C.c_int, [C.c_void_p, PD, PD, PD]),
    "mpd_set_sentinel": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "mpd_sentinel_check": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong)]),
    "mpd_set_path": (C.c_int, [C.c_void_p, C.c_int]),
    "mpd_set_tile": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "mpd_set_zchunk": (C.c_int, [C.c_void_p, C.c_int]),
    "mpd_set_pass_zchunk": (C.c_int, [C.c_void_p, C.c_int]),
    "mpd_get_path": (C.c_int, [C.c_void_p, PI]),
    "mpd_get_pass_zchunk": (C.c_int, [C.c_void_p, PI]),
    "mpd_set_limiter_tile": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "mpd_get_tiles": (C.c_int, [C.c_void_p, PI]),
    "mpd_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "mpd_autotune": (C.c_int, [C.c_void_p, C.c_int, PI, PI, PI]),
    "mpd_set_arith": (C.c_int, [C.c_void_p, C.c_int]),
    "mpd_step_kernel_times": (C.c_int, [C.c_void_p, C.c_int, PD, C.c_int, PI]),
    "mpd_calibration_copy": (C.c_int, [C.c_size_t, C.c_int, PD]),
    "mpd_generate_wind": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mpd_generate_blobs": (C.c_int, [C.c_void_p, C.c_int, C.c_int, PD, PD, PD, C.c_int]),
    "mpd_set_adjoint_scatter": (C.c_int, [C.c_void_p, C.c_int]),
    "mpd_set_adjoint_tiled": (C.c_int, [C.c_void_p, C.c_int]),
    "mpd_set_adjoint_tape": (C.c_int, [C.c_void_p, C.c_longlong]),
    "mpd_get_adjoint_tape": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), PI]),
    "mpd_sample_receptors": (C.c_int, [C.c_void_p, C.c_int, C.c_int, PI, C.c_int, PD]),
    "mpd_adjoint_step": (C.c_int, [C.c_void_p, PD, PD, PD, PD]),
    "mpd_adjoint_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, PI, C.c_int, PD, PD,
                                  PD, PD, PD]),
    "mpd_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, PI, C.c_int, C.c_int]),
    "mpd_batch_destroy": (C.c_int, [C.c_void_p]),
    "mpd_batch_instance": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "mpd_batch_cells": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong)]),
    "mpd_batch_set_arith": (C.c_int, [C.c_void_p, C.c_int]),
    "mpd_batch_set_tile": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "mpd_batch_set_limiter_tile": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "mpd_batch_get_tiles": (C.c_int, [C.c_void_p, PI]),
    "mpd_batch_set_path": (C.c_int, [C.c_void_p, C.c_int]),
    "mpd_batch_get_path": (C.c_int, [C.c_void_p, PI]),
    "mpd_batch_set_pass_zchunk": (C.c_int, [C.c_void_p, C.c_int]),
    "mpd_batch_autotune": (C.c_int, [C.c_void_p, C.c_int, PI, PI, PI, PI]),
    "mpd_batch_step": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "mpd_batch_step_timed": (C.c_int, [C.c_void_p, C.c_int, C.c_int, PD]),
}

_lib = None


def lib():
    """The loaded libsmolar_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: build it with "
                              "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise MpdError(rc, lib().mpd_error_string().decode())


def _p(a):
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]):
        raise TypeError("arrays must be C-contiguous float64 numpy arrays")
    return a.ctypes.data_as(PD)


def device_count():
    n = C.c_int(0)
    _check(lib().mpd_device_count(C.byref(n)))
    return n.value


def calibration_copy(n, reps=10):
    """ms per launch of a plain device copy of n doubles."""
    ms = C.c_double(0)
    _check(lib().mpd_calibration_copy(n, reps, C.byref(ms)))
    return ms.value


def smolar(nx, ny, ns, dx, dy, ds, dt, u0, u1, v0, v1, w0, w1, c0, c1, flow_type=DIVERGENT,
           order=3):
    """One frame-mode step, reference smolar() semantics (smolar.h:45-48).  Returns
    (c0, c1): the arrays the reference's *c0/*c1 point to on return -- c1 holds the
    new field, c0 the previous pass's field (swapped for even order)."""
    p0, p1 = PD(_p(c0).contents), PD(_p(c1).contents)
    pp0, pp1 = C.pointer(p0), C.pointer(p1)
    _check(lib().smolar(nx, ny, ns, _p(dx), _p(dy), _p(ds), float(dt), _p(u0), _p(u1),
                        _p(v0), _p(v1), _p(w0), _p(w1), pp0, pp1, flow_type, order))
    by_addr = {c0.ctypes.data: c0, c1.ctypes.data: c1}
    return (by_addr[C.addressof(pp0.contents.contents)],
            by_addr[C.addressof(pp1.contents.contents)])


class MSmolar:
    """msmolar()/msmolar_gather() (smolar.h:95-136) over a device-resident descriptor."""

    def __init__(self):
        self.desc = C.c_void_p(None)

    def step(self, nx, ny, ns, dx, dy, ds, dt, u0, u1, v0, v1, w0, w1, c0, c1,
             flow_type=DIVERGENT, order=3, sx=1, sy=1, ss=1, mode=MSMOLAR_SERIAL):
        pp0, pp1 = C.pointer(PD(_p(c0).contents)), C.pointer(PD(_p(c1).contents))
        _check(lib().msmolar(C.byref(self.desc), nx, ny, ns, _p(dx), _p(dy), _p(ds),
                             float(dt), _p(u0), _p(u1), _p(v0), _p(v1), _p(w0), _p(w1),
                             pp0, pp1, flow_type, order, sx, sy, ss, mode))
        by_addr = {c0.ctypes.data: c0, c1.ctypes.data: c1}
        return (by_addr[C.addressof(pp0.contents.contents)],
                by_addr[C.addressof(pp1.contents.contents)])

    def gather(self, c0, c1, nx=0, ny=0, ns=0):
        pp0, pp1 = C.pointer(PD(_p(c0).contents)), C.pointer(PD(_p(c1).contents))
        _check(lib().msmolar_gather(C.byref(self.desc), nx, ny, ns, None, None, None, 0.0,
                                    None, None, None, None, None, None, pp0, pp1, 0, 0, 0,
                                    0, 0, 0))

    def close(self):
        if self.desc:
            lib().msmolar_free(C.byref(self.desc))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Grid:
    """A device-resident advection instance (mpd_grid)."""

    def __init__(self, nx, ny, nz, boundary=PERIODIC, iord=2, nonosc=True, _handle=None):
        self.shape = (nx, ny, nz)
        self.n = nx * ny * nz
        self.boundary = boundary
        self.owned = _handle is None
        self.h = C.c_void_p(None)
        if _handle is not None:
            self.h = _handle
        else:
            _check(lib().mpd_create(C.byref(self.h), nx, ny, nz, boundary, iord,
                                    int(nonosc)))

    def face_sizes(self):
        nx, ny, nz = self.shape
        if self.boundary == FRAME:
            return ((nx - 1) * (ny - 2) * (nz - 2), (nx - 2) * (ny - 1) * (nz - 2),
                    (nx - 2) * (ny - 2) * (nz - 1))
        return (self.n,) * 3

    def set_winds(self, dx, dy, dz, dt, u0, u1, v0, v1, w0, w1):
        _check(lib().mpd_set_winds(self.h, _p(dx), _p(dy), _p(dz), float(dt), _p(u0),
                                   _p(u1), _p(v0), _p(v1), _p(w0), _p(w1)))

    def set_courant(self, U, V, W):
        _check(lib().mpd_set_courant(self.h, _p(U), _p(V), _p(W)))

    def get_courant(self):
        out = [np.empty(s) for s in self.face_sizes()]
        _check(lib().mpd_get_courant(self.h, *(_p(a) for a in out)))
        return out

    def set_field(self, c):
        c = np.ascontiguousarray(c, dtype=np.float64)
        assert c.size == self.n
        _check(lib().mpd_set_field(self.h, _p(c)))

    def field(self):
        out = np.empty(self.n)
        _check(lib().mpd_get_field(self.h, _p(out)))
        return out

    def previous_field(self):
        out = np.empty(self.n)
        _check(lib().mpd_get_previous_field(self.h, _p(out)))
        return out

    def set_emission(self, q):
        if q is None:
            _check(lib().mpd_set_emission(self.h, None))
        else:
            q = np.ascontiguousarray(q, dtype=np.float64)
            _check(lib().mpd_set_emission(self.h, _p(q)))

    def emission(self):
        out = np.empty(self.n)
        _check(lib().mpd_get_emission(self.h, _p(out)))
        return out

    def step(self, nsteps=1):
        _check(lib().mpd_step(self.h, nsteps))

    def step_timed(self, nsteps):
        ms = C.c_double(0)
        _check(lib().mpd_step_timed(self.h, nsteps, C.byref(ms)))
        return ms.value

    def synchronize(self):
        _check(lib().mpd_synchronize(self.h))

    def stats(self):
        m, lo, hi = C.c_double(), C.c_double(), C.c_double()
        _check(lib().mpd_field_stats(self.h, C.byref(m), C.byref(lo), C.byref(hi)))
        return m.value, lo.value, hi.value

    def set_sentinel(self, every, nonnegative=False):
        """Device check of the field after every `every` steps (0: off) for non-finite
        (and, with nonnegative, negative) values."""
        _check(lib().mpd_set_sentinel(self.h, int(every), int(nonnegative)))

    def sentinel_check(self):
        """The first step count (since arming) after which the sentinel saw a bad value,
        or None."""
        step = C.c_longlong(-1)
        rc = lib().mpd_sentinel_check(self.h, C.byref(step))
        if rc == 5:  # MPD_ERR_NUMERIC: the sentinel fired
            return step.value
        _check(rc)
        return None

    def set_path(self, path):
        _check(lib().mpd_set_path(self.h, path))

    def set_tile(self, tx, ty):
        _check(lib().mpd_set_tile(self.h, tx, ty))

    def set_pass_zchunk(self, zc):
        """z-chunk of the one-kernel-per-pass kernel (0: chosen from the size)."""
        _check(lib().mpd_set_pass_zchunk(self.h, zc))

    def pass_zchunk(self):
        """The z-chunk the one-kernel-per-pass kernel uses."""
        v = C.c_int(0)
        _check(lib().mpd_get_pass_zchunk(self.h, C.byref(v)))
        return v.value

    def path(self):
        """The kernel path the steps take (1 staged, 2 two kernels per pass, 3 one, 4 one
        in its single-plane form)."""
        v = C.c_int(0)
        _check(lib().mpd_get_path(self.h, C.byref(v)))
        return v.value

    def set_zchunk(self, zc):
        _check(lib().mpd_set_zchunk(self.h, zc))

    def set_limiter_tile(self, tx, ty, zc=0):
        """Launch shape of the limiter kernel alone (0, 0: that of the first kernel)."""
        _check(lib().mpd_set_limiter_tile(self.h, tx, ty, zc))

    def tiles(self):
        """((tile_x, tile_y, zc) of the first kernel, the same of the limiter kernel);
        zeros mean the defaults / the first kernel's."""
        v = (C.c_int * 6)()
        _check(lib().mpd_get_tiles(self.h, v))
        return tuple(v[0:3]), tuple(v[3:6])

    def set_graph(self, on):
        _check(lib().mpd_set_graph(self.h, int(on)))

    def autotune(self, nsteps=2):
        """Picks, for each of the two fused kernels, the fastest tile shape and z-chunk
        length on this grid (state kept); returns the first kernel's (tile_x, tile_y, zc)
        (tiles() gives both)."""
        v = [C.c_int() for _ in range(3)]
        _check(lib().mpd_autotune(self.h, nsteps, *(C.byref(x) for x in v)))
        return tuple(x.value for x in v)

    def kernel_times(self, nsteps, max_kernels=16):
        """Average duration (ms) of each kernel of a step over nsteps steps."""
        ms = (C.c_double * max_kernels)()
        n = C.c_int(0)
        _check(lib().mpd_step_kernel_times(self.h, nsteps, ms, max_kernels, C.byref(n)))
        return [ms[i] for i in range(min(n.value, max_kernels))]

    def set_arith(self, arith):
        _check(lib().mpd_set_arith(self.h, arith))

    def adjoint_step(self, c, lout, lq=None):
        """lin = A'(c)^T lout (one step); lq += lout in place when given."""
        c = np.ascontiguousarray(c, dtype=np.float64)
        lout = np.ascontiguousarray(lout, dtype=np.float64)
        lin = np.empty(self.n)
        _check(lib().mpd_adjoint_step(self.h, _p(c), _p(lout), _p(lin), _p(lq)))
        return lin

    def set_adjoint_scatter(self, atomic):
        _check(lib().mpd_set_adjoint_scatter(self.h, int(atomic)))

    def set_adjoint_tiled(self, tiled):
        _check(lib().mpd_set_adjoint_tiled(self.h, int(tiled)))

    def set_adjoint_tape(self, max_bytes):
        """Bytes receptor_gradient may use for keeping the forward records of whole
        checkpoint segments beyond the last one (-1: the free device memory less a
        reserve; 0: the last segment only)."""
        _check(lib().mpd_set_adjoint_tape(self.h, int(max_bytes)))

    def adjoint_tape(self):
        """(tape budget in bytes, segment records the last gradient run kept)."""
        b, k = C.c_longlong(0), C.c_int(0)
        _check(lib().mpd_get_adjoint_tape(self.h, C.byref(b), C.byref(k)))
        return b.value, k.value

    def sample_receptors(self, nsteps, receptors, sample_every):
        rec = np.ascontiguousarray(receptors, dtype=np.int32)
        out = np.empty((nsteps // sample_every) * len(rec))
        _check(lib().mpd_sample_receptors(self.h, nsteps, len(rec), rec.ctypes.data_as(PI),
                                          sample_every, _p(out)))
        return out.reshape(nsteps // sample_every, len(rec))

    def receptor_gradient(self, nsteps, checkpoint_every, receptors, sample_every, obs,
                          grads=True):
        """(J, dJ/dc0, dJ/dq, (forward_ms, backward_ms, recompute_ms)) of the
        receptor-misfit loss from the current field (recompute_ms: the part of the backward
        spent recomputing checkpoint segments); the grid ends holding c^nsteps."""
        rec = np.ascontiguousarray(receptors, dtype=np.int32)
        obs = np.ascontiguousarray(obs, dtype=np.float64).ravel()
        J = C.c_double(0)
        g0 = np.empty(self.n) if grads else None
        gq = np.empty(self.n) if grads else None
        t = (C.c_double * 3)()
        _check(lib().mpd_adjoint_run(self.h, nsteps, checkpoint_every, len(rec),
                                     rec.ctypes.data_as(PI), sample_every, _p(obs),
                                     C.byref(J), _p(g0), _p(gq), t))
        return J.value, g0, gq, (t[0], t[1], t[2])

    def close(self):
        if self.h and self.owned:
            lib().mpd_destroy(self.h)
        self.h = C.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_PER_INSTANCE, BATCH_FUSED, BATCH_GRAPH = 0, 1, 2


class Batch:
    """An ensemble of periodic instances of uneven sizes (mpd_batch)."""

    def __init__(self, dims, iord=2, nonosc=True):
        self.dims = [tuple(int(x) for x in d) for d in dims]
        flat = (C.c_int * (3 * len(self.dims)))(*[x for d in self.dims for x in d])
        self.h = C.c_void_p(None)
        _check(lib().mpd_batch_create(C.byref(self.h), len(self.dims), flat, iord,
                                      int(nonosc)))
        self.instances = []
        for i, d in enumerate(self.dims):
            gh = C.c_void_p(None)
            _check(lib().mpd_batch_instance(self.h, i, C.byref(gh)))
            self.instances.append(Grid(*d, PERIODIC, iord, nonosc, _handle=gh))

    def cells(self):
        n = C.c_longlong(0)
        _check(lib().mpd_batch_cells(self.h, C.byref(n)))
        return n.value

    def set_arith(self, arith):
        _check(lib().mpd_batch_set_arith(self.h, arith))

    def set_tile(self, tx, ty, zc, persistent=True):
        _check(lib().mpd_batch_set_tile(self.h, tx, ty, zc, int(persistent)))

    def set_limiter_tile(self, tx, ty, zc=0):
        """Launch shape of the limiter kernel alone (0, 0: that of the first kernel)."""
        _check(lib().mpd_batch_set_limiter_tile(self.h, tx, ty, zc))

    def tiles(self):
        """((tile_x, tile_y, zc, persistent) of the first kernel, (tile_x, tile_y, zc) of
        the limiter kernel; zeros: the first kernel's)."""
        v = (C.c_int * 7)()
        _check(lib().mpd_batch_get_tiles(self.h, v))
        return tuple(v[0:4]), tuple(v[4:7])

    def set_path(self, path):
        """0 automatic, 2 two kernels per corrective pass, 3 one kernel per pass."""
        _check(lib().mpd_batch_set_path(self.h, path))

    def path(self):
        v = C.c_int(0)
        _check(lib().mpd_batch_get_path(self.h, C.byref(v)))
        return v.value

    def set_pass_zchunk(self, zc):
        _check(lib().mpd_batch_set_pass_zchunk(self.h, zc))

    def autotune(self, nsteps=2):
        v = [C.c_int() for _ in range(4)]
        _check(lib().mpd_batch_autotune(self.h, nsteps, *(C.byref(x) for x in v)))
        return tuple(x.value for x in v)

    def step(self, nsteps=1, mode=BATCH_GRAPH):
        _check(lib().mpd_batch_step(self.h, nsteps, mode))

    def step_timed(self, nsteps, mode=BATCH_GRAPH):
        ms = C.c_double(0)
        _check(lib().mpd_batch_step_timed(self.h, nsteps, mode, C.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            for g in self.instances:
                g.h = C.c_void_p(None)
            lib().mpd_batch_destroy(self.h)
            self.h = C.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
