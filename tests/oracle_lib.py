"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and, where it was built, the
reference's own core compiled from its sources (oracle/_ref/libsmolar_ref.so, and
oracle/_ref/libmsmolar_ref.so with its decomposed step).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg as the checker, never by the product.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libsmolar_ref.so")
MS_REF_SO = os.path.join(ROOT, "oracle", "_ref", "libmsmolar_ref.so")

PD = C.POINTER(C.c_double)
PPD = C.POINTER(PD)


def ptr(a):
    """Pointer to the data of a C-contiguous float64 array (None passes NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(PD)


_SMOLAR_ARGS = [C.c_int, C.c_int, C.c_int, PD, PD, PD, C.c_double,
                PD, PD, PD, PD, PD, PD, PPD, PPD, C.c_int, C.c_int]


def _bind(lib):
    if hasattr(lib, "orc_smolar"):
        lib.orc_smolar.argtypes = _SMOLAR_ARGS
        lib.orc_smolar.restype = C.c_int
        for name in ("orc_frame_courant",):
            getattr(lib, name).restype = None
        lib.orc_per_step.argtypes = [C.c_int] * 3 + [PD] * 3 + [C.c_int, C.c_int, PD, PD, PD]
        lib.orc_per_step.restype = C.c_int
        lib.orc_per_run.argtypes = [C.c_int] * 3 + [PD] * 3 + [C.c_int, C.c_int, PD, PD, C.c_int]
        lib.orc_per_run.restype = C.c_int
        lib.orc_per_courant.argtypes = [C.c_int] * 3 + [PD] * 3 + [C.c_double] + [PD] * 9
        lib.orc_per_courant.restype = None
        lib.orc_per_donor.argtypes = [C.c_int] * 3 + [PD] * 5
        lib.orc_per_antidiff.argtypes = [C.c_int] * 3 + [PD] * 7
        lib.orc_per_limiter_beta.argtypes = [C.c_int] * 3 + [PD] * 7
        lib.orc_per_limiter_apply.argtypes = [C.c_int] * 3 + [PD] * 5
        lib.orc_set_threads.argtypes = [C.c_int]
        lib.orc_set_threads.restype = None
        lib.orc_get_threads.argtypes = []
        lib.orc_get_threads.restype = C.c_int
        lib.orc_tlm_step.argtypes = [C.c_int] * 3 + [PD] * 3 + [C.c_int] + [PD] * 4
        lib.orc_tlm_step.restype = C.c_int
        lib.orc_adj_step.argtypes = [C.c_int] * 3 + [PD] * 3 + [C.c_int] + [PD] * 4
        lib.orc_adj_step.restype = C.c_int
    if hasattr(lib, "smolar"):
        lib.smolar.argtypes = _SMOLAR_ARGS
        lib.smolar.restype = C.c_int
    if hasattr(lib, "msmolar"):
        for name in ("msmolar", "msmolar_gather"):
            getattr(lib, name).argtypes = [C.POINTER(C.c_void_p)] + _SMOLAR_ARGS + [C.c_int] * 4
            getattr(lib, name).restype = C.c_int
    return lib


_oracle = None
_ref = None
_ms_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        _oracle = _bind(C.CDLL(ORACLE_SO))
    return _oracle


def have_ref():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = _bind(C.CDLL(REF_SO))
    return _ref


def have_ms_ref():
    return os.path.exists(MS_REF_SO)


def ms_ref():
    global _ms_ref
    if _ms_ref is None:
        _ms_ref = _bind(C.CDLL(MS_REF_SO))
    return _ms_ref


def call_smolar(fn, nx, ny, ns, dx, dy, ds, dt, u0, u1, v0, v1, w0, w1, c0, c1, order,
                flow_type=1):
    """Calls a smolar()-signature function on numpy buffers; returns (new_c0, new_c1)
    as the arrays the output pointers point to (the reference permutes them)."""
    p0 = ptr(c0)
    p1 = ptr(c1)
    pp0 = C.pointer(p0)
    pp1 = C.pointer(p1)
    rc = fn(nx, ny, ns, ptr(dx), ptr(dy), ptr(ds), dt, ptr(u0), ptr(u1), ptr(v0), ptr(v1),
            ptr(w0), ptr(w1), pp0, pp1, flow_type, order)
    a0 = C.addressof(pp0.contents.contents)
    a1 = C.addressof(pp1.contents.contents)
    by_addr = {c0.ctypes.data: c0, c1.ctypes.data: c1}
    return rc, by_addr[a0], by_addr[a1]


def set_threads(n):
    """Threads of the oracle's periodic loops (OpenMP; results do not depend on it)."""
    oracle().orc_set_threads(int(n))


def get_threads():
    return oracle().orc_get_threads()


def per_step(nx, ny, nz, U, V, W, iord, nonosc, cin, q=None):
    out = np.empty_like(cin)
    rc = oracle().orc_per_step(nx, ny, nz, ptr(U), ptr(V), ptr(W), iord, int(nonosc),
                               ptr(q), ptr(cin), ptr(out))
    assert rc == 0
    return out


def per_run(nx, ny, nz, U, V, W, iord, nonosc, c, nsteps, q=None):
    c = np.array(c, dtype=np.float64, copy=True, order="C")
    rc = oracle().orc_per_run(nx, ny, nz, ptr(U), ptr(V), ptr(W), iord, int(nonosc),
                              ptr(q), ptr(c), nsteps)
    assert rc == 0
    return c


def per_courant(nx, ny, nz, dx, dy, dz, dt, u0, u1, v0, v1, w0, w1):
    U = np.empty(nx * ny * nz)
    V = np.empty_like(U)
    W = np.empty_like(U)
    oracle().orc_per_courant(nx, ny, nz, ptr(dx), ptr(dy), ptr(dz), dt, ptr(u0), ptr(u1),
                             ptr(v0), ptr(v1), ptr(w0), ptr(w1), ptr(U), ptr(V), ptr(W))
    return U, V, W


def tlm_step(nx, ny, nz, U, V, W, nonosc, c, dc, dq=None):
    out = np.empty_like(c)
    assert oracle().orc_tlm_step(nx, ny, nz, ptr(U), ptr(V), ptr(W), int(nonosc), ptr(c),
                                 ptr(dc), ptr(dq), ptr(out)) == 0
    return out


def adj_step(nx, ny, nz, U, V, W, nonosc, c, lout, lq=None):
    lin = np.empty_like(c)
    assert oracle().orc_adj_step(nx, ny, nz, ptr(U), ptr(V), ptr(W), int(nonosc), ptr(c),
                                 ptr(lout), ptr(lin), ptr(lq)) == 0
    return lin


def receptor_loss_and_gradient(nx, ny, nz, U, V, W, nonosc, c0, q, nsteps, receptors,
                               sample_every, obs):
    """J = 0.5 sum over samples s (steps sample_every, 2*sample_every, ...) and receptor
    cells r of (c^s_r - obs[s-index, r])^2 for c^{n+1} = A(c^n) + q; returns
    (J, dJ/dc0, dJ/dq) by a full forward sweep and the oracle adjoint backwards."""
    states = [np.array(c0, dtype=np.float64)]
    for _ in range(nsteps):
        states.append(per_step(nx, ny, nz, U, V, W, 2, nonosc, states[-1], q))
    J = 0.0
    lam = np.zeros_like(c0)
    lq = np.zeros_like(c0)
    for n in range(nsteps, 0, -1):
        if n % sample_every == 0:
            s = n // sample_every - 1
            r = states[n][receptors] - obs[s]
            J += 0.5 * float(r @ r)
            np.add.at(lam, receptors, r)
        lam = adj_step(nx, ny, nz, U, V, W, nonosc, states[n - 1], lam, lq)
    return J, lam, lq
