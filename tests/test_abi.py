"""The C ABI library loads without a GPU and exports every function that include/*.h
declares; host-only entry points (timing, argument checks, device count) behave."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import mpdata_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("smolar.h", "timing.h", "mpdata.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name in ("if", "defined", "sizeof", "return"):
                continue
            start = text.rfind("\n", 0, m.start())
            line = text[start:m.start()]
            if "#" in line:  # preprocessor lines (#define, #error)
                continue
            names.add(name)
    return names


def test_library_exports_every_declared_function():
    lib = M.lib()
    declared = declared_functions()
    assert {"smolar", "msmolar", "msmolar_gather", "smolar_get_time", "mpd_step"} <= declared
    out = subprocess.run(["nm", "-D", "--defined-only", M.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = declared - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    for name in declared:
        getattr(lib, name)
    assert set(M.SIGNATURES) == declared


def test_library_does_not_link_the_oracle():
    out = subprocess.run(["readelf", "-d", M.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    assert "oracle" not in out and "smolar_ref" not in out
    syms = subprocess.run(["nm", "-D", M.LIB_PATH], capture_output=True, text=True,
                          check=True).stdout
    assert "orc_" not in syms


class TimeT(C.Structure):
    _fields_ = [("seconds", C.c_long), ("nanoseconds", C.c_long)]


def test_host_timing_api():
    lib = M.lib()
    a, b, r = TimeT(), TimeT(), TimeT()
    lib.smolar_get_timer_resolution(C.byref(r))
    assert r.seconds == 0 and 0 < r.nanoseconds < 1_000_000
    lib.smolar_get_time(C.byref(a))
    lib.smolar_get_time(C.byref(b))
    d = lib.smolar_get_time_diff(C.byref(a), C.byref(b))
    assert 0.0 <= d < 1.0
    # borrow across a second boundary (timing.c:53-57)
    a.seconds, a.nanoseconds, b.seconds, b.nanoseconds = 10, 900_000_000, 12, 100_000_000
    assert abs(lib.smolar_get_time_diff(C.byref(a), C.byref(b)) - 1.2) < 1e-12


def test_argument_errors_without_compute():
    lib = M.lib()
    n = C.c_int(-1)
    assert lib.mpd_device_count(C.byref(n)) == 0 and n.value >= 0
    h = C.c_void_p(None)
    assert lib.mpd_create(C.byref(h), 2, 5, 5, M.FRAME, 2, 0) == 1  # frame needs >= 3
    assert "frame grid" in lib.mpd_error_string().decode()
    assert lib.mpd_create(C.byref(h), 8, 8, 8, M.FRAME, 2, 1) == 4  # limiter: periodic only
    assert lib.mpd_create(C.byref(h), 8, 8, 8, M.PERIODIC, 0, 0) == 1
    assert lib.mpd_step(None, 1) == 1
    x = np.zeros(27)
    assert lib.smolar(3, 3, 3, None, None, None, 1.0, *([None] * 6), None, None, 1, 2) == 1
    with pytest.raises(M.MpdError):
        M._check(1)
