"""The reference's own driver programs, unchanged, on this library.

tests/ring2d/test.c and tests/ring3d/test.c of the reference are compiled against
include/smolar.h (their own init.h beside them) and this project's implementation of the
reference's init.h interface (mpdata_amd/drivers/test_config.c), and linked against
libsmolar_hip.so (oracle/Makefile target `callers`, outputs in oracle/_ref).

* CPU: where the reference tree is mounted, both programs compile with the signature
  checks as errors and link; the built programs load the library and print their usage.
* GPU: the reference's `ring2d 128 1024` run -- its own main program, on the HIP
  library -- writes the reference's ring2d.bin and standard output hash for hash (the
  fixture recorded from the reference's sources); `ring3d 24 4` follows the oracle's
  smolar() step by step, bit for bit.
"""
import hashlib
import os
import subprocess

import numpy as np
import pytest

import mpdata_amd as M
import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("MPD_REFERENCE_DIR", "/root/reference")
BUILT = os.path.join(ROOT, "oracle", "_ref")
GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.mark.parametrize("prog", ["ring2d", "ring3d"])
def test_reference_callers_compile_and_link(prog, tmp_path):
    src = os.path.join(REF, "tests", prog, "test.c")
    if not os.path.exists(src):
        pytest.skip("reference tree not mounted")
    exe = tmp_path / prog
    drv = os.path.join(ROOT, "mpdata_amd", "drivers")
    r = subprocess.run(
        ["gcc", "-std=c99", "-O1", "-DDOUBLE", "-I", os.path.join(ROOT, "include"),
         "-Werror=implicit-function-declaration", "-Werror=incompatible-pointer-types",
         "-Werror=int-conversion", "-o", str(exe), src, os.path.join(drv, "test_config.c"),
         os.path.join(drv, "ring_setup.c"), "-L", os.path.join(ROOT, "mpdata_amd"),
         "-lsmolar_hip", "-Wl,-rpath," + os.path.join(ROOT, "mpdata_amd"), "-lm"],
        capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, cwd=tmp_path)
    assert out.returncode == 0 and out.stdout == f"{prog} <n> <nt>\n"


def caller(prog):
    exe = os.path.join(BUILT, f"{prog}_caller")
    if not os.path.exists(exe):
        pytest.skip("the reference's callers were not built (reference tree absent at build)")
    return exe


@pytest.mark.gpu
def test_reference_ring2d_program_reproduces_its_own_run(tmp_path):
    if M.device_count() < 1:
        pytest.skip("no GPU")
    exe = caller("ring2d")
    d = dict(np.load(os.path.join(GOLD, "ring2d_128_1024.npz")))
    print("case reference ring2d/test.c 128 1024 on libsmolar_hip", flush=True)
    r = subprocess.run([exe, "128", str(int(d["nt"]))], capture_output=True, text=True,
                       cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    h = hashlib.sha256()
    with open(tmp_path / "ring2d.bin", "rb") as f:
        for chunk in iter(lambda: f.read(1 << 24), b""):
            h.update(chunk)
    assert hashlib.sha256(r.stdout.encode()).hexdigest() == str(d["stdout_sha256"])
    assert h.hexdigest() == str(d["bin_sha256"])
    for f in ("ring2d.ctl", "ring2d.gs", "ring2d.pl"):
        assert (tmp_path / f).exists()


@pytest.mark.gpu
def test_reference_ring3d_program_follows_the_oracle(tmp_path):
    if M.device_count() < 1:
        pytest.skip("no GPU")
    exe = caller("ring3d")
    n, nt = 24, 4
    print(f"case reference ring3d/test.c {n} {nt} on libsmolar_hip", flush=True)
    r = subprocess.run([exe, str(n), str(nt)], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout == "".join(f"step {i}\n" for i in range(nt + 1))
    nn = n ** 3
    raw = np.fromfile(tmp_path / "ring3d.raw").reshape(nt + 1, 4, nn)
    u, v, w, c = (raw[0, i].copy() for i in range(4))
    one = np.ones(n)
    c0, c1 = c.copy(), np.zeros(nn)
    for s in range(nt):
        _, a0, a1 = O.call_smolar(O.oracle().orc_smolar, n, n, n, one, one, one, 0.0125, u, u,
                                  v, v, w, w, c0, c1, 3)
        c0, c1 = a1, a0
        np.testing.assert_array_equal(raw[s + 1, 3], c0)
        np.testing.assert_array_equal(raw[s + 1, 0], u)
