"""The benchmark configurations: seeded field parameters and their set-up on the device.

Every random choice comes from splitmix64 (Steele, Lea & Flood 2014) so that the same
seed gives the same configuration on any host.  Heavy per-cell work runs on the device
(mpd_generate_wind / mpd_generate_blobs); the host only draws a few dozen parameters.

  ring2d(n=128)       2-D doubly periodic grid, u = -w(y - n/2), v = w(x - n/2) scaled to
                      max Courant 0.5, cone of height 1 and radius 15 at (64, 96)
  ring3d(n=128)       periodic n^3, solid-body rotation about (1,1,1)/sqrt(3) through the
                      centre at max Courant 0.4, Gaussian blob sigma 8 cells, 30 cells
                      off-axis along (1,-1,0)/sqrt(2)
  deform(shape, seed) periodic grid, divergence-free wind curl(A) of 4 Fourier modes per
                      potential component (seeded), max Courant 0.5; 32 seeded Gaussians
                      (sigma in [4, 16], amplitude in [0.5, 1], periodic images)
  ensemble(n, seed)   n instances drawn from {64^3, 96^3, 128^3, 160x160x96, 192x192x64};
                      each the ring3d rotation about its own centre plus a seeded
                      divergence-free perturbation of 10% of the rotation speed, max
                      Courant 0.4, its own Gaussian blob
  inversion_grid(n)   source inversion: periodic n^3, the deform wind and field of seed 5
                      with an emission of 8 seeded positive Gaussians; with `factors`,
                      the twin run whose emission amplitudes are scaled by them
  inversion_receptors(ncell)  64 seeded distinct receptor cells and the twin's factors
                      1 + 0.2 xi, xi ~ U[-1, 1] per source
"""
import ctypes as C
import math

import numpy as np

from . import (Batch, Grid, PERIODIC, _check, lib)

MAX_MODES = 8
MAX_BLOBS = 64
TARGET_FIELD, TARGET_EMISSION = 0, 1
MASK = (1 << 64) - 1


class WindSpec(C.Structure):
    _fields_ = [("axis", C.c_double * 3), ("centre", C.c_double * 3), ("omega", C.c_double),
                ("nmodes", C.c_int),
                ("k", ((C.c_double * 3) * MAX_MODES) * 3),
                ("amp", (C.c_double * MAX_MODES) * 3),
                ("phase", (C.c_double * MAX_MODES) * 3),
                ("cmax", C.c_double)]


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * 2.0 ** -53)

    def randint(self, n):
        return self.next() % n


def wind_spec(axis=(0, 0, 1), centre=(0, 0, 0), omega=0.0, modes=None, cmax=0.5):
    """modes: list (per potential component d) of lists of (kx, ky, kz, amp, phase)."""
    s = WindSpec()
    a = np.asarray(axis, dtype=float)
    n = np.linalg.norm(a)
    for e in range(3):
        s.axis[e] = a[e] / n if n > 0 else 0.0
        s.centre[e] = centre[e]
    s.omega = omega
    modes = modes or [[], [], []]
    s.nmodes = max(len(m) for m in modes) if modes else 0
    assert s.nmodes <= MAX_MODES
    for d in range(3):
        for m, (kx, ky, kz, amp, ph) in enumerate(modes[d]):
            s.k[d][m][0], s.k[d][m][1], s.k[d][m][2] = kx, ky, kz
            s.amp[d][m] = amp
            s.phase[d][m] = ph
    s.cmax = cmax
    return s


def random_modes(rng, nmodes, scale):
    """nmodes Fourier modes per potential component: wavenumbers in {-3..3}^3 minus the
    zero vector, amplitude scale*U[0.5,1]/|k|, phase U[0, 2 pi)."""
    out = []
    for _ in range(3):
        comp = []
        for _ in range(nmodes):
            while True:
                k = [rng.randint(7) - 3 for _ in range(3)]
                if any(k):
                    break
            kn = math.sqrt(sum(x * x for x in k))
            comp.append((k[0], k[1], k[2], scale * rng.uniform(0.5, 1.0) / kn,
                         rng.uniform(0.0, 2 * math.pi)))
        out.append(comp)
    return out


def generate_wind(g, spec):
    _check(lib().mpd_generate_wind(g.h, C.addressof(spec)))


def generate_blobs(g, centres, sigma, amp, periodic=True, target=TARGET_FIELD):
    c = np.ascontiguousarray(np.asarray(centres, dtype=np.float64).ravel())
    s = np.ascontiguousarray(np.asarray(sigma, dtype=np.float64))
    a = np.ascontiguousarray(np.asarray(amp, dtype=np.float64))
    P = C.POINTER(C.c_double)
    _check(lib().mpd_generate_blobs(g.h, target, len(s), c.ctypes.data_as(P),
                                    s.ctypes.data_as(P), a.ctypes.data_as(P), int(periodic)))


# ------------------------------------------------------------------ configs ---

def ring3d_wind(shape, cmax=0.4):
    nx, ny, nz = shape
    return wind_spec(axis=(1, 1, 1), centre=(nx / 2, ny / 2, nz / 2), omega=1.0, cmax=cmax)


def ring3d_blob(shape, offset=30.0, sigma=8.0):
    nx, ny, nz = shape
    c = np.array([nx / 2, ny / 2, nz / 2]) + offset * np.array([1.0, -1.0, 0.0]) / math.sqrt(2)
    return c, sigma


def ring3d(g, n=128):
    generate_wind(g, ring3d_wind((n, n, n)))
    c, s = ring3d_blob((n, n, n))
    generate_blobs(g, [c], [s], [1.0], periodic=False)


def ring2d(n=128):
    """Returns (U, V, W, c) host arrays of the 2-D rotation + cone case (nz = 1)."""
    x = np.arange(n, dtype=np.float64)
    X, Y = np.meshgrid(x, x, indexing="xy")
    u = -(Y - n / 2)
    v = X - n / 2
    U = 0.25 * ((u + np.roll(u, -1, axis=1)) * 2.0)
    V = 0.25 * ((v + np.roll(v, -1, axis=0)) * 2.0)
    s = 0.5 / max(np.abs(U).max(), np.abs(V).max())
    r = np.sqrt((X - 64.0 * n / 128) ** 2 + (Y - 96.0 * n / 128) ** 2)
    c = np.maximum(0.0, 1.0 - r / (15.0 * n / 128))
    return ((U * s).ravel().copy(), (V * s).ravel().copy(), np.zeros(n * n), c.ravel().copy())


def deform(g, shape, seed, cmax=0.5, nblobs=32, emission_blobs=0, emission_scale=1e-3):
    """Config-3-style wind and field on grid g; with emission_blobs > 0 also an emission
    increment of that many positive Gaussians (amplitude emission_scale*U[0.5,1])."""
    nx, ny, nz = shape
    rng = SplitMix64(seed)
    generate_wind(g, wind_spec(modes=random_modes(rng, 4, 1.0), cmax=cmax))
    cen = [(rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(0, nz)) for _ in range(nblobs)]
    sig = [rng.uniform(4.0, 16.0) for _ in range(nblobs)]
    amp = [rng.uniform(0.5, 1.0) for _ in range(nblobs)]
    generate_blobs(g, cen, sig, amp, periodic=True)
    if emission_blobs:
        cen = [(rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(0, nz))
               for _ in range(emission_blobs)]
        sig = [rng.uniform(4.0, 16.0) for _ in range(emission_blobs)]
        amp = [emission_scale * rng.uniform(0.5, 1.0) for _ in range(emission_blobs)]
        generate_blobs(g, cen, sig, amp, periodic=True, target=TARGET_EMISSION)
        return cen, sig, amp
    return None


ENSEMBLE_SHAPES = [(64, 64, 64), (96, 96, 96), (128, 128, 128), (160, 160, 96), (192, 192, 64)]


def ensemble_dims(n=192, seed=4):
    rng = SplitMix64(seed)
    return [ENSEMBLE_SHAPES[rng.randint(len(ENSEMBLE_SHAPES))] for _ in range(n)]


def ensemble(n=192, seed=4, iord=2):
    """A Batch of the config-4 ensemble with its fields generated on the device."""
    dims = ensemble_dims(n, seed)
    b = Batch(dims, iord=iord, nonosc=True)
    rng = SplitMix64(seed * 1000003 + 1)
    for g, (nx, ny, nz) in zip(b.instances, dims):
        r0 = min(nx, ny, nz) / 2
        # curl of a mode of amplitude a and wavenumber |k| peaks near a*2*pi*|k|/L:
        # scale the modes to about 10% of the rotation speed at half radius
        scale = 0.1 * (r0 / 2) * min(nx, ny, nz) / (2 * math.pi) / 2
        spec = wind_spec(axis=(1, 1, 1), centre=(nx / 2, ny / 2, nz / 2), omega=1.0,
                         modes=random_modes(rng, 2, scale), cmax=0.4)
        generate_wind(g, spec)
        ang = rng.uniform(0, 2 * math.pi)
        e1 = np.array([1.0, -1.0, 0.0]) / math.sqrt(2)
        e2 = np.array([1.0, 1.0, -2.0]) / math.sqrt(6)
        off = (r0 / 2) * (math.cos(ang) * e1 + math.sin(ang) * e2)
        c = np.array([nx / 2, ny / 2, nz / 2]) + off
        generate_blobs(g, [c], [8.0], [1.0], periodic=True)
    return b


INVERSION_SEED = 5


def inversion_receptors(ncell, nrec=64, nsources=8, spread=0.2, seed=5005):
    """(receptor cells: nrec distinct seeded indices below ncell, sorted; the twin run's
    emission factors 1 + spread*xi, xi ~ U[-1, 1], one per source)."""
    rng = SplitMix64(seed)
    rec = sorted({rng.randint(ncell) for _ in range(nrec * 2)})[:nrec]
    factors = [1.0 + spread * rng.uniform(-1.0, 1.0) for _ in range(nsources)]
    return np.array(rec, np.int32), factors


def inversion_grid(n, nonosc=True, factors=None, iord=2):
    """A periodic n^3 grid holding the source-inversion case: the deformational wind and
    32 Gaussians of seed 5 and an emission of 8 seeded positive Gaussians (amplitudes
    scaled by `factors` for the twin run that makes the observations)."""
    shape = (n, n, n)
    g = Grid(n, n, n, PERIODIC, iord=iord, nonosc=nonosc)
    cen, sig, amp = deform(g, shape, seed=INVERSION_SEED, emission_blobs=8)
    if factors is not None:
        generate_blobs(g, cen, sig, [a * f for a, f in zip(amp, factors)], periodic=True,
                       target=TARGET_EMISSION)
    return g
