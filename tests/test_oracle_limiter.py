"""Pins the oracle's nonoscillatory limiter from both sides.

The limiter is not in the reference (Smolarkiewicz & Grabowski 1990, in the form of
Smolarkiewicz & Margolin 1998, J. Comput. Phys. 140:459-480, eqs. (29)-(31)).  The
invariant tests of test_oracle_periodic.py catch a limiter that is too weak (new
extrema); a limiter that is too strong -- in the extreme, one that zeroes every
antidiffusive velocity and leaves the donor-cell scheme -- passes them all.  These tests
catch that side:

* a hand-worked four-cell periodic case: every intermediate of one limited corrective
  pass (donor field, antidiffusive velocities, bounds, ratios, limited velocities, the
  result) against values worked out in exact rational arithmetic from the published
  formulas, written here as literals;
* accuracy on a smooth profile: over one revolution of a 1-D periodic Gaussian the
  limited scheme's error stays within 10 % of the unlimited scheme's and below 30 % of
  the donor scheme's, and it converges at the unlimited scheme's second order under grid
  refinement where the donor scheme does not.
"""
import numpy as np
import pytest

import oracle_lib as O

# c^n = (1, 2, 7, 8) on a periodic line of four cells, uniform Courant number 1/2.
#
#   donor pass (upwind flux u c_i through face i+1/2):
#       c1_i = c_i - (c_i - c_{i-1})/2                      = (9/2, 3/2, 9/2, 15/2)
#   antidiffusive velocity of face i+1/2 (uniform u: no divergence and no cross terms):
#       A = |u|(1-|u|)(c1_{i+1} - c1_i)/(c1_{i+1} + c1_i)   = (-1/8, 1/8, 1/16, -1/16)
#   bounds: max / min of c^n and c1 over {i-1, i, i+1}:
#       hi = (8, 7, 8, 8),   lo = (1, 1, 3/2, 1)
#   pass fluxes f = A c1_upwind:  (-3/16, 3/16, 9/32, -9/32)
#       A_in  = pos(f_{i-1/2}) - neg(f_{i+1/2})            = (3/16, 0, 3/16, 9/16)
#       A_out = pos(f_{i+1/2}) - neg(f_{i-1/2})            = (9/32, 3/8, 9/32, 0)
#   ratios (eq. 30):
#       b_up = (hi - c1)/A_in   = (56/3, +inf, 56/3, 8/9)
#       b_dn = (c1 - lo)/A_out  = (112/9, 4/3, 32/3, +inf)
#   limited velocities (eq. 31):
#       A'_{i+1/2} = min(1, b_dn_i, b_up_{i+1}) [A]+ + min(1, b_up_i, b_dn_{i+1}) [A]-
#                  = (-1/8, 1/8, (8/9)(1/16), (8/9)(-1/16)) = (-1/8, 1/8, 1/18, -1/18)
#   result c2 = c1 - (f'_{i+1/2} - f'_{i-1/2})              = (71/16, 9/8, 71/16, 8)
#   unlimited: c2 = c1 - (f_{i+1/2} - f_{i-1/2})            = (141/32, 9/8, 141/32, 129/16)
# Cell 3 receives through both of its faces; the limiter scales both by its b_up = 8/9
# and the cell lands exactly on its bound 8, where the unlimited pass overshoots to
# 8 + 1/16.
CN = np.array([1.0, 2.0, 7.0, 8.0])
C1 = np.array([9 / 2, 3 / 2, 9 / 2, 15 / 2])
A = np.array([-1 / 8, 1 / 8, 1 / 16, -1 / 16])
BUP = np.array([56 / 3, np.inf, 56 / 3, 8 / 9])
BDN = np.array([112 / 9, 4 / 3, 32 / 3, np.inf])
AL = np.array([-1 / 8, 1 / 8, 1 / 18, -1 / 18])
C2 = np.array([71 / 16, 9 / 8, 71 / 16, 8.0])
C2_UNLIMITED = np.array([141 / 32, 9 / 8, 141 / 32, 129 / 16])


def four_cell_case():
    U = np.full(4, 0.5)
    z = np.zeros(4)
    return U, z.copy(), z.copy()


def close(got, want, rel=4e-16):
    """Each finite entry within a few ulp of the exact value; 'infinite' entries (a zero
    flux sum: the oracle divides by DBL_MIN) must be at least 1e300."""
    fin = np.isfinite(want)
    assert np.all(np.abs(got[fin] - want[fin]) <= rel * np.abs(want[fin])), (got, want)
    assert np.all(got[~fin] >= 1e300), (got, want)


def test_hand_worked_four_cell_limited_pass():
    U, V, W = four_cell_case()
    lib = O.oracle()
    c1 = np.empty(4)
    lib.orc_per_donor(4, 1, 1, O.ptr(U), O.ptr(V), O.ptr(W), O.ptr(CN), O.ptr(c1))
    np.testing.assert_array_equal(c1, C1)  # dyadic values: exact
    a, b, c = np.empty(4), np.empty(4), np.empty(4)
    lib.orc_per_antidiff(4, 1, 1, O.ptr(c1), O.ptr(U), O.ptr(V), O.ptr(W), O.ptr(a),
                         O.ptr(b), O.ptr(c))
    np.testing.assert_array_equal(a, A)
    np.testing.assert_array_equal(b, 0.0)
    np.testing.assert_array_equal(c, 0.0)
    bup, bdn = np.empty(4), np.empty(4)
    lib.orc_per_limiter_beta(4, 1, 1, O.ptr(CN), O.ptr(c1), O.ptr(a), O.ptr(b), O.ptr(c),
                             O.ptr(bup), O.ptr(bdn))
    close(bup, BUP)
    close(bdn, BDN)
    lib.orc_per_limiter_apply(4, 1, 1, O.ptr(bup), O.ptr(bdn), O.ptr(a), O.ptr(b), O.ptr(c))
    close(a, AL)
    # the whole step (donor pass + limited corrective pass) and its unlimited counterpart
    close(O.per_step(4, 1, 1, U, V, W, 2, True, CN), C2, rel=1e-15)
    close(O.per_step(4, 1, 1, U, V, W, 2, False, CN), C2_UNLIMITED, rel=1e-15)


def test_hand_worked_case_mirrored():
    """The same case flowing the other way (u = -1/2 on the reversed line) gives the
    mirrored values: the limiter's two branches ([A]+ and [A]-) and both ratios are
    exercised in the other roles."""
    U = np.full(4, -0.5)
    z = np.zeros(4)
    cn = CN[::-1].copy()
    got = O.per_step(4, 1, 1, U, z, z, 2, True, cn)
    close(got, C2[::-1].copy(), rel=1e-15)


def revolution_error(n, sigma, courant, iord, nonosc):
    """Relative L2 error after one revolution of a Gaussian on a periodic line."""
    x = np.arange(n, dtype=np.float64)
    c = np.exp(-0.5 * ((x - n / 2) / sigma) ** 2)
    U = np.full(n, courant)
    z = np.zeros(n)
    steps = int(round(n / courant))
    out = O.per_run(n, 1, 1, U, z, z, iord, nonosc, c, steps)
    return float(np.linalg.norm(out - c) / np.linalg.norm(c))


def test_limited_scheme_keeps_second_order_accuracy():
    # 64 cells, sigma 4, Courant 0.4, 160 steps: donor 0.44, unlimited 0.111, limited 0.112
    donor = revolution_error(64, 4.0, 0.4, 1, False)
    unlim = revolution_error(64, 4.0, 0.4, 2, False)
    lim = revolution_error(64, 4.0, 0.4, 2, True)
    assert lim <= 1.1 * unlim
    assert lim <= 0.3 * donor
    # refinement (cells and width doubled, same Courant number): second order for both
    # MPDATA variants, well below first order for donor cell
    lim2 = revolution_error(128, 8.0, 0.4, 2, True)
    lim4 = revolution_error(256, 16.0, 0.4, 2, True)
    donor4 = revolution_error(256, 16.0, 0.4, 1, False)
    donor2 = revolution_error(128, 8.0, 0.4, 1, False)
    order_lim = np.log2(lim2 / lim4)
    order_donor = np.log2(donor2 / donor4)
    assert order_lim >= 1.7, order_lim
    assert order_donor <= 1.0, order_donor


@pytest.mark.parametrize("iord", [2, 3])
def test_limiter_is_inactive_on_a_smooth_wide_profile(iord):
    """Where no cell approaches its local bounds (a wide, smooth, strictly positive
    profile, one step) every ratio exceeds 1 and the limited step equals the unlimited
    one bit for bit: the limiter must not touch what does not need limiting."""
    n = 48
    x = np.arange(n, dtype=np.float64)
    c = 2.0 + np.sin(2 * np.pi * x / n) * 0.01
    U = np.full(n, 0.3)
    z = np.zeros(n)
    lim = O.per_step(n, 1, 1, U, z, z, iord, True, c)
    raw = O.per_step(n, 1, 1, U, z, z, iord, False, c)
    # ratios > 1 everywhere except near the two extrema of the sine; compare away from them
    far = np.ones(n, bool)
    for e in (n // 4, 3 * n // 4):
        far[e - 4:e + 5] = False
    np.testing.assert_array_equal(lim[far], raw[far])
