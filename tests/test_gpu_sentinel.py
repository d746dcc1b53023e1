"""Step sentinel (mpd_set_sentinel / mpd_sentinel_check): a device check of the field
every k steps, with no host synchronisation in between.  The conditions are planted in
the input data from the host (a NaN, negative values), so every kernel runs normally."""
import numpy as np
import pytest

import mpdata_amd as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if M.device_count() < 1:
        pytest.skip("no GPU")


def case(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    U, V, W = (rng.uniform(-0.15, 0.15, n) for _ in range(3))
    return U, V, W, rng.uniform(0.0, 1.0, n) ** 2


def grid(dims, nonosc, U, V, W, c):
    g = M.Grid(*dims, M.PERIODIC, iord=2, nonosc=nonosc)
    g.set_courant(U, V, W)
    g.set_field(c)
    return g


def test_sentinel_quiet_on_a_healthy_run_and_changes_nothing():
    dims = (24, 20, 16)
    U, V, W, c = case(*dims, 1)
    ref = grid(dims, True, U, V, W, c)
    ref.step(12)
    want = ref.field()
    ref.close()
    g = grid(dims, True, U, V, W, c)
    g.set_sentinel(1, nonnegative=True)
    g.step(12)
    assert g.sentinel_check() is None
    np.testing.assert_array_equal(g.field(), want)
    g.close()


def test_sentinel_reports_the_first_step_with_a_nan():
    dims = (24, 20, 16)
    U, V, W, c = case(*dims, 2)
    c = c.copy()
    c[1234] = np.nan
    g = grid(dims, True, U, V, W, c)
    g.set_sentinel(1)
    g.step(3)
    assert g.sentinel_check() == 1
    # re-arming on a healthy field clears it
    g.set_field(case(*dims, 3)[3])
    g.set_sentinel(2)
    g.step(6)
    assert g.sentinel_check() is None
    g.close()


def test_sentinel_negative_values_only_when_asked():
    dims = (20, 18, 12)
    U, V, W, c = case(*dims, 4)
    c = c - 0.5  # a signed field (the unlimited scheme transports it as it is)
    for nonneg, want in ((True, 2), (False, None)):
        g = grid(dims, False, U, V, W, c)
        g.set_sentinel(2, nonnegative=nonneg)
        g.step(5)
        assert g.sentinel_check() == want
        g.close()


def test_sentinel_on_a_frame_grid_and_with_graph_replay_requested():
    nx, ny, nz = 13, 11, 9
    rng = np.random.default_rng(5)
    n = nx * ny * nz
    g = M.Grid(nx, ny, nz, M.FRAME, iord=3, nonosc=False)
    one = [np.ones(nx), np.ones(ny), np.ones(nz)]
    winds = [rng.uniform(-0.3, 0.3, n) for _ in range(6)]
    g.set_winds(*one, 0.9, *winds)
    c = rng.uniform(0.0, 1.0, n)
    c[77] = np.inf
    g.set_field(c)
    g.set_sentinel(1)
    g.step(2)
    assert g.sentinel_check() == 1
    g.close()
    dims = (16, 16, 16)
    U, V, W, c = case(*dims, 6)
    g = grid(dims, True, U, V, W, c)
    g.set_graph(True)  # replay is bypassed while the sentinel is armed
    g.set_sentinel(4, nonnegative=True)
    g.step(16)
    assert g.sentinel_check() is None
    g.close()


def test_autotune_trial_steps_do_not_count():
    """The autotuner's trial steps neither run the sentinel nor advance its step count:
    armed every 4 steps, 3 steps, a tuning (many trial steps), a NaN planted from the
    host, 1 step -> flagged after step 4."""
    dims = (36, 20, 11)  # a shape no other test tunes
    U, V, W, c = case(*dims, 7)
    g = grid(dims, True, U, V, W, c)
    g.set_path(2)  # the two-kernel path: the one with launch shapes to tune
    g.set_sentinel(4, nonnegative=True)
    g.step(3)
    before = g.field()
    g.autotune(1)
    np.testing.assert_array_equal(g.field(), before)
    assert g.sentinel_check() is None
    bad = before.copy()
    bad[123] = np.nan
    g.set_field(bad)
    g.step(1)
    assert g.sentinel_check() == 4
    g.close()
