"""The batched-ensemble and source-inversion configurations on the GPU.

* Ensemble (192 periodic instances drawn from {64^3, 96^3, 128^3, 160x160x96,
  192x192x64}, seed 4, iord 2 nonoscillatory, ~322 M cells): the hipGraph-replayed
  batched step (one kernel per pass) and the per-instance launch sequences (two kernels
  per pass) give the same fields -- bit for bit in exact arithmetic, to 1e-12 in fast
  arithmetic -- for every instance; so do the first instance of each size stepped alone
  and, for the 64^3 instance, the oracle; mass is conserved to rounding and no instance
  goes negative.
* Source inversion at reduced size (32^3, the inversion generators: deformational wind
  and field of seed 5, 8 emission sources, 64 seeded receptors sampled every 16 steps
  against a twin run with perturbed source amplitudes; 64 steps, checkpoints every 32):
  J, dJ/dc0 and dJ/dq against the oracle's full-sweep gradient to 1e-11, with and
  without kept segment records.
* Source inversion at full size (256^3): one adjoint step of a state of the run against
  the oracle's adjoint (<= 1e-12 normwise) and the dot-product test against the oracle's
  tangent-linear model in both controls, <TLM(dc, dq), lam> = <dc, ADJ lam> + <dq, lq>
  to 1e-12.
"""
import numpy as np
import pytest

import mpdata_amd as M
import mpdata_amd.configs as K
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if M.device_count() < 1:
        pytest.skip("no GPU")


def relerr(got, want):
    return float(np.abs(got - want).max() / np.abs(want).max())


@pytest.mark.parametrize("arith", [M.ARITH_EXACT, M.ARITH_FAST])
def test_ensemble_192_graph_equals_per_instance_single_grids_and_oracle(arith):
    """The batched graph takes one kernel per pass at this size (automatic), the
    per-instance launches two (each instance is small): in exact arithmetic both equal the
    oracle and each other bit for bit; in fast arithmetic they agree to 1e-12."""
    n, nsteps = 192, 5
    dims = K.ensemble_dims(n, seed=4)
    b = K.ensemble(n, seed=4)
    b.set_arith(arith)
    assert b.path() == 3
    print(f"case ensemble {n} instances, {b.cells()} cells, {nsteps} steps, arith {arith}",
          flush=True)

    def same(a, c):
        if arith == M.ARITH_EXACT:
            return bool(np.array_equal(a, c))
        return relerr(a, c) <= 1e-12

    start = [g.field() for g in b.instances]
    mass0 = [c.sum() for c in start]
    firsts = sorted({d: i for i, d in reversed(list(enumerate(dims)))}.values())
    courant = {i: b.instances[i].get_courant() for i in firsts}
    b.step(nsteps, M.BATCH_GRAPH)
    graph = [g.field() for g in b.instances]
    print("case ensemble: per-instance launches", flush=True)
    for g, c in zip(b.instances, start):
        g.set_field(c)
    b.step(nsteps, M.BATCH_PER_INSTANCE)
    for i, g in enumerate(b.instances):
        got = g.field()
        if not same(got, graph[i]):
            raise AssertionError(f"instance {i} {dims[i]}: graph and per-instance differ "
                                 f"(rel {relerr(got, graph[i]):.3e})")
        assert abs(got.sum() - mass0[i]) <= 1e-12 * mass0[i], f"instance {i} mass"
        assert got.min() >= 0.0, f"instance {i} negative"
    b.close()
    smallest = min(firsts, key=lambda i: dims[i][0] * dims[i][1] * dims[i][2])
    for i in firsts:
        print(f"case ensemble: instance {i} {dims[i]} against a single grid", flush=True)
        U, V, W = courant[i]
        g = M.Grid(*dims[i], M.PERIODIC, iord=2, nonosc=True)
        g.set_arith(arith)
        g.set_courant(U, V, W)
        g.set_field(start[i])
        g.step(nsteps)
        assert same(g.field(), graph[i]), f"instance {i} against a single grid"
        g.close()
        if i == smallest:
            want = O.per_run(*dims[i], U, V, W, 2, True, start[i], nsteps)
            assert same(graph[i], want)


def inversion_case(m, nsteps, every):
    rec, factors = K.inversion_receptors(m ** 3)
    twin = K.inversion_grid(m, factors=factors)
    obs = twin.sample_receptors(nsteps, rec, every)
    twin.close()
    return rec, obs


@pytest.mark.parametrize("nonosc", [True, False])
def test_inversion_gradient_matches_oracle_at_32cubed(nonosc):
    m, nsteps, every, ck = 32, 64, 16, 32
    rec, obs = inversion_case(m, nsteps, every)
    g = K.inversion_grid(m, nonosc=nonosc)
    U, V, W = g.get_courant()
    c0, q = g.field(), g.emission()
    assert q.min() >= 0.0 and q.max() > 0.0
    Jc, g0c, gqc = O.receptor_loss_and_gradient(m, m, m, U, V, W, nonosc, c0, q, nsteps, rec,
                                                every, obs)
    assert Jc > 0.0
    for tape in (0, -1):
        print(f"case inversion 32^3 nonosc={nonosc} tape={tape}", flush=True)
        g.set_field(c0)
        g.set_adjoint_tape(tape)
        J, g0, gq, _ = g.receptor_gradient(nsteps, ck, rec, every, obs)
        assert abs(J - Jc) <= 1e-13 * Jc
        assert relerr(g0, g0c) <= 1e-11
        assert relerr(gq, gqc) <= 1e-11
    g.close()


def test_inversion_full_size_adjoint_step_and_dot_product():
    n = 256
    g = K.inversion_grid(n)
    U, V, W = g.get_courant()
    q = g.emission()
    print("case inversion 256^3: 8 steps to a state of the run", flush=True)
    g.step(8)
    c = g.field()
    rng = np.random.default_rng(256)
    lam = rng.standard_normal(c.size)
    lq = np.zeros(c.size)
    print("case inversion 256^3: GPU adjoint step", flush=True)
    lin = g.adjoint_step(c, lam, lq)
    g.close()
    np.testing.assert_array_equal(lq, lam)  # the emission enters the step as the identity
    print("case inversion 256^3: oracle adjoint step", flush=True)
    want = O.adj_step(n, n, n, U, V, W, True, c, lam)
    assert relerr(lin, want) <= 1e-12
    del want
    print("case inversion 256^3: oracle tangent-linear step", flush=True)
    dc = rng.standard_normal(c.size) * c
    dq = rng.standard_normal(c.size) * q
    t = O.tlm_step(n, n, n, U, V, W, True, c, dc, dq)
    lhs = float(t @ lam)
    rhs = float(dc @ lin) + float(dq @ lq)
    print(f"dot-product test: <TLM d, lam> {lhs:.15e}, <d, ADJ lam> {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
