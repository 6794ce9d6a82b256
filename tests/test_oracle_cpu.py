"""Pins the CPU oracle: against the reference's known answers, against the reference's own C KLU
build, and against what the mathematics fixes (dense linear algebra, finite differences,
independent numpy restatements).  Runs without a GPU."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from gridoxide_amd.grid import PQ, PV, SLACK, Network, config_grid, generate_grid, scenario_specs


@pytest.fixture(scope="module")
def known(golden_dir):
    with open(os.path.join(golden_dir, "known_answers.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def three_bus(golden_dir):
    with open(os.path.join(golden_dir, "network.json")) as f:
        return Network.from_json(f.read())


def dense_ybus(net):
    """Independent dense restatement of the pi-model / transformer stamps."""
    n = net.n_bus
    Y = np.zeros((n, n), complex)
    for f, t, r, x, b, g in zip(net.line_from, net.line_to, net.line_r, net.line_x, net.line_b, net.line_g):
        if f == t:
            Y[f, f] += g + 1j * b
            continue
        y = 1 / (r + 1j * x)
        Y[f, f] += y + (g + 1j * b) / 2
        Y[t, t] += y + (g + 1j * b) / 2
        Y[f, t] -= y
        Y[t, f] -= y
    for f, t, ys, ysh, a in zip(net.tr_from, net.tr_to, net.tr_ys, net.tr_ysh, net.tr_tap):
        Y[f, f] += (ys + ysh / 2) / abs(a) ** 2
        Y[f, t] += -ys / np.conj(a)
        Y[t, f] += -ys / a
        Y[t, t] += ys + ysh / 2
    return Y


def test_three_bus_known_answers(oracle_mod, three_bus, known):
    ka = known["three_bus"]
    o = oracle_mod.Oracle(three_bus)
    vm, va, rc = o.linear_initial_guess(three_bus.p_spec, three_bus.q_spec, three_bus.vm, three_bus.va)
    assert rc == 0
    np.testing.assert_allclose(vm, ka["linear_guess_vm"], atol=ka["linear_guess_tol"])
    np.testing.assert_allclose(va, ka["linear_guess_va"], atol=ka["linear_guess_tol"])
    p, q = o.injections(vm, va)
    r, c, v = o.jacobian(vm, va, p, q)
    J = np.zeros((3, 3))
    J[r, c] = v
    np.testing.assert_allclose(J, ka["jacobian_iter1"], atol=ka["printed_tol"])
    f = np.r_[three_bus.p_spec[1:] - p[1:], three_bus.q_spec[2:] - q[2:]]
    np.testing.assert_allclose(f, ka["mismatch_iter1"], atol=ka["printed_tol"] * 1e-2)
    np.testing.assert_allclose(np.linalg.solve(J, f), ka["dx_iter1"], atol=ka["printed_tol"] * 1e-2)
    vmo, vao, st, it, hist = o.batch_solve(three_bus.p_spec, three_bus.q_spec, vm, va, tol=1e-6, max_iter=20)
    assert st[0] == 0
    for i, (m, a) in enumerate(ka["golden_voltages"]):
        assert abs(vmo[0, i] - m) < ka["golden_tol"] and abs(vao[0, i] - a) < ka["golden_tol"]


def test_three_bus_reference_klu_factors(oracle_mod, three_bus, known):
    """The reference's C KLU on the documented Jacobian reproduces klu.md:436-446."""
    if not oracle_mod.ref_available():
        pytest.skip("reference KLU build not present")
    ka = known["three_bus"]
    J = np.array(ka["jacobian_iter1"])
    r, c = np.nonzero(J)
    ap, ai, ax = oracle_mod.triplets_to_csc(3, r.astype(np.int32), c.astype(np.int32), J[r, c])
    k = oracle_mod.KluRef(ap, ai, ax, scale=0)
    ex = k.extract()
    assert list(ex["R"]) == [0, 3] and list(ex["P"]) == [0, 1, 2] and list(ex["Q"]) == [0, 1, 2]
    # U diagonal stored last in each column
    udiag = [ex["Ux"][ex["Up"][j + 1] - 1] for j in range(3)]
    np.testing.assert_allclose(udiag, ka["udiag"], atol=1e-6)
    for j in range(3):
        got = [(int(i), x) for i, x in zip(ex["Li"][ex["Lp"][j] + 1:ex["Lp"][j + 1]], ex["Lx"][ex["Lp"][j] + 1:ex["Lp"][j + 1]])]
        assert [g[0] for g in got] == [e[0] for e in ka["L"][j]]
        np.testing.assert_allclose([g[1] for g in got], [e[1] for e in ka["L"][j]], atol=1e-6)


def test_ybus_matches_dense_restatement(oracle_mod):
    net = config_grid(1)
    rowptr, col, y = oracle_mod.ybus(net)
    Y = dense_ybus(net)
    S = sp.csr_matrix((y, col, rowptr), shape=(net.n_bus, net.n_bus)).toarray()
    np.testing.assert_allclose(S, Y, rtol=1e-14, atol=1e-12)
    assert np.all(np.diff(col[rowptr[0]:rowptr[1]]) > 0)


def test_injections_match_dense(oracle_mod):
    net = config_grid(1)
    o = oracle_mod.Oracle(net)
    rng = np.random.default_rng(0)
    vm = 1 + 0.05 * rng.standard_normal(net.n_bus)
    va = 0.1 * rng.standard_normal(net.n_bus)
    p, q = o.injections(vm, va)
    V = vm * np.exp(1j * va)
    S = V * np.conj(dense_ybus(net) @ V)
    np.testing.assert_allclose(p, S.real, rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(q, S.imag, rtol=1e-12, atol=1e-10)


def test_jacobian_is_derivative_of_mismatch(oracle_mod):
    """The H/N/M/L formulas are the exact derivatives: compare with central differences."""
    net = config_grid(1)
    o = oracle_mod.Oracle(net)
    rng = np.random.default_rng(1)
    vm = 1 + 0.03 * rng.standard_normal(net.n_bus)
    va = 0.05 * rng.standard_normal(net.n_bus)
    ns = np.nonzero(net.bus_type != SLACK)[0]
    pq = np.nonzero(net.bus_type == PQ)[0]
    n = ns.size + pq.size

    def g(x):
        a = va.copy()
        m = vm.copy()
        a[ns] = x[:ns.size]
        m[pq] = x[ns.size:]
        p, q = o.injections(m, a)
        return np.r_[p[ns], q[pq]]

    x0 = np.r_[va[ns], vm[pq]]
    p, q = o.injections(vm, va)
    r, c, v = o.jacobian(vm, va, p, q)
    J = np.zeros((n, n))
    J[r, c] = v
    h = 1e-6
    cols = rng.choice(n, 25, replace=False)
    for j in cols:
        e = np.zeros(n)
        e[j] = h
        fd = (g(x0 + e) - g(x0 - e)) / (2 * h)
        np.testing.assert_allclose(J[:, j], fd, atol=2e-6)


def _xorshift_matrices(trials=40):
    """Random diagonally dominant asymmetric matrices, n = 3..14, like the reference's differential
    tests (src/klu_native/kernel.rs:815-885)."""
    rng = np.random.default_rng(0xA24BAED4)
    for trial in range(trials):
        n = 3 + trial % 12
        A = np.zeros((n, n))
        mask = rng.random((n, n)) < 0.35
        A[mask] = rng.uniform(-1, 1, mask.sum())
        np.fill_diagonal(A, 0.0)
        A += np.diag(np.abs(A).sum(1) + 1 + rng.random(n))
        yield n, A


def test_oracle_lu_vs_dense_and_reference_klu(oracle_mod):
    for n, A in _xorshift_matrices():
        r, c = np.nonzero(A)
        ap, ai, ax = oracle_mod.triplets_to_csc(n, r.astype(np.int32), c.astype(np.int32), A[r, c])
        b = np.arange(1, n + 1, dtype=float)
        x, rc = oracle_mod.sparse_solve(ap, ai, ax, b)
        assert rc == 0
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
        if oracle_mod.ref_available():
            k = oracle_mod.KluRef(ap, ai, ax, scale=0)
            np.testing.assert_allclose(k.solve(b), x, rtol=1e-10, atol=1e-12)


def test_oracle_lu_pivots_off_diagonal(oracle_mod):
    """A tiny diagonal forces partial pivoting; the solve stays exact."""
    A = np.array([[1e-9, 2.0, 0.0], [3.0, 1e-9, 1.0], [0.0, 1.0, 4.0]])
    r, c = np.nonzero(A)
    ap, ai, ax = oracle_mod.triplets_to_csc(3, r.astype(np.int32), c.astype(np.int32), A[r, c])
    b = np.array([1.0, 2.0, 3.0])
    x, rc = oracle_mod.sparse_solve(ap, ai, ax, b)
    assert rc == 0
    np.testing.assert_allclose(A @ x, b, rtol=1e-12)


def test_linear_initial_guess_matches_scipy(oracle_mod):
    net = config_grid(1)
    o = oracle_mod.Oracle(net)
    vm, va, rc = o.linear_initial_guess(net.p_spec, net.q_spec, net.vm, net.va)
    assert rc == 0
    Y = dense_ybus(net)
    pq = np.nonzero(net.bus_type == PQ)[0]
    known_ = np.nonzero(net.bus_type != PQ)[0]
    Yp = Y.copy()
    Yp[pq, pq] += -np.conj(net.p_spec[pq] + 1j * net.q_spec[pq])
    Uk = net.vm[known_] * np.exp(1j * net.va[known_])
    U = spla.spsolve(sp.csc_matrix(Yp[np.ix_(pq, pq)]), -Yp[np.ix_(pq, known_)] @ Uk)
    np.testing.assert_allclose(vm[pq], np.abs(U), rtol=1e-10)
    np.testing.assert_allclose(va[pq], np.angle(U), atol=1e-10)
    assert np.array_equal(vm[known_], net.vm[known_])


def test_newton_converges_quadratically_and_meets_tolerance(oracle_mod):
    net = config_grid(1)
    o = oracle_mod.Oracle(net)
    vm, va, st, it, hist = o.batch_solve(net.p_spec, net.q_spec, net.vm, net.va, tol=1e-8, max_iter=20)
    assert st[0] == 0 and hist[0, it[0] - 1] < 1e-8
    h = hist[0, :it[0]]
    # quadratic convergence: log-ratio roughly doubles once in the basin
    assert h[-1] < h[-2] ** 1.5
    p, q = o.injections(vm[0], va[0])
    ns = net.bus_type != SLACK
    pqm = net.bus_type == PQ
    assert np.abs(net.p_spec[ns] - p[ns]).max() < 1e-8 and np.abs(net.q_spec[pqm] - q[pqm]).max() < 1e-8


def test_batch_thread_invariance(oracle_mod):
    """src: tests/batch_solver_test.rs:128-147."""
    net = config_grid(1)
    p, q = scenario_specs(net, 16, seed=4)
    o = oracle_mod.Oracle(net)
    a = o.batch_solve(p, q, net.vm, net.va, threads=1)
    b = o.batch_solve(p, q, net.vm, net.va, threads=4)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_generator_shape():
    net = generate_grid(500, 250, seed=3)
    assert net.n_branch == 499 + 250
    assert (net.bus_type == SLACK).sum() == 1 and net.bus_type[0] == SLACK
    assert abs((net.bus_type == PV).sum() - 0.15 * 499) <= 1
    ends = np.r_[net.line_from, net.tr_from], np.r_[net.line_to, net.tr_to]
    assert np.all(np.abs(ends[0] - ends[1]) <= 50)
    pairs = set(zip(np.minimum(*ends), np.maximum(*ends)))
    assert len(pairs) == net.n_branch
    assert abs(net.tr_from.size - 0.1 * net.n_branch) <= 1
    assert np.all((np.abs(net.tr_tap) >= 0.95) & (np.abs(net.tr_tap) <= 1.05))
    pv = net.bus_type == PV
    assert abs(net.p_spec[pv].sum() - 0.95 * net.load_p.sum()) < 1e-9
    # JSON round trip
    back = Network.from_json(net.to_json())
    assert np.array_equal(back.tr_tap, net.tr_tap) and np.array_equal(back.line_r, net.line_r)


def test_scenario_model():
    net = config_grid(1)
    p, q = scenario_specs(net, 4, seed=1)
    p2, _ = scenario_specs(net, 8, seed=1)
    assert np.array_equal(p, p2[:4])  # a scenario does not depend on the batch size
    pv = net.bus_type == PV
    pq = net.bus_type == PQ
    for s in range(4):
        assert abs(p[s, pv].sum() - 0.95 * (-p[s, pq]).sum()) < 1e-9
        ratio = q[s, pq] / p[s, pq]
        assert np.allclose(ratio, net.q_spec[pq] / net.p_spec[pq])


def test_failed_first_factorization_is_not_sticky(oracle_mod):
    """src/solver.rs:711-716: when the first Jacobian a solver meets has no valid pivot order, that
    solve is Singular after one mismatch and the cache stays empty; the next scenario factors its own
    first Jacobian.  On one worker the later scenarios therefore equal a batch without the failed one."""
    net = config_grid(1)
    S = 4
    p, q = scenario_specs(net, S, seed=21)
    vm0 = np.repeat(net.vm[None], S, 0)
    va0 = np.repeat(net.va[None], S, 0)
    vm0[0, 17] = np.nan
    o = oracle_mod.Oracle(net)
    vm, va, st, it, hist = o.batch_solve(p, q, vm0, va0, init=0, threads=1)
    assert st[0] == 2 and it[0] == 1
    assert np.all(st[1:] == 0)
    ref = o.batch_solve(p[1:], q[1:], vm0[1:], va0[1:], init=0, threads=1)
    assert np.array_equal(vm[1:], ref[0]) and np.array_equal(va[1:], ref[1]) and np.array_equal(it[1:], ref[3])


def test_oracle_elimination_order_option(oracle_mod):
    """The oracle's first factorization can take a given elimination order (the device's BTF + AMD
    orders): same statuses and iterations as bus-major order, voltages within rounding, and an L+U
    fill no larger than bus-major's."""
    from gridoxide_amd import Topology
    net = config_grid(1)
    S = 6
    p, q = scenario_specs(net, S, seed=4)
    vm0 = np.repeat(net.vm[None], S, 0)
    va0 = np.repeat(net.va[None], S, 0)
    topo = Topology.from_network(net)
    topo.factor_at(net.vm, net.va, net.p_spec, net.q_spec)
    pp, qq, _, _ = topo.ordering()
    o = oracle_mod.Oracle(net)
    a = o.batch_solve(p, q, vm0, va0, init=1)
    b = o.batch_solve(p, q, vm0, va0, init=1, order=(qq, pp))
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    conv = a[2] == 0
    assert conv.all()
    assert np.abs(a[0] - b[0]).max() < 1e-12 and np.abs(a[1] - b[1]).max() < 1e-12
