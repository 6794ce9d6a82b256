"""The product's host symbolic phase (own BTF + AMD + pivoting factor) against the reference's own
C KLU (vendor/suitesparse, compiled into oracle/_ref), bit for bit: symbolic row/column orders
P/Q, block boundaries R, final pivot rows Pnum and the L/U patterns.  Runs without a GPU (the
symbolic phase never touches the device)."""
import json
import os

import numpy as np
import pytest

from gridoxide_amd import Topology, build_ybus, config_grid, generate_grid
from gridoxide_amd import _lib


def _ref(oracle_mod):
    if not oracle_mod.ref_available():
        pytest.skip("reference KLU build (oracle/_ref) not present")


def _sys_orders(n, rows, cols, vals):
    L = _lib.lib()
    r, c, v = _lib.i32(rows), _lib.i32(cols), _lib.f64(vals)
    h = L.gx_sys_create(n, r.size, _lib.ptr(r, _lib._i32p), _lib.ptr(c, _lib._i32p), _lib.ptr(v))
    if not h:
        return None
    try:
        p, q, pn = (np.zeros(n, np.int32) for _ in range(3))
        rr = np.zeros(n + 1, np.int32)
        nb = np.zeros(1, np.int32)
        nnz = L.gx_sys_ordering(h, _lib.ptr(p, _lib._i32p), _lib.ptr(q, _lib._i32p), _lib.ptr(rr, _lib._i32p),
                                _lib.ptr(pn, _lib._i32p), _lib.ptr(nb, _lib._i32p))
        return p, q, rr[:nb[0] + 1], pn, nnz
    finally:
        L.gx_sys_destroy(h)


def _klu_orders(oracle_mod, n, rows, cols, vals):
    ap, ai, ax = oracle_mod.triplets_to_csc(n, rows, cols, vals)
    k = oracle_mod.KluRef(ap, ai, ax, scale=0)
    P, Q, R = k.symbolic()
    ex = k.extract()
    return P, Q, R, ex


def _klu_lu_pattern(ex, n):
    """Per pivot column: sorted strictly-lower L rows and strictly-upper U rows (F included)."""
    lcols, ucols = [], []
    for j in range(n):
        li = ex["Li"][ex["Lp"][j]:ex["Lp"][j + 1]]
        ui = ex["Ui"][ex["Up"][j]:ex["Up"][j + 1]]
        lcols.append(sorted(int(i) for i in li if i != j))
        ucols.append(sorted(int(i) for i in ui if i != j))
    return lcols, ucols


@pytest.mark.parametrize("config", [1, 2, 3])
def test_power_flow_jacobian_symbolic_bit_exact(oracle_mod, config):
    _ref(oracle_mod)
    net = config_grid(config)
    o = oracle_mod.Oracle(net)
    p, q = o.injections(net.vm, net.va)
    r, c, v = o.jacobian(net.vm, net.va, p, q)
    topo = Topology(net.bus_type, build_ybus(net))
    topo.factor_at(net.vm, net.va, net.p_spec, net.q_spec)
    P, Q, R, Pnum = topo.ordering()
    n = topo.sizes()["n"]
    kP, kQ, kR, ex = _klu_orders(oracle_mod, n, r, c, v)
    assert np.array_equal(R, kR)
    assert np.array_equal(P, kP) and np.array_equal(Q, kQ)
    assert np.array_equal(Pnum, ex["P"]) and np.array_equal(Q, ex["Q"])
    if R.size == 2:  # one block: the L/U patterns coincide entry for entry
        cp, ulen, row = topo.lu_pattern()
        lcols, ucols = _klu_lu_pattern(ex, n)
        for k in range(n):
            seg = row[cp[k]:cp[k + 1]]
            assert list(seg[:ulen[k]]) == ucols[k]
            assert list(seg[ulen[k] + 1:]) == lcols[k]


def _random_sparse(rng, n, density, symmetric):
    A = np.zeros((n, n))
    mask = rng.random((n, n)) < density
    if symmetric:
        mask = mask | mask.T
    A[mask] = rng.uniform(-1, 1, mask.sum())
    np.fill_diagonal(A, 0.0)
    A += np.diag(np.abs(A).sum(1) + 1.0)
    return A


@pytest.mark.parametrize("symmetric", [True, False])
def test_random_matrices_orders_bit_exact(oracle_mod, symmetric):
    """AMD tie-breaking, BTF blocks (unsymmetric patterns are reducible) and pivots on 60 random
    matrices of 4..80 unknowns."""
    _ref(oracle_mod)
    rng = np.random.default_rng(12345 + symmetric)
    for trial in range(60):
        n = int(rng.integers(4, 81))
        A = _random_sparse(rng, n, float(rng.uniform(0.02, 0.15)), symmetric)
        r, c = np.nonzero(A)
        mine = _sys_orders(n, r, c, A[r, c])
        assert mine is not None
        P, Q, R, Pnum, _ = mine
        kP, kQ, kR, ex = _klu_orders(oracle_mod, n, r.astype(np.int32), c.astype(np.int32), A[r, c])
        assert np.array_equal(R, kR), f"trial {trial} blocks"
        assert np.array_equal(P, kP) and np.array_equal(Q, kQ), f"trial {trial} orders"
        assert np.array_equal(Pnum, ex["P"]), f"trial {trial} pivots"


def test_off_diagonal_pivots_bit_exact(oracle_mod):
    """Weak diagonals make partial pivoting pick off-diagonal rows; the pivot rows still match."""
    _ref(oracle_mod)
    rng = np.random.default_rng(99)
    hit = 0
    for trial in range(40):
        n = int(rng.integers(5, 40))
        A = _random_sparse(rng, n, 0.2, True)
        d = rng.random(n) < 0.4
        A[np.arange(n)[d], np.arange(n)[d]] *= 1e-5  # weak diagonal -> pivoting
        r, c = np.nonzero(A)
        mine = _sys_orders(n, r, c, A[r, c])
        kP, kQ, kR, ex = _klu_orders(oracle_mod, n, r.astype(np.int32), c.astype(np.int32), A[r, c])
        assert np.array_equal(mine[0], kP) and np.array_equal(mine[1], kQ)
        assert np.array_equal(mine[3], ex["P"]), f"trial {trial}"
        hit += int(not np.array_equal(ex["P"], kP))
    assert hit > 0  # the fixture really exercised off-diagonal pivots


def test_doc_btf_example(golden_dir):
    """docs/src/solvers/klu.md:336-373: three BTF blocks, p = [0,2,3,1], q = [1,3,0,2]."""
    with open(os.path.join(golden_dir, "known_answers.json")) as f:
        ex = json.load(f)["btf_4x4"]
    e = np.array(ex["before"])
    P, Q, R, Pnum, _ = _sys_orders(4, e[:, 0].astype(int), e[:, 1].astype(int), e[:, 2])
    assert R.size - 1 == ex["btf_blocks"]
    assert list(P) == ex["p"] and list(Q) == ex["q"]


def test_generator_grids_of_other_sizes(oracle_mod):
    """Fresh grids (not the benchmark seeds) at 50..3000 buses: orders still identical."""
    _ref(oracle_mod)
    for nb, ch, seed in [(50, 10, 1), (300, 150, 2), (1200, 600, 3), (3000, 1500, 4)]:
        net = generate_grid(nb, ch, seed)
        o = oracle_mod.Oracle(net)
        p, q = o.injections(net.vm, net.va)
        r, c, v = o.jacobian(net.vm, net.va, p, q)
        topo = Topology(net.bus_type, build_ybus(net))
        topo.factor_at(net.vm, net.va, net.p_spec, net.q_spec)
        P, Q, R, Pnum = topo.ordering()
        kP, kQ, kR, ex = _klu_orders(oracle_mod, topo.sizes()["n"], r, c, v)
        assert np.array_equal(P, kP) and np.array_equal(Q, kQ) and np.array_equal(Pnum, ex["P"]), nb


def test_external_pivot_order_reproduces_pattern():
    """gx_topo_set_pivot_order with the order found by factor_at yields the same L+U pattern."""
    net = config_grid(1)
    a = Topology.from_network(net)
    a.factor_at(net.vm, net.va, net.p_spec, net.q_spec)
    P, Q, R, Pnum = a.ordering()
    b = Topology.from_network(net)
    b.set_pivot_order(Q, Pnum)
    for x, y in zip(a.lu_pattern(), b.lu_pattern()):
        assert np.array_equal(x, y)


def test_dense_rows_and_larger_matrices_bit_exact(oracle_mod):
    """The ordering's dense-row rule (rows of A + A' with more than max(16, 10 sqrt n) neighbours
    are left out and ordered last), mass elimination and supervariables on larger patterns: 2-D
    grid Laplacians (many indistinguishable neighbourhoods), grids with a few dense rows and
    columns, and random patterns of 300-1,500 unknowns, against the reference's C KLU."""
    _ref(oracle_mod)
    rng = np.random.default_rng(2024)
    cases = []
    for side in (12, 25):
        n = side * side
        A = np.zeros((n, n))
        for i in range(side):
            for j in range(side):
                k = i * side + j
                if i + 1 < side:
                    A[k, k + side] = A[k + side, k] = -1.0
                if j + 1 < side:
                    A[k, k + 1] = A[k + 1, k] = -1.0
        A += np.diag(np.abs(A).sum(1) + 1.0)
        cases.append(A)
        D = A.copy()
        for d in rng.choice(n, 3, replace=False):  # dense row and column
            D[d, :] = D[:, d] = rng.uniform(-1, 1, n)
            D[d, d] = n
        cases.append(D)
    for n in (300, 800, 1500):
        cases.append(_random_sparse(rng, n, 3.0 / n, bool(n % 2)))
    for t, A in enumerate(cases):
        n = A.shape[0]
        r, c = np.nonzero(A)
        P, Q, R, Pnum, _ = _sys_orders(n, r, c, A[r, c])
        kP, kQ, kR, ex = _klu_orders(oracle_mod, n, r.astype(np.int32), c.astype(np.int32), A[r, c])
        assert np.array_equal(R, kR), f"case {t} blocks"
        assert np.array_equal(P, kP) and np.array_equal(Q, kQ), f"case {t} orders"
        assert np.array_equal(Pnum, ex["P"]), f"case {t} pivots"
