"""GPU: PowerFlowModel over PGM input and ZIP injections, through the C ABI.

Every golden PGM fixture of the reference (tests/klu_native_jacobian_test.rs, tests/pgm_*_test.rs)
solved on the device against the fixture's expected node voltages at the reference's tolerances, and
against the CPU oracle on the same converted network (status equal, iterations equal for FP64 and at
most +1 for mixed precision, |V| and theta within 1e-8).
"""
import json
import os

import numpy as np
import pytest

from gridoxide_amd import (BatchSolver, BusOverride, Mode, PersistentSolver, Scenario, SolveStatus, Topology,
                           generate_grid)
from gridoxide_amd import pgm
from gridoxide_amd.grid import PQ, PV
from gridoxide_amd.model import PowerFlowModel

pytestmark = pytest.mark.gpu

V_TOL = 1e-8
SINGLE = [("basic-node", 1e-5), ("node", 1e-5), ("line", 1e-5), ("shunt", 1e-5), ("source", 1e-5),
          ("sym_load", 1e-5), ("sym_gen", 1e-5), ("transmission-case", 1e-5)]
BATCH = [("transformer", 1e-4), ("distribution-case", 1e-5), ("three_winding_transformer", 1e-4)]


def _say(msg):
    print(msg, flush=True)


def _path(golden_dir, case, name="input.json"):
    return os.path.join(golden_dir, "pgm", case, name)


def _load(golden_dir, case, name):
    with open(_path(golden_dir, case, name)) as f:
        return json.load(f)


def _check_golden(model, expected_nodes, tol):
    vm, va = np.array(model.voltage_mag()), np.array(model.voltage_ang())
    for nd in expected_nodes:
        i = model.node_index(nd["id"])
        assert abs(vm[i] - nd["u_pu"]) < tol, (nd["id"], vm[i], nd["u_pu"])
        assert abs(va[i] - nd["u_angle"]) < tol, (nd["id"], va[i], nd["u_angle"])


def _check_oracle(oracle_mod, model, pm, mode):
    """FP64: the oracle's iterates, so within 1e-8.  Mixed precision may stop one iteration later
    than the oracle at this 1e-6 mismatch tolerance, so its voltages are held to 1e-6."""
    net = pm.network
    rvm, rva, rst, rit, _ = oracle_mod.Oracle(net).persistent_solve(net.p_spec, net.q_spec, net.vm, net.va,
                                                                     tol=1e-6, max_iter=20)
    assert rst == 0
    tol = V_TOL if mode == Mode.FP64 else 1e-6
    vm, va = np.array(model.voltage_mag()), np.array(model.voltage_ang())
    assert np.abs(vm - rvm).max() <= tol and np.abs(va - rva).max() <= tol


@pytest.mark.parametrize("mode", [Mode.FP64, Mode.MIXED])
@pytest.mark.parametrize("case,tol", SINGLE)
def test_model_golden_single(oracle_mod, golden_dir, case, tol, mode):
    _say(f"pgm {case} mode={mode.name}")
    model = PowerFlowModel.from_pgm_json(_path(golden_dir, case), mode=mode)
    model.solve()
    _check_golden(model, _load(golden_dir, case, "sym_output.json")["data"]["node"], tol)
    _check_oracle(oracle_mod, model, model._model, mode)


@pytest.mark.parametrize("case,tol", BATCH)
def test_model_golden_batch(oracle_mod, golden_dir, case, tol):
    doc = _load(golden_dir, case, "input.json")
    updates = _load(golden_dir, case, "update_batch.json")["data"]
    expected = _load(golden_dir, case, "sym_output_batch.json")["data"]
    _say(f"pgm batch {case}: {len(updates)} scenarios")
    for upd, exp in zip(updates, expected):
        model = PowerFlowModel.from_pgm(pgm.apply_batch_update(doc, upd))
        model.solve()
        _check_golden(model, exp["node"], tol)
        _check_oracle(oracle_mod, model, model._model, Mode.FP64)


def test_solve_twice_and_reset(golden_dir):
    """Repeated solves reuse the cached analysis and give the same voltages; reset() re-analyses."""
    model = PowerFlowModel.from_pgm_json(_path(golden_dir, "transmission-case"))
    model.solve()
    a = model.voltage_mag()
    model.solve()
    assert model.voltage_mag() == a
    model.reset()
    model.solve()
    assert np.abs(np.array(model.voltage_mag()) - np.array(a)).max() <= 1e-12
    kv = model.voltage_kv()
    assert np.allclose(kv, np.array(a) * model._model.u_rated / 1e3, rtol=0, atol=0)


def test_solve_raises_when_not_converged(golden_dir):
    model = PowerFlowModel.from_pgm_json(_path(golden_dir, "transmission-case"), max_iter=1)
    with pytest.raises(RuntimeError, match="did not converge within 1 iterations"):
        model.solve()


@pytest.mark.parametrize("case", ["transmission-case", "sym_load"])
def test_solve_batch_scaled_vs_oracle(oracle_mod, golden_dir, case):
    """solve_batch_scaled == the oracle on each scaled scenario (uniform_load_scaling leaves the
    ZIP terms unscaled, as the reference's)."""
    model = PowerFlowModel.from_pgm_json(_path(golden_dir, case))
    scales = [0.5, 0.8, 1.0, 1.3, 1.6]
    res = model.solve_batch_scaled(scales)
    assert len(res) == len(scales)
    net = model._model.network
    o = oracle_mod.Oracle(net)
    for f, r in zip(scales, res):
        p = np.where((net.p_spec != 0) | (net.q_spec != 0), net.p_spec * f, net.p_spec)
        q = np.where((net.p_spec != 0) | (net.q_spec != 0), net.q_spec * f, net.q_spec)
        vm, va, st, it, hist = o.persistent_solve(p, q, net.vm, net.va, tol=1e-6, max_iter=20)
        assert r.converged == (st == 0) and r.converged
        assert r.iterations == it
        assert abs(r.max_mismatch - hist[it - 1]) <= 1e-9 and r.max_mismatch < 1e-6
        assert np.abs(np.array(r.voltage_mag) - vm).max() <= V_TOL
        assert np.abs(np.array(r.voltage_ang) - va).max() <= V_TOL


def test_solve_batch_overrides_and_errors(oracle_mod, golden_dir):
    model = PowerFlowModel.from_pgm_json(_path(golden_dir, "transmission-case"))
    kinds, p, q = model.bus_spec()
    pq = [i for i, k in enumerate(kinds) if k == 2 and p[i] != 0.0]
    scen = [[], [(pq[0], 1.5 * p[pq[0]], 1.5 * q[pq[0]])], [(b, 0.7 * p[b], 0.7 * q[b]) for b in pq]]
    res = model.solve_batch(scen, threads=8)
    net = model._model.network
    o = oracle_mod.Oracle(net)
    for ov, r in zip(scen, res):
        pp, qq = net.p_spec.copy(), net.q_spec.copy()
        for b, a, c in ov:
            pp[b], qq[b] = a, c
        vm, va, st, it, _ = o.persistent_solve(pp, qq, net.vm, net.va, tol=1e-6, max_iter=20)
        assert r.converged and st == 0 and r.iterations == it
        assert np.abs(np.array(r.voltage_mag) - vm).max() <= V_TOL
    with pytest.raises(RuntimeError, match="overrides bus 99"):
        model.solve_batch([[(99, 0.0, 0.0)]])
    assert model.solve_batch([]) == []


def test_initial_guess_and_ybus_vs_oracle(oracle_mod, golden_dir):
    model = PowerFlowModel.from_pgm_json(_path(golden_dir, "transmission-case"))
    net = model._model.network
    o = oracle_mod.Oracle(net)
    gvm, gva, rc = o.linear_initial_guess(net.p_spec, net.q_spec, net.vm, net.va)
    assert rc == 0
    vm, va = model.initial_guess()
    assert np.abs(np.array(vm) - gvm).max() <= 1e-12 and np.abs(np.array(va) - gva).max() <= 1e-12
    rows, cols, g, b = model.ybus_triplets()
    orows = np.repeat(np.arange(net.n_bus), np.diff(o.rowptr))
    assert rows == orows.tolist() and cols == o.col.tolist()
    assert np.array_equal(np.array(g), o.y.real) and np.array_equal(np.array(b), o.y.imag)


@pytest.mark.parametrize("mode", [Mode.FP64, Mode.MIXED])
def test_zip_batch_vs_oracle(oracle_mod, mode):
    """ZIP terms of every kind on a meshed 300-bus grid, 40 scenarios through BatchSolver (linear
    initial guess, device mismatch with effective_injection) against the oracle."""
    net = generate_grid(300, 150, 11)
    rng = np.random.default_rng(5)
    z = np.zeros((net.n_bus, 6))
    loaded = np.nonzero(net.bus_type == PQ)[0]
    z[loaded] = rng.uniform(-0.02, 0.0, (loaded.size, 6))
    pv = np.nonzero(net.bus_type == PV)[0]
    z[pv, 0] = rng.uniform(0.0, 0.01, pv.size)  # constant-power terms on generator buses too
    net.zip_coef = z
    S = 40
    scen = [Scenario([BusOverride(int(i)).p(net.p_spec[i] * (0.5 + 0.01 * s)).q(net.q_spec[i] * (0.5 + 0.01 * s))
                      for i in loaded]) for s in range(S)]
    _say(f"zip batch mode={mode.name}")
    reports = BatchSolver(mode=mode).solve(net, scen, 1e-8, 20)
    p, q, vm0, va0 = BatchSolver.scenario_arrays(net, scen)
    rvm, rva, rst, rit, _ = oracle_mod.Oracle(net).batch_solve(p, q, vm0, va0, tol=1e-8, max_iter=20, init=1,
                                                               threads=4)
    st = np.array([int(r.stats.status) for r in reports])
    it = np.array([r.stats.iterations() for r in reports])
    assert np.array_equal(st, rst) and (rst == 0).all()
    if mode == Mode.FP64:
        assert np.array_equal(it, rit)
    else:
        assert (it >= rit).all() and (it <= rit + 1).all()
    vm = np.array([r.voltage_mag for r in reports])
    va = np.array([r.voltage_ang for r in reports])
    assert np.abs(vm - rvm).max() <= V_TOL and np.abs(va - rva).max() <= V_TOL


def test_zip_persistent_solver_and_removal(oracle_mod):
    """PersistentSolver picks the ZIP terms up from the network; clearing them on the topology
    returns the constant-power answer."""
    net = generate_grid(80, 40, 2)
    z = np.zeros((net.n_bus, 6))
    z[net.bus_type == PQ, 4] = -0.05
    net.zip_coef = z
    work = net.copy()
    stats = PersistentSolver().solve(work, 1e-8, 20)
    assert stats.status == SolveStatus.Converged
    rvm, rva, rst, rit, _ = oracle_mod.Oracle(net).batch_solve(net.p_spec, net.q_spec, net.vm, net.va, tol=1e-8,
                                                               max_iter=20, init=0)
    assert rst[0] == 0 and stats.iterations() == rit[0]
    assert np.abs(work.vm - rvm[0]).max() <= V_TOL
    topo = Topology.from_network(net)
    topo.set_zip(None)
    vm, va, st, it, _ = topo.batch_solve(net.p_spec[None], net.q_spec[None], net.vm[None], net.va[None], 1e-8, 20)
    bare = net.copy()
    bare.zip_coef = None
    bvm, bva, bst, bit, _ = oracle_mod.Oracle(bare).batch_solve(net.p_spec, net.q_spec, net.vm, net.va, tol=1e-8,
                                                                max_iter=20, init=0)
    assert st[0] == bst[0] == 0 and np.abs(vm[0] - bvm[0]).max() <= V_TOL
    assert np.abs(vm[0] - rvm[0]).max() > 1e-4
