"""GPU parity: the HIP path (through the C ABI) against the CPU oracle and the reference's
known answers.

Bars: identical convergence status; identical iteration count for
the FP64 path and at most +1 for mixed precision; converged |V| and theta within 1e-8 pu / rad of
the FP64 oracle on the same inputs.
"""
import json
import os

import numpy as np
import pytest

from gridoxide_amd import (BatchSolver, BusOverride, Mode, Network, PersistentSolver, Scenario, SolveStatus,
                           Topology, config_grid, scenario_specs)
from gridoxide_amd import _lib

pytestmark = pytest.mark.gpu

V_TOL = 1e-8      # pu and rad, against the FP64 oracle
MIS_TOL = 1e-8    # convergence tolerance used by the workloads


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def known(golden_dir):
    with open(os.path.join(golden_dir, "known_answers.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def three_bus(golden_dir):
    with open(os.path.join(golden_dir, "network.json")) as f:
        return Network.from_json(f.read())


def _oracle_batch(oracle_mod, net, p, q, vm0, va0, tol=MIS_TOL, max_iter=20):
    o = oracle_mod.Oracle(net)
    return o.batch_solve(p, q, vm0, va0, tol=tol, max_iter=max_iter, init=0, threads=4)


def _compare(gpu, ref, mixed=False, check_hist=True):
    vm, va, st, it, hist = gpu
    rvm, rva, rst, rit, rhist = ref
    assert np.array_equal(st, rst), f"status differs: {np.nonzero(st != rst)[0][:10]}"
    if mixed:
        assert np.all(it >= rit) and np.all(it <= rit + 1), f"iterations {it} vs {rit}"
    else:
        assert np.array_equal(it, rit), f"iterations differ at {np.nonzero(it != rit)[0][:10]}"
    conv = st == 0
    if conv.any():
        dvm = np.abs(vm[conv] - rvm[conv]).max()
        dva = np.abs(va[conv] - rva[conv]).max()
        _say(f"  max|dVm|={dvm:.3e} max|dVa|={dva:.3e}")
        assert dvm <= V_TOL and dva <= V_TOL
    if check_hist and not mixed:
        # a diverging Newton run amplifies rounding differences chaotically, so the mismatch
        # history is compared where the iteration converged
        for s in np.nonzero(conv)[0]:
            n = it[s]
            np.testing.assert_allclose(hist[s, :n], rhist[s, :n], rtol=1e-6, atol=1e-12)


def test_three_bus_golden_voltages(three_bus, known):
    _say("case: 3-bus network.json, flat start, FP64")
    net = three_bus
    topo = Topology.from_network(net)
    vm, va, st, it, hist = topo.batch_solve(net.p_spec[None], net.q_spec[None], net.vm[None], net.va[None],
                                            tol=1e-6, max_iter=20)
    assert st[0] == 0
    for i, (m, a) in enumerate(known["three_bus"]["golden_voltages"]):
        assert abs(vm[0, i] - m) < known["three_bus"]["golden_tol"]
        assert abs(va[0, i] - a) < known["three_bus"]["golden_tol"]


def test_three_bus_linear_start_matches_doc_step(three_bus, known, oracle_mod):
    """From the documented linear initial guess the first mismatch and step match klu.md:420-447."""
    _say("case: 3-bus from the documented linear initial guess")
    net = three_bus
    o = oracle_mod.Oracle(net)
    vm0, va0, rc = o.linear_initial_guess(net.p_spec, net.q_spec, net.vm, net.va)
    assert rc == 0
    topo = Topology.from_network(net)
    vm, va, st, it, hist = topo.batch_solve(net.p_spec[None], net.q_spec[None], vm0[None], va0[None],
                                            tol=1e-6, max_iter=1)
    ka = known["three_bus"]
    assert abs(hist[0, 0] - max(abs(x) for x in ka["mismatch_iter1"])) < ka["printed_tol"]
    dva = va[0] - va0
    dvm = vm[0] - vm0
    assert abs(dva[1] - ka["dx_iter1"][0]) < ka["printed_tol"]
    assert abs(dva[2] - ka["dx_iter1"][1]) < ka["printed_tol"]
    assert abs(dvm[2] - ka["dx_iter1"][2]) < ka["printed_tol"]


# launch shapes: (scenarios per tile, solve ring rows); 0 = automatic (2-scenario tiles at these
# batch sizes).  A 64-row ring gathers every triangular-solve entry that reaches farther than 31 rows.
SHAPES_C1 = [(0, 0), (4, 0), (8, 0), (16, 64), (64, 0), (64, 64)]
SHAPES_C2 = [(0, 0), (2, 64), (16, 0), (4, 64), (8, 0), (64, 0)]


@pytest.mark.parametrize("shape", SHAPES_C1)
@pytest.mark.parametrize("mode", [Mode.FP64, Mode.MIXED])
def test_config1_batch_vs_oracle(oracle_mod, mode, shape):
    _say(f"case: config-1 grid (200 buses), 70 scenarios, flat start, mode={mode.name}, tile/ring={shape}")
    net = config_grid(1)
    p, q = scenario_specs(net, 70, seed=7)
    vm0 = np.repeat(net.vm[None], 70, 0)
    va0 = np.repeat(net.va[None], 70, 0)
    topo = Topology.from_network(net)
    topo.set_launch(*shape)
    gpu = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20, mode=mode)
    ref = _oracle_batch(oracle_mod, net, p, q, vm0, va0)
    assert np.all(ref[2] == 0)
    _compare(gpu, ref, mixed=(mode == Mode.MIXED))


@pytest.mark.parametrize("shape", SHAPES_C2)
@pytest.mark.parametrize("mode", [Mode.FP64, Mode.MIXED])
def test_config2_subset_vs_oracle(oracle_mod, mode, shape):
    _say(f"case: config-2 grid (2000 buses), 128 scenarios, flat start, mode={mode.name}, tile/ring={shape}")
    net = config_grid(2)
    p, q = scenario_specs(net, 128, seed=11)
    vm0 = np.repeat(net.vm[None], 128, 0)
    va0 = np.repeat(net.va[None], 128, 0)
    topo = Topology.from_network(net)
    topo.set_launch(*shape)
    gpu = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20, mode=mode)
    ref = _oracle_batch(oracle_mod, net, p, q, vm0, va0)
    _compare(gpu, ref, mixed=(mode == Mode.MIXED))


def test_divergent_scenario_isolated():
    """src: tests/batch_solver_test.rs:153-183 -- an unsolvable scenario does not perturb others."""
    _say("case: config-1 grid, scenario 4 infeasible (P=Q=-1e6)")
    net = config_grid(1)
    p, q = scenario_specs(net, 8, seed=3)
    vm0 = np.repeat(net.vm[None], 8, 0)
    va0 = np.repeat(net.va[None], 8, 0)
    topo = Topology.from_network(net)
    clean = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20)
    bad_p, bad_q = p.copy(), q.copy()
    bad_p[4, 5] = -1e6
    bad_q[4, 5] = -1e6
    got = topo.batch_solve(bad_p, bad_q, vm0, va0, tol=MIS_TOL, max_iter=20)
    assert got[2][4] != 0
    keep = [k for k in range(8) if k != 4]
    assert np.all(got[2][keep] == 0)
    assert np.array_equal(got[0][keep], clean[0][keep]) and np.array_equal(got[1][keep], clean[1][keep])


def test_nan_start_is_singular_and_isolated():
    """A non-finite solve is reported as Singular (src/klu_native/mod.rs:192, src/solver.rs:717-720)."""
    _say("case: config-1 grid, scenario 2 starts with a NaN voltage")
    net = config_grid(1)
    p, q = scenario_specs(net, 5, seed=5)
    vm0 = np.repeat(net.vm[None], 5, 0)
    va0 = np.repeat(net.va[None], 5, 0)
    vm0[2, 17] = np.nan
    topo = Topology.from_network(net)
    vm, va, st, it, hist = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20)
    assert st[2] == 2 and it[2] == 1
    assert np.all(st[[0, 1, 3, 4]] == 0)


def test_singular_first_jacobian_factors_the_next_scenario(oracle_mod):
    """src/solver.rs:711-716: the first Jacobian of the batch (scenario 0, NaN start) has no valid
    pivot order; scenario 0 is Singular after one mismatch and the pivot order comes from scenario 1,
    so scenarios 1.. equal, bit for bit, a batch that starts at scenario 1, and the oracle (whose
    failed first factorization is not sticky either) within 1e-8."""
    _say("case: config-1 grid, scenario 0 starts with a NaN voltage (first factorization fails)")
    net = config_grid(1)
    S = 5
    p, q = scenario_specs(net, S, seed=5)
    vm0 = np.repeat(net.vm[None], S, 0)
    va0 = np.repeat(net.va[None], S, 0)
    vm0[0, 17] = np.nan
    for mode in (Mode.FP64, Mode.MIXED):
        topo = Topology.from_network(net)
        vm, va, st, it, hist = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20, mode=mode)
        assert st[0] == 2 and it[0] == 1 and np.all(st[1:] == 0), (mode, st, it)
        rest = Topology.from_network(net).batch_solve(p[1:], q[1:], vm0[1:], va0[1:], tol=MIS_TOL, max_iter=20,
                                                      mode=mode)
        assert np.array_equal(vm[1:], rest[0]) and np.array_equal(va[1:], rest[1])
        assert np.array_equal(it[1:], rest[3])
        ref = oracle_mod.Oracle(net).batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20, init=0, threads=1)
        _compare((vm[1:], va[1:], st[1:], it[1:], hist[1:]), tuple(a[1:] for a in ref), mixed=mode == Mode.MIXED)
        assert ref[2][0] == 2 and ref[3][0] == 1


def test_no_valid_first_jacobian_returns_starting_points():
    """Every scenario's first Jacobian is singular (NaN starts): no pivot order exists, every scenario
    is Singular after its first mismatch and keeps its starting point; nothing is an error."""
    _say("case: config-1 grid, every scenario starts with a NaN voltage")
    net = config_grid(1)
    S = 3
    p, q = scenario_specs(net, S, seed=6)
    vm0 = np.repeat(net.vm[None], S, 0)
    va0 = np.repeat(net.va[None], S, 0)
    vm0[:, 17] = np.nan
    topo = Topology.from_network(net)
    vm, va, st, it, hist = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20)
    assert np.all(st == 2) and np.all(it == 1)
    assert np.array_equal(va, va0) and np.array_equal(np.isnan(vm), np.isnan(vm0))
    assert np.array_equal(vm[~np.isnan(vm)], vm0[~np.isnan(vm0)])
    # a later batch with valid starts factors normally
    ok = topo.batch_solve(p, q, np.repeat(net.vm[None], S, 0), va0, tol=MIS_TOL, max_iter=20)
    assert np.all(ok[2] == 0)


def test_batch_equals_sequential_persistent_solver():
    """src: tests/batch_solver_test.rs:68-97 -- batch == sequential solves, bit for bit."""
    _say("case: config-1 grid, batch of 6 vs 6 sequential single-scenario solves")
    net = config_grid(1)
    scen = [Scenario([BusOverride(int(i)).p(float(net.p_spec[i] * f)).q(float(net.q_spec[i] * f))
                      for i in np.nonzero(net.p_spec != 0)[0]]) for f in (0.6, 0.8, 0.9, 1.0, 1.1, 1.2)]
    topo = Topology.from_network(net)
    reports = BatchSolver().solve(net, scen, MIS_TOL, 20, topology=topo)
    p, q, vm0, va0 = BatchSolver.scenario_arrays(net, scen)
    for s in range(len(scen)):
        one = topo.batch_solve(p[s:s + 1], q[s:s + 1], vm0[s:s + 1], va0[s:s + 1], tol=MIS_TOL, max_iter=20,
                               init="linear")
        assert reports[s].stats.status == SolveStatus.Converged
        assert np.array_equal(reports[s].voltage_mag, one[0][0])
        assert np.array_equal(reports[s].voltage_ang, one[1][0])


def test_persistent_solver_warm_equals_fresh(three_bus):
    """src: tests/persistent_solver_test.rs:24-65 -- a warm solve matches a fresh one."""
    _say("case: 3-bus persistent solver warm vs fresh")
    ps = PersistentSolver()
    a = three_bus.copy()
    ps.solve(a, 1e-6, 20)
    b = three_bus.copy()
    b.p_spec[2] *= 1.5
    b.q_spec[2] *= 1.5
    ps.solve(b, 1e-6, 20)
    c = three_bus.copy()
    c.p_spec[2] *= 1.5
    c.q_spec[2] *= 1.5
    PersistentSolver().solve(c, 1e-6, 20)
    assert np.abs(b.vm - c.vm).max() < 1e-9 and np.abs(b.va - c.va).max() < 1e-9


def test_device_pointer_path_matches_host_path():
    import torch
    _say("case: config-1 grid, 100 scenarios, device-resident inputs")
    net = config_grid(1)
    S = 100
    p, q = scenario_specs(net, S, seed=9)
    vm0 = np.repeat(net.vm[None], S, 0)
    va0 = np.repeat(net.va[None], S, 0)
    topo = Topology.from_network(net)
    host = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20)
    dev = torch.device("cuda:0")
    tp, tq, tvm0, tva0 = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (p, q, vm0, va0))
    vm = torch.empty_like(tvm0)
    va = torch.empty_like(tva0)
    st = torch.empty(S, dtype=torch.int32, device=dev)
    it = torch.empty(S, dtype=torch.int32, device=dev)
    hist = torch.empty((S, 20), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    topo.batch_solve_device(tp, tq, tvm0, tva0, vm, va, st, it, hist, tol=MIS_TOL, max_iter=20)
    assert np.array_equal(vm.cpu().numpy(), host[0]) and np.array_equal(va.cpu().numpy(), host[1])
    assert np.array_equal(st.cpu().numpy(), host[2]) and np.array_equal(it.cpu().numpy(), host[3])


def _sys_solve(entries, n, vals_list, rhs):
    L = _lib.lib()
    r = np.array([e[0] for e in entries], np.int32)
    c = np.array([e[1] for e in entries], np.int32)
    v0 = np.array([e[2] for e in entries], np.float64)
    h = L.gx_sys_create(n, len(entries), _lib.ptr(r, _lib._i32p), _lib.ptr(c, _lib._i32p), _lib.ptr(v0))
    assert h
    out = []
    try:
        for vals in vals_list:
            x = np.zeros(n)
            rc = L.gx_sys_factor_and_solve_values(h, _lib.ptr(np.asarray(vals, np.float64)),
                                                  _lib.ptr(np.asarray(rhs, np.float64)), _lib.ptr(x))
            out.append((rc, x))
    finally:
        L.gx_sys_destroy(h)
    return out


def test_sys_btf_example_refactor(known):
    """docs/src/solvers/klu.md:336-405: factor, then refactor with new values on the same pattern."""
    _say("case: gx_sys on the documented 4x4 BTF example")
    ex = known["btf_4x4"]
    before = ex["before"]
    after = ex["after"]
    res = _sys_solve(before, 4, [[e[2] for e in before], [e[2] for e in after]], ex["b"])
    assert res[0][0] == 0 and res[1][0] == 0
    np.testing.assert_allclose(res[0][1], ex["x_before"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(res[1][1], ex["x_after"], rtol=0, atol=1e-14)


def test_sys_refactor_new_values_and_zero_pivot(known):
    """src/klu_native/kernel.rs:741-812: refactor matches a dense solve; a new zero pivot is reported."""
    _say("case: gx_sys refactor with new values, then a new zero pivot")
    ex = known["refactor_new_values"]
    n = ex["n"]
    b = np.arange(1, n + 1, dtype=np.float64)
    res = _sys_solve(ex["original"], n, [[e[2] for e in ex["updated"]]], b)
    A = np.zeros((n, n))
    for r, c, v in ex["updated"]:
        A[r, c] += v
    assert res[0][0] == 0
    np.testing.assert_allclose(res[0][1], np.linalg.solve(A, b), rtol=1e-13, atol=1e-14)
    ez = known["refactor_new_zero_pivot"]
    res = _sys_solve(ez["entries"], 2, [[e[2] for e in ez["singular"]]], [1.0, 1.0])
    assert res[0][0] == _lib.GX_SINGULAR


def test_islands_marked_and_reported(oracle_mod):
    """src/network.rs:121-223, src/solver.rs:234-312: an island without a Slack bus is pinned to
    V = 0; an island with two is solved and reported Ambiguous; the main island is unaffected."""
    _say("case: config-1 grid plus an unreferenced and an ambiguous island")
    from test_abi_cpu import _with_islands
    from gridoxide_amd import IslandStatus
    net, n0 = _with_islands()
    plain = config_grid(1)
    rep = BatchSolver(init="given").solve(net, [Scenario()], MIS_TOL, 20)[0]
    ref = _oracle_batch(oracle_mod, plain, plain.p_spec[None], plain.q_spec[None], plain.vm[None], plain.va[None])
    assert rep.stats.status == SolveStatus.Converged
    assert np.abs(rep.voltage_mag[:n0] - ref[0][0]).max() <= V_TOL
    assert np.abs(rep.voltage_ang[:n0] - ref[1][0]).max() <= V_TOL
    assert np.all(rep.voltage_mag[n0:n0 + 3] == 0) and np.all(rep.voltage_ang[n0:n0 + 3] == 0)
    assert [r.status for r in rep.islands] == [IslandStatus.Converged, IslandStatus.NoReferenceBus,
                                               IslandStatus.AmbiguousReferenceBus]
    assert rep.islands[1].slack_indices == [] and len(rep.islands[2].slack_indices) == 2


def test_linear_initial_guess_three_bus_doc_values(three_bus, known):
    """docs/src/solvers/klu.md:420-422: the linear guess of network.json (max_iter 0 returns it)."""
    _say("case: 3-bus device linear initial guess")
    net = three_bus
    topo = Topology.from_network(net)
    vm, va, st, it, _ = topo.batch_solve(net.p_spec[None], net.q_spec[None], net.vm[None], net.va[None],
                                         tol=1e-6, max_iter=0, init="linear")
    ka = known["three_bus"]
    assert st[0] == 1 and it[0] == 0
    np.testing.assert_allclose(vm[0], ka["linear_guess_vm"], atol=ka["linear_guess_tol"])
    np.testing.assert_allclose(va[0], ka["linear_guess_va"], atol=ka["linear_guess_tol"])


@pytest.mark.parametrize("mode", [Mode.FP64, Mode.MIXED])
def test_linear_initial_guess_batch_vs_oracle(oracle_mod, mode):
    """BatchSolver semantics (src/batch.rs:268): linear guess per scenario, then Newton."""
    _say(f"case: config-1 grid, 40 scenarios, linear initial guess, mode={mode.name}")
    net = config_grid(1)
    p, q = scenario_specs(net, 40, seed=21)
    vm0 = np.repeat(net.vm[None], 40, 0)
    va0 = np.repeat(net.va[None], 40, 0)
    topo = Topology.from_network(net)
    # the guess alone
    g = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=0, init="linear")
    o = oracle_mod.Oracle(net)
    for s in (0, 17, 39):
        gvm, gva, rc = o.linear_initial_guess(p[s], q[s], vm0[s], va0[s])
        assert rc == 0
        assert np.abs(g[0][s] - gvm).max() < 1e-12 and np.abs(g[1][s] - gva).max() < 1e-12
    gpu = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20, mode=mode, init="linear")
    ref = o.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20, init=1, threads=4)
    _compare(gpu, ref, mixed=(mode == Mode.MIXED))


@pytest.mark.parametrize("tile", [0, 8])
def test_mixed_fallback_reproduces_fp64_for_failing_scenarios(tile):
    """A scenario whose FP32 factorization fails is re-solved with the FP64 path: its reported
    result is then exactly the FP64 mode's, and the others keep their mixed-precision results.
    Scenario 3 starts with |V| = 1e20 at one PQ bus: its Jacobian entries (~1e41) overflow FP32,
    so the FP32 factor is non-finite (Singular), while FP64 represents them.  At 8 scenarios per
    tile the mixed mode runs pair tiles (and its FP64 re-solve pairs of doubles), the FP64 mode
    8-lane groups: the same bits."""
    _say(f"case: config-1 grid, 16 scenarios, scenario 3 overflows FP32, mixed vs FP64, tile {tile}")
    net = config_grid(1)
    p, q = scenario_specs(net, 16, seed=13)
    vm0 = np.repeat(net.vm[None], 16, 0)
    va0 = np.repeat(net.va[None], 16, 0)
    bus = int(np.nonzero(net.bus_type == 2)[0][0])
    vm0[3, bus] = 1e20
    topo = Topology.from_network(net)
    topo.set_launch(tile, 0)
    mixed = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20, mode=Mode.MIXED)
    assert topo.fallbacks() >= 1
    fp64 = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20, mode=Mode.FP64)
    assert mixed[2][3] == fp64[2][3] != 0 and mixed[3][3] == fp64[3][3]
    assert np.array_equal(mixed[0][3], fp64[0][3]) and np.array_equal(mixed[1][3], fp64[1][3])
    ok = [s for s in range(16) if s != 3]
    assert np.all(mixed[2][ok] == 0)
    assert np.abs(mixed[0][ok] - fp64[0][ok]).max() <= V_TOL


@pytest.mark.parametrize("groups", [16, 32])
def test_pair_tiles_equal_four_scenario_tiles(groups):
    """Mixed precision at 8 scenarios per tile runs pair tiles: 4-lane groups with two scenarios
    per lane, on the schedule of 4-scenario tiles; with 32 refactor groups (quad tiles) the refactor
    runs 2-lane groups with four scenarios per lane on the schedule of 2-scenario tiles.  Every
    entry takes its updates in the same order, so the results equal the 4-scenario tiles' bit for
    bit, including a scenario whose FP32 factor is singular (its lane partners are unaffected) and
    the padding of a last, partial tile.  The 32-group refactor also splits partner columns (partial
    visits written back undivided, finished in their own parts), which the 16-group one does not."""
    _say(f"case: config-2 grid (2000 buses), 70 scenarios, mixed, tile 8 ({groups} refactor groups) vs tile 4")
    net = config_grid(2)
    p, q = scenario_specs(net, 70, seed=31)
    vm0 = np.repeat(net.vm[None], 70, 0)
    va0 = np.repeat(net.va[None], 70, 0)
    vm0[5, int(np.nonzero(net.bus_type == 2)[0][0])] = np.nan
    got = {}
    for tile in (8, 4):
        topo = Topology.from_network(net)
        topo.set_launch(tile, 0)
        topo.set_refactor_groups(groups)
        got[tile] = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20, mode=Mode.MIXED)
    for a, b in zip(got[8][:4], got[4][:4]):
        assert np.array_equal(a, b, equal_nan=True)
    assert got[8][2][5] == 2 and (got[8][2] == 0).sum() > 0
    # the lane partner of the singular scenario (scenario 4) converged like its tile-4 self
    assert got[8][2][4] == got[4][2][4]


def test_mixed_fallback_with_unreferenced_island_equals_fp64():
    """The FP64 re-solve of a mixed-precision fallback starts from the point the FP64 path starts
    from: buses of an island without a Slack bus are pinned to V = 0 in both."""
    _say("case: config-1 grid plus islands, scenario 1 overflows FP32, mixed vs FP64")
    from test_abi_cpu import _with_islands
    net, n0 = _with_islands()
    S = 4
    p = np.repeat(net.p_spec[None], S, 0)
    q = np.repeat(net.q_spec[None], S, 0)
    vm0 = np.repeat(net.vm[None], S, 0)
    va0 = np.repeat(net.va[None], S, 0)
    vm0[:, n0:n0 + 3] = 1.05  # the unreferenced island starts away from 0
    bus = int(np.nonzero(net.bus_type[:n0] == 2)[0][0])
    vm0[1, bus] = 1e20
    topo = Topology.from_network(net)
    mixed = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20, mode=Mode.MIXED)
    assert topo.fallbacks() >= 1
    fp64 = Topology.from_network(net).batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20, mode=Mode.FP64)
    assert mixed[2][1] == fp64[2][1] and mixed[3][1] == fp64[3][1]
    assert np.array_equal(mixed[0][1], fp64[0][1]) and np.array_equal(mixed[1][1], fp64[1][1])
    assert np.all(mixed[0][:, n0:n0 + 3] == 0) and np.all(fp64[0][:, n0:n0 + 3] == 0)


@pytest.mark.parametrize("shape", [(0, 0), (64, 0)])
@pytest.mark.parametrize("mode", [Mode.FP64, Mode.MIXED])
def test_n1_outages_vs_oracle_on_reduced_networks(oracle_mod, mode, shape):
    """Half the scenarios take one random non-bridge branch out of service on the shared pattern;
    each is compared with the oracle solving the network without that branch."""
    from gridoxide_amd import without_branch
    _say(f"case: config-1 grid, 48 scenarios, half N-1, mode={mode.name}, tile/ring={shape}")
    net = config_grid(1)
    S = 48
    p, q = scenario_specs(net, S, seed=21)
    vm0 = np.repeat(net.vm[None], S, 0)
    va0 = np.repeat(net.va[None], S, 0)
    topo = Topology.from_network(net)
    topo.set_launch(*shape)
    n_br = net.line_from.size + net.tr_from.size
    rng = np.random.default_rng(5)
    candidates = [e for e in range(n_br) if not topo.branch_is_bridge(e)]
    outage = np.full(S, -1, np.int32)
    outage[1::2] = rng.choice(candidates, S // 2)
    gpu = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20, mode=mode, outage=outage)
    ref = [np.zeros_like(gpu[0]), np.zeros_like(gpu[1]), np.zeros(S, np.int32), np.zeros(S, np.int32),
           np.zeros_like(gpu[4])]
    for e in np.unique(outage):
        idx = np.nonzero(outage == e)[0]
        rnet = net if e < 0 else without_branch(net, int(e))
        r = _oracle_batch(oracle_mod, rnet, p[idx], q[idx], vm0[idx], va0[idx])
        for a in range(5):
            ref[a][idx] = r[a]
    assert np.all(ref[2] == 0)
    _compare(gpu, tuple(ref), mixed=(mode == Mode.MIXED))
    # an outage changes the answer: the N-1 scenarios differ from their intact counterparts
    base = topo.batch_solve(p, q, vm0, va0, tol=MIS_TOL, max_iter=20, mode=mode)
    assert np.abs(gpu[0][1::2] - base[0][1::2]).max() > 1e-6
    assert np.array_equal(gpu[0][0::2], base[0][0::2])
