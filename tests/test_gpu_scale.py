"""GPU parity at the sizes of the benchmark workloads: the 10,000-bus grid of the mixed-precision
workload, the uneven multi-topology job list through the request engine with captured hipGraph
steps, and the largest grid of the serving pool.

The FP64 oracle factors with the device's BTF + AMD elimination order (the order KLU would use), so
that it finishes these sizes in seconds.  Bars: identical statuses; identical iteration counts for
FP64 and at most one more for mixed precision; converged |V| and theta within 1e-8 of the oracle.
The generator's own loads make the 10,000-bus grid infeasible (every scenario diverges, in the
oracle too); there the first mismatches are compared, and converged results are checked on the same
scenarios at 2% and 3% of their loads (all converge; a mix converges).
"""
import numpy as np
import pytest

from gridoxide_amd import Engine, Mode, Topology, config_grid, scenario_loads
from gridoxide_amd.grid import without_branch
from gridoxide_amd.workloads import serving_pool, uneven_batch

pytestmark = pytest.mark.gpu

V_TOL = 1e-8


def _say(msg):
    print(msg, flush=True)


def _order(topo):
    p, q, _, _ = topo.ordering()
    return q, p


def _check(got, ref, mixed, where="", rerun=None):
    """rerun(k, updates): the oracle's (vm, va) for scenario k after exactly `updates` Newton
    updates from the same start.  Where mixed precision took one iteration more than the oracle, its
    voltages are compared with the oracle's after the same number of updates: two iterates on
    either side of the mismatch tolerance can differ by more than 1e-8 on an ill-conditioned grid."""
    vm, va, st, it, _ = got
    rvm, rva, rst, rit, _ = ref
    assert np.array_equal(st, rst), f"{where} status {st} vs {rst}"
    if mixed:
        assert np.all((it >= rit) & (it <= rit + 1)), f"{where} iterations {it} vs {rit}"
    else:
        assert np.array_equal(it, rit), f"{where} iterations {it} vs {rit}"
    rvm, rva = rvm.copy(), rva.copy()
    if rerun is not None:
        for k in np.nonzero((st == 0) & (it == rit + 1))[0]:
            rvm[k], rva[k] = rerun(int(k), int(it[k]) - 1)
    conv = st == 0
    if conv.any():
        dvm = np.abs(vm[conv] - rvm[conv]).max()
        dva = np.abs(va[conv] - rva[conv]).max()
        _say(f"  {where} {int(conv.sum())} converged: max|dVm|={dvm:.2e} max|dVa|={dva:.2e}")
        assert dvm <= V_TOL and dva <= V_TOL, where
    return int(conv.sum())


@pytest.fixture(scope="module")
def grid10k():
    net = config_grid(3)
    topo = Topology.from_network(net)
    topo.factor_at(net.vm, net.va, net.p_spec, net.q_spec)
    S = 16
    p, q = scenario_loads(net, S, seed=2026)  # the benchmark's first 16 scenarios
    vm0 = np.repeat(net.vm[None], S, 0)
    va0 = np.repeat(net.va[None], S, 0)
    return net, topo, p, q, vm0, va0


@pytest.mark.parametrize("mode", [Mode.FP64, Mode.MIXED])
def test_10k_bus_workload_loads_vs_oracle(oracle_mod, grid10k, mode):
    """The benchmark's own scenarios: every one diverges on both sides; statuses and iteration
    counts agree, and so does the first mismatch (before any step) and, in FP64, the second."""
    net, topo, p, q, vm0, va0 = grid10k
    _say(f"case: 10,000-bus grid, 16 scenarios of the workload, {mode.name}, 4-scenario tiles")
    topo.set_launch(4, 0)
    got = topo.batch_solve(p, q, vm0, va0, tol=1e-8, max_iter=20, mode=mode)
    ref = oracle_mod.Oracle(net).batch_solve(p, q, vm0, va0, tol=1e-8, max_iter=20, threads=8, order=_order(topo))
    assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3])
    np.testing.assert_allclose(got[4][:, 0], ref[4][:, 0], rtol=1e-12)
    if mode == Mode.FP64:
        np.testing.assert_allclose(got[4][:, 1], ref[4][:, 1], rtol=1e-6)
    topo.set_launch(0, 0)


@pytest.mark.parametrize("tile", [0, 4])
@pytest.mark.parametrize("mode", [Mode.FP64, Mode.MIXED])
def test_10k_bus_converging_loads_vs_oracle(oracle_mod, grid10k, mode, tile):
    """The same scenarios at 2% of their loads (all converge) and at 3% (some diverge)."""
    net, topo, p, q, vm0, va0 = grid10k
    topo.set_launch(tile, 0)
    for scale in (0.02, 0.03):
        _say(f"case: 10,000-bus grid, 16 scenarios at {scale:.0%} loads, {mode.name}, tile {tile or 'auto'}")
        got = topo.batch_solve(scale * p, scale * q, vm0, va0, tol=1e-8, max_iter=20, mode=mode)
        ref = oracle_mod.Oracle(net).batch_solve(scale * p, scale * q, vm0, va0, tol=1e-8, max_iter=20, threads=8,
                                                 order=_order(topo))
        n = _check(got, ref, mode == Mode.MIXED, f"{scale:.0%}")
        if scale == 0.02:
            assert n == 16
        else:
            assert 0 < n < 16
    topo.set_launch(0, 0)


def test_10k_bus_full_launch_shape_mixed(oracle_mod, grid10k):
    """8,192 scenarios at 2% of the workload's loads in one mixed-precision batch (the benchmark's
    launch: pair tiles of 8 FP32 scenarios, 1,024 wavefronts, the LDS staging and ring sizes of the
    full grid): 16 scenarios spread over the batch equal the oracle, and equal, bit for bit, the same
    scenarios solved as a 16-scenario batch of 4-scenario tiles."""
    net, topo, _, _, _, _ = grid10k
    S = 8192
    p, q = scenario_loads(net, S, seed=2026)
    p *= 0.02
    q *= 0.02
    vm0 = np.repeat(net.vm[None], S, 0)
    va0 = np.repeat(net.va[None], S, 0)
    _say("case: 10,000-bus grid, 8,192 scenarios at 2% loads, mixed, automatic tile")
    got = topo.batch_solve(p, q, vm0, va0, tol=1e-8, max_iter=20, mode=Mode.MIXED)
    assert np.all(got[2] == 0)
    pick = np.linspace(0, S - 1, 16).astype(int)
    ref = oracle_mod.Oracle(net).batch_solve(p[pick], q[pick], vm0[pick], va0[pick], tol=1e-8, max_iter=20,
                                             threads=8, order=_order(topo))
    _check(tuple(a[pick] for a in got), ref, True, "8192-batch sample")
    topo.set_launch(4, 0)
    alone = topo.batch_solve(p[pick], q[pick], vm0[pick], va0[pick], tol=1e-8, max_iter=20, mode=Mode.MIXED)
    topo.set_launch(0, 0)
    for a, b in zip(alone[:4], got[:4]):
        assert np.array_equal(a, b[pick])


def test_uneven_job_list_engine_vs_oracle(oracle_mod):
    """The uneven multi-topology workload (64 topologies of 103-14,737 buses, half of each one's
    scenarios N-1) in one request engine with captured hipGraph steps, mixed precision: four
    scenarios of every request (two intact, one N-1 where present) against the oracle, N-1 ones on
    the network without that branch."""
    jobs = uneven_batch()
    topos = [Topology.from_network(j.net) for j in jobs]
    for topo, j in zip(topos, jobs):
        topo.factor_at(j.vm0[0], j.va0[0], j.p[0], j.q[0])
        topo.set_launch(4, 0)
    _say(f"case: uneven job list, {len(jobs)} topologies, {sum(j.p.shape[0] for j in jobs)} scenarios, "
         f"engine with hipGraph steps, mixed")
    eng = Engine(tile=4, graphs=True)
    ids = [eng.submit(topo, j.p, j.q, j.vm0, j.va0, max_iter=20, mode=Mode.MIXED, outage=j.outage)
           for topo, j in zip(topos, jobs)]
    assert eng.run() == 0
    results = [eng.collect(i) for i in ids]
    eng.close()
    conv = 0
    for k, (topo, j, res) in enumerate(zip(topos, jobs, results)):
        intact = np.nonzero(j.outage < 0)[0][:2]
        n1 = np.nonzero(j.outage >= 0)[0][:1]
        ref = oracle_mod.Oracle(j.net).batch_solve(j.p[intact], j.q[intact], j.vm0[intact], j.va0[intact], tol=1e-8,
                                                   max_iter=20, threads=8, order=_order(topo))
        conv += _check(tuple(a[intact] for a in res), ref, True, f"request {k} ({j.net.n_bus} buses) intact")
        for s in n1:
            red = without_branch(j.net, int(j.outage[s]))
            r1 = oracle_mod.Oracle(red).batch_solve(j.p[s:s + 1], j.q[s:s + 1], j.vm0[s:s + 1], j.va0[s:s + 1],
                                                    tol=1e-8, max_iter=20, threads=1, order=_order(topo))
            conv += _check(tuple(a[s:s + 1] for a in res), r1, True, f"request {k} N-1 scenario {s}")
    assert conv > 0


@pytest.mark.parametrize("mode", [Mode.FP64, Mode.MIXED])
def test_largest_serving_grid_request_vs_oracle(oracle_mod, mode):
    """One request of the serving workload's kind on the largest grid of the pool (49,785 buses,
    92,100 unknowns, 1.84 M L+U entries) through the request engine: no scenario converges there (on
    either side: grids of the pool above ~14,000 buses diverge from the flat start at any load), so
    statuses, iteration counts and the first mismatch are compared, and in FP64 the second."""
    pool = serving_pool()
    net = max(pool, key=lambda g: g.n_bus)
    assert net.n_bus > 40000
    topo = Topology.from_network(net)
    topo.factor_at(net.vm, net.va, net.p_spec, net.q_spec)
    S = 8
    p, q = scenario_loads(net, S, seed=77)
    vm0 = np.repeat(net.vm[None], S, 0)
    va0 = np.repeat(net.va[None], S, 0)
    _say(f"case: serving pool's largest grid ({net.n_bus} buses), {S} scenarios, engine, {mode.name}")
    eng = Engine(tile=4, graphs=True)
    rid = eng.submit(topo, p, q, vm0, va0, max_iter=20, mode=mode)
    eng.run()
    got = eng.collect(rid)
    eng.close()
    ref = oracle_mod.Oracle(net).batch_solve(p, q, vm0, va0, tol=1e-8, max_iter=20, threads=8, order=_order(topo))
    assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3])
    np.testing.assert_allclose(got[4][:, 0], ref[4][:, 0], rtol=1e-12)
    if mode == Mode.FP64:
        np.testing.assert_allclose(got[4][:, 1], ref[4][:, 1], rtol=1e-6)


def _pool_grid_that_converges():
    """The largest grid of the serving pool whose scenarios converge at 0.2% of their loads."""
    return max((g for g in serving_pool() if g.n_bus < 10000), key=lambda g: g.n_bus)


def test_warm_started_request_equals_oracle_from_the_same_start(oracle_mod):
    """Device-resident warm starts: request 1 leaves its first converged solution on the device;
    request 2 (init="warm") starts every scenario from it.  Request 2 equals, bit for bit, the same
    request started explicitly from that solution (the first converged scenario of request 1, as
    collected), and the oracle from the same start within 1e-8; the batch path's warm start agrees
    with the engine's.  Before any solution is kept, "warm" starts from vm0/va0."""
    net = _pool_grid_that_converges()
    topo = Topology.from_network(net)
    topo.factor_at(net.vm, net.va, net.p_spec, net.q_spec)
    S = 4
    p1, q1 = scenario_loads(net, S, seed=81)
    p2, q2 = scenario_loads(net, S, seed=82)
    p1, q1, p2, q2 = 0.002 * p1, 0.002 * q1, 0.002 * p2, 0.002 * q2
    vm0 = np.repeat(net.vm[None], S, 0)
    va0 = np.repeat(net.va[None], S, 0)
    _say(f"case: {net.n_bus}-bus pool grid, warm start kept on the device by request 1, used by request 2")
    eng = Engine(tile=4, graphs=True)
    cold = eng.submit(topo, p2, q2, vm0, va0, max_iter=20, mode=Mode.MIXED, init="warm")  # nothing kept yet
    eng.run()
    cold = eng.collect(cold)
    flat = eng.submit(topo, p2, q2, vm0, va0, max_iter=20, mode=Mode.MIXED)
    eng.run()
    flat = eng.collect(flat)
    for a, b in zip(cold, flat):
        assert np.array_equal(a, b)
    r1 = eng.submit(topo, p1, q1, vm0, va0, max_iter=20, mode=Mode.MIXED, keep_warm=True)
    eng.run()
    r1 = eng.collect(r1)
    s = int(np.nonzero(r1[2] == 0)[0][0])
    wvm = np.repeat(r1[0][s][None], S, 0)
    wva = np.repeat(r1[1][s][None], S, 0)
    r2 = eng.submit(topo, p2, q2, vm0, va0, max_iter=20, mode=Mode.MIXED, init="warm")
    eng.run()
    r2 = eng.collect(r2)
    r3 = eng.submit(topo, p2, q2, wvm, wva, max_iter=20, mode=Mode.MIXED)
    eng.run()
    r3 = eng.collect(r3)
    eng.close()
    for a, b in zip(r2, r3):
        assert np.array_equal(a, b)
    assert np.all(r2[2] == 0) and r2[3].mean() < flat[3].mean()
    o = oracle_mod.Oracle(net)
    ref = o.batch_solve(p2, q2, wvm, wva, tol=1e-8, max_iter=20, threads=4, order=_order(topo))

    def rerun(k, updates):
        r = o.batch_solve(p2[k:k + 1], q2[k:k + 1], wvm[k:k + 1], wva[k:k + 1], tol=0.0, max_iter=updates,
                          order=_order(topo))
        return r[0][0], r[1][0]

    _say(f"  iterations device {r2[3]} oracle {ref[3]}")
    _check(r2, ref, True, "warm-started request", rerun)
    b = topo.batch_solve(p2, q2, vm0, va0, tol=1e-8, max_iter=20, mode=Mode.MIXED, init="warm")
    topo.set_launch(4, 0)
    b4 = topo.batch_solve(p2, q2, vm0, va0, tol=1e-8, max_iter=20, mode=Mode.MIXED, init="warm")
    for x, y in zip(b4, r2):
        assert np.array_equal(x, y)
    assert np.all(b[2] == 0)
