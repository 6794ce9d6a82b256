"""The request engine (many topologies in flight, one persistent LU launch per step) against the
one-batch path: identical results for the same inputs, tile width and pivot order, in step mode
and with captured hipGraphs; a request's result does not depend on what else is in flight."""
import numpy as np
import pytest

from gridoxide_amd import Engine, Mode, Topology, config_grid, generate_grid, scenario_loads, scenario_specs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _requests(tile=4):
    """Three topologies of different sizes; one with N-1 scenarios."""
    out = []
    for k, (nb, S, seed) in enumerate(((200, 40, 1), (500, 24, 2), (120, 70, 3))):
        net = config_grid(1) if k == 0 else generate_grid(nb, nb // 2, seed=seed)
        p, q = (scenario_specs if k != 1 else scenario_loads)(net, S, seed=10 + k)
        vm0 = np.repeat(net.vm[None], S, 0)
        va0 = np.repeat(net.va[None], S, 0)
        topo = Topology.from_network(net)
        topo.set_launch(tile, 0)
        outage = None
        if k == 0:
            n_br = net.line_from.size + net.tr_from.size
            cand = [e for e in range(n_br) if not topo.branch_is_bridge(e)]
            outage = np.full(S, -1, np.int32)
            outage[::3] = np.random.default_rng(1).choice(cand, len(outage[::3]))
        out.append((net, topo, p, q, vm0, va0, outage))
    return out


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("tile", [4, 8])
@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("mode", [Mode.FP64, Mode.MIXED])
def test_engine_equals_batch_solve(mode, graphs, tile):
    """Tile 8 runs pair tiles (and quad-tile FP32 refactors) in both paths."""
    _say(f"case: engine, 3 topologies, mode={mode.name}, graphs={graphs}, tile={tile}")
    reqs = _requests(tile)
    # the one-batch path first (it also fixes each topology's pivot order)
    ref = [topo.batch_solve(p, q, vm0, va0, tol=1e-8, max_iter=20, mode=mode, outage=o)
           for _, topo, p, q, vm0, va0, o in reqs]
    eng = Engine(tile=tile, graphs=graphs)
    ids = [eng.submit(topo, p, q, vm0, va0, tol=1e-8, max_iter=20, mode=mode, outage=o)
           for _, topo, p, q, vm0, va0, o in reqs]
    assert eng.run() == 0
    for rid, r in zip(ids, ref):
        got = eng.collect(rid)
        assert got is not None
        _same(got, r)
    assert eng.bytes_in_use() == 0


def test_engine_result_alone_equals_result_in_the_mix():
    """Continuous batching: a request admitted while others are mid-solve gets the same result as
    when it runs alone."""
    _say("case: engine alone vs mixed admission")
    reqs = _requests()
    net, topo, p, q, vm0, va0, o = reqs[0]
    alone = Engine(tile=4)
    a = alone.submit(topo, p, q, vm0, va0, mode=Mode.MIXED, outage=o)
    alone.run()
    res_alone = alone.collect(a)
    mix = Engine(tile=4)
    others = [mix.submit(t2, p2, q2, v2, a2, mode=Mode.MIXED, outage=o2) for _, t2, p2, q2, v2, a2, o2 in reqs[1:]]
    for _ in range(3):
        mix.step()
    b = mix.submit(topo, p, q, vm0, va0, mode=Mode.MIXED, outage=o)
    c = mix.submit(topo, p[::-1].copy(), q[::-1].copy(), vm0, va0, mode=Mode.MIXED)
    mix.run()
    _same(mix.collect(b), res_alone)
    assert mix.collect(c) is not None
    for rid in others:
        assert mix.collect(rid) is not None


def test_engine_budget_admission():
    _say("case: engine budget admission")
    net, topo, p, q, vm0, va0, _ = _requests()[2]
    need = Engine.request_bytes(topo, p.shape[0], 20, Mode.MIXED)
    eng = Engine(tile=4, budget_bytes=int(need * 1.5))
    first = eng.submit(topo, p, q, vm0, va0, mode=Mode.MIXED)
    assert first is not None
    assert eng.submit(topo, p, q, vm0, va0, mode=Mode.MIXED) is None  # over budget while the first is held
    eng.run()
    assert eng.collect(first) is not None
    second = eng.submit(topo, p, q, vm0, va0, mode=Mode.MIXED)
    assert second is not None
    eng.run()
    assert eng.collect(second) is not None
