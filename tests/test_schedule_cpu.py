"""The LU kernel's back-substitution level schedule (gx_topo_back_levels), run on the CPU by an
emulation of the kernel's item loop with its memory timing: the values of a step are read at the end
of the step four before it (the first four steps' before the loop), LDS values at the start of the
step, and every END writes x to y and to its LDS ring slot as the kernel does.  A dependency the schedule got wrong reads a stale value
and changes the answer.  The answer must equal, bit for bit, the column-oriented back substitution
(x_k = y_k / U(k,k), then y_r -= U(r,k) x_k for the U rows r of column k, k descending), which
applies every row's subtractions in the same order."""
import numpy as np
import pytest

from gridoxide_amd import Topology, config_grid, generate_grid
from gridoxide_amd import _lib
from gridoxide_amd._lib import check, ptr

P = 4                   # items per lane group and step (kBackItems)
D = 4                   # steps between loading a step's values and running it (kBackAhead)
POS = (1 << 29) - 1
ROW = (1 << 19) - 1
SLOT = (1 << 12) - 1
ENTRY, START, END, PAD = 0, 1, 2, 3


def _factored(net):
    t = Topology.from_network(net)
    t.factor_at(net.vm, net.va, net.p_spec, net.q_spec)
    return t


def _schedule(topo, tile, ring_rows):
    L = _lib.lib()
    steps = np.zeros(1, np.int32)
    need = L.gx_topo_back_levels(topo._h, tile, ring_rows, None, 0, ptr(steps, _lib._i32p))
    assert need < 0
    items = np.zeros(-need, np.int32)
    got = check(L.gx_topo_back_levels(topo._h, tile, ring_rows, ptr(items, _lib._i32p), items.size,
                                      ptr(steps, _lib._i32p)))
    assert got == items.size
    G = 64 // tile
    return items.reshape(int(steps[0]), P, G, 2), G


def _column_oriented(cp, ulen, row, lu, y):
    x = y.copy()
    for k in range(len(ulen) - 1, -1, -1):
        x[k] = x[k] / lu[cp[k] + ulen[k]]
        for q in range(cp[k], cp[k] + ulen[k]):
            r = row[q]
            x[r] = x[r] - lu[q] * x[k]
    return x


def _emulate(items, G, lu, y, ring_rows):
    steps = items.shape[0]
    assert steps % D == 0
    mem = y.copy()
    lds = np.full(ring_rows, np.nan)

    def values(q):
        d = items[min(q, steps - 1)]
        u = lu[d[..., 0] & POS]
        idx = np.where(d[..., 1] >= 0, d[..., 1] & ROW, 0)
        return u, mem[idx]

    acc = np.zeros(G)
    loaded = {q: values(q) for q in range(min(D, steps))}
    for q in range(steps):
        d = items[q]
        lx = lds[np.where(d[..., 1] < 0, d[..., 1] & SLOT, 0)]
        u, yv = loaded.pop(q)
        for i in range(P):
            for g in range(G):
                a, b = int(d[i, g, 0]), int(d[i, g, 1])
                kind = (a & 0xFFFFFFFF) >> 29
                if kind == START:
                    acc[g] = yv[i, g]
                elif kind == ENTRY:
                    acc[g] = acc[g] - u[i, g] * (lx[i, g] if b < 0 else yv[i, g])
                elif kind == END:
                    xk = acc[g] / u[i, g]
                    mem[b & ROW] = xk
                    lds[(b >> 19) & SLOT] = xk
        loaded[q + D] = values(q + D)  # the set refilled after this step
    return mem


@pytest.mark.parametrize("grid,tile,ring_rows", [("c1", 4, 64), ("c1", 8, 512), ("c1", 16, 256), ("c2", 4, 512),
                                                 ("c2", 8, 64), ("g", 4, 64)])
def test_back_levels_equal_column_oriented(grid, tile, ring_rows):
    net = {"c1": lambda: config_grid(1), "c2": lambda: config_grid(2), "g": lambda: generate_grid(700, 900, 5)}[grid]()
    topo = _factored(net)
    cp, ulen, row = topo.lu_pattern()
    items, G = _schedule(topo, tile, ring_rows)
    n = ulen.size
    rng = np.random.default_rng(11)
    lu = rng.uniform(-0.3, 0.3, cp[-1])
    for k in range(n):
        lu[cp[k] + ulen[k]] = rng.uniform(1.0, 2.0) * rng.choice([-1.0, 1.0])
    y = rng.uniform(-1.0, 1.0, n)
    want = _column_oriented(cp, ulen, row, lu, y)
    got = _emulate(items, G, lu, y, ring_rows)
    assert np.array_equal(got, want)
    # every row ends exactly once, every U entry is used exactly once
    kinds = (items[..., 0].astype(np.int64) & 0xFFFFFFFF) >> 29
    ends = items[..., 1][kinds == END] & ROW
    assert np.array_equal(np.sort(ends), np.arange(n))
    assert np.count_nonzero(kinds == ENTRY) == int(ulen.sum())
    # on the workload grids, fewer steps than columns: the rows of a level run in parallel lane groups
    if grid != "g":
        assert items.shape[0] < n


def test_back_levels_arguments():
    topo = _factored(config_grid(1))
    L = _lib.lib()
    assert L.gx_topo_back_levels(topo._h, 3, 64, None, 0, None) == _lib.GX_ERR_ARG
    fresh = Topology.from_network(config_grid(1))
    assert L.gx_topo_back_levels(fresh._h, 4, 64, None, 0, None) == _lib.GX_ERR_STATE
