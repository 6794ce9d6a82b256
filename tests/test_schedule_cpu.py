"""The LU kernel's back-substitution level schedule (gx_topo_back_levels), run on the CPU by an
emulation of the kernel's item loop with its memory timing: the values of a step are read at the end
of the step four before it (the first four steps' before the loop), LDS values at the start of the
step, and every row's x goes to y and to the LDS ring row its START named at the end of its END's
step, as the kernel does.  A dependency the schedule got wrong reads a stale value
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
POS = (1 << 29) - 1     # payload of the first word
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
    """Runs the items as the kernel does for one scenario (offsets are row * TW and slot * S)."""
    steps = items.shape[0]
    assert steps % (2 * D) == 0
    TW = 64 // G
    S = TW + 1
    mem = y.copy()
    lds = np.full(ring_rows * S, np.nan)
    lds[(ring_rows - 1) * S] = 0.0  # the zero row
    first = items[..., 0].astype(np.int64) & 0xFFFFFFFF
    kinds = first >> 29
    payload = first & POS

    def values(q):
        q = min(q, steps - 1)
        u = lu[np.minimum(payload[q], lu.size - 1)]
        b = items[q][..., 1]
        return u, mem[np.where(b >= 0, b, 0) // TW]

    acc = np.zeros(G)
    xslot = np.zeros(G, np.int64)
    xrow = np.zeros(G, np.int64)
    loaded = {q: values(q) for q in range(min(D, steps))}
    for q in range(steps):
        b_all = items[q][..., 1]
        lx = lds[np.where(b_all < 0, b_all & POS, 0)]
        u, yv = loaded.pop(q)
        pend = {}
        for i in range(P):
            for g in range(G):
                kind, b = int(kinds[q, i, g]), int(b_all[i, g])
                if kind == START:
                    xslot[g] = payload[q, i, g]
                    xrow[g] = b
                v = acc[g] - u[i, g] * (lx[i, g] if b < 0 else yv[i, g])
                acc[g] = yv[i, g] if kind == START else v
                if kind == END:
                    assert g not in pend
                    pend[g] = (acc[g], u[i, g], xrow[g], xslot[g])
        for g, (a, pv, row, sl) in pend.items():  # the step's row ends
            xk = a / pv
            mem[row // TW] = xk
            lds[sl] = xk
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
    starts = items[..., 1][kinds == START] // (64 // G)
    assert np.array_equal(np.sort(starts), np.arange(n))
    assert np.count_nonzero(kinds == END) == n
    assert np.count_nonzero(kinds == ENTRY) == int(ulen.sum())
    # at most one END per lane group and step (the kernel divides once per step)
    assert (np.count_nonzero(kinds == END, axis=1) <= 1).all()
    # on the workload grids, fewer steps than columns: the rows of a level run in parallel lane groups
    if grid != "g":
        assert items.shape[0] < n


def test_back_levels_arguments():
    topo = _factored(config_grid(1))
    L = _lib.lib()
    assert L.gx_topo_back_levels(topo._h, 3, 64, None, 0, None) == _lib.GX_ERR_ARG
    fresh = Topology.from_network(config_grid(1))
    assert L.gx_topo_back_levels(fresh._h, 4, 64, None, 0, None) == _lib.GX_ERR_STATE
