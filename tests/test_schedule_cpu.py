"""The LU kernel's back-substitution level schedule (gx_topo_back_levels), run on the CPU by an
emulation of the kernel's item loop with its memory timing: the values of a step are read at the end
of the step four before it (the first four steps' before the loop), LDS values at the start of the
step, and every row's x goes to y and to the LDS ring row its START named at the end of its END's
step, as the kernel does.  A dependency the schedule got wrong reads a stale value
and changes the answer.  The answer must equal, bit for bit, the column-oriented back substitution
(x_k = y_k / U(k,k), then y_r -= U(r,k) x_k for the U rows r of column k, k descending), which
applies every row's subtractions in the same order.

The forward substitution's levels (gx_topo_forward_levels) run through the same emulation with no
division, against the column-oriented forward substitution; the refactor-by-parts schedule
(gx_topo_refactor_parts) runs through an emulation of its part loop (stage, update rounds with the
value loads a chunk ahead, division, write-back) against the column-by-column left-looking refactor."""
import numpy as np
import pytest

from gridoxide_amd import Topology, config_grid, generate_grid
from gridoxide_amd import _lib
from gridoxide_amd._lib import check, ptr

P = 4                   # items per lane group and step (kBackItems)
PART_REGS = 4           # refactor part entries per lane group staged through registers (kPartRegs)
MIN_CHUNKS = 2          # chunks per part at least = chunks between a value's load and its use
STAGE_ZERO = 1 << 30    # pent_pos flag of a structural-fill entry (kStageZero): staged as zero
JOINT = 1               # rounds whose reads the kernel issues before their writes (kJointRounds)
D = 4                   # steps between loading a step's values and running it (kBackAhead)
POS = (1 << 29) - 1     # payload of the first word
ENTRY, START, END, PAD, COMBINE = 0, 1, 2, 3, 4


def _factored(net):
    t = Topology.from_network(net)
    t.factor_at(net.vm, net.va, net.p_spec, net.q_spec)
    return t


def _schedule(topo, tile, ring_rows, forward=False):
    L = _lib.lib()
    fn = L.gx_topo_forward_levels if forward else L.gx_topo_back_levels
    steps = np.zeros(1, np.int32)
    need = fn(topo._h, tile, ring_rows, None, 0, ptr(steps, _lib._i32p))
    assert need < 0
    items = np.zeros(-need, np.int32)
    got = check(fn(topo._h, tile, ring_rows, ptr(items, _lib._i32p), items.size, ptr(steps, _lib._i32p)))
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


def _emulate(items, G, lu, y, ring_rows, divide=True):
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
                xv = lx[i, g] if b < 0 else yv[i, g]
                if kind == START:
                    acc[g] = xv  # y_k, or zero (a partial piece starts from the zero row)
                elif kind == COMBINE:
                    assert b < 0 and not np.isnan(xv)  # a partial, from the ring
                    acc[g] = acc[g] + xv
                else:
                    acc[g] = acc[g] - u[i, g] * xv
                if kind == END:
                    assert g not in pend
                    pend[g] = (acc[g], u[i, g], xrow[g], xslot[g])
        for g, (a, pv, row, sl) in pend.items():  # the step's piece ends
            if row < 0:  # a partial: the ring only
                lds[sl] = a
                continue
            xk = a / pv if divide else a
            mem[row // TW] = xk
            lds[sl] = xk
        loaded[q + D] = values(q + D)  # the set refilled after this step
    return mem


def _check_pieces(items, G, n, n_entries):
    """Every row starts from y_k once (its last piece) and ends once into y; every entry is used
    once; each partial piece ends once and is combined once; one END per lane group and step."""
    kinds = (items[..., 0].astype(np.int64) & 0xFFFFFFFF) >> 29
    b = items[..., 1]
    own = (kinds == START) & (b >= 0)
    assert np.array_equal(np.sort(b[own] // (64 // G)), np.arange(n))
    partial = np.count_nonzero((kinds == START) & (b < 0))
    assert np.count_nonzero(kinds == END) == n + partial
    assert np.count_nonzero(kinds == COMBINE) == partial
    assert np.count_nonzero(kinds == ENTRY) == n_entries
    assert (np.count_nonzero(kinds == END, axis=1) <= 1).all()


@pytest.mark.parametrize("forward", [False, True])
def test_solve_pieces_independent_of_tile(forward):
    """A row's pieces and their order depend on the pattern only: the solves give the same bits for
    every tile width (so a batch equals its scenarios solved one by one whatever tile each picks)."""
    net = config_grid(2)
    topo = _factored(net)
    cp, ulen, row = topo.lu_pattern()
    n = ulen.size
    rng = np.random.default_rng(14)
    lu = rng.uniform(-0.3, 0.3, cp[-1])
    for k in range(n):
        lu[cp[k] + ulen[k]] = rng.uniform(1.0, 2.0) * rng.choice([-1.0, 1.0])
    y = rng.uniform(-1.0, 1.0, n)
    got = []
    for tile in (2, 4, 8, 16):
        items, G = _schedule(topo, tile, 512, forward=forward)
        got.append(_emulate(items, G, lu, y, 512, divide=not forward))
        kinds = (items[..., 0].astype(np.int64) & 0xFFFFFFFF) >> 29
        if tile == 4:
            assert np.count_nonzero(kinds == COMBINE) > 0  # the workload grid has rows cut into pieces
    for g in got[1:]:
        assert np.array_equal(g, got[0])


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
    # (the sums of long rows are taken in pieces: the same terms, another rounding order)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13)
    _check_pieces(items, G, n, int(ulen.sum()))
    # on the workload grids, fewer steps than columns: the rows of a level run in parallel lane groups
    if grid != "g":
        assert items.shape[0] < n


def test_back_levels_arguments():
    topo = _factored(config_grid(1))
    L = _lib.lib()
    assert L.gx_topo_back_levels(topo._h, 3, 64, None, 0, None) == _lib.GX_ERR_ARG
    fresh = Topology.from_network(config_grid(1))
    assert L.gx_topo_back_levels(fresh._h, 4, 64, None, 0, None) == _lib.GX_ERR_STATE


def _forward_column_oriented(cp, ulen, row, lu, f):
    y = f.copy()
    for j in range(len(ulen)):
        for q in range(cp[j] + ulen[j] + 1, cp[j + 1]):
            y[row[q]] = y[row[q]] - lu[q] * y[j]
    return y


@pytest.mark.parametrize("grid,tile,ring_rows", [("c1", 4, 64), ("c1", 16, 256), ("c2", 2, 512), ("c2", 4, 512),
                                                 ("c2", 8, 64), ("g", 4, 64)])
def test_forward_levels_equal_column_oriented(grid, tile, ring_rows):
    net = {"c1": lambda: config_grid(1), "c2": lambda: config_grid(2), "g": lambda: generate_grid(700, 900, 5)}[grid]()
    topo = _factored(net)
    cp, ulen, row = topo.lu_pattern()
    items, G = _schedule(topo, tile, ring_rows, forward=True)
    n = ulen.size
    rng = np.random.default_rng(12)
    lu = rng.uniform(-0.3, 0.3, cp[-1])
    f = rng.uniform(-1.0, 1.0, n)
    want = _forward_column_oriented(cp, ulen, row, lu, f)
    got = _emulate(items, G, lu, f, ring_rows, divide=False)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13)
    _check_pieces(items, G, n, int((np.diff(cp) - ulen - 1).sum()))


def _chunk_rounds(G):
    """Update rounds per chunk of the part round stream (part_chunk_rounds): 8 at 32 lane groups."""
    return 8 if G >= 32 else 16


def _parts(topo, tile):
    L = _lib.lib()
    sizes = np.zeros(4, np.int64)
    none = ctypes_null = None
    check(L.gx_topo_refactor_parts(topo._h, tile, ptr(sizes, _lib._i64p), none, none, none, ctypes_null))
    n_parts, n_ent, n_slot, cap = (int(v) for v in sizes)
    pa = np.zeros(4 * n_parts, np.int32)
    pos = np.zeros(n_ent, np.int32)
    div = np.zeros(n_ent, np.int32)
    slots = np.zeros(n_slot, np.int32)
    check(L.gx_topo_refactor_parts(topo._h, tile, ptr(sizes, _lib._i64p), ptr(pa, _lib._i32p), ptr(pos, _lib._i32p),
                                   ptr(div, _lib._i32p), ptr(slots, _lib._i32p)))
    G = 64 // tile
    return pa.reshape(n_parts, 4), pos, div, slots.reshape(-1, G, 2), cap


def _sequential_refactor(cp, ulen, row, a):
    """Left-looking refactor with the pivots in place: column k gathers A(:, k), subtracts
    L(:, j) x_j for its U entries j in ascending order, then divides its L entries by the pivot."""
    n = ulen.size
    lu = a.copy()
    x = np.zeros(n)
    for k in range(n):
        c0, c1 = cp[k], cp[k + 1]
        rows = row[c0:c1]
        x[rows] = lu[c0:c1]
        for u in range(ulen[k]):
            j = row[c0 + u]
            l0, l1 = cp[j] + ulen[j] + 1, cp[j + 1]
            x[row[l0:l1]] = x[row[l0:l1]] - lu[l0:l1] * x[j]
        vals = x[rows].copy()
        piv = c0 + ulen[k]
        vals[ulen[k] + 1:] = vals[ulen[k] + 1:] / vals[ulen[k]]
        lu[c0:c1] = vals
        assert row[piv] == k
        x[rows] = 0.0
    return lu


def _emulate_parts(pa, pos, div, slots, a, cap, G):
    """The part loop of the kernel for one scenario.  The first G * kPartRegs entries of a part are
    staged from values loaded a part ahead (after the write-back of part p - 2), or as zero where
    flagged; the others start at zero and take their values from staging items (x_o -= value * x_m,
    row m = cap + 1 holding -1).  The rounds run JOINT at a time, every read of the group before its
    writes; a chunk that holds any of the part's rounds runs all NB of them.  The values of chunk c
    of part p are loaded kPartMinChunks = 2 chunks ahead: after the write-back of part p - 2 (c < 2)
    or p - 1 (c >= 2).  A load sees the last write-back of its position by then, else the Jacobian
    value.  A position is written back at most twice (a partner column's partial visit, undivided,
    then its own).  The rows of the part buffers start as NaN."""
    big = np.iinfo(np.int64).max
    w1p, w2p = np.full(a.size, big), np.full(a.size, big)  # the parts that wrote each position back
    w1v, w2v = np.full(a.size, np.nan), np.full(a.size, np.nan)

    def load(x, seen):
        return np.where(w2p[x] <= seen, w2v[x], np.where(w1p[x] <= seen, w1v[x], a[x]))

    buf = [np.full(cap + 2, np.nan), np.full(cap + 2, np.nan)]
    mone = cap + 1
    for b in buf:
        b[mone] = -1.0
    regs = G * PART_REGS
    NB = _chunk_rounds(G)
    for p, (pb, ln, ch0, nr) in enumerate(pa):
        assert 0 < ln <= cap and nr % JOINT == 0
        xc, xprev = buf[p & 1], buf[(p & 1) ^ 1]
        xc[:mone] = np.nan
        ent = np.arange(pb, pb + ln)
        zero = (pos[ent] & STAGE_ZERO) != 0
        epos = pos[ent] & ~STAGE_ZERO
        assert (w2p[epos] == big).all()  # a third write-back of a position would be a schedule error
        k = min(ln, regs)
        xc[:k] = np.where(zero[:k], 0.0, load(epos[:k], p - 2))
        xc[k:ln] = 0.0
        nch = max(MIN_CHUNKS, -(-nr // NB))
        for r0 in range(0, nch * NB, JOINT):
            seen = p - 2 if r0 < MIN_CHUNKS * NB else p - 1
            x = slots[ch0 * NB + r0:ch0 * NB + r0 + JOINT, :, 0].astype(np.int64).ravel()
            wd = slots[ch0 * NB + r0:ch0 * NB + r0 + JOINT, :, 1].astype(np.int64).ravel() & 0xFFFFFFFF
            tg, su, off, fl = wd & 0x3FF, (wd >> 10) & 0x3FF, (wd >> 20) & 0x3FF, (wd >> 31) & 1
            assert not ((wd >> 30) & 1).any()
            real = tg != ln  # empty slots target the row past the part
            if r0 >= nr:  # (rounds past the part's last one: empty slots only)
                assert not real.any()
                if r0 // NB * NB >= nr:  # the kernel skips a chunk with none of the part's rounds
                    continue
            assert (tg <= ln).all()
            assert ((su[real] < ln) | (su[real] == mone)).all()
            assert np.unique(tg[real]).size == np.count_nonzero(real)
            assert not np.isin(tg[real], su[real]).any()
            staging = real & (su == mone)  # registered rows are staged only when flagged as zero
            assert (zero[np.minimum(tg[staging], ln - 1)] | (tg[staging] >= regs)).all()
            lv = load(x, seen)
            lv = np.where(fl == 1, xprev[np.where(fl == 1, off, 0)], lv)
            nv = xc[tg] - lv * xc[su]  # every read of the joint round, then its writes
            xc[tg] = nv
        d = div[ent]
        v = xc[:ln].copy()
        assert not np.isnan(v).any()
        m = d >= 0
        v[m] = v[m] / xc[d[m]]
        xc[:ln][m] = v[m]
        first = w1p[epos] == big
        w1p[epos[first]], w1v[epos[first]] = p, v[first]
        w2p[epos[~first]], w2v[epos[~first]] = p, v[~first]
    return np.where(w2p != big, w2v, w1v)


def _first_staging_fill(pos, n_pos):
    """Structural fill: the positions whose first staging (in part order) is as zero.  A later
    staging flagged as zero takes a partner column's partial values from the previous part's
    buffer instead of memory."""
    p = pos & ~STAGE_ZERO
    _, first = np.unique(p, return_index=True)
    fill = np.zeros(n_pos, bool)
    fill[p[first]] = (pos[first] & STAGE_ZERO) != 0
    return fill


@pytest.mark.parametrize("grid,tile", [("c1", 2), ("c1", 4), ("c1", 8), ("c1", 16), ("c2", 2), ("c2", 4), ("g", 8)])
def test_refactor_parts_equal_column_by_column(grid, tile):
    net = {"c1": lambda: config_grid(1), "c2": lambda: config_grid(2), "g": lambda: generate_grid(700, 900, 5)}[grid]()
    topo = _factored(net)
    cp, ulen, row = topo.lu_pattern()
    pa, pos, div, slots, cap = _parts(topo, tile)
    n = ulen.size
    # every entry of L+U is staged by one part, or by two: a partner column's partial visit, whose
    # entries are written back undivided, then its own part
    staged = np.bincount(pos & ~STAGE_ZERO, minlength=cp[-1])
    assert staged.min() == 1 and staged.max() <= 2
    twice = np.nonzero(staged == 2)[0]
    p_first = np.array([np.nonzero((pos & ~STAGE_ZERO) == q)[0][0] for q in twice[:200]], np.int64)
    assert (div[p_first] == -1).all()
    if grid == "c2" and 64 // tile >= 32:
        assert twice.size > 0  # partner visits run on 32 lane groups, and this grid has partners
    rng = np.random.default_rng(13)
    a = rng.uniform(-0.2, 0.2, cp[-1])
    a[cp[:-1] + ulen] = rng.uniform(3.0, 4.0, n) * rng.choice([-1.0, 1.0], n)
    fill = _first_staging_fill(pos, cp[-1])
    assert fill.any()  # the pattern has structural fill, staged as zero
    a[fill] = 0.0
    want = _sequential_refactor(cp, ulen, row, a)
    a_dev = a.copy()
    a_dev[fill] = rng.uniform(-9.0, 9.0, fill.sum())  # stale values the refactor must not read
    got = _emulate_parts(pa, pos, div, slots, a_dev, cap, 64 // tile)
    assert np.array_equal(got, want)
    # fewer parts than columns: the columns of a level are factored together
    assert pa.shape[0] < n
