"""Host-side mirror of the reference's solver / batch interface over the HIP backend.

The names, argument meanings and error behaviour follow the reference so that a caller of
``PersistentSolver`` / ``BatchSolver`` finds the same surface:

* ``BusType``, ``SolveStatus``, ``SolveStats``            src/types.rs:5-9, src/solver.rs:118-171
* ``JacobianBackend`` (adds ``HIP``)                      src/solver.rs:61-108
* ``BusOverride``, ``Scenario``, ``uniform_load_scaling`` src/batch.rs:39-105
* ``BatchError`` and its variants                         src/batch.rs:107-134
* ``PersistentSolver`` (solve / solve_with_stats / reset) src/solver.rs:395-494
* ``BatchSolver`` (solve)                                 src/batch.rs:152-289

Every numeric step runs in ``libgridox_hip.so`` on the GPU; this module only marshals arrays.
"""
from __future__ import annotations

import dataclasses
import enum
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, f64, i32, ptr
from .grid import PQ, PV, SLACK, Network


class BusType(enum.IntEnum):
    Slack = SLACK
    PV = PV
    PQ = PQ


class SolveStatus(enum.IntEnum):
    Converged = _lib.GX_CONVERGED
    MaxIterationsReached = _lib.GX_MAX_ITER
    Singular = _lib.GX_SINGULAR


class IslandStatus(enum.IntEnum):
    """src/solver.rs:179-212"""
    Converged = _lib.GX_CONVERGED
    MaxIterationsReached = _lib.GX_MAX_ITER
    Singular = _lib.GX_SINGULAR
    NoReferenceBus = 3
    AmbiguousReferenceBus = 4


@dataclasses.dataclass
class IslandReport:
    bus_indices: List[int]
    slack_indices: List[int]
    status: IslandStatus


class JacobianBackend(enum.Enum):
    """Backend selector; only ``HIP`` is implemented here (the reference's CPU backends are not)."""
    HIP = "hip"


class Mode(enum.IntEnum):
    FP64 = _lib.GX_MODE_FP64
    MIXED = _lib.GX_MODE_MIXED


@dataclasses.dataclass
class SolveStats:
    status: SolveStatus
    mismatch_history: List[float]

    def iterations(self) -> int:
        return len(self.mismatch_history)

    def final_mismatch(self) -> float:
        """Max |mismatch| at the last iteration, 0.0 if none ran (src/solver.rs:153-156)."""
        return self.mismatch_history[-1] if self.mismatch_history else 0.0


@dataclasses.dataclass
class PowerFlowReport:
    voltage_mag: np.ndarray
    voltage_ang: np.ndarray
    stats: SolveStats
    islands: List[IslandReport] = dataclasses.field(default_factory=list)


# ---------------------------------------------------------------- scenarios and errors
@dataclasses.dataclass
class BusOverride:
    bus: int
    p_spec: Optional[float] = None
    q_spec: Optional[float] = None
    voltage_mag: Optional[float] = None

    def p(self, v: float) -> "BusOverride":
        self.p_spec = v
        return self

    def q(self, v: float) -> "BusOverride":
        self.q_spec = v
        return self

    def vm(self, v: float) -> "BusOverride":
        self.voltage_mag = v
        return self


@dataclasses.dataclass
class Scenario:
    bus_overrides: List[BusOverride] = dataclasses.field(default_factory=list)
    branch_outages: List[int] = dataclasses.field(default_factory=list)


def uniform_load_scaling(template: Network, factor: float) -> Scenario:
    """Scales every bus with a nonzero p_spec or q_spec by ``factor`` (src/batch.rs:98-105)."""
    idx = np.nonzero((template.p_spec != 0.0) | (template.q_spec != 0.0))[0]
    return Scenario([BusOverride(int(i), float(template.p_spec[i] * factor), float(template.q_spec[i] * factor))
                     for i in idx])


class BatchError(Exception):
    pass


class OutagesUnsupported(BatchError):
    def __init__(self, scenario: int):
        super().__init__(f"scenario {scenario} requests branch outages, which are not implemented yet "
                         f"(see batch::Scenario::branch_outages)")
        self.scenario = scenario


class BusOutOfRange(BatchError):
    def __init__(self, scenario: int, bus: int, n_buses: int):
        super().__init__(f"scenario {scenario} overrides bus {bus}, but the template has only {n_buses} bus(es)")
        self.scenario, self.bus, self.n_buses = scenario, bus, n_buses


# ---------------------------------------------------------------- Y-bus and topology
@dataclasses.dataclass
class YBus:
    n: int
    rowptr: np.ndarray
    col: np.ndarray
    y: np.ndarray  # complex128


def _branch_args(net: Network):
    """The line/transformer argument block of gx_ybus_build and gx_topo_set_branches."""
    cplx = lambda a: f64(np.stack([a.real, a.imag], 1).ravel()) if a.size else np.zeros(2)
    lf, lt = i32(net.line_from), i32(net.line_to)
    lr, lx, lb, lg = f64(net.line_r), f64(net.line_x), f64(net.line_b), f64(net.line_g)
    tf, tt = i32(net.tr_from), i32(net.tr_to)
    ys, ysh, tap = cplx(net.tr_ys), cplx(net.tr_ysh), cplx(net.tr_tap)
    keep = (lf, lt, lr, lx, lb, lg, tf, tt, ys, ysh, tap)  # the arrays must outlive the call
    args = (lf.size, ptr(lf, _lib._i32p), ptr(lt, _lib._i32p), ptr(lr), ptr(lx), ptr(lb), ptr(lg), tf.size,
            ptr(tf, _lib._i32p), ptr(tt, _lib._i32p), ptr(ys), ptr(ysh), ptr(tap))
    return keep, args


def build_ybus(net: Network) -> YBus:
    """Merged Y-bus of a network (gx_ybus_build; build_ybus + stamp_transformers + finish)."""
    L = _lib.lib()
    cap = 4 * (net.line_from.size + net.tr_from.size) + net.n_bus
    rowptr = np.zeros(net.n_bus + 1, np.int32)
    col = np.zeros(cap, np.int32)
    yre = np.zeros(cap)
    yim = np.zeros(cap)
    keep, args = _branch_args(net)
    nnz = check(L.gx_ybus_build(net.n_bus, *args, cap, ptr(rowptr, _lib._i32p), ptr(col, _lib._i32p), ptr(yre),
                                ptr(yim)))
    del keep
    return YBus(net.n_bus, rowptr, col[:nnz].copy(), yre[:nnz] + 1j * yim[:nnz])


class Topology:
    """One topology + bus-type assignment on the device (gx_topo)."""

    def __init__(self, bus_type: np.ndarray, ybus: YBus):
        L = _lib.lib()
        self.n_bus = int(ybus.n)
        self.bus_type = i32(bus_type)
        self.ybus = ybus
        self._yre = f64(ybus.y.real)
        self._yim = f64(ybus.y.imag)
        err = np.zeros(1, np.int32)
        h = L.gx_topo_create(self.n_bus, ptr(self.bus_type, _lib._i32p), ptr(i32(ybus.rowptr), _lib._i32p),
                             ptr(i32(ybus.col), _lib._i32p), ptr(self._yre), ptr(self._yim), ptr(err, _lib._i32p))
        if not h:
            raise _lib.GxError(int(err[0]), L.gx_last_error().decode())
        self._h = h

    @classmethod
    def from_network(cls, net: Network) -> "Topology":
        topo = cls(net.bus_type, build_ybus(net))
        topo.set_branches(net)
        if net.zip_coef is not None:
            topo.set_zip(net.zip_coef)
        return topo

    def set_zip(self, zip_coef: Optional[np.ndarray]):
        """Per-bus ZIP terms ([n_bus, 6]: (P, Q) of the constant-power, -current and -impedance
        parts) used by every later solve's mismatch; None removes them."""
        if zip_coef is None:
            check(_lib.lib().gx_topo_set_zip(self._h, None))
            self._zip = None
            return
        z = f64(zip_coef).reshape(-1)
        if z.size != 6 * self.n_bus:
            raise ValueError(f"zip_coef must have shape [{self.n_bus}, 6]")
        check(_lib.lib().gx_topo_set_zip(self._h, ptr(z)))
        self._zip = z

    def set_branches(self, net: Network):
        """Registers the network's branches (lines, then transformers) for per-scenario outages."""
        keep, args = _branch_args(net)
        check(_lib.lib().gx_topo_set_branches(self._h, *args))
        del keep
        self.n_branches = int(net.line_from.size + net.tr_from.size)

    def branch_is_bridge(self, branch: int) -> bool:
        """True when taking ``branch`` out of service would split its island (such outages are
        rejected)."""
        return bool(check(_lib.lib().gx_topo_branch_is_bridge(self._h, int(branch))))

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().gx_topo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sizes(self) -> dict:
        s = np.zeros(8, np.int64)
        check(_lib.lib().gx_topo_sizes(self._h, ptr(s, _lib._i64p)))
        keys = ["n_bus", "nnz_y", "n", "nnz_j", "nnz_lu", "n_blocks", "factored", "n_update"]
        return {k: int(v) for k, v in zip(keys, s)}

    def jacobian_pattern(self):
        nj = self.sizes()["nnz_j"]
        r = np.zeros(nj, np.int32)
        c = np.zeros(nj, np.int32)
        check(_lib.lib().gx_topo_jacobian_pattern(self._h, ptr(r, _lib._i32p), ptr(c, _lib._i32p)))
        return r, c

    def factor_at(self, vm, va, p_spec, q_spec):
        check(_lib.lib().gx_topo_factor_at(self._h, ptr(f64(vm)), ptr(f64(va)), ptr(f64(p_spec)), ptr(f64(q_spec))))

    def set_pivot_order(self, q, pnum):
        check(_lib.lib().gx_topo_set_pivot_order(self._h, ptr(i32(q), _lib._i32p), ptr(i32(pnum), _lib._i32p)))

    def ordering(self):
        s = self.sizes()
        n = s["n"]
        p, q, r = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n + 1, np.int32)
        pnum = np.zeros(n, np.int32) if s["factored"] else None
        check(_lib.lib().gx_topo_ordering(self._h, ptr(p, _lib._i32p), ptr(q, _lib._i32p), ptr(r, _lib._i32p),
                                          ptr(pnum, _lib._i32p) if pnum is not None else None))
        return p, q, r[:s["n_blocks"] + 1].copy(), pnum

    def lu_pattern(self):
        s = self.sizes()
        cp = np.zeros(s["n"] + 1, np.int32)
        ulen = np.zeros(s["n"], np.int32)
        row = np.zeros(s["nnz_lu"], np.int32)
        check(_lib.lib().gx_topo_lu_pattern(self._h, ptr(cp, _lib._i32p), ptr(ulen, _lib._i32p), ptr(row, _lib._i32p)))
        return cp, ulen, row

    def fallbacks(self) -> int:
        """Scenarios of the last mixed-precision solve that were re-solved with the FP64 path."""
        return check(_lib.lib().gx_batch_fallbacks(self._h))

    def islands(self):
        """(component of each bus, verdict of each component: 0 solvable, 1 no reference, 2 ambiguous)."""
        comp = np.zeros(self.n_bus, np.int32)
        verdict = np.zeros(self.n_bus, np.int32)
        nc = check(_lib.lib().gx_topo_islands(self._h, ptr(comp, _lib._i32p), ptr(verdict, _lib._i32p)))
        return comp, verdict[:nc].copy()

    def island_status(self, n_scen: int) -> np.ndarray:
        """[n_scen, n_islands] IslandStatus codes of the last batch solve."""
        nc = check(_lib.lib().gx_topo_islands(self._h, None, None))
        out = np.zeros((n_scen, nc), np.int32)
        check(_lib.lib().gx_batch_island_status(self._h, ptr(out, _lib._i32p)))
        return out

    def island_reports(self, n_scen: int) -> List[List[IslandReport]]:
        comp, verdict = self.islands()
        st = self.island_status(n_scen)
        members = [np.nonzero(comp == c)[0] for c in range(verdict.size)]
        slacks = [[int(i) for i in m if self.bus_type[i] == SLACK] for m in members]
        return [[IslandReport([int(i) for i in members[c]], slacks[c], IslandStatus(int(st[s, c])))
                 for c in range(verdict.size)] for s in range(n_scen)]

    def profile(self, enable: bool = True):
        """Returns {kernel class: (total ms, launches)} accumulated since the last call and turns
        per-launch event timing on/off for later solves."""
        ms = np.zeros(len(_lib.KERNEL_CLASSES))
        n = np.zeros(len(_lib.KERNEL_CLASSES), np.int32)
        check(_lib.lib().gx_topo_profile(self._h, int(enable), ptr(ms), ptr(n, _lib._i32p)))
        return {k: (float(a), int(b)) for k, a, b in zip(_lib.KERNEL_CLASSES, ms, n)}

    def set_launch(self, tile: int = 0, ring_rows: int = 0):
        """Scenarios per tile (2, 4, 8, 16; 0 = automatic) and LDS solve-window rows (power of two,
        0 = default) for later solves."""
        check(_lib.lib().gx_topo_set_launch(self._h, int(tile), int(ring_rows)))

    def set_refactor_groups(self, groups: int = 0):
        """Lane groups of the refactor of mixed-precision 8-scenario tiles (0 = automatic, 16, 32)."""
        check(_lib.lib().gx_topo_set_refactor_groups(self._h, int(groups)))

    def batch_solve(self, p_spec, q_spec, vm0, va0, tol=1e-8, max_iter=20, mode=Mode.FP64, init="given",
                    outage=None, keep_warm=False):
        """Host arrays [n_scen, n_bus] -> (vm, va, status, iters, history).  init: "given" starts
        from vm0/va0; "linear" first applies linear_initial_guess on the device; "warm" starts from
        the topology's resident warm start (kept by an earlier keep_warm solve), else vm0/va0.
        outage: optional [n_scen] branch index out of service per scenario (-1: none)."""
        p_spec = f64(np.atleast_2d(p_spec))
        S = p_spec.shape[0]
        q_spec = f64(np.broadcast_to(q_spec, (S, self.n_bus)))
        vm0 = f64(np.broadcast_to(vm0, (S, self.n_bus)))
        va0 = f64(np.broadcast_to(va0, (S, self.n_bus)))
        vm = np.zeros((S, self.n_bus))
        va = np.zeros((S, self.n_bus))
        st = np.zeros(S, np.int32)
        it = np.zeros(S, np.int32)
        hist = np.zeros((S, max(max_iter, 1)))
        out = None if outage is None else i32(np.broadcast_to(outage, (S,)))
        check(_lib.lib().gx_batch_solve_outages(
            self._h, S, p_spec.ctypes.data, q_spec.ctypes.data, vm0.ctypes.data, va0.ctypes.data,
            None if out is None else out.ctypes.data, float(tol), int(max_iter), int(mode), _init_flag(init, keep_warm),
            vm.ctypes.data, va.ctypes.data, st.ctypes.data, it.ctypes.data, hist.ctypes.data, None))
        return vm, va, st, it, hist[:, :max_iter]

    def batch_solve_device(self, p_spec, q_spec, vm0, va0, vm, va, status, iters, history, tol=1e-8, max_iter=20,
                           mode=Mode.FP64, stream=0, init="given", outage=None, keep_warm=False):
        """Device-resident variant: every argument is a CUDA/HIP tensor (torch) already in HBM
        (outage: optional int32 tensor [n_scen])."""
        S = int(p_spec.shape[0])
        check(_lib.lib().gx_batch_solve_outages(self._h, S, p_spec.data_ptr(), q_spec.data_ptr(), vm0.data_ptr(),
                                        va0.data_ptr(), None if outage is None else outage.data_ptr(),
                                        float(tol), int(max_iter), int(mode),
                                        _lib.GX_DEVICE_PTRS | _init_flag(init, keep_warm),
                                        vm.data_ptr(), va.data_ptr(), status.data_ptr(), iters.data_ptr(),
                                        history.data_ptr() if history is not None else None, stream or None))


def _init_flag(init: str, keep_warm: bool = False) -> int:
    """init: "given" (vm0/va0 as passed), "linear" (linear_initial_guess first) or "warm" (the
    topology's resident warm start when it holds one, else vm0/va0); keep_warm: leave the first
    converged scenario's solution on the device as the topology's warm start."""
    flags = _lib.GX_KEEP_WARM if keep_warm else 0
    if init == "given":
        return flags
    if init == "linear":
        return flags | _lib.GX_INIT_LINEAR
    if init == "warm":
        return flags | _lib.GX_INIT_WARM
    raise ValueError(f"unknown init {init!r} (expected 'given', 'linear' or 'warm')")


# ---------------------------------------------------------------- reference-shaped solvers
def _stats(status: int, iters: int, hist_row: np.ndarray) -> SolveStats:
    return SolveStats(SolveStatus(int(status)), [float(x) for x in hist_row[:int(iters)]])


class PersistentSolver:
    """One topology's cached pattern + pivot order, reused across solves (src/solver.rs:395-494)."""

    def __init__(self, backend: JacobianBackend = JacobianBackend.HIP, mode: Mode = Mode.FP64):
        if backend is not JacobianBackend.HIP:
            raise ValueError("only the HIP backend is available")
        self.mode = mode
        self._topo: Optional[Topology] = None
        self._key = None

    def solve_with_stats(self, net: Network, tol: float, max_iter: int, ybus: Optional[YBus] = None):
        """Solves ``net`` from its own vm/va; updates net.vm / net.va in place like the reference's
        ``&mut [Bus]`` and returns ``SolveStats``."""
        if self._topo is None:
            self._topo = Topology(net.bus_type, ybus if ybus is not None else build_ybus(net))
            if net.zip_coef is not None:
                self._topo.set_zip(net.zip_coef)
        vm, va, st, it, hist = self._topo.batch_solve(net.p_spec[None], net.q_spec[None], net.vm[None], net.va[None],
                                                      tol, max_iter, self.mode)
        net.vm[:] = vm[0]
        net.va[:] = va[0]
        return _stats(st[0], it[0], hist[0])

    def solve(self, net: Network, tol: float, max_iter: int, ybus: Optional[YBus] = None) -> SolveStats:
        return self.solve_with_stats(net, tol, max_iter, ybus)

    def reset(self):
        if self._topo is not None:
            self._topo.close()
        self._topo = None


class BatchSolver:
    """Scenario batch over one shared topology (src/batch.rs:152-289), all scenarios on the GPU.

    Like the reference, every scenario starts from linear_initial_guess (src/batch.rs:268) unless
    ``init="given"`` asks for the template/override voltages as they are (flat start).

    ``outages=True`` accepts N-1 scenarios: ``Scenario.branch_outages`` with at most one branch
    (lines first, then transformers), solved on the shared pattern with that branch's stamps taken
    out of the scenario's Y-bus.  The default keeps the reference's behaviour of rejecting them."""

    def __init__(self, backend: JacobianBackend = JacobianBackend.HIP, mode: Mode = Mode.FP64, init: str = "linear",
                 outages: bool = False):
        if backend is not JacobianBackend.HIP:
            raise ValueError("only the HIP backend is available")
        _init_flag(init)
        self.mode = mode
        self.init = init
        self.outages = outages

    @staticmethod
    def scenario_arrays(template: Network, scenarios: Sequence[Scenario], outages: bool = False):
        """Validates like the reference and expands overrides into dense [S, n_bus] arrays (with
        ``outages``, also returns the [S] outaged branch per scenario, -1 for none)."""
        n = template.n_bus
        n_br = template.line_from.size + template.tr_from.size
        for i, sc in enumerate(scenarios):
            if sc.branch_outages and not (outages and len(sc.branch_outages) == 1):
                raise OutagesUnsupported(i)
            for e in sc.branch_outages:
                if e < 0 or e >= n_br:
                    raise ValueError(f"scenario {i}: branch {e} out of range ({n_br} branches)")
            for ov in sc.bus_overrides:
                if ov.bus >= n or ov.bus < 0:
                    raise BusOutOfRange(i, ov.bus, n)
        S = len(scenarios)
        p = np.repeat(template.p_spec[None], S, 0)
        q = np.repeat(template.q_spec[None], S, 0)
        vm = np.repeat(template.vm[None], S, 0)
        for s, sc in enumerate(scenarios):
            for ov in sc.bus_overrides:
                if ov.p_spec is not None:
                    p[s, ov.bus] = ov.p_spec
                if ov.q_spec is not None:
                    q[s, ov.bus] = ov.q_spec
                if ov.voltage_mag is not None:
                    vm[s, ov.bus] = ov.voltage_mag
        va = np.repeat(template.va[None], S, 0)
        if outages:
            out = np.array([sc.branch_outages[0] if sc.branch_outages else -1 for sc in scenarios], np.int32)
            return p, q, vm, va, out
        return p, q, vm, va

    def solve(self, template: Network, scenarios: Sequence[Scenario], tol: float, max_iter: int,
              topology: Optional[Topology] = None) -> List[PowerFlowReport]:
        """Starting point: the template voltages with overrides applied, then (by default) the
        linear initial guess."""
        arrays = self.scenario_arrays(template, scenarios, self.outages)
        p, q, vm0, va0 = arrays[:4]
        out = arrays[4] if self.outages and (arrays[4] >= 0).any() else None
        if len(scenarios) == 0:
            return []
        topo = topology or Topology.from_network(template)
        vm, va, st, it, hist = topo.batch_solve(p, q, vm0, va0, tol, max_iter, self.mode, init=self.init,
                                                outage=out)
        isl = topo.island_reports(len(scenarios))
        return [PowerFlowReport(vm[s], va[s], _stats(st[s], it[s], hist[s]), isl[s]) for s in range(len(scenarios))]


class Engine:
    """Many (topology, scenarios) requests in flight on one GPU (gx_engine): each step is one
    Newton iteration of every active request with one persistent LU launch over all their tiles.
    Requests join between steps, subject to the device-memory budget, and leave when finished."""

    def __init__(self, tile: int = 4, budget_bytes: int = 0, graphs: bool = False):
        err = np.zeros(1, np.int32)
        self._h = _lib.lib().gx_engine_create(int(tile), 0, int(budget_bytes), None, ptr(err, _lib._i32p))
        if not self._h:
            check(int(err[0]))
        self.tile = int(tile)
        self.graphs = graphs
        self._meta = {}
        self._keep = {}

    def close(self):
        if getattr(self, "_h", None):
            try:
                _lib.lib().gx_engine_destroy(self._h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    def __del__(self):
        self.close()

    @staticmethod
    def request_bytes(topo: "Topology", n_scen: int, max_iter: int = 20, mode: Mode = Mode.MIXED,
                      init: str = "given", tile: int = 4) -> int:
        return int(check(_lib.lib().gx_engine_request_bytes(topo._h, int(n_scen), int(max_iter), int(mode),
                                                            _init_flag(init), int(tile))))

    def bytes_in_use(self) -> int:
        return int(check(_lib.lib().gx_engine_bytes(self._h)))

    def submit(self, topo: "Topology", p_spec, q_spec, vm0, va0, tol=1e-8, max_iter=20, mode=Mode.MIXED,
               init="given", outage=None, keep_warm=False):
        """Host arrays [n_scen, n_bus].  Returns the request id, or None when the request does not
        fit the budget now (retry after collecting others).  init="warm" starts from the topology's
        resident warm start; keep_warm leaves this request's first converged solution there."""
        p_spec = f64(np.atleast_2d(p_spec))
        S = p_spec.shape[0]
        q_spec = f64(np.broadcast_to(q_spec, (S, topo.n_bus)))
        vm0 = f64(np.broadcast_to(vm0, (S, topo.n_bus)))
        va0 = f64(np.broadcast_to(va0, (S, topo.n_bus)))
        out = None if outage is None else i32(np.broadcast_to(outage, (S,)))
        rc = _lib.lib().gx_engine_submit(self._h, topo._h, S, p_spec.ctypes.data, q_spec.ctypes.data,
                                         vm0.ctypes.data, va0.ctypes.data, None if out is None else out.ctypes.data,
                                         float(tol), int(max_iter), int(mode), _init_flag(init, keep_warm))
        if rc == _lib.GX_ERR_BUDGET:
            return None
        rid = check(rc)
        self._meta[rid] = (S, topo.n_bus, int(max_iter))
        return rid

    def step(self) -> int:
        """One Newton step of every active request; returns how many remain active."""
        return int(check(_lib.lib().gx_engine_step(self._h)))

    def run(self, max_steps: int = 1 << 30) -> int:
        """Steps until every request finished (or max_steps), each step a replayed hipGraph when the
        engine was created with graphs=True."""
        if self.graphs:
            return int(check(_lib.lib().gx_engine_run(self._h, int(max_steps))))
        active = 1
        for _ in range(max_steps):
            active = self.step()
            if active == 0:
                break
        return active

    def collect(self, rid: int):
        """(vm, va, status, iters, history) of a finished request (it is then released), else None."""
        S, nb, mi = self._meta[rid]
        vm = np.zeros((S, nb))
        va = np.zeros((S, nb))
        st = np.zeros(S, np.int32)
        it = np.zeros(S, np.int32)
        hist = np.zeros((S, max(mi, 1)))
        done = check(_lib.lib().gx_engine_collect(self._h, int(rid), vm.ctypes.data, va.ctypes.data, st.ctypes.data,
                                                  it.ctypes.data, hist.ctypes.data))
        if not done:
            return None
        del self._meta[rid]
        return vm, va, st, it, hist[:, :mi]

    def fallbacks(self, rid: int) -> int:
        return int(check(_lib.lib().gx_engine_fallbacks(self._h, int(rid))))
