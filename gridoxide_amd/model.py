"""``PowerFlowModel`` / ``BatchResult``: the reference's Python API (``src/python.rs``) over the HIP backend.

Same names, arguments and error behaviour as the reference's PyO3 classes, so a script written against
``gridoxide.PowerFlowModel`` runs unchanged apart from the backend name:

* ``PowerFlowModel.from_pgm_json(path, backend, tol, max_iter, s_base_va, freq_hz)``
  (src/python.rs:134-168): PGM JSON -> buses, branches and shunts (``gridoxide_amd.pgm``);
* ``solve()`` (src/python.rs:256-274): template voltages, the linear initial guess, then the Newton
  loop with unreferenced islands pinned to V = 0; raises ``RuntimeError`` when an island did not
  converge or met a singular Jacobian;
* ``solve_batch(scenarios, threads)`` / ``solve_batch_scaled(scales, threads)``
  (src/python.rs:310-350, 441-472): one ``BatchResult`` per scenario, in scenario order; failures are
  reported in ``converged``, never raised;
* ``reset``, ``n_nodes``, ``default_threads``, ``ybus_triplets``, ``bus_spec``, ``initial_guess``,
  ``zip_term_counts``, ``voltage_mag``, ``voltage_ang``, ``voltage_kv`` (src/python.rs:225-438).

Every solve runs on the GPU through ``libgridox_hip.so``; there is no CPU path.  ``backend`` accepts
``"hip"`` only, and ``threads`` is accepted for compatibility and has no effect: a batch is one set
of device launches over all its scenarios.  ``mode`` (an addition) selects the FP64 or the mixed
precision Newton loop.  CGMES input (``from_cgmes``, ``bus_index_for_mrid``) is not part of this
package.
"""
from __future__ import annotations

import dataclasses
import json
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import pgm as _pgm
from .solver import (BatchError, BatchSolver, BusOverride, IslandStatus, Mode, Scenario, SolveStatus, Topology,
                     uniform_load_scaling)

_REFERENCE_BACKENDS = ("scalar", "block", "klu", "klu_native", "pardiso")


def _parse_backend(name: str) -> str:
    if name == "hip":
        return name
    if name in _REFERENCE_BACKENDS:
        raise ValueError(f"backend '{name}' is a CPU backend of the reference; this package runs 'hip' only")
    raise ValueError(f"unknown backend '{name}', expected 'hip'")


@dataclasses.dataclass
class BatchResult:
    """One scenario's outcome (src/python.rs:62-79)."""

    voltage_mag: List[float]
    voltage_ang: List[float]
    iterations: int
    converged: bool
    max_mismatch: float


class PowerFlowModel:
    """A power flow model over one topology, solved on the GPU; repeated ``solve()`` calls reuse the
    cached symbolic analysis and pivot order until ``reset()`` (src/python.rs:103-127)."""

    def __init__(self, model: _pgm.PgmModel, backend: str = "hip", tol: float = 1e-6, max_iter: int = 20,
                 mode: Mode = Mode.FP64):
        self._backend = _parse_backend(backend)
        self._model = model
        self._template = model.network
        self._tol = float(tol)
        self._max_iter = int(max_iter)
        self._mode = Mode(mode)
        self._vm = self._template.vm.copy()
        self._va = self._template.va.copy()
        self._topo: Optional[Topology] = None

    # ---- construction
    @staticmethod
    def from_pgm_json(path: str, backend: str = "hip", tol: float = 1e-6, max_iter: int = 20,
                      s_base_va: float = 1e6, freq_hz: float = 50.0, mode: Mode = Mode.FP64) -> "PowerFlowModel":
        """Loads a PGM-format JSON file (src/python.rs:134-168).  Unreadable files raise
        ``RuntimeError``, malformed documents ``ValueError``, like the reference."""
        try:
            with open(path) as f:
                raw = f.read()
        except OSError as e:
            raise RuntimeError(f"reading {path}: {e}") from None
        try:
            doc = json.loads(raw)
        except json.JSONDecodeError as e:
            raise ValueError(f"parsing {path} as PGM JSON: {e}") from None
        return PowerFlowModel.from_pgm(doc, backend, tol, max_iter, s_base_va, freq_hz, mode)

    @staticmethod
    def from_pgm(doc: dict, backend: str = "hip", tol: float = 1e-6, max_iter: int = 20, s_base_va: float = 1e6,
                 freq_hz: float = 50.0, mode: Mode = Mode.FP64) -> "PowerFlowModel":
        """``from_pgm_json`` for an already parsed document (e.g. one scenario of a PGM batch update
        applied with ``gridoxide_amd.pgm.apply_batch_update``)."""
        return PowerFlowModel(_pgm.from_pgm(doc, s_base_va, freq_hz), backend, tol, max_iter, mode)

    # ---- topology on the device
    def _topology(self) -> Topology:
        if self._topo is None:
            self._topo = Topology.from_network(self._template)
        return self._topo

    def reset(self) -> None:
        """Discards the cached symbolic analysis and pivot order (src/python.rs:276-282)."""
        if self._topo is not None:
            self._topo.close()
        self._topo = None

    @property
    def n_nodes(self) -> int:
        """Buses, including star buses of three-winding transformers and each active source's Slack
        bus (src/python.rs:221-227)."""
        return self._template.n_bus

    @staticmethod
    def default_threads() -> int:
        """Host workers a batch would use: 1, since the batch runs as device launches."""
        return 1

    # ---- solves
    def solve(self) -> None:
        """Solves from the template with the linear initial guess (src/python.rs:256-274).  Islands
        without a source are pinned to V = 0; an island that did not converge within ``max_iter``
        iterations or whose Jacobian is singular raises ``RuntimeError`` (the voltages reached are
        kept, as in the reference)."""
        t = self._template
        topo = self._topology()
        vm, va, st, it, hist = topo.batch_solve(t.p_spec[None], t.q_spec[None], t.vm[None], t.va[None], self._tol,
                                                self._max_iter, self._mode, init="linear")
        self._vm = vm[0].copy()
        self._va = va[0].copy()
        for island in topo.island_reports(1)[0]:
            if island.status == IslandStatus.MaxIterationsReached:
                raise RuntimeError(f"power flow did not converge within {self._max_iter} iterations "
                                   f"(component with buses {island.bus_indices})")
            if island.status == IslandStatus.Singular:
                raise RuntimeError(f"Jacobian is singular (component with buses {island.bus_indices})")

    def _run_batch(self, scenarios: Sequence[Scenario]) -> List[BatchResult]:
        try:
            reports = BatchSolver(mode=self._mode).solve(self._template, scenarios, self._tol, self._max_iter,
                                                         topology=self._topology() if scenarios else None)
        except BatchError as e:
            raise RuntimeError(str(e)) from None
        return [BatchResult([float(v) for v in r.voltage_mag], [float(v) for v in r.voltage_ang],
                            r.stats.iterations(), r.stats.status == SolveStatus.Converged, r.stats.final_mismatch())
                for r in reports]

    def solve_batch(self, scenarios: Sequence[Sequence[Tuple[int, float, float]]],
                    threads: Optional[int] = None) -> List[BatchResult]:
        """One scenario per list of ``(bus_index, p_spec, q_spec)`` overrides in per unit; buses not
        named keep the model's values (src/python.rs:284-328)."""
        del threads
        return self._run_batch([Scenario([BusOverride(int(b)).p(float(p)).q(float(q)) for b, p, q in ov])
                                for ov in scenarios])

    def solve_batch_scaled(self, scales: Sequence[float], threads: Optional[int] = None) -> List[BatchResult]:
        """One scenario per factor, every nonzero ``p_spec``/``q_spec`` multiplied by it
        (src/python.rs:330-345, uniform_load_scaling src/batch.rs:98-105)."""
        del threads
        return self._run_batch([uniform_load_scaling(self._template, float(f)) for f in scales])

    # ---- inspection
    def ybus_triplets(self) -> Tuple[List[int], List[int], List[float], List[float]]:
        """``(rows, cols, g, b)`` of every stored Y-bus entry, row-major with ascending columns
        (src/python.rs:352-373)."""
        y = self._topology().ybus
        rows = np.repeat(np.arange(y.n), np.diff(y.rowptr))
        return ([int(r) for r in rows], [int(c) for c in y.col], [float(v) for v in y.y.real],
                [float(v) for v in y.y.imag])

    def bus_spec(self) -> Tuple[List[int], List[float], List[float]]:
        """``(bus_type, p_spec, q_spec)`` per bus of the template, 0 = Slack, 1 = PV, 2 = PQ
        (src/python.rs:375-396)."""
        t = self._template
        return [int(k) for k in t.bus_type], [float(v) for v in t.p_spec], [float(v) for v in t.q_spec]

    def initial_guess(self) -> Tuple[List[float], List[float]]:
        """``(voltage_mag, voltage_ang)`` after the linear initial guess, the point the Newton loop
        starts from (src/python.rs:398-410), computed on the device.  Buses of an island without a
        source are reported at their pinned V = 0."""
        t = self._template
        vm, va, _, _, _ = self._topology().batch_solve(t.p_spec[None], t.q_spec[None], t.vm[None], t.va[None],
                                                       self._tol, 0, self._mode, init="linear")
        return [float(v) for v in vm[0]], [float(v) for v in va[0]]

    def zip_term_counts(self) -> List[int]:
        """Voltage-dependent ZIP terms per bus (src/python.rs:412-419)."""
        return self._model.zip_term_counts()

    def voltage_mag(self) -> List[float]:
        return [float(v) for v in self._vm]

    def voltage_ang(self) -> List[float]:
        return [float(v) for v in self._va]

    def voltage_kv(self) -> List[float]:
        """Line-to-line voltage in kV: ``voltage_mag * u_rated / 1e3`` (src/python.rs:431-437)."""
        return [float(v) for v in self._vm * self._model.u_rated / 1e3]

    # ---- PGM node ids (no counterpart: the reference's callers keep node_id_to_idx themselves)
    def node_index(self, node_id: int) -> int:
        """Bus index of a PGM node id (ascending-id numbering, src/pgm.rs:261-265)."""
        return self._model.id_to_idx[int(node_id)]
