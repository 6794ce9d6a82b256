"""Network model, the synthetic meshed-grid generator and the scenario load model.

The network model mirrors the reference's input types: ``Bus`` (``src/types.rs:27-43``),
``Line`` (``src/types.rs:45-54``) and ``Transformer`` (``src/types.rs:60-69``), held here as
structure-of-arrays numpy columns because the device wants columns, not records.

Bus types use the reference's three kinds (``src/types.rs:5-9``) with the integer codes the C ABI
takes: ``SLACK = 0``, ``PV = 1``, ``PQ = 2``.

The generator is the meshed test grid the benchmark workloads are defined on:

* bus ``i >= 1`` links to a uniformly drawn bus in ``[max(0, i-20), i-1]`` (a spanning tree),
  plus ``n_chords`` extra branches between uniform pairs with ``|i-j| <= 50`` that duplicate no
  existing branch;
* every branch has ``r ~ U[0.002, 0.02]``, ``x = r * U[5, 12]``, total charging
  ``b ~ U[0, 0.05]``; 10% of branches (uniformly chosen) are transformers with a real tap
  ``k ~ U[0.95, 1.05]`` and no phase shift;
* bus 0 is the slack at ``1.0∠0``; 15% of the other buses are PV with ``Vset ~ U[1.00, 1.05]``;
* every PQ bus draws ``P ~ U[0.05, 0.5]`` pu of load on a 100 MVA base and ``Q = P * U[0.1, 0.4]``;
* PV generation totals ``0.95 *`` the total load, split in proportion to ``U[0.5, 1.5]`` weights.

A transformer is stamped as the reference's ``branch_calc_param`` (1,1) branch
(``src/network.rs:339-360``) with ``y_series = 1/(r + jx)``, ``y_shunt = j*b`` and ``tap = k + 0j``.

The scenario model scales each PQ load by ``g * m_i`` with a global ``g ~ U[0.7, 1.2]`` and per-bus
``m_i ~ LogNormal(0, 0.1)``, then rescales PV generation to 0.95 of that scenario's total load with
the grid's own weights.  A scenario is expressed the reference's way, as ``BusOverride`` values of
``p_spec``/``q_spec`` for every loaded or PV bus (``src/batch.rs:39-68``).
"""
from __future__ import annotations

import dataclasses
import json
from typing import Optional

import numpy as np

SLACK, PV, PQ = 0, 1, 2
_TYPE_NAMES = {SLACK: "Slack", PV: "PV", PQ: "PQ"}
_TYPE_CODES = {v: k for k, v in _TYPE_NAMES.items()}
_ZIP_KINDS = ("ConstPower", "ConstCurrent", "ConstImpedance")  # ZipKind (src/types.rs:14-18)


@dataclasses.dataclass
class Network:
    """One topology plus a template operating point (the reference's ``&[Bus]`` template)."""

    bus_type: np.ndarray          # int32 [n_bus]
    vm: np.ndarray                # float64 [n_bus] voltage_mag (initial guess / setpoint)
    va: np.ndarray                # float64 [n_bus] voltage_ang
    p_spec: np.ndarray            # float64 [n_bus]
    q_spec: np.ndarray            # float64 [n_bus]
    line_from: np.ndarray         # int32 [n_line]
    line_to: np.ndarray
    line_r: np.ndarray            # float64
    line_x: np.ndarray
    line_b: np.ndarray            # total line charging
    line_g: np.ndarray            # total shunt conductance
    tr_from: np.ndarray           # int32 [n_tr]
    tr_to: np.ndarray
    tr_ys: np.ndarray             # complex128 series admittance
    tr_ysh: np.ndarray            # complex128 shunt admittance (split half to each end)
    tr_tap: np.ndarray            # complex128 tap ratio k*exp(j*shift)
    # ZIP terms (Bus::zip_terms, src/types.rs:14-24, 40-41) summed per bus and kind: [n_bus, 6] =
    # (P, Q) of the constant-power, constant-current and constant-impedance terms; None when no bus has any
    zip_coef: Optional[np.ndarray] = None
    # generator bookkeeping used by the scenario model (not part of the reference's types)
    load_p: Optional[np.ndarray] = None   # float64 [n_bus] load P at the template (PQ buses)
    load_q: Optional[np.ndarray] = None
    pv_weight: Optional[np.ndarray] = None  # float64 [n_bus] generation weights (PV buses)

    @property
    def n_bus(self) -> int:
        return int(self.bus_type.shape[0])

    @property
    def n_branch(self) -> int:
        return int(self.line_from.shape[0] + self.tr_from.shape[0])

    def copy(self) -> "Network":
        return dataclasses.replace(self, **{f.name: (None if getattr(self, f.name) is None else np.array(getattr(self, f.name), copy=True))
                                            for f in dataclasses.fields(self)})

    # ---- JSON (the reference's native {buses, lines} layout, extended with transformers) ----
    def to_json(self) -> str:
        buses = []
        for i in range(self.n_bus):
            bus = {"idx": i, "bus_type": _TYPE_NAMES[int(self.bus_type[i])],
                   "voltage_mag": float(self.vm[i]), "voltage_ang": float(self.va[i]),
                   "p_spec": float(self.p_spec[i]), "q_spec": float(self.q_spec[i]),
                   "q_min": -999.0, "q_max": 999.0}
            if self.zip_coef is not None:
                z = self.zip_coef[i]
                bus["zip_terms"] = [{"s_const": [float(z[2 * k]), float(z[2 * k + 1])], "kind": _ZIP_KINDS[k]}
                                    for k in range(3) if z[2 * k] != 0.0 or z[2 * k + 1] != 0.0]
            buses.append(bus)
        lines = [{"from": int(f), "to": int(t), "r": float(r), "x": float(x), "b_shunt": float(b), "g_shunt": float(g)}
                 for f, t, r, x, b, g in zip(self.line_from, self.line_to, self.line_r, self.line_x, self.line_b, self.line_g)]
        trs = [{"from": int(f), "to": int(t), "from_status": 1, "to_status": 1,
                "y_series": [float(ys.real), float(ys.imag)], "y_shunt": [float(ysh.real), float(ysh.imag)],
                "tap": [float(tp.real), float(tp.imag)]}
               for f, t, ys, ysh, tp in zip(self.tr_from, self.tr_to, self.tr_ys, self.tr_ysh, self.tr_tap)]
        return json.dumps({"buses": buses, "lines": lines, "transformers": trs})

    @staticmethod
    def from_json(text: str) -> "Network":
        d = json.loads(text)
        buses = sorted(d["buses"], key=lambda b: b["idx"])
        if [b["idx"] for b in buses] != list(range(len(buses))):
            raise ValueError("bus idx values must be 0..n-1")
        lines = d.get("lines", [])
        trs = d.get("transformers", [])
        for t in trs:
            if t.get("from_status", 1) != 1 or t.get("to_status", 1) != 1:
                raise ValueError("only closed (1,1) transformer branches are supported")
        c = lambda v: complex(v[0], v[1])
        zip_coef = None
        if any(b.get("zip_terms") for b in buses):
            zip_coef = np.zeros((len(buses), 6))
            for i, b in enumerate(buses):
                for term in b.get("zip_terms", []):
                    if term["kind"] not in _ZIP_KINDS:
                        raise ValueError(f"bus {i}: unknown ZIP kind {term['kind']!r}")
                    k = _ZIP_KINDS.index(term["kind"])
                    zip_coef[i, 2 * k] += float(term["s_const"][0])
                    zip_coef[i, 2 * k + 1] += float(term["s_const"][1])
        return Network(
            bus_type=np.array([_TYPE_CODES[b["bus_type"]] for b in buses], dtype=np.int32),
            vm=np.array([b["voltage_mag"] for b in buses], dtype=np.float64),
            va=np.array([b["voltage_ang"] for b in buses], dtype=np.float64),
            p_spec=np.array([b["p_spec"] for b in buses], dtype=np.float64),
            q_spec=np.array([b["q_spec"] for b in buses], dtype=np.float64),
            line_from=np.array([l["from"] for l in lines], dtype=np.int32),
            line_to=np.array([l["to"] for l in lines], dtype=np.int32),
            line_r=np.array([l["r"] for l in lines], dtype=np.float64),
            line_x=np.array([l["x"] for l in lines], dtype=np.float64),
            line_b=np.array([l.get("b_shunt", 0.0) for l in lines], dtype=np.float64),
            line_g=np.array([l.get("g_shunt", 0.0) for l in lines], dtype=np.float64),
            tr_from=np.array([t["from"] for t in trs], dtype=np.int32),
            tr_to=np.array([t["to"] for t in trs], dtype=np.int32),
            tr_ys=np.array([c(t["y_series"]) for t in trs], dtype=np.complex128),
            tr_ysh=np.array([c(t["y_shunt"]) for t in trs], dtype=np.complex128),
            tr_tap=np.array([c(t["tap"]) for t in trs], dtype=np.complex128),
            zip_coef=zip_coef,
        )


def generate_grid(n_bus: int, n_chords: int, seed: int) -> Network:
    """The meshed test grid described in the module docstring, deterministic in ``seed``."""
    if n_bus < 2:
        raise ValueError("need at least two buses")
    rng = np.random.default_rng(seed)
    edges = set()
    frm, to = [], []
    for i in range(1, n_bus):
        j = int(rng.integers(max(0, i - 20), i))  # uniform in [max(0,i-20), i-1]
        frm.append(j)
        to.append(i)
        edges.add((j, i))
    max_chords = sum(min(50, n_bus - 1 - a) for a in range(n_bus)) - len(edges)
    if n_chords > max_chords:
        raise ValueError("too many chords for this bus count")
    added = 0
    while added < n_chords:
        a = int(rng.integers(0, n_bus))
        b = int(rng.integers(max(0, a - 50), min(n_bus - 1, a + 50) + 1))
        if a == b:
            continue
        key = (min(a, b), max(a, b))
        if key in edges:
            continue
        edges.add(key)
        frm.append(key[0])
        to.append(key[1])
        added += 1
    nbr = len(frm)
    frm = np.array(frm, dtype=np.int32)
    to = np.array(to, dtype=np.int32)
    r = rng.uniform(0.002, 0.02, nbr)
    x = r * rng.uniform(5.0, 12.0, nbr)
    b = rng.uniform(0.0, 0.05, nbr)
    is_tr = np.zeros(nbr, dtype=bool)
    is_tr[rng.choice(nbr, size=int(round(0.1 * nbr)), replace=False)] = True
    tap = rng.uniform(0.95, 1.05, nbr)

    bus_type = np.full(n_bus, PQ, dtype=np.int32)
    bus_type[0] = SLACK
    n_pv = int(round(0.15 * (n_bus - 1)))
    pv = rng.choice(np.arange(1, n_bus), size=n_pv, replace=False)
    bus_type[pv] = PV
    vm = np.ones(n_bus)
    va = np.zeros(n_bus)
    vm[pv] = rng.uniform(1.00, 1.05, n_pv)
    load_p = np.zeros(n_bus)
    load_q = np.zeros(n_bus)
    pq = np.nonzero(bus_type == PQ)[0]
    load_p[pq] = rng.uniform(0.05, 0.5, pq.size)
    load_q[pq] = load_p[pq] * rng.uniform(0.1, 0.4, pq.size)
    w = np.zeros(n_bus)
    w[pv] = rng.uniform(0.5, 1.5, n_pv)
    p_spec = -load_p.copy()
    q_spec = -load_q.copy()
    if n_pv:
        p_spec[pv] = 0.95 * load_p.sum() * w[pv] / w[pv].sum()

    ys = 1.0 / (r + 1j * x)
    return Network(
        bus_type=bus_type, vm=vm, va=va, p_spec=p_spec, q_spec=q_spec,
        line_from=frm[~is_tr], line_to=to[~is_tr], line_r=r[~is_tr], line_x=x[~is_tr],
        line_b=b[~is_tr], line_g=np.zeros(int((~is_tr).sum())),
        tr_from=frm[is_tr], tr_to=to[is_tr], tr_ys=ys[is_tr], tr_ysh=1j * b[is_tr],
        tr_tap=tap[is_tr] + 0j,
        load_p=load_p, load_q=load_q, pv_weight=w,
    )


def scenario_specs(net: Network, n_scen: int, seed: int):
    """Per-scenario ``(p_spec, q_spec)`` arrays, shape ``[n_scen, n_bus]``, for the load model in the
    module docstring.  Scenario ``s`` draws from its own generator seeded ``(seed, s)``, so a scenario's
    values do not depend on how many scenarios are requested."""
    n = net.n_bus
    pq = net.bus_type == PQ
    pv = net.bus_type == PV
    p = np.empty((n_scen, n))
    q = np.empty((n_scen, n))
    wsum = net.pv_weight[pv].sum() if pv.any() else 1.0
    for s in range(n_scen):
        rng = np.random.default_rng([seed, s])
        g = rng.uniform(0.7, 1.2)
        m = rng.lognormal(0.0, 0.1, n)
        lp = net.load_p * g * m
        lq = net.load_q * g * m
        p[s] = np.where(pq, -lp, 0.0)
        q[s] = np.where(pq, -lq, 0.0)
        if pv.any():
            p[s, pv] = 0.95 * lp[pq].sum() * net.pv_weight[pv] / wsum
    # slack and untouched buses keep the template's own values
    other = ~(pq | pv)
    p[:, other] = net.p_spec[other]
    q[:, other] = net.q_spec[other]
    q[:, pv] = net.q_spec[pv]
    return p, q


def scenario_loads(net: Network, n_scen: int, seed: int):
    """Per-scenario ``(p_spec, q_spec)`` drawn from the grid generator's own load distribution
    (PQ loads P~U[0.05,0.5] pu, Q = P*U[0.1,0.4]; PV generation 0.95 of the total load split in
    proportion to fresh U[0.5,1.5] weights), the scenario model of the mixed-precision, uneven-batch
    and serving workloads.  Scenario ``s`` draws from its own generator seeded ``(seed, s)``."""
    n = net.n_bus
    pq = np.nonzero(net.bus_type == PQ)[0]
    pv = np.nonzero(net.bus_type == PV)[0]
    p = np.tile(net.p_spec, (n_scen, 1))
    q = np.tile(net.q_spec, (n_scen, 1))
    for s in range(n_scen):
        rng = np.random.default_rng([seed, s])
        lp = rng.uniform(0.05, 0.5, pq.size)
        lq = lp * rng.uniform(0.1, 0.4, pq.size)
        p[s, pq] = -lp
        q[s, pq] = -lq
        if pv.size:
            w = rng.uniform(0.5, 1.5, pv.size)
            p[s, pv] = 0.95 * lp.sum() * w / w.sum()
    return p, q


def without_branch(net: Network, branch: int) -> Network:
    """The network with branch ``branch`` (lines first, then transformers) removed: the reference
    view of an N-1 scenario, whose Y-bus equals the base one minus that branch's stamps."""
    import dataclasses
    nl = net.line_from.size
    if branch < nl:
        keep = np.arange(nl) != branch
        return dataclasses.replace(net, line_from=net.line_from[keep], line_to=net.line_to[keep],
                                   line_r=net.line_r[keep], line_x=net.line_x[keep], line_b=net.line_b[keep],
                                   line_g=net.line_g[keep])
    keep = np.arange(net.tr_from.size) != branch - nl
    return dataclasses.replace(net, tr_from=net.tr_from[keep], tr_to=net.tr_to[keep], tr_ys=net.tr_ys[keep],
                               tr_ysh=net.tr_ysh[keep], tr_tap=net.tr_tap[keep])


def config_grid(config: int, seed: int = 20261203) -> Network:
    """The grid of workload ``config`` (1: 200 buses, 2: 2,000 buses, 3: 10,000 buses)."""
    sizes = {1: (200, 100), 2: (2000, 1000), 3: (10000, 5000)}
    n, c = sizes[config]
    return generate_grid(n, c, seed + config)


def load_reference_network_json(text: str) -> Network:
    """Parses the reference's native JSON (``src/json.rs:4-8``: ``{buses, lines}``, no transformers)."""
    return Network.from_json(text)
