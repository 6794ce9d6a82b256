"""Power Grid Model (PGM) JSON input -> the bus/branch network the solver consumes.

Restates the reference's conversion (``src/pgm.rs``, ``src/network.rs``) on the host, where it runs
once per topology; nothing here is on the per-iteration path.

* every PGM ``node`` becomes a PQ bus, numbered in ascending node id (``node_id_to_idx``,
  ``src/pgm.rs:261-265``);
* active loads (negative) and generators (positive) add their injection ``/ s_base`` to their node;
  ``asym_*`` components add the sum of their three phases; type 0 is constant power, type 1 becomes a
  constant-impedance ZIP term (S proportional to |V|^2) and type 2 a constant-current one (S
  proportional to |V|) (``src/pgm.rs:359-399``);
* an active ``voltage_regulator`` turns its generator's node into a PV bus at ``u_ref``
  (``src/pgm.rs:426-440``);
* lines are taken to per unit on the ``from`` node's voltage base; a line open at one end leaves its
  full charging as a shunt at the closed end; a line open at both ends is dropped
  (``src/pgm.rs:441-474``);
* two-winding transformers get the complex tap of ``transformer_tap`` and the series/shunt admittance
  of ``transformer_admittances`` (``src/pgm.rs:475-490``, ``src/network.rs:519-563``);
* each three-winding transformer adds a star bus (after the nodes) and three transformer legs
  (``src/pgm.rs:492-561``, ``src/network.rs:572-594``);
* each active source adds a Slack bus at ``u_ref`` (after the star buses) joined to its node by the
  source impedance (``src/pgm.rs:563-580``, ``src/network.rs:495-499``);
* active ``shunt`` components are self-admittances on the diagonal (``src/pgm.rs:270-284``).

Branches are expressed in the solver's ``Network`` columns.  Shunts of every kind (``shunt``
components, charging of half-open lines, the connected end of a half-open transformer) are self-loop
lines, which the Y-bus builder stamps on the diagonal alone (``build_ybus``,
``src/network.rs:228-232``).  A transformer open at one or both ends still stamps its four Y-bus
positions in the reference (``stamp_transformers``, ``src/network.rs:365-375``), with zeros where
``branch_calc_param`` (``src/network.rs:339-360``) gives none; it is kept here as a transformer with
zero admittance so the Y-bus pattern (and so the connected components and the Jacobian pattern) is
the reference's, while its effective shunt goes in as a self-loop line.
"""
from __future__ import annotations

import cmath
import dataclasses
import json
import math
from typing import Dict, List, Optional, Tuple

import numpy as np

from .grid import PQ, PV, SLACK, Network

# PGM load/generator type codes (src/pgm.rs:78-80)
CONST_POWER, CONST_IMPEDANCE, CONST_CURRENT = 0, 1, 2


class PgmError(ValueError):
    """A PGM document that cannot be converted (missing fields, unknown node ids)."""


@dataclasses.dataclass
class PgmModel:
    """A converted PGM document: the network (buses in the reference's order: nodes by ascending
    id, then three-winding star buses, then one Slack bus per active source) plus the per-bus data
    the solver itself does not need."""

    network: Network
    node_ids: np.ndarray          # int64 [n_nodes] PGM node id of bus i (ascending)
    u_rated: np.ndarray           # float64 [n_bus] rated line-to-line voltage in V
    zip_terms: List[List[Tuple[int, complex]]]  # per bus: (CONST_IMPEDANCE | CONST_CURRENT, S pu)

    @property
    def id_to_idx(self) -> Dict[int, int]:
        return {int(i): k for k, i in enumerate(self.node_ids)}

    def zip_term_counts(self) -> List[int]:
        return [len(t) for t in self.zip_terms]


# ---------------------------------------------------------------- reference helpers (src/network.rs)
def tap_ratio_from_voltages(u_num: float, u_denom: float, clock: int) -> complex:
    """``(u_num / u_denom) * exp(j * clock * pi / 6)`` (src/network.rs:305-307)."""
    return cmath.rect(u_num / u_denom, clock * math.pi / 6.0)


def transformer_tap(u1, u2, tap_side, tap_pos, tap_min, tap_max, tap_nom, tap_size, clock) -> complex:
    """Off-nominal complex tap; the tap position is clamped to the changer's range first
    (src/network.rs:519-530)."""
    pos = min(max(tap_pos, min(tap_min, tap_max)), max(tap_min, tap_max))
    delta = (pos - tap_nom) * tap_size
    if tap_side == 0:
        return tap_ratio_from_voltages(u1 + delta, u1, clock)
    return tap_ratio_from_voltages(u2, u2 + delta, clock)


def transformer_admittances_ex(u_field, u_base, sn, uk, pk, i0, p0, s_base_va) -> Tuple[complex, complex]:
    """Per-unit series and magnetising admittance on the ``u_base`` side; the sign of ``uk`` is kept
    on the series reactance (src/network.rs:542-558)."""
    base_y = s_base_va / (u_base * u_base)
    sign = 1.0 if uk >= 0.0 else -1.0
    z_abs = abs(uk) * u_field * u_field / sn
    r = pk * u_field * u_field / (sn * sn)
    x2 = z_abs * z_abs - r * r
    x = sign * (math.sqrt(x2) if x2 > 0.0 else 0.0)
    y_series = (1.0 + 0j) / complex(r, x) / base_y
    g = p0 / (u_field * u_field)
    y_abs = i0 * sn / (u_field * u_field)
    b2 = y_abs * y_abs - g * g
    b = -math.sqrt(b2) if b2 > 0.0 else 0.0
    return y_series, complex(g, b) / base_y


def transformer_admittances(u2, sn, uk, pk, i0, p0, s_base_va) -> Tuple[complex, complex]:
    """src/network.rs:560-563."""
    return transformer_admittances_ex(u2, u2, sn, uk, pk, i0, p0, s_base_va)


def three_winding_star_params(sn_1, sn_2, sn_3, uk_12, uk_13, uk_23, pk_12, pk_13, pk_23):
    """Delta -> star short-circuit voltages and losses on side 1's power base
    (src/network.rs:572-594)."""
    uk12 = uk_12 * sn_1 / min(sn_1, sn_2)
    uk13 = uk_13 * sn_1 / min(sn_1, sn_3)
    uk23 = uk_23 * sn_1 / min(sn_2, sn_3)
    uk = (0.5 * (uk12 + uk13 - uk23),
          0.5 * (uk12 + uk23 - uk13) * (sn_2 / sn_1),
          0.5 * (uk13 + uk23 - uk12) * (sn_3 / sn_1))
    pk12 = pk_12 * (sn_1 / min(sn_1, sn_2)) ** 2
    pk13 = pk_13 * (sn_1 / min(sn_1, sn_3)) ** 2
    pk23 = pk_23 * (sn_1 / min(sn_2, sn_3)) ** 2
    pk = (0.5 * (pk12 + pk13 - pk23),
          0.5 * (pk12 + pk23 - pk13) * (sn_2 / sn_1) ** 2,
          0.5 * (pk13 + pk23 - pk12) * (sn_3 / sn_1) ** 2)
    return uk, pk


def source_impedance_pu(sk: float, rx_ratio: float, s_base_va: float) -> Tuple[float, float]:
    """(r, x) of a source's internal impedance, |z| = s_base / sk (src/network.rs:495-499)."""
    z = s_base_va / sk
    x = z / math.sqrt(rx_ratio * rx_ratio + 1.0)
    return rx_ratio * x, x


def half_open_branch_shunt(y_series: complex, y_shunt: complex) -> complex:
    """Shunt seen at the closed end of a branch open at the other end; zero without a magnetising
    branch (src/network.rs:320-327)."""
    if y_shunt == 0:
        return 0j
    return y_shunt * 0.5 + 1.0 / (1.0 / y_series + 2.0 / y_shunt)


# ---------------------------------------------------------------- document access
def _get(obj: dict, key: str, comp: str):
    try:
        return obj[key]
    except KeyError:
        raise PgmError(f"{comp} entry {obj.get('id', '?')}: missing field '{key}'") from None


def _data(doc: dict) -> dict:
    if "data" not in doc or not isinstance(doc["data"], dict):
        raise PgmError("PGM input must be an object with a 'data' object")
    d = doc["data"]
    for req in ("node", "source"):
        if req not in d:
            raise PgmError(f"PGM input has no '{req}' list")
    return d


def apply_batch_update(doc: dict, update: dict) -> dict:
    """One scenario of a PGM batch update merged into a copy of the input document: every updated
    component is found by id and its listed attributes replaced (tests/common/mod.rs:35-57)."""
    out = json.loads(json.dumps(doc))
    data = out["data"]
    for comp, entries in update.items():
        if comp not in data:
            raise PgmError(f"update names component '{comp}', which the input does not have")
        by_id = {e["id"]: e for e in data[comp]}
        for upd in entries:
            if upd["id"] not in by_id:
                raise PgmError(f"no {comp} entry with id {upd['id']} in the input")
            by_id[upd["id"]].update({k: v for k, v in upd.items() if k != "id"})
    return out


# ---------------------------------------------------------------- conversion
def from_pgm(doc: dict, s_base_va: float = 1e6, freq_hz: float = 50.0) -> PgmModel:
    """Converts a parsed PGM input document (``pgm_to_buses_and_branches`` + ``pgm_shunts_1ph``,
    src/pgm.rs:347-583 and 270-284)."""
    d = _data(doc)
    nodes = d["node"]
    ids = sorted(int(_get(n, "id", "node")) for n in nodes)
    if len(set(ids)) != len(ids):
        raise PgmError("duplicate node id")
    idx = {i: k for k, i in enumerate(ids)}
    u_node = {int(n["id"]): float(_get(n, "u_rated", "node")) for n in nodes}

    def node_of(e, key, comp):
        nid = int(_get(e, key, comp))
        if nid not in idx:
            raise PgmError(f"{comp} {e.get('id', '?')}: unknown node {nid}")
        return nid

    n_nodes = len(ids)
    p_inj: Dict[int, float] = {}
    q_inj: Dict[int, float] = {}
    zips: Dict[int, List[Tuple[int, complex]]] = {}

    def accumulate(node, kind, p, q, sign):
        s = complex(p, q) * sign / s_base_va
        if kind == CONST_IMPEDANCE or kind == CONST_CURRENT:
            zips.setdefault(node, []).append((kind, s))
        else:
            p_inj[node] = p_inj.get(node, 0.0) + s.real
            q_inj[node] = q_inj.get(node, 0.0) + s.imag

    for comp, sign, asym in (("sym_load", -1.0, False), ("sym_gen", 1.0, False), ("asym_load", -1.0, True),
                             ("asym_gen", 1.0, True)):
        for e in d.get(comp, []):
            if int(_get(e, "status", comp)) == 0:
                continue
            p, q = _get(e, "p_specified", comp), _get(e, "q_specified", comp)
            if asym:
                p, q = sum(float(v) for v in p), sum(float(v) for v in q)
            accumulate(node_of(e, "node", comp), int(e.get("type", CONST_POWER)), float(p), float(q), sign)

    bus_type = [PQ] * n_nodes
    vm = [1.0] * n_nodes
    p_spec = [p_inj.get(i, 0.0) for i in ids]
    q_spec = [q_inj.get(i, 0.0) for i in ids]
    u_rated = [u_node[i] for i in ids]
    zip_terms = [zips.get(i, []) for i in ids]

    gen_node = {int(g["id"]): int(g["node"]) for g in d.get("sym_gen", [])}
    for vr in d.get("voltage_regulator", []):
        if int(_get(vr, "status", "voltage_regulator")) == 0:
            continue
        node = gen_node.get(int(_get(vr, "regulated_object", "voltage_regulator")))
        if node is None:
            continue
        b = idx[node]
        if bus_type[b] == PQ:
            bus_type[b] = PV
            vm[b] = float(_get(vr, "u_ref", "voltage_regulator"))

    lf, lt, lr, lx, lb, lg = [], [], [], [], [], []

    def line(f, t, r, x, b, g=0.0):
        lf.append(f), lt.append(t), lr.append(r), lx.append(x), lb.append(b), lg.append(g)

    omega = 2.0 * math.pi * freq_hz
    for e in d.get("line", []):
        fs, ts = int(_get(e, "from_status", "line")), int(_get(e, "to_status", "line"))
        fn, tn = node_of(e, "from_node", "line"), node_of(e, "to_node", "line")
        c1 = float(_get(e, "c1", "line"))
        if fs == 1 and ts == 1:
            zb = u_node[fn] ** 2 / s_base_va
            line(idx[fn], idx[tn], float(_get(e, "r1", "line")) / zb, float(_get(e, "x1", "line")) / zb,
                 omega * c1 * zb)
        elif fs == 1 and ts == 0:
            zb = u_node[fn] ** 2 / s_base_va
            line(idx[fn], idx[fn], 0.0, 0.0, omega * c1 * zb)
        elif fs == 0 and ts == 1:
            zb = u_node[tn] ** 2 / s_base_va
            line(idx[tn], idx[tn], 0.0, 0.0, omega * c1 * zb)

    tf, tt, tys, tysh, ttap = [], [], [], [], []
    shunt_loops = []  # (bus, y) of half-open transformers, stamped after the lines

    def transformer(f, t, fs, ts, ys, ysh, tap):
        if fs == 1 and ts == 1:
            tf.append(f), tt.append(t), tys.append(ys), tysh.append(ysh), ttap.append(tap)
            return
        # the reference still stamps the four positions (zeros where branch_calc_param gives none)
        tf.append(f), tt.append(t), tys.append(0j), tysh.append(0j), ttap.append(1 + 0j)
        if fs == 1 and ts == 0:
            shunt_loops.append((f, half_open_branch_shunt(ys, ysh) / (abs(tap) ** 2)))
        elif fs == 0 and ts == 1:
            shunt_loops.append((t, half_open_branch_shunt(ys, ysh)))

    def tr_field(e, k, comp="transformer"):
        return _get(e, k, comp)

    for e in d.get("transformer", []):
        tap = transformer_tap(float(tr_field(e, "u1")), float(tr_field(e, "u2")), int(tr_field(e, "tap_side")),
                              int(tr_field(e, "tap_pos")), int(tr_field(e, "tap_min")), int(tr_field(e, "tap_max")),
                              int(tr_field(e, "tap_nom")), float(tr_field(e, "tap_size")), int(tr_field(e, "clock")))
        ys, ysh = transformer_admittances(float(tr_field(e, "u2")), float(tr_field(e, "sn")), float(tr_field(e, "uk")),
                                          float(tr_field(e, "pk")), float(tr_field(e, "i0")), float(tr_field(e, "p0")),
                                          s_base_va)
        transformer(idx[node_of(e, "from_node", "transformer")], idx[node_of(e, "to_node", "transformer")],
                    int(tr_field(e, "from_status")), int(tr_field(e, "to_status")), ys, ysh, tap)

    three = d.get("three_winding_transformer", [])
    for i, e in enumerate(three):
        c = "three_winding_transformer"
        g = lambda k: _get(e, k, c)  # noqa: E731
        star = n_nodes + i
        n1, n2, n3 = node_of(e, "node_1", c), node_of(e, "node_2", c), node_of(e, "node_3", c)
        u1r, u2r, u3r = u_node[n1], u_node[n2], u_node[n3]
        direction = 1.0 if int(g("tap_max")) > int(g("tap_min")) else -1.0
        delta = direction * (int(g("tap_pos")) - int(g("tap_nom"))) * float(g("tap_size"))
        u1, u2, u3 = float(g("u1")), float(g("u2")), float(g("u3"))
        side = int(g("tap_side"))
        if side == 0:
            u1 += delta
        elif side == 1:
            u2 += delta
        else:
            u3 += delta
        (uk1, uk2, uk3), (pk1, pk2, pk3) = three_winding_star_params(
            float(g("sn_1")), float(g("sn_2")), float(g("sn_3")), float(g("uk_12")), float(g("uk_13")),
            float(g("uk_23")), float(g("pk_12")), float(g("pk_13")), float(g("pk_23")))
        bus_type.append(PQ), vm.append(1.0), p_spec.append(0.0), q_spec.append(0.0), u_rated.append(u1r)
        zip_terms.append([])
        ys1, ysh1 = transformer_admittances_ex(u1, u1r, float(g("sn_1")), uk1, pk1, float(g("i0")), float(g("p0")),
                                               s_base_va)
        transformer(idx[n1], star, int(g("status_1")), 1, ys1, ysh1, tap_ratio_from_voltages(u1 * u1r, u1 * u1r, 0))
        ys2, ysh2 = transformer_admittances_ex(u1, u1r, float(g("sn_2")), uk2, pk2, 0.0, 0.0, s_base_va)
        transformer(idx[n2], star, int(g("status_2")), 1, ys2, ysh2,
                    tap_ratio_from_voltages(u2 * u1r, u1 * u2r, 12 - int(g("clock_12"))))
        ys3, ysh3 = transformer_admittances_ex(u1, u1r, float(g("sn_3")), uk3, pk3, 0.0, 0.0, s_base_va)
        transformer(idx[n3], star, int(g("status_3")), 1, ys3, ysh3,
                    tap_ratio_from_voltages(u3 * u1r, u1 * u3r, 12 - int(g("clock_13"))))

    n_src = 0
    for e in d["source"]:
        if int(_get(e, "status", "source")) == 0:
            continue
        v = n_nodes + len(three) + n_src
        n_src += 1
        node = node_of(e, "node", "source")
        r, x = source_impedance_pu(float(_get(e, "sk", "source")), float(_get(e, "rx_ratio", "source")), s_base_va)
        bus_type.append(SLACK), vm.append(float(_get(e, "u_ref", "source"))), p_spec.append(0.0)
        q_spec.append(0.0), u_rated.append(u_node[node]), zip_terms.append([])
        line(v, idx[node], r, x, 0.0)

    for e in d.get("shunt", []):
        if int(_get(e, "status", "shunt")) == 0:
            continue
        node = node_of(e, "node", "shunt")
        y = complex(float(_get(e, "g1", "shunt")), float(_get(e, "b1", "shunt"))) / (s_base_va / u_node[node] ** 2)
        line(idx[node], idx[node], 0.0, 0.0, y.imag, y.real)
    for b, y in shunt_loops:
        line(b, b, 0.0, 0.0, y.imag, y.real)

    n_bus = len(bus_type)
    zip_coef = None
    if any(zip_terms):
        zip_coef = np.zeros((n_bus, 6))
        for b, terms in enumerate(zip_terms):
            for kind, s in terms:  # columns: constant power (none from PGM), current, impedance
                k = 2 if kind == CONST_CURRENT else 4
                zip_coef[b, k] += s.real
                zip_coef[b, k + 1] += s.imag
    net = Network(
        bus_type=np.array(bus_type, np.int32), vm=np.array(vm, np.float64), va=np.zeros(n_bus),
        p_spec=np.array(p_spec, np.float64), q_spec=np.array(q_spec, np.float64),
        line_from=np.array(lf, np.int32), line_to=np.array(lt, np.int32), line_r=np.array(lr, np.float64),
        line_x=np.array(lx, np.float64), line_b=np.array(lb, np.float64), line_g=np.array(lg, np.float64),
        tr_from=np.array(tf, np.int32), tr_to=np.array(tt, np.int32), tr_ys=np.array(tys, np.complex128),
        tr_ysh=np.array(tysh, np.complex128), tr_tap=np.array(ttap, np.complex128), zip_coef=zip_coef)
    return PgmModel(net, np.array(ids, np.int64), np.array(u_rated, np.float64), zip_terms)


def load_pgm_json(path: str, s_base_va: float = 1e6, freq_hz: float = 50.0) -> PgmModel:
    """Reads and converts a PGM input file."""
    with open(path) as f:
        try:
            doc = json.load(f)
        except json.JSONDecodeError as e:
            raise PgmError(f"parsing {path} as PGM JSON: {e}") from None
    return from_pgm(doc, s_base_va, freq_hz)
