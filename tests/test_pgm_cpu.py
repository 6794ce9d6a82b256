"""PGM ingest and ZIP injections on the CPU: the reference's own golden PGM power-flow fixtures
(tests/klu_native_jacobian_test.rs, tests/pgm_*_test.rs, with their tolerances) solved by the oracle
from the converted networks, the conversion's closed forms, and the PowerFlowModel surface that needs
no GPU."""
import json
import math
import os

import numpy as np
import pytest

from gridoxide_amd import pgm
from gridoxide_amd.grid import PQ, SLACK, Network, generate_grid
from gridoxide_amd.model import PowerFlowModel

# (case, tolerance) as in the reference's tests
SINGLE = [("basic-node", 1e-5), ("node", 1e-5), ("line", 1e-5), ("shunt", 1e-5), ("source", 1e-5),
          ("sym_load", 1e-5), ("sym_gen", 1e-5), ("transmission-case", 1e-5)]
BATCH = [("transformer", 1e-4), ("distribution-case", 1e-5), ("three_winding_transformer", 1e-4)]


def pgm_dir(golden_dir, case):
    return os.path.join(golden_dir, "pgm", case)


def load(golden_dir, case, name):
    with open(os.path.join(pgm_dir(golden_dir, case), name)) as f:
        return json.load(f)


def check_nodes(model, vm, va, expected_nodes, tol):
    idx = model.id_to_idx
    for nd in expected_nodes:
        i = idx[nd["id"]]
        assert abs(vm[i] - nd["u_pu"]) < tol, (nd["id"], vm[i], nd["u_pu"])
        assert abs(va[i] - nd["u_angle"]) < tol, (nd["id"], va[i], nd["u_angle"])


def oracle_solve(oracle_mod, model):
    net = model.network
    o = oracle_mod.Oracle(net)
    return o.persistent_solve(net.p_spec, net.q_spec, net.vm, net.va, tol=1e-6, max_iter=20)


@pytest.mark.parametrize("case,tol", SINGLE)
def test_golden_single(oracle_mod, golden_dir, case, tol):
    model = pgm.from_pgm(load(golden_dir, case, "input.json"))
    vm, va, st, it, _ = oracle_solve(oracle_mod, model)
    assert st == 0
    check_nodes(model, vm, va, load(golden_dir, case, "sym_output.json")["data"]["node"], tol)


@pytest.mark.parametrize("case,tol", BATCH)
def test_golden_batch(oracle_mod, golden_dir, case, tol):
    doc = load(golden_dir, case, "input.json")
    updates = load(golden_dir, case, "update_batch.json")["data"]
    expected = load(golden_dir, case, "sym_output_batch.json")["data"]
    assert len(updates) == len(expected) > 1
    for upd, exp in zip(updates, expected):
        model = pgm.from_pgm(pgm.apply_batch_update(doc, upd))
        vm, va, st, _, _ = oracle_solve(oracle_mod, model)
        assert st == 0
        check_nodes(model, vm, va, exp["node"], tol)


def test_zip_terms_change_the_answer(oracle_mod, golden_dir):
    """The sym_load fixture's constant-current/-impedance loads: dropping the ZIP terms moves the
    solution off the golden one by more than the tolerance, and the converged state satisfies
    V conj(Y V) = p_spec + ZIP(|V|) on a dense restatement."""
    model = pgm.from_pgm(load(golden_dir, "sym_load", "input.json"))
    net = model.network
    assert model.zip_term_counts() == [0, 2, 0] and net.zip_coef is not None
    vm, va, st, _, _ = oracle_solve(oracle_mod, model)
    assert st == 0
    o = oracle_mod.Oracle(net)
    Y = np.zeros((net.n_bus, net.n_bus), complex)
    rows = np.repeat(np.arange(net.n_bus), np.diff(o.rowptr))
    Y[rows, o.col] = o.y
    V = vm * np.exp(1j * va)
    S = V * np.conj(Y @ V)
    z = net.zip_coef
    pe = net.p_spec + z[:, 0] + z[:, 2] * vm + z[:, 4] * vm * vm
    qe = net.q_spec + z[:, 1] + z[:, 3] * vm + z[:, 5] * vm * vm
    pq = net.bus_type == PQ
    assert np.max(np.abs(S.real[pq] - pe[pq])) < 1e-6
    assert np.max(np.abs(S.imag[pq] - qe[pq])) < 1e-6
    bare = Network(**{**net.__dict__, "zip_coef": None})
    vb, ab, stb, _, _ = oracle_mod.Oracle(bare).persistent_solve(net.p_spec, net.q_spec, net.vm, net.va, 1e-6, 20)
    assert stb == 0
    exp = load(golden_dir, "sym_load", "sym_output.json")["data"]["node"]
    worst = max(abs(vb[model.id_to_idx[nd["id"]]] - nd["u_pu"]) for nd in exp)
    assert worst > 1e-5


def test_zip_mismatch_synthetic(oracle_mod):
    """Random ZIP terms of every kind on a meshed grid: the oracle's converged state satisfies the
    dense effective-injection balance at every PV (P) and PQ (P, Q) bus."""
    net = generate_grid(60, 30, 7)
    rng = np.random.default_rng(3)
    z = np.zeros((net.n_bus, 6))
    loaded = np.nonzero(net.bus_type == PQ)[0]
    z[loaded] = rng.uniform(-0.05, 0.0, (loaded.size, 6))
    net.zip_coef = z
    o = oracle_mod.Oracle(net)
    vm, va, st, it, _ = o.batch_solve(net.p_spec, net.q_spec, net.vm, net.va, tol=1e-10, max_iter=30, init=1)
    assert st[0] == 0
    Y = np.zeros((net.n_bus, net.n_bus), complex)
    rows = np.repeat(np.arange(net.n_bus), np.diff(o.rowptr))
    Y[rows, o.col] = o.y
    V = vm[0] * np.exp(1j * va[0])
    S = V * np.conj(Y @ V)
    v = vm[0]
    pe = net.p_spec + z[:, 0] + z[:, 2] * v + z[:, 4] * v * v
    qe = net.q_spec + z[:, 1] + z[:, 3] * v + z[:, 5] * v * v
    ns = net.bus_type != SLACK
    assert np.max(np.abs(S.real[ns] - pe[ns])) < 1e-9
    assert np.max(np.abs(S.imag[net.bus_type == PQ] - qe[net.bus_type == PQ])) < 1e-9


# ---------------------------------------------------------------- conversion closed forms
def test_source_impedance():
    r, x = pgm.source_impedance_pu(1e10, 0.1, 1e6)
    assert math.isclose(math.hypot(r, x), 1e6 / 1e10, rel_tol=1e-14)
    assert math.isclose(r / x, 0.1, rel_tol=1e-14)


def test_transformer_tap_clamp_and_side():
    # tap_pos clamped into [min, max] whichever way round they are given
    a = pgm.transformer_tap(10500.0, 400.0, 0, 9, -5, 5, 0, 105.0, 0)
    assert math.isclose(a.real, (10500.0 + 5 * 105.0) / 10500.0, rel_tol=1e-15) and a.imag == 0.0
    b = pgm.transformer_tap(10500.0, 400.0, 0, -9, 5, -5, 0, 105.0, 0)
    assert math.isclose(b.real, (10500.0 - 5 * 105.0) / 10500.0, rel_tol=1e-15)
    # tap on side 2 divides; the clock rotates by 30 degrees per step
    c = pgm.transformer_tap(10500.0, 400.0, 1, 2, -5, 5, 0, 4.0, 11)
    assert math.isclose(abs(c), 400.0 / 408.0, rel_tol=1e-15)
    assert math.isclose(np.angle(c), 11 * math.pi / 6 - 2 * math.pi, rel_tol=1e-14)


def test_three_winding_star_sums():
    """With equal ratings the star legs add up to the pairwise short-circuit values."""
    (u1, u2, u3), (p1, p2, p3) = pgm.three_winding_star_params(1e6, 1e6, 1e6, 0.1, 0.12, 0.08, 3e3, 4e3, 2e3)
    assert math.isclose(u1 + u2, 0.1) and math.isclose(u1 + u3, 0.12) and math.isclose(u2 + u3, 0.08)
    assert math.isclose(p1 + p2, 3e3) and math.isclose(p1 + p3, 4e3) and math.isclose(p2 + p3, 2e3)


def test_half_open_shunt_is_kron_reduction():
    """The shunt seen at the closed end equals eliminating the floating far node of the pi model."""
    ys, ysh = 1 / (0.01 + 0.1j), 0.002 - 0.03j
    yff = ys + ysh / 2
    reduced = yff - ys * ys / (ys + ysh / 2)
    assert abs(pgm.half_open_branch_shunt(ys, ysh) - reduced) < 1e-12 * abs(reduced)
    assert pgm.half_open_branch_shunt(ys, 0j) == 0j


def test_transformer_admittance_magnitudes():
    ys, ysh = pgm.transformer_admittances(400.0, 1e6, 0.1, 1e4, 0.01, 1e3, 1e6)
    base_y = 1e6 / 400.0 ** 2
    z = 1 / (ys * base_y)  # ohms on the 400 V side
    assert math.isclose(abs(z), 0.1 * 400.0 ** 2 / 1e6, rel_tol=1e-12)
    assert math.isclose(z.real, 1e4 * 400.0 ** 2 / 1e12, rel_tol=1e-12)
    ym = ysh * base_y
    assert math.isclose(abs(ym), 0.01 * 1e6 / 400.0 ** 2, rel_tol=1e-12)
    assert math.isclose(ym.real, 1e3 / 400.0 ** 2, rel_tol=1e-12) and ym.imag < 0


def _two_node_doc(**tr):
    t = {"id": 3, "from_node": 1, "to_node": 2, "from_status": 1, "to_status": 1, "u1": 10500.0, "u2": 400.0,
         "sn": 1e6, "uk": 0.1, "pk": 1e4, "i0": 0.01, "p0": 1e3, "clock": 0, "tap_side": 0, "tap_pos": 0,
         "tap_min": -5, "tap_max": 5, "tap_nom": 0, "tap_size": 105.0}
    t.update(tr)
    return {"data": {"node": [{"id": 1, "u_rated": 10500.0}, {"id": 2, "u_rated": 400.0}], "transformer": [t],
                     "source": [{"id": 4, "node": 1, "status": 1, "u_ref": 1.0, "sk": 1e10, "rx_ratio": 0.1}]}}


@pytest.mark.parametrize("fs,ts", [(1, 0), (0, 1), (0, 0)])
def test_half_open_transformer_keeps_pattern(oracle_mod, fs, ts):
    """A transformer open at an end still stamps its four Y positions (zeros off the diagonal), and
    the closed end carries branch_calc_param's shunt (divided by k^2 on the from side)."""
    def entries(doc):
        rp, col, y = oracle_mod.ybus(pgm.from_pgm(doc).network)
        return {(r, c): v for r, c, v in zip(np.repeat(np.arange(3), np.diff(rp)), col, y)}

    pos = entries(_two_node_doc(tap_pos=2, from_status=fs, to_status=ts))
    none = entries(_two_node_doc(tap_pos=2, from_status=0, to_status=0))  # the source line alone
    assert (0, 1) in pos and (1, 0) in pos and pos[(0, 1)] == 0 and pos[(1, 0)] == 0
    t = pgm.from_pgm(_two_node_doc(tap_pos=2)).network
    k2 = abs(t.tr_tap[0]) ** 2
    sh = pgm.half_open_branch_shunt(t.tr_ys[0], t.tr_ysh[0])
    want_ff = sh / k2 if (fs, ts) == (1, 0) else 0
    want_tt = sh if (fs, ts) == (0, 1) else 0
    assert abs(pos[(0, 0)] - none[(0, 0)] - want_ff) < 1e-9 * abs(none[(0, 0)])
    assert abs(pos[(1, 1)] - want_tt) < 1e-12 * max(1, abs(want_tt))


def test_bus_order_and_sources(golden_dir):
    """Nodes by ascending id, then star buses, then one Slack bus per active source."""
    m = pgm.from_pgm(load(golden_dir, "three_winding_transformer", "input.json"))
    net = m.network
    n_nodes = len(m.node_ids)
    assert list(m.node_ids) == sorted(m.node_ids)
    assert net.n_bus == n_nodes + 2
    assert net.bus_type[n_nodes] == PQ and net.bus_type[n_nodes + 1] == SLACK
    assert np.count_nonzero(net.bus_type == SLACK) == 1
    src = pgm.from_pgm(load(golden_dir, "source", "input.json")).network
    assert np.count_nonzero(src.bus_type == SLACK) == 2


def test_document_errors(golden_dir):
    doc = load(golden_dir, "line", "input.json")
    with pytest.raises(pgm.PgmError, match="no 'source'"):
        pgm.from_pgm({"data": {"node": doc["data"]["node"]}})
    bad = json.loads(json.dumps(doc))
    del bad["data"]["line"][0]["r1"]
    with pytest.raises(pgm.PgmError, match="missing field 'r1'"):
        pgm.from_pgm(bad)
    bad = json.loads(json.dumps(doc))
    bad["data"]["line"][0]["to_node"] = 999
    with pytest.raises(pgm.PgmError, match="unknown node 999"):
        pgm.from_pgm(bad)
    with pytest.raises(pgm.PgmError, match="no line entry with id 77"):
        pgm.apply_batch_update(doc, {"line": [{"id": 77, "r1": 1.0}]})


def test_native_json_zip_round_trip():
    net = generate_grid(12, 4, 1)
    z = np.zeros((net.n_bus, 6))
    z[3] = [0.01, 0.0, -0.02, -0.01, -0.03, 0.005]
    net.zip_coef = z
    back = Network.from_json(net.to_json())
    np.testing.assert_array_equal(back.zip_coef, z)
    assert Network.from_json(generate_grid(12, 4, 1).to_json()).zip_coef is None


# ---------------------------------------------------------------- PowerFlowModel without the GPU
def test_model_surface(golden_dir):
    path = os.path.join(pgm_dir(golden_dir, "transmission-case"), "input.json")
    m = PowerFlowModel.from_pgm_json(path)
    assert m.n_nodes == 12
    kinds, p, q = m.bus_spec()
    assert kinds[-1] == 0 and kinds.count(0) == 1 and all(k == 2 for k in kinds[:-1])
    assert len(p) == len(q) == 12
    assert m.zip_term_counts() == [0] * 12
    assert m.voltage_mag() == [1.0] * 11 + [1.03]  # before any solve: the template (source u_ref 1.03)
    assert PowerFlowModel.default_threads() == 1
    sym = PowerFlowModel.from_pgm_json(os.path.join(pgm_dir(golden_dir, "sym_load"), "input.json"))
    assert sym.zip_term_counts() == [0, 2, 0]
    with pytest.raises(ValueError, match="CPU backend"):
        PowerFlowModel.from_pgm_json(path, backend="klu_native")
    with pytest.raises(ValueError, match="unknown backend"):
        PowerFlowModel.from_pgm_json(path, backend="cuda")
    with pytest.raises(RuntimeError, match="reading"):
        PowerFlowModel.from_pgm_json(path + ".missing")


def test_model_malformed_json(tmp_path):
    bad = tmp_path / "bad.json"
    bad.write_text("{not json")
    with pytest.raises(ValueError, match="parsing"):
        PowerFlowModel.from_pgm_json(str(bad))
