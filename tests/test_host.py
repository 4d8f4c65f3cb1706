"""CPU tests of the host side: the C ABI library loads and exports every declared symbol, the
instance generators reproduce the reference's fingerprints, the draws reproduce the stored DSR
record's fingerprints, Hong-Li's decisions reproduce the stored p-value, and the RUM surface
validates its input with the reference's messages."""
import json
import os
import re
import ctypes

import numpy as np
import pandas as pd
import pytest

from conftest import GOLDEN, ROOT
from solver import _lib, defaults, instances, params
from solver.inference import Draws, MenuDraws, cells_from_panel, hong_li
from solver.inference.bootstrap import _sha
from solver.rum_solver import RUM, choose_method, _plan_flow_cg, _plan_bm_admm


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "rumgpu.h")).read()
    declared = set(re.findall(r"\b(rumgpu_\w+)\s*\(", header))
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)
    assert _lib.load().rumgpu_version() == 1


def test_generators_reproduce_reference_fingerprints():
    refs = json.load(open(os.path.join(GOLDEN, "opt_refs_subset.json")))
    for key in ("n18_sparse", "n18_m2", "n18_m39", "n18_d1pct", "n18_medium", "n19_sparse", "n20_sparse",
                "n20_d1pct", "n25_sparse"):
        n, cat = int(key.split("_")[0][1:]), key.split("_", 1)[1]
        cd, mp = instances.get_instance(n, cat)
        assert instances.fingerprint(cd) == refs[key]["cd_hash"] and mp == refs[key]["mp"], key


def test_benchmark_configs_are_deterministic():
    cd, n = instances.config1()
    assert n == 8 and cd.shape == (1016, 3)
    sums = {}
    for m, p in zip(cd[:, 0], cd[:, 2]):
        sums[int(m)] = sums.get(int(m), 0.0) + p
    assert len(sums) == 247 and all(abs(v - 1) < 1e-12 for v in sums.values())
    assert instances.fingerprint(cd) == instances.fingerprint(instances.config1()[0])
    cd2, n2 = instances.config2()
    assert n2 == 18 and abs(cd2.shape[0] - 23593) < 40


def _dsr():
    panel = pd.read_csv(os.path.join(GOLDEN, "dsr_overload_data_25_long.csv"))
    menu, alt, p = cells_from_panel(panel, 25, alternatives=[0])
    return panel, menu, alt, p


def test_dsr_draws_reproduce_stored_fingerprints():
    rec = json.load(open(os.path.join(GOLDEN, "dsr_alt0_record.json")))
    panel, menu, alt, p = _dsr()
    assert menu.size == 301
    assert _sha(np.column_stack([menu.astype(np.float64), alt.astype(np.float64), p])) == rec["fingerprints"]["data"]
    D = Draws(panel, menu, alt, p, B=10000, seed=0)
    assert D.input_fingerprint == rec["fingerprints"]["panel"]
    for _ in D.blocks():
        pass
    assert D.fingerprint == rec["fingerprints"]["draws"]
    assert D.n_redraws == rec["panel"]["redraws"]


def _ai_luna():
    z = np.load(os.path.join(GOLDEN, "ai_luna_replies.npz"))          # allow_pickle=False (numpy's default)
    table = pd.DataFrame({c: z[c].astype(np.int64) for c in ("respondent", "menu", "chosen")})
    menu, alt, p = cells_from_panel(table, 18)
    return table, menu, alt, p


def test_ai_luna_panel_draws_reproduce_stored_fingerprints():
    """The reference's AI bootstrap record (gpt-5.6-luna, n = 18, 100 replicates as respondents,
    B = 10,000 multinomial_reject draws, seed 0): the cells (73,492 rows), the panel and the whole
    B x M draw matrix hash to the stored fingerprints."""
    rec = json.load(open(os.path.join(GOLDEN, "ai_luna_record.json")))["panel"]
    table, menu, alt, p = _ai_luna()
    assert menu.size == 73492
    assert _sha(np.column_stack([menu.astype(np.float64), alt.astype(np.float64), p])) == rec["fingerprints"]["data"]
    D = Draws(table, menu, alt, p, B=rec["settings"]["B"], seed=rec["settings"]["seed"])
    assert D.input_fingerprint == rec["fingerprints"]["panel"]
    for _ in D.blocks():
        pass
    assert D.fingerprint == rec["fingerprints"]["draws"]
    assert D.n_redraws == rec["panel"]["redraws"]


def test_ai_luna_menu_draws_and_distance_interval_reproduce_stored_record():
    """The same data resampled menu by menu (MenuDraws, B = 10,000, seed 0): the replies and the draw
    matrix hash to the stored fingerprints, and the two-sided 95 % distance interval from the stored
    dist(p_hat) bracket is the stored one to the last bit (c = the 95 % quantile of ||p_b - p_hat||)."""
    rec = json.load(open(os.path.join(GOLDEN, "ai_luna_record.json")))
    table, menu, alt, p = _ai_luna()
    D = MenuDraws(table, menu, alt, p, B=rec["replies"]["settings"]["B"], seed=rec["replies"]["settings"]["seed"])
    assert D.input_fingerprint == rec["replies"]["fingerprints"]["replies"]
    dev = np.concatenate([np.sqrt(((blk - p) ** 2).sum(1)) for _lo, _hi, blk in D.blocks()])
    assert D.fingerprint == rec["replies"]["fingerprints"]["draws"]
    di = rec["replies"]["distance_interval"]
    lo, hi, c = hong_li.distance_interval(rec["p_hat"]["lb"], rec["p_hat"]["ub"], dev, 0.05)
    assert (lo, hi, c) == (di["lower"], di["upper"], di["half_width"])


def test_hong_li_reproduces_stored_decisions():
    rec = json.load(open(os.path.join(GOLDEN, "dsr_alt0_record.json")))
    s = rec["settings"]
    r_n, eps_n, sc = hong_li.scaling(s["n"])
    assert abs(r_n - s["r_n"]) < 1e-12 and abs(eps_n - s["eps_n"]) < 1e-15
    L0, U0 = rec["p_hat"]["lb"], rec["p_hat"]["ub"]
    tau = (sc * L0, sc * U0)
    assert np.allclose(tau, rec["tau"], rtol=0, atol=1e-12)
    dec = hong_li.decide(rec["draws"]["lb"], rec["draws"]["ub"], *tau)
    p_lo, p_hi = hong_li.p_bracket(dec)
    pv = rec["results"]["p_value"]["0.05"]
    assert (int((dec == 0).sum()), int((dec == 1).sum()), int((dec == 2).sum())) == (pv["n_below"], pv["n_above"],
                                                                                    pv["n_undecided"])
    assert p_hi == pv["p"]


def _key_tree(d):
    """{dotted path: kind} of a record's dict levels (lists and scalars are leaves)"""
    out = {}
    for k, v in d.items():
        if isinstance(v, dict):
            out[k] = "dict"
            out.update({k + "." + kk: vv for kk, vv in _key_tree(v).items()})
        else:
            out[k] = "none" if v is None else "list" if isinstance(v, list) else "str" if isinstance(v, str) \
                else "bool" if isinstance(v, bool) else "number"
    return out


def test_record_layout_matches_reference_record(tmp_path):
    """A record written here has the layout of the reference's stored record of the same run (its
    bootstrap.py:356-368 writes both): the same keys at every level with the same kinds of value,
    and every draws field as a list of B values. Levels filled only by a solve (the plan, p_hat's
    certificates, the results) are compared by their keys once present, values the screen fills
    by their keys. Settings also carry
    `input` ('panel' or 'replies'), which the reference's writer emits and its stored run predates."""
    ref = json.load(open(os.path.join(GOLDEN, "dsr_alt0_record.json")))
    panel, menu, alt, p = _dsr()
    rum = RUM(menu, alt, p, num_alternatives=25)
    boot = rum.bootstrap(panel=panel, B=10000)           # nothing solved: no device is touched
    path = str(tmp_path / "record.json")
    boot.save(path)
    ours = json.load(open(path))
    assert set(ours) == set(ref)
    assert set(ours["settings"]) == set(ref["settings"]) | {"input"} and ours["settings"]["input"] == "panel"
    t_ours, t_ref = _key_tree(ours), _key_tree(ref)
    solved_only = ("plan", "p_hat.", "results.", "tau", "when")
    for k, kind in t_ref.items():
        if k.startswith(solved_only) or k in ("p_hat.lb", "p_hat.ub"):
            continue
        # values filled by the screen (draw fingerprint, redraw counts) are None until it runs
        assert k in t_ours and t_ours[k] in (kind, "none"), (k, t_ours.get(k), kind)
    for k in ref["draws"]:
        assert len(ours["draws"][k]) == len(ref["draws"][k]) == 10000
    assert set(ours["p_hat"]) == set(ref["p_hat"]) and set(ours["results"]) >= set(ref["results"])
    assert set(ours["scaling"]) == set(ref["scaling"]) and ours["settings"]["r_n"] == ref["settings"]["r_n"]
    # and a checkpoint of another run is refused with the reference's message
    other = dict(ref, settings=dict(ref["settings"], input="panel", B=500))
    with open(path, "w") as f:
        json.dump(other, f)
    with pytest.raises(ValueError, match="belongs to another run"):
        rum.bootstrap(panel=panel, B=10000, checkpoint=path)


def test_menu_draws_statistics():
    rng = np.random.default_rng(7)
    rows = []
    for m in range(1, 32):
        items = [a for a in range(5) if (m >> a) & 1]
        if len(items) < 2:
            continue
        for r in range(60):
            rows.append((m, int(rng.choice(items))))
    rep = pd.DataFrame(rows, columns=["menu", "chosen"])
    menu, alt, p = cells_from_panel(rep, 5)
    D = MenuDraws(rep, menu, alt, p, B=2000, seed=0)
    X = np.vstack([b for _, _, b in D.blocks()])
    inv = np.unique(menu, return_inverse=True)[1]
    sums = np.zeros((X.shape[0], inv.max() + 1))
    np.add.at(sums.T, inv, X.T)
    assert np.abs(sums - 1).max() < 1e-12
    N = D.N[inv]
    assert np.abs(X * N - np.round(X * N)).max() < 1e-9
    D2 = MenuDraws(rep, menu, alt, p, B=2000, seed=0)
    for _ in D2.blocks():
        pass
    assert D2.fingerprint == D.fingerprint


def test_rum_input_errors_match_reference_messages():
    with pytest.raises(ValueError, match="no observations"):
        RUM([], [], [])
    with pytest.raises(ValueError, match="row 0: empty menu"):
        RUM([0], [0], [1.0])
    with pytest.raises(ValueError, match="row 0: alternative 2 is not in its menu"):
        RUM([3], [2], [1.0])
    with pytest.raises(ValueError, match=r"row 0: \(menu 3, alternative 0\) appears more than once"):
        RUM([3, 3], [0, 0], [0.5, 0.5])
    with pytest.raises(ValueError, match="outside"):
        RUM([3, 3], [0, 1], [1.5, 0.5])
    rum = RUM([3, 3, 1], [0, 1, 0], [0.5, 0.5, 1.0])
    assert rum.n == 2 and rum._single.size == 1
    with pytest.raises(ValueError, match="unknown method"):
        rum.calculate_distance(method="nope")


def test_planner_rules_and_beta():
    prof = dict(use_gpu=True, gpu=dict(supported=True, name="stub"), ram_bytes=64 << 30)
    import solver.machine as machine
    saved = machine.gpu_free_bytes
    machine.gpu_free_bytes = lambda device=0: 250 << 30
    try:
        assert choose_method(18, 23568, False, prof)[0] == "bm_admm"
        assert choose_method(18, 306, False, prof)[0] == "flow_cg"
        assert choose_method(10, 5000, False, prof)[0] == "flow_cg"
        a = _plan_bm_admm(18, 23568, False, prof, dict(defaults.METHOD_ARGS["bm_admm"]))
        assert a["dtype"] == "float64" and a["use_graph"] is True
        assert abs(a["beta"] - params.bm_beta_default(23568)) < 1e-15
        a = _plan_flow_cg(25, 301, False, prof, dict(defaults.METHOD_ARGS["flow_cg"]))
        assert (a["pricing"], a["dtype"]) == ("mask", "float32")
        # many rows per alternative: the slab form (no per-subset row loop)
        a = _plan_flow_cg(22, 230691, False, prof, dict(defaults.METHOD_ARGS["flow_cg"]))
        assert (a["pricing"], a["dtype"]) == ("slab", "float32")
        a = _plan_flow_cg(22, 230691, False, prof, dict(defaults.METHOD_ARGS["flow_cg"], dtype="uint24"))
        assert (a["pricing"], a["dtype"]) == ("stream", "uint24")
        a = _plan_flow_cg(16, 250, False, prof, dict(defaults.METHOD_ARGS["flow_cg"]))
        assert a["pricing"] == "mask"
        a = _plan_flow_cg(25, 301, False, prof, dict(defaults.METHOD_ARGS["flow_cg"], pricing="stream_u16"))
        assert (a["pricing"], a["dtype"]) == ("stream", "uint16")
        with pytest.raises(ValueError, match="contradicts"):
            _plan_flow_cg(20, 10, False, prof, dict(defaults.METHOD_ARGS["flow_cg"], pricing="stream_u16",
                                                    dtype="float32"))
        machine.gpu_free_bytes = lambda device=0: 1 << 20
        with pytest.raises(RuntimeError, match="needs"):
            _plan_bm_admm(22, 10000, False, prof, dict(defaults.METHOD_ARGS["bm_admm"]))
    finally:
        machine.gpu_free_bytes = saved
    # the reference's beta anchors (params.py of the reference: mp = 306 -> 0.6, 24k -> 5)
    assert abs(params.bm_beta_default(306) - 0.6) < 0.05 and abs(params.bm_beta_default(24000) - 5.0) < 1.0


def test_adaptive_penalty_cadence_in_iterations():
    """beta="adaptive" certifies every params.BM_BETA_ADAPT_CHECK iterations unless method_args names
    check, and its rule acts every BM_BETA_ADAPT["every_iters"] iterations, counted in checks"""
    prof = dict(use_gpu=True, gpu=dict(supported=True, name="stub"), ram_bytes=64 << 30)
    import solver.machine as machine
    saved = machine.gpu_free_bytes
    machine.gpu_free_bytes = lambda device=0: 250 << 30
    try:
        rum = RUM.from_rows(instances.warmup_instance(12))
        _, _, a = rum._plan("bm_admm", {"beta": "adaptive"}, False, prof)
        assert a["check"] == params.BM_BETA_ADAPT_CHECK == 100
        assert a["_beta_adapt"]["every"] * a["check"] == params.BM_BETA_ADAPT["every_iters"] == 500
        _, _, a = rum._plan("bm_admm", {"beta": "adaptive", "check": 50}, False, prof)
        assert a["check"] == 50 and a["_beta_adapt"]["every"] == 10
        _, _, a = rum._plan("bm_admm", {"beta": "cert"}, False, prof)
        assert a["check"] == defaults.METHOD_ARGS["bm_admm"]["check"] == 50 and "_beta_adapt" not in a
    finally:
        machine.gpu_free_bytes = saved


def test_no_gpu_means_no_method():
    rum = RUM.from_rows(instances.warmup_instance(6))
    prof = dict(use_gpu=False, gpu=None, ram_bytes=1 << 34)
    with pytest.raises(RuntimeError, match="no method can run"):
        rum.calculate_distance(machine=prof)


def test_native_master_qp_matches_numpy_and_oracle():
    """rumgpu_qp_simplex (host code in the library) against the numpy active set and the oracle's
    restatement, on Gram matrices of random 0/1 columns (the master QP's form), including
    duplicated and dependent columns"""
    from oracle import oracle as O
    from solver.common.qp_native import solve_qp_native, solve_qp_numpy
    rng = np.random.default_rng(4)
    for trial in range(40):
        M, K = int(rng.integers(3, 40)), int(rng.integers(1, 60))
        C = (rng.random((K, M)) < 0.4).astype(np.float64)
        if trial % 5 == 0 and K > 3:
            C[1] = C[0]                               # a duplicate column
            C[2] = np.clip(C[0] + C[3], 0, 1)         # often dependent
        p = rng.random(M)
        Q, h = C @ C.T, C @ p
        lam, sup = solve_qp_native(Q, h)
        lam2, sup2 = solve_qp_numpy(Q, h)
        lam3, _ = O.qp_simplex(Q, h)
        assert abs(lam.sum() - 1.0) < 1e-12 and (lam >= 0).all()
        f = lambda l: 0.5 * l @ Q @ l - h @ l      # noqa: E731
        # the same optimum value (the weights may differ on degenerate masters)
        assert abs(f(lam) - f(lam2)) <= 1e-9 * max(1.0, abs(f(lam2))), trial
        assert abs(f(lam) - f(lam3)) <= 1e-9 * max(1.0, abs(f(lam3))), trial
        # KKT on the support: gradient = mu there, >= mu elsewhere
        g = Q @ lam - h
        on = lam > 1e-9
        mu = g[on].mean()
        assert np.abs(g[on] - mu).max() <= 1e-7 * max(1.0, np.abs(g).max()), trial
        assert (g[~on] >= mu - 1e-7 * max(1.0, np.abs(g).max())).all(), trial
        # warm start from a support gives the same optimum
        lam4, _ = solve_qp_native(Q, h, A_init=sup2)
        assert abs(f(lam4) - f(lam2)) <= 1e-9 * max(1.0, abs(f(lam2)))


def test_column_store_planning_and_packed_columns():
    """The flow_cg plan charges a single run's device column store (bitsets of max_iters + 2
    columns, p, p_out) to the GPU budget, and keeps the columns on the host when the store does not
    fit beside the sweep; packed columns read back as the 0/1 columns they were packed from."""
    from solver.flow_cg.engine import PackedColumns
    prof = dict(use_gpu=True, gpu=dict(supported=True, name="stub"), ram_bytes=64 << 30)
    import solver.machine as machine
    saved = machine.gpu_free_bytes
    mp = 104885
    try:
        machine.gpu_free_bytes = lambda device=0: 250 << 30
        a = _plan_flow_cg(20, mp, False, prof, dict(defaults.METHOD_ARGS["flow_cg"]))
        k = int(a["max_iters"]) + 2
        cols = 8 * k * ((mp + 63) // 64) + 16 * mp + 20 * k
        assert "_cols" not in a and a["pricing"] == "slab"
        need = a["_gpu_bytes"] - cols
        # a budget that holds the sweep but not the store beside it
        frac = defaults.AUTO["gpu_memory_fraction"]
        machine.gpu_free_bytes = lambda device=0: int((need + cols // 2) / frac)
        b = _plan_flow_cg(20, mp, False, prof, dict(defaults.METHOD_ARGS["flow_cg"]))
        assert b["pricing"] == "slab" and b["_cols"] == "host" and b["_gpu_bytes"] == need
        # at most 4,096 rows the columns stay on the host and nothing is charged
        machine.gpu_free_bytes = lambda device=0: 250 << 30
        c = _plan_flow_cg(16, 250, False, prof, dict(defaults.METHOD_ARGS["flow_cg"]))
        assert "_cols" not in c
    finally:
        machine.gpu_free_bytes = saved
    rng = np.random.default_rng(5)
    M = 1000
    C = (rng.random((7, M)) < 0.3).astype(np.uint8)
    packed = []
    for col in C:
        pad = np.zeros(((M + 63) // 64) * 64, np.uint8)
        pad[:M] = col
        packed.append(np.packbits(pad, bitorder="little").view(np.uint64))
    pc = PackedColumns(packed, M)
    assert len(pc) == 7 and all(np.array_equal(pc[i], C[i]) for i in range(7))
    assert [x.tolist() for x in pc[2:4]] == [C[2].tolist(), C[3].tolist()]
    assert all(np.array_equal(x, y) for x, y in zip(pc, C))


def test_adaptive_penalty_rule():
    """solve_kadmm's beta_adapt on a scripted handle: doubles while ub - phi dominates, halves while
    phi - lb dominates, a reversal takes sqrt(2) and freezes, at most max_changes changes"""
    from solver.bm_admm.engine import solve_kadmm

    class Fake:
        dtype, graph_len = np.dtype(np.float64), 0

        def __init__(self, script):
            self.beta, self.k, self.script, self.log = 1.0, 0, script, []

        def run(self, k):
            self.k += 1

        def phi(self):
            return self.script(self.k)[1]

        def set_beta(self, b):
            self.beta = b
            self.log.append(b)

    def drive(script, **kw):
        h = Fake(script)
        rule = dict(every=2, ratio=10.0, factor=2.0, freeze_on_flip=True, max_changes=3)
        rule.update(kw)
        solve_kadmm(h, max_iter=40, check=1, push_fn=lambda hh: script(hh.k)[0], ub_arm_immediately=True,
                    lb_fn=lambda hh: script(hh.k)[2], gap_abs_tol=1e-12, beta_adapt=rule)
        return h.log

    infeasible = lambda k: (1.0 + 1.0 / k, 0.5, 0.5 - 1e-3 / k)     # noqa: E731  ub - phi >> phi - lb
    assert drive(infeasible) == [2.0, 4.0, 8.0]                      # capped at 3 changes
    optimal = lambda k: (0.5 + 1e-3 / k, 0.5, 0.5 - 1.0 / k)         # noqa: E731
    assert drive(optimal) == [0.5, 0.25, 0.125]
    flip = lambda k: (infeasible if k < 3 else optimal)(k)           # noqa: E731
    log = drive(flip, max_changes=10)
    assert log[0] == 2.0 and abs(log[1] - 2.0 / np.sqrt(2.0)) < 1e-15 and len(log) == 2
    balanced = lambda k: (0.6 + 1.0 / k, 0.5, 0.4 - 1.0 / k)         # noqa: E731
    assert drive(balanced) == []


def test_master_qp_state_sequence():
    """a growing master solved with a QPState (its last working set's factor reused) reaches the
    optimum of the stateless solve at every step"""
    from solver.common.qp_native import QPState, solve_qp_native
    rng = np.random.default_rng(11)
    M = 60
    C = (rng.random((40, M)) < 0.35).astype(np.float64)
    p = rng.random(M)
    st = QPState()
    A1 = A2 = None
    for K in range(3, 41):
        Q, h = C[:K] @ C[:K].T, C[:K] @ p
        Qbig = np.zeros((64, 64))
        Qbig[:K, :K] = Q                             # a view with a wider row stride, as the engine has
        lam1, A1 = solve_qp_native(Qbig[:K, :K], h, A_init=A1, state=st)
        lam2, A2 = solve_qp_native(Q, h, A_init=A2)
        f = lambda l: 0.5 * l @ Q @ l - h @ l      # noqa: E731
        assert abs(f(lam1) - f(lam2)) <= 1e-10 * max(1.0, abs(f(lam2))), K
    st.close()


_QP_SEQUENCE = r"""
import hashlib, sys
sys.path[:0] = sys.argv[1:3]
import numpy as np
from solver.common.qp_native import QPState, solve_qp_native
rng = np.random.default_rng(7)
M, P, K_MAX = 1500, 3000, 700
p = rng.random(M)
p *= 50.0 / p.sum()               # far from every column's reach: the support grows to hundreds
pool = (rng.random((P, M)) < 0.12).astype(np.float64)
C = np.zeros((K_MAX, M))
Q = np.zeros((K_MAX + 8, K_MAX + 8))
h = np.zeros(K_MAX)
st, A, p_out = QPState(), [0], np.zeros(M)
dig, worst, kmax = hashlib.sha256(), 0.0, 0
for k in range(K_MAX):
    j = int(np.argmax(pool @ (p - p_out))) if k else 0
    C[k] = pool[j]
    row = C[:k + 1] @ C[k]
    Q[k, :k + 1] = row
    Q[:k + 1, k] = row
    h[k] = C[k] @ p
    Qk, hk = Q[:k + 1, :k + 1], h[:k + 1]
    A_prev = A
    lam, A = solve_qp_native(Qk, hk, A_init=A, state=st)
    dig.update(lam.tobytes())
    kmax = max(kmax, len(A))
    p_out = lam @ C[:k + 1]
    if k % 50 == 49:
        lam2, _ = solve_qp_native(np.ascontiguousarray(Qk), hk, A_init=A_prev, max_outer=2000)
        f = lambda l: 0.5 * l @ Qk @ l - hk @ l
        worst = max(worst, abs(f(lam) - f(lam2)) / max(1.0, abs(f(lam2))))
print(dig.hexdigest(), worst, kmax)
"""


def test_master_qp_gradient_threads_do_not_change_the_weights():
    """A growing master large enough for the inactive-column gradient to run on host threads (K up
    to 700, a working set of a few hundred: above the library's threshold) and to read the state's
    float32 copy of Q: every step's weights are the same bit for bit with 1 and 8 threads, and
    reach the optimum of a stateless solve from the same warm working set (every 50 steps)."""
    import subprocess
    import sys
    outs = []
    for t in (1, 8):
        env = dict(os.environ, OMP_NUM_THREADS=str(t))
        r = subprocess.run([sys.executable, "-c", _QP_SEQUENCE, os.path.join(ROOT, "rum-testing-gpu_amd"), ROOT],
                           env=env, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        outs.append(r.stdout.split())
    assert outs[0][0] == outs[1][0], outs
    assert float(outs[0][1]) <= 1e-10, outs
    assert int(outs[0][2]) >= 200, outs       # K x k above the threaded gradient's threshold


def test_out_of_memory_gives_exit_flag_minus_3(monkeypatch):
    """A device allocation that fails inside a runner ends the call with exit flag -3 and the
    library's message (rum_solver.py:1217-1218 of the reference maps MemoryError and "out of memory"
    to it); the C ABI's status 2 becomes that MemoryError. Forced on the host: the runner raises as
    the engine would, so no GPU is needed."""
    import solver.rum_solver as RS
    from solver import machine
    assert RS._is_oom(MemoryError("x")) and RS._is_oom(RuntimeError("hipMalloc: out of memory"))
    assert not RS._is_oom(RuntimeError("invalid argument"))
    with pytest.raises(MemoryError, match="out of memory"):
        _lib.check(_lib.ENOMEM, "slab of 3.7 GB")
    prof = dict(use_gpu=True, gpu=dict(name="test", arch="gfx950", memory_total=1 << 38, supported=True, reason=""),
                cpu_count=8, ram_bytes=1 << 36, gurobi=None)
    monkeypatch.setattr(machine, "gpu_free_bytes", lambda device=0: 1 << 38)

    def fails(task):
        raise MemoryError("out of memory: bm_admm slab needs 3.700 GB, 0.100 GB free")

    monkeypatch.setitem(RS._RUN, "bm_admm", fails)
    menu = np.array([3, 3, 5, 5, 6, 6], dtype=np.int64)
    alt = np.array([0, 1, 0, 2, 1, 2], dtype=np.int64)
    rum = RUM(menu, alt, np.array([0.6, 0.4, 0.5, 0.5, 0.3, 0.7]))
    r = rum.calculate_distance(method="bm_admm", machine=prof)
    assert r.exit_flag == -3 and r.status == "out of memory" and "3.700 GB" in r.message
    with rum.session(method="bm_admm", machine=prof) as s:
        r2 = s.solve(rum.choice_probability)
    assert r2.exit_flag == -3


class _OraclePricer:
    """pricing through the oracle's mask DP (CPU), for the host-side column generation tests"""
    layout, dtype = "mask", "float32"

    def __init__(self, n, masks_x, offsets):
        self.n, self.masks_x, self.offsets = n, masks_x, offsets

    def price(self, scores):
        from oracle import oracle as O
        vmax, rank, _V, _pred = O.price_mask(self.n, self.masks_x, np.asarray(scores, dtype=np.float32), self.offsets)
        return float(vmax), rank, 0.0

    def info(self):
        return 0, 0.0


def _qp_vectors():
    """the reference's master QP and p_out golden vectors (tests/golden/make_qp_native_vectors.py)"""
    return np.load(os.path.join(GOLDEN, "qp_native_vectors.npz"))


def test_master_qp_matches_the_reference_vectors():
    """The reference's own solve_qp_native (common/qp_native.py:37-100) on 40 random and degenerate
    masters (duplicated columns, more columns than rows, ties in h, p a mixture of two columns, warm
    working sets valid, out of range and empty): the numpy form returns its lam and support bit for
    bit; the library's form (a growing Cholesky factor) the same support in the same order, and lam
    within 1e-13."""
    from solver.common.qp_native import solve_qp_native, solve_qp_numpy
    z = _qp_vectors()
    assert z["K"].size == 40
    for i, K in enumerate(z["K"]):
        K = int(K)
        Q = z["Q"][z["off_q"][i]:z["off_q"][i + 1]].reshape(K, K)
        h = z["h"][z["off_h"][i]:z["off_h"][i + 1]]
        A0 = None if z["a_none"][i] else [int(a) for a in z["a_init"][z["off_a"][i]:z["off_a"][i + 1]]]
        lam = z["lam"][z["off_h"][i]:z["off_h"][i + 1]]
        sup = [int(a) for a in z["support"][z["off_s"][i]:z["off_s"][i + 1]]]
        ln, sn = solve_qp_numpy(Q, h, A_init=A0)
        assert np.array_equal(ln, lam) and list(sn) == sup, i
        lc, sc = solve_qp_native(Q, h, A_init=A0)
        assert [int(a) for a in sc] == sup and np.abs(lc - lam).max() < 1e-13, (i, np.abs(lc - lam).max())


def test_pout_accumulation_matches_the_reference_vectors():
    """Two growing masters (12 and 20 columns) through the reference's BitsetGramMaintainer and
    PoutMaintainer (flow_cg/qp_bitset_incremental.py:36-141), warm from the last support at every
    step: Columns gives the same Q and h, the numpy master the same lam, and p_out (moved by the
    change of lam, columns ascending) the same bits at all 32 steps. The library's rounds (p_out
    summed afresh from lam over the support) stay within 1e-15 of it, lam within 1e-13."""
    from solver.common.qp_native import solve_qp_numpy
    from solver.flow_cg.engine import Columns, NativeRounds
    z = _qp_vectors()
    o_cols = o_lam = 0                      # start of the sequence in pout_cols / pout_steps, pout_lam
    for j, (M, K) in enumerate(zip(z["pout_M"], z["pout_K"])):
        M, K = int(M), int(K)
        C = z["pout_cols"][o_cols:o_cols + K * M].reshape(K, M).astype(np.uint8)
        p = z["pout_p"][z["pout_off"][j]:z["pout_off"][j + 1]]
        cols = Columns(M, p, K)
        nat = NativeRounds(2, np.full(M, 3, dtype=np.uint64), np.zeros(M, dtype=np.int32), p, K)
        lam_prev, sup = np.zeros(0), None
        for k in range(1, K + 1):
            cols.add(C[k - 1])
            nat.add_column(C[k - 1])
            Qv, hv = cols.view()
            Cf = C[:k].astype(np.float64)
            assert np.array_equal(Qv, Cf @ Cf.T)
            assert np.array_equal(hv, np.array([p[np.flatnonzero(c)].sum() for c in C[:k]]))
            lam, sup = solve_qp_numpy(Qv, hv, A_init=sup)
            ref_lam = z["pout_lam"][o_lam + (k - 1) * k // 2:][:k]
            assert np.array_equal(lam, ref_lam), (j, k)
            d = lam.copy()
            d[:lam_prev.size] -= lam_prev
            cols.apply_delta(d)
            lam_prev = lam
            want = z["pout_steps"][o_cols + (k - 1) * M:][:M]
            assert np.array_equal(cols.p_out, want), (j, k)
            if k == 1:
                nat.set_lambda(np.array([1.0]), [0])
            nat.master()
            lam_n, pout_n, _c = nat.result()
            assert np.abs(pout_n - want).max() <= 1e-15 and np.abs(lam_n - lam).max() < 1e-13, (j, k)
        nat.close()
        o_cols += K * M
        o_lam += K * (K + 1) // 2


def test_master_batch_equals_single_masters():
    """rumgpu_cgr_master_batch (a bootstrap half's masters on host threads) gives every run the
    result of its own rumgpu_cgr_master bit for bit, cold and warm, from 5 to 900 columns, with and
    without a new ranking to add first; a run that fails gets its own error while the others get
    their results."""
    from solver.flow_cg.engine import MasterCall, NativeRounds, master_batch, xmajor
    rng = np.random.default_rng(5)
    n, M = 12, 180
    menus = set()
    while len(menus) < M:
        k = int(rng.integers(2, 7))
        menus.add(int(sum(1 << int(i) for i in rng.choice(np.arange(1, n), k - 1, replace=False))) | 1)
    menu = np.array(sorted(menus), dtype=np.uint64)
    _perm, masks_x, alt_x, _off = xmajor(menu, np.zeros(M, dtype=np.int32), n)

    def runs(K, seed, D=9):
        r2, out = np.random.default_rng(seed), []
        for _ in range(D):
            r = NativeRounds(n, masks_x, alt_x, r2.uniform(0.05, 0.95, M), K + 2)
            for _k in range(K):
                r.add_ranking(r2.permutation(n).astype(np.int32))
            r.set_lambda(np.array([1.0]), [0])
            out.append(r)
        return out

    for K, D in ((5, 9), (40, 9), (900, 4)):
        a, b = runs(K, K, D), runs(K, K, D)
        r3 = np.random.default_rng(K + 1)
        for rnd in range(3):                # warm from the last factor; from the second round on, a new
            new = [r3.permutation(n).astype(np.int32) for _ in a] if rnd else [None] * len(a)
            one = []                        # ranking's column first (the batch adds it on its thread)
            for r, q in zip(a, new):
                if q is not None:
                    r.add_ranking(q)
                one.append(r.master())
            many = master_batch([MasterCall(r, q) for r, q in zip(b, new)])
            for x, (y, e) in zip(one, many):
                assert e is None and x[0] == y[0] and x[1] == y[1] and np.array_equal(x[2], y[2])
        for r in a + b:
            r.close()
    good = runs(5, 1, D=3)
    empty = NativeRounds(n, masks_x, alt_x, np.full(M, 0.5), 4)        # no column: the master refuses
    assert master_batch([]) == []
    res = master_batch([MasterCall(good[0]), MasterCall(empty), MasterCall(good[1])])
    assert res[0][1] is None and res[2][1] is None and isinstance(res[1][1], ValueError) and res[1][0] is None
    for r in good + [empty]:
        r.close()


def test_native_rounds_reach_the_exact_projection():
    """Column generation with the library's host rounds (rumgpu_cgr_*: bitset columns, integer Gram,
    warm master QP, p_out, scores) against the numpy master, both priced by the oracle's mask DP on
    the CPU: both reach the exact distance (oracle projection), with the same columns' Gram counts,
    and a warm restart from the kept columns lands on the same distance."""
    from oracle import oracle as O
    from solver.flow_cg.engine import cg_steps, drive, xmajor
    rng = np.random.default_rng(11)
    n = 6
    menus = [m for m in range(1 << n) if bin(m).count("1") >= 2 and rng.random() < 0.5]
    om, ox = [], []
    for m in menus:
        for a in range(n):
            if (m >> a) & 1:
                om.append(m)
                ox.append(a)
    menu, alt = np.array(om, dtype=np.uint64), np.array(ox, dtype=np.int32)
    p = rng.random(menu.size) + 0.05
    _, inv = np.unique(menu, return_inverse=True)
    p = p / np.bincount(inv, weights=p)[inv]
    p = p.astype(np.float32).astype(np.float64)            # the engine works on float32(p)
    d_exact, _lb, _x = O.exact_projection(n, menu.astype(np.int64), alt.astype(np.int64), p, tol=1e-14)
    perm, masks_x, _ax, offsets = xmajor(menu, alt, n)
    pricer = _OraclePricer(n, masks_x, offsets)
    res = {}
    for qp in ("native", "numpy"):
        r = drive(cg_steps(n, menu, alt, p, pricer, gap_abs_tol=1e-7, max_iters=800, qp=qp), pricer)
        res[qp] = r
        assert r["stop_reason"] in ("gap_abs", "rc"), r["stop_reason"]
        # the lower bound is limited by the float32 scores (about 1e-8 here), hence abs 1e-7
        assert abs(r["distance"] - d_exact) < 1e-7 and r["lb"] <= d_exact + 1e-9, (qp, r["distance"], d_exact)
        # p_out is the mixture of the columns with weights lam
        C = np.asarray(r["columns"], dtype=np.float64)
        assert np.allclose(r["lam"] @ C, r["p_out"], atol=1e-12) and abs(r["lam"].sum() - 1) < 1e-12
    # warm restart from the native run's columns: the same distance within its tolerance
    rn = res["native"]
    r2 = drive(cg_steps(n, menu, alt, p, pricer, gap_abs_tol=1e-7, max_iters=800, qp="native",
                        init_columns=rn["columns"], init_lambda=rn["lam"]), pricer)
    assert abs(r2["distance"] - d_exact) < 1e-7 and r2["iters"] <= 3
    # the two masters follow the same trajectory (same QP; h and p_out summed in other orders)
    assert res["native"]["n_cols"] == res["numpy"]["n_cols"]
    assert abs(res["native"]["distance"] - res["numpy"]["distance"]) < 1e-12


def test_autotuner_screens_then_times_the_fastest_few(tmp_path, monkeypatch):
    """The tuner's selection on a handle stand-in with a modelled clock: every candidate gets a short
    run, the three fastest a full run, and the full-run times pick the plan (a finalist that screened
    fastest but runs slower loses; a plan the handle rejects is skipped); the graph length is restored."""
    from solver.bm_admm import autotune
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "plans.json"))
    monkeypatch.setattr(autotune, "_MEM", {})
    clock = [0.0]
    monkeypatch.setattr(autotune, "time", type("T", (), {"perf_counter": staticmethod(lambda: clock[0])}))
    nb = 15
    cands = autotune.candidates(nb)
    cost = {c: 100.0 + i for i, c in enumerate(cands)}        # steady per-iteration microseconds
    screen_bias = {cands[5]: -50.0, cands[7]: -40.0, cands[9]: -30.0}   # fast on the short run only
    cost[cands[7]] = 99.0                                     # the fastest on the full run
    rejected = cands[0]

    class FakeHandle:
        n, dtype = nb + 1, "float64"

        def __init__(self):
            self.graph_len, self.plan, self.log = 0, None, []

        def set_graph(self, length):
            self.graph_len = length

        def set_plan(self, *plan):
            if plan == rejected:
                raise ValueError("too much LDS")
            self.plan = plan

        def run(self, iters):
            full = self.graph_len == 50
            self.log.append((self.plan, iters, full))
            clock[0] += iters * 1e-6 * (cost[self.plan] + (0.0 if full else screen_bias.get(self.plan, 0.0)))

    h = FakeHandle()
    plan, us = autotune.tune(h, device="fake")
    assert plan == cands[7] and h.plan == cands[7] and h.graph_len == 0
    full_runs = {p for p, it, full in h.log if full and it == 200}
    assert full_runs == {cands[5], cands[7], cands[9]}
    assert "%d,%d,%d" % rejected not in us and len(us) == len(cands) - 1
    assert us["%d,%d,%d" % cands[7]] == pytest.approx(99.0)
