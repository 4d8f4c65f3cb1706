"""GPU parity of the HIP kernels against the CPU oracle, through the C ABI.

Bit-exact where the operations and their order are the same (every kernel is built with
-ffp-contract=off, the oracle too): the butterfly (window passes apply the bit stages in ascending
order), the per-menu prox, the plain (reference-schedule) iteration, the mask pricing sweep and the
stream sweeps. Tolerance-based where the order differs by design: the fused iteration (window
rounds in alternating direction, z = b2 (x2 - xk)) within 1e-12 relative after 1 and 20
iterations in float64; the certificates (compensated device reductions) within 1e-12 relative.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from solver import _lib, instances
from solver.bm_admm.engine import Handle
from solver.flow_cg.engine import Pricer, xmajor

pytestmark = pytest.mark.gpu


def _rows(n, frac, rng, kmin=2):
    menus = [m for m in range(1 << n) if bin(m).count("1") >= kmin and rng.random() < frac]
    om, ox = [], []
    for m in menus:
        for a in range(n):
            if (m >> a) & 1:
                om.append(m)
                ox.append(a)
    menu, alt = np.array(om, dtype=np.uint64), np.array(ox, dtype=np.int32)
    p = rng.random(menu.size)
    _, inv = np.unique(menu, return_inverse=True)
    p = p / np.bincount(inv, weights=p)[inv]
    return menu, alt, p


def _medium(n, cells, rng):
    """random menus of size >= 2 until about `cells` cells, all members observed"""
    seen, om, ox = set(), [], []
    while len(om) < cells:
        m = int(rng.integers(0, 1 << n))
        if bin(m).count("1") < 2 or m in seen:
            continue
        seen.add(m)
        for a in range(n):
            if (m >> a) & 1:
                om.append(m)
                ox.append(a)
    menu, alt = np.array(om, dtype=np.uint64), np.array(ox, dtype=np.int32)
    p = rng.random(menu.size)
    _, inv = np.unique(menu, return_inverse=True)
    return menu, alt, p / np.bincount(inv, weights=p)[inv]


@pytest.mark.parametrize("n", [2, 3, 6, 11, 12, 15, 18, 19])
def test_butterfly64_bit_exact(n):
    ctx = _lib.context(0)
    rng = np.random.default_rng(n)
    for coefs in [(1.0, -1.0, 0.0, 1.0), (1.0, 0.0, -1.0, math.sqrt(2.0)), (0.3, -1.7, 2.1, 0.9)]:
        v = rng.standard_normal(n << (n - 1))
        want = O.butterfly(v, n, *coefs)
        got = v.copy()
        _lib.check(ctx.lib.rumgpu_butterfly64(ctx.h, n, _lib.ptr(got), *coefs))
        assert np.array_equal(got, want), (n, coefs, np.abs(got - want).max())


@pytest.mark.parametrize("n", [4, 12, 18])
def test_butterfly32_bit_exact(n):
    ctx = _lib.context(0)
    v = np.random.default_rng(n).standard_normal(n << (n - 1)).astype(np.float32)
    want = O.butterfly(v, n, 1.0, 1.0, 0.0, 1.0)
    got = v.copy()
    _lib.check(ctx.lib.rumgpu_butterfly32(ctx.h, n, _lib.ptr(got), 1.0, 1.0, 0.0, 1.0))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("n", [3, 7, 12])
def test_prox_bit_exact(n, dtype):
    rng = np.random.default_rng(5 + n)
    menu, alt, p = _rows(n, 0.5, rng)
    keep = rng.random(menu.size) < 0.75                   # partly observed menus: mixed curvature
    menu, alt, p = menu[keep], alt[keep], p[keep]
    h = Handle(n, menu, alt, p, dtype=dtype, beta=1.3)
    dt = np.dtype(dtype)
    xk = rng.standard_normal(n << (n - 1)).astype(dt).astype(np.float64)
    got = h.prox_slab(xk)
    obs = O.obs_index(n, menu, alt)
    want = O.prox(n, xk.astype(dt), O.scatter(n, obs, p.astype(dt), dt), O.scatter(n, obs, np.ones(obs.size, dt), dt),
                  dt.type(1.3))
    assert np.array_equal(got.astype(dt), want)
    h.close()


def _oracle_state(n, menu, alt, p, beta, dtype):
    return O.AdmmState(n, menu, alt, p, beta, dtype=dtype)


@pytest.mark.parametrize("n", [5, 9, 12, 14])
def test_plain_iterations_bit_exact(n):
    rng = np.random.default_rng(50 + n)
    menu, alt, p = _medium(n, min(4000, n << (n - 2)), rng)
    beta = 0.9
    h = Handle(n, menu, alt, p, dtype="float64", beta=beta, plain=True)
    st = _oracle_state(n, menu, alt, p, beta, np.float64)
    cold = h.get_state()
    assert np.array_equal(cold["x2"], st.x2)
    h.run(7)
    st.run(7)
    got = h.get_state()
    for k in ("x2", "s", "y", "z"):
        assert np.array_equal(got[k], getattr(st, k)), (k, np.abs(got[k] - getattr(st, k)).max())
    h.close()


@pytest.mark.parametrize("n", [4, 8, 11, 13, 18])
def test_fused_iterations_match_oracle(n):
    rng = np.random.default_rng(70 + n)
    menu, alt, p = _medium(n, min(6000, n << (n - 2)), rng)
    beta = 1.1
    h = Handle(n, menu, alt, p, dtype="float64", beta=beta)
    st = _oracle_state(n, menu, alt, p, beta, np.float64)
    for k_it in (1, 19):
        h.run(k_it)
        st.run(k_it)
        got = h.get_state()
        for k in ("x2", "s", "y", "z"):
            ref = getattr(st, k)
            scale = max(1.0, float(np.abs(ref).max()))
            err = float(np.abs(got[k] - ref).max()) / scale
            assert err < 1e-11, (n, k_it, k, err)
    h.close()


def test_fused_graph_replay_matches_eager():
    n = 14
    rng = np.random.default_rng(3)
    menu, alt, p = _medium(n, 3000, rng)
    a = Handle(n, menu, alt, p, dtype="float64", beta=1.0)
    b = Handle(n, menu, alt, p, dtype="float64", beta=1.0)
    b.set_graph(10)
    for _ in range(3):                          # replays with host work in between
        a.run(10)
        b.run(10)
        sa, sb = a.get_state(), b.get_state()
        for k in ("x2", "s", "y", "z"):
            assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()


@pytest.mark.parametrize("n,dtype", [(7, "float64"), (12, "float32"), (16, "float64")])
def test_certificates_match_oracle(n, dtype):
    rng = np.random.default_rng(90 + n)
    menu, alt, p = _medium(n, min(5000, n << (n - 2)), rng)
    h = Handle(n, menu, alt, p, dtype=dtype, beta=1.0)
    h.run(40)
    x2 = h.get_state()["x2"]
    obs = O.obs_index(n, menu, alt)
    ub, fit = h.push(want_fit=True)
    ub_o, fit_o = O.push_ub64(n, x2, obs, p)
    assert abs(ub - ub_o) <= 1e-12 * max(1.0, ub_o)
    assert np.allclose(fit, fit_o, rtol=0, atol=1e-12)
    lb = h.lb()
    lb_o, _ = O.fw_lb64(n, obs, x2[obs], p)
    assert abs(lb - lb_o) <= 1e-10 * max(1.0, lb_o)
    phi = h.phi()
    assert abs(phi - math.sqrt(math.fsum((x2[obs] - p) ** 2))) <= 1e-12 * max(1.0, phi)
    ctx = _lib.context(0)
    import ctypes
    lbv = ctypes.c_double()
    y_obs = np.ascontiguousarray(x2[obs])           # held: ctypes keeps only the address
    _lib.check(ctx.lib.rumgpu_lb64(ctx.h, n, _lib.ptr(menu), _lib.ptr(alt), menu.size, _lib.ptr(y_obs),
                                   _lib.ptr(p), ctypes.byref(lbv)))
    assert abs(lbv.value - lb_o) <= 1e-10 * max(1.0, lb_o)
    h.close()


# ------------------------------------------------------------------------------------- pricing
@pytest.mark.parametrize("n", [6, 10, 14, 18])
def test_mask_pricing_bit_exact(n):
    rng = np.random.default_rng(n)
    menu, alt, p = _medium(n, min(8000, n << (n - 3)), rng)
    perm, masks_x, _alts, offsets = xmajor(menu, alt, n)
    pr = Pricer(n, masks_x, offsets, layout="mask")
    for trial in range(3):
        scores = rng.standard_normal(menu.size).astype(np.float32)
        vmax, rank, margin = pr.price(scores)
        vo, rank_o, V, pred = O.price_mask(n, masks_x, scores, offsets)
        assert vmax == vo and np.array_equal(rank, rank_o) and margin == 0.0
    pr.close()


@pytest.mark.parametrize("n", [8, 13, 18])
def test_slab_pricing_equals_mask_on_quantized_duals(n):
    """scores = k 2^-12 with |k| < 2^7 and mp 2^7 < 2^24: every float32 partial sum is exact, so the
    butterfly-order sums equal the row-order sums and the columns are identical."""
    rng = np.random.default_rng(100 + n)
    menu, alt, p = _medium(n, min(20000, n << (n - 3)), rng)
    perm, masks_x, _alts, offsets = xmajor(menu, alt, n)
    pm = Pricer(n, masks_x, offsets, layout="mask")
    ps = Pricer(n, masks_x, offsets, layout="slab")
    for trial, bits in enumerate((0, 3, 6, 10)):
        if bits <= n:
            ps.set_slab_tile(bits)              # tiles of the layered max: any size, the same result
        scores = (rng.integers(-127, 128, size=menu.size) / 4096.0).astype(np.float32)
        a = pm.price(scores)
        b = ps.price(scores)
        assert a[0] == b[0] and np.array_equal(a[1], b[1]), (bits, a, b)
    pm.close()
    ps.close()


@pytest.mark.parametrize("layout", ["mask", "slab"])
@pytest.mark.parametrize("n", [9, 16])
def test_sweep_graph_replays_match_oracle(n, layout):
    """The layered sweep is captured at a pricer's first sweep and replayed after: eight replays on
    alternating scores, with eager device work (float64 lower-bound sweeps) and host syncs between
    them and a change of slab tiles (a new capture) halfway, each equal to the oracle's mask sweep
    (quantized scores for the slab layout, where every float32 partial sum is exact)."""
    rng = np.random.default_rng(300 + n)
    menu, alt, p = _medium(n, min(6000, n << (n - 3)), rng)
    perm, masks_x, _alts, offsets = xmajor(menu, alt, n)
    pr = Pricer(n, masks_x, offsets, layout=layout)
    pm = Pricer(n, masks_x, offsets, layout="mask")       # its float64 sweeps: eager work on the stream
    S = [(rng.integers(-127, 128, size=menu.size) / 4096.0).astype(np.float32) for _ in range(2)]
    if layout == "mask":
        S[1] = rng.standard_normal(menu.size).astype(np.float32)
    ref = [O.price_mask(n, masks_x, s, offsets)[:2] for s in S]
    for rep in range(8):
        if layout == "slab" and rep == 4:
            pr.set_slab_tile(5)
        vmax, rank, _m = pr.price(S[rep % 2])
        assert vmax == ref[rep % 2][0] and np.array_equal(rank, ref[rep % 2][1]), (rep, vmax, ref[rep % 2][0])
        pm.vmax64(rng.standard_normal(menu.size))          # eager kernels between replays
    pr.close()
    pm.close()


@pytest.mark.parametrize("mode", ["float32", "uint16", "uint24"])
@pytest.mark.parametrize("n", [9, 14, 17, 18])
def test_stream_pricing_bit_exact(n, mode):
    """Stream pricing (float32, uint16, uint24 layers): vmax, margin and the ranking (the column that
    enters the master) equal the oracle's bit for bit: the stored predecessors walked from the full
    set, then the sub-lattice DP (pricing_stream.py:348-405). n = 18 prices config2's rows (23,595
    cells). Both random scores and quantized duals (multiples of 1/4, where ties are everywhere and
    the lowest-member rule decides) are priced."""
    rng = np.random.default_rng(200 + n)
    if n == 18:
        cd, _n = instances.config2()
        menu, alt = cd[:, 0].astype(np.uint64), cd[:, 1].astype(np.int32)
    else:
        menu, alt, p = _medium(n, min(3000, n << (n - 3)), rng)
    perm, masks_x, _alts, offsets = xmajor(menu, alt, n)
    pr = Pricer(n, masks_x, offsets, layout="stream", dtype=mode)
    omode = {"float32": "f32", "uint16": "u16", "uint24": "u24"}[mode]
    for quant in (False, True):
        scores = rng.standard_normal(menu.size).astype(np.float32)
        if quant:
            scores = (np.round(scores * 4) / 4).astype(np.float32)
        print("[case] stream %s n=%d M=%d quantized=%s" % (mode, n, menu.size, quant), flush=True)
        vmax, rank, margin = pr.price(scores)
        vo, mo, pred_top = O.price_stream(n, omode, masks_x, scores, offsets)
        assert vmax == vo and abs(margin - mo) <= 1e-12 * max(1.0, mo)
        rank_o = O.stream_ranking(n, masks_x, scores, offsets, pred_top)
        assert np.array_equal(rank, rank_o), (quant, rank, rank_o)
        if mode == "float32":
            vmask, rank_mask, _V, _p = O.price_mask(n, masks_x, scores, offsets)
            assert vmax == vmask and np.array_equal(rank, rank_mask)
    pr.close()


@pytest.mark.parametrize("layout", ["mask", "slab", "stream"])
@pytest.mark.parametrize("n,cells", [(9, 1000), (14, 7001)])
def test_column_store_matches_oracle(n, cells, layout):
    """rumgpu_cg_cols_*: each ranking's column bitset equals the oracle's ranking column on the
    x-major rows, its Gram row equals the integer overlaps, its h = col'p equals the exactly rounded
    sum to 1e-12 (relative), p_out = sum_t lam_t col_(idx_t) equals the
    same sum taken in t order on the host bit for bit, and pricing the resident scores equals pricing
    those scores passed from the host. 70 columns from an initial room of 2 also cover the store's
    growth; a reload of the first 5 columns continues at index 5."""
    rng = np.random.default_rng(300 + n)
    menu, alt, p = _medium(n, cells, rng)
    perm, masks_x, alt_x, offsets = xmajor(menu, alt, n)
    M = masks_x.size
    W = (M + 63) // 64
    pr = Pricer(n, masks_x, offsets, layout=layout)
    p_x = p[perm].astype(np.float32).astype(np.float64)
    pr.cols_init(p_x, 2)
    cols, bits_all = [], []
    for k in range(70):
        rank = rng.permutation(n).astype(np.int32)
        gram, hk, bits = pr.cols_add(rank, k)
        col = O.ranking_column(n, masks_x, alt_x, rank).astype(bool)
        assert abs(hk - math.fsum(p_x[col])) <= 1e-12 * max(1.0, abs(hk)), k
        assert np.array_equal(np.unpackbits(bits.view(np.uint8), bitorder="little")[:M].astype(bool), col), k
        assert bits.size == W and not np.any(np.unpackbits(bits.view(np.uint8), bitorder="little")[M:])
        cols.append(col)
        bits_all.append(bits)
        assert np.array_equal(gram, [float((c & col).sum()) for c in cols]), k
    for trial in range(3):
        idx = np.sort(rng.choice(70, size=int(rng.integers(1, 70)), replace=False)).astype(np.int32)
        lam = rng.random(idx.size)
        lam /= lam.sum()
        pr.cols_scores(idx, lam)
        acc = np.zeros(M)
        for i, v in zip(idx, lam):
            acc = acc + np.where(cols[i], v, 0.0)
        assert np.array_equal(pr.cols_pout(), acc)
        scores = (p_x - acc).astype(np.float32)
        a = pr.price(None)
        b = pr.price(scores)
        assert a[0] == b[0] and np.array_equal(a[1], b[1]) and a[2] == b[2]
    pr.cols_load(np.stack(bits_all[:5]))
    rank = rng.permutation(n).astype(np.int32)
    gram, _hk, bits = pr.cols_add(rank, 5)
    col = O.ranking_column(n, masks_x, alt_x, rank).astype(bool)
    assert np.array_equal(gram, [float((c & col).sum()) for c in cols[:5] + [col]])
    pr.cols_scores(np.array([5], dtype=np.int32), np.array([1.0]))
    assert np.array_equal(pr.cols_pout(), np.where(col, 1.0, 0.0))
    with pytest.raises(ValueError):
        pr.cols_add(np.zeros(n, dtype=np.int32), 6)          # not a permutation
    with pytest.raises(ValueError):
        pr.cols_scores(np.array([6], dtype=np.int32), np.array([1.0]))   # only columns 0..5 are stored
    pr.close()


@pytest.mark.parametrize("n,alt0", [(7, False), (12, False), (16, False), (13, True)])
def test_mask_lb64_matches_slab_bound(n, alt0):
    """flow_cg's float64 lower bound on the mask lattice and on the stream layout's two float64
    layers equals the slab form (fw_lb_rows_gpu) and the oracle's within float64 rounding"""
    import ctypes
    rng = np.random.default_rng(600 + n)
    menu, alt, p = _medium(n, min(5000, n << (n - 3)), rng)
    if alt0:                                           # rows of alternative 0 only
        keep = alt == 0
        menu, alt, p = menu[keep], alt[keep], p[keep]
    perm, masks_x, _alts, offsets = xmajor(menu, alt, n)
    pr = Pricer(n, masks_x, offsets, layout="mask")
    y = np.clip(p + 0.05 * rng.standard_normal(p.size), 0.0, 1.0)
    lb_mask = pr.lb64(perm, y, p)
    obs = O.obs_index(n, menu, alt)
    lb_o, _ = O.fw_lb64(n, obs, y, p)
    ctx = _lib.context(0)
    lbv = ctypes.c_double()
    _lib.check(ctx.lib.rumgpu_lb64(ctx.h, n, _lib.ptr(menu), _lib.ptr(alt), menu.size, _lib.ptr(y), _lib.ptr(p),
                                   ctypes.byref(lbv)))
    assert abs(lb_mask - lb_o) <= 1e-10 * max(1.0, lb_o)
    assert abs(lb_mask - lbv.value) <= 1e-10 * max(1.0, lb_o)
    # the two-layer float64 sweep of the stream layout (sweep_max64): the same bound
    ps = Pricer(n, masks_x, offsets, layout="stream", dtype="uint16")
    lb_stream = ps.lb64(perm, y, p)
    assert abs(lb_stream - lb_mask) <= 1e-12 * max(1.0, lb_o)
    ps.close()
    pr.close()


@pytest.mark.parametrize("n,D,alt0", [(7, 3, False), (12, 8, False), (13, 11, True), (16, 2, False)])
def test_batched_vmax64_equals_single(n, D, alt0):
    """the float64 longest paths of D residual vectors in one batched sweep equal the single-draw
    sweep of each bit for bit (D above 8 runs in chunks)"""
    rng = np.random.default_rng(700 + n)
    menu, alt, p = _medium(n, min(5000, n << (n - 3)), rng)
    if alt0:
        keep = alt == 0
        menu, alt, p = menu[keep], alt[keep], p[keep]
    perm, masks_x, _alts, offsets = xmajor(menu, alt, n)
    pr = Pricer(n, masks_x, offsets, layout="mask")
    R = rng.standard_normal((D, menu.size))
    vb = pr.vmax64_batch(R)
    for d in range(D):
        assert vb[d] == pr.vmax64(R[d]), d
    pr.close()


@pytest.mark.parametrize("graph", [0, 5])
def test_set_beta_carries_the_state(graph):
    """rumgpu_bm_set_beta settles the owed s, y update with the old penalty, then iterates with the
    new one: equal, bit for bit, to restarting a new object at the new penalty from the state"""
    n = 12
    rng = np.random.default_rng(21)
    menu, alt, p = _medium(n, 2500, rng)
    a = Handle(n, menu, alt, p, dtype="float64", beta=0.7)
    b = Handle(n, menu, alt, p, dtype="float64", beta=0.7)
    if graph:
        a.set_graph(graph)
    a.run(10)
    b.run(10)
    a.set_beta(1.9)
    st = b.get_state()
    c = Handle(n, menu, alt, p, dtype="float64", beta=1.9)
    c.set_state(st)
    a.run(10)
    c.run(10)
    sa, sc = a.get_state(), c.get_state()
    for k in ("x2", "s", "y", "z"):
        assert np.array_equal(sa[k], sc[k]), k
    for h in (a, b, c):
        h.close()


@pytest.mark.parametrize("n,plan", [(13, (6, 2, 3)), (13, (12, 4, 4)), (18, (11, 3, 4)), (18, (8, 2, 4)),
                                    (14, (13, 4, 3)), (18, (13, 4, 4))])
def test_window_plans_match_oracle(n, plan):
    """any window plan (rumgpu_bm_set_plan) gives the reference iteration within rounding; a 13-bit
    window at bit 0 (68 KiB of float64 LDS per tile) is one window at n = 14 and two at n = 18"""
    rng = np.random.default_rng(80 + n)
    menu, alt, p = _medium(n, min(6000, n << (n - 2)), rng)
    h = Handle(n, menu, alt, p, dtype="float64", beta=1.1)
    h.set_plan(*plan)
    h.set_graph(4)
    st = _oracle_state(n, menu, alt, p, 1.1, np.float64)
    h.run(13)
    st.run(13)
    got = h.get_state()
    for k in ("x2", "s", "y", "z"):
        ref = getattr(st, k)
        err = float(np.abs(got[k] - ref).max()) / max(1.0, float(np.abs(ref).max()))
        assert err < 1e-11, (plan, k, err)
    h.close()


def test_autotuner_picks_a_valid_plan(tmp_path, monkeypatch):
    from solver.bm_admm import autotune
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "plans.json"))
    monkeypatch.setattr(autotune, "_MEM", {})
    n = 14
    rng = np.random.default_rng(5)
    menu, alt, p = _medium(n, 3000, rng)
    h = Handle(n, menu, alt, p, dtype="float64", beta=1.0)
    plan, us = autotune.tune(h, device="test", iters=20, warm=5)
    assert tuple(plan) in autotune.candidates(n - 1) and len(us) >= 2 and h.graph_len == 0
    assert (tmp_path / "plans.json").exists()
    # a second object of the same n and dtype takes the cached plan without timing
    h2 = Handle(n, menu, alt, p, dtype="float64", beta=1.0)
    plan2, _ = autotune.tune(h2, device="test")
    assert plan2 == plan
    h.close()
    h2.close()
