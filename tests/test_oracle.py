"""Pins the CPU oracle (oracle/rum_oracle.cpp + oracle/oracle.py) before it is trusted as the checker.

Each oracle function and what pins it:
  butterfly          explicit Kronecker product of the 2x2 matrix (numpy kron, n <= 9); Mobius o zeta = I
  prox               KKT conditions of the (weighted) simplex projection; brute force of the support on
                     small menus; the equal-curvature case equals the sort-based projection
  admm_iters         invariants over many iterations: x2 stays in the menu simplexes; the certified loop's
                     bracket contains the exact distance (n = 5, 6)
  push_ub64          the pushed flow is a probability distribution over rankings: menu sums are 1 and
                     BM polynomials (Mobius of x_ub) are >= 0; ub >= exact distance
  fw_lb64            lb <= exact distance; at the exact projection lb = dist (tight)
  price_mask         brute force over all n! rankings (n <= 7), value and a maximizing ranking
  price_stream       f32 mode bit-identical to the mask sweep (values and ranking); u16 / u24 within
                     their returned margins of the float32 value
  qp_simplex         KKT conditions; agreement with brute-force support enumeration (K <= 8)
  exact_projection   brute force pricing, KKT of the master; distance 0 for a point built as a mixture
"""
import itertools
import math

import numpy as np
import pytest

from oracle import oracle as O


def _rows_all_menus(n, rng, kmin=2, frac=1.0):
    menus = [m for m in range(1 << n) if bin(m).count("1") >= kmin and rng.random() < frac]
    om, ox = [], []
    for m in menus:
        for a in range(n):
            if (m >> a) & 1:
                om.append(m)
                ox.append(a)
    return np.array(om, dtype=np.uint64), np.array(ox, dtype=np.int32)


def _random_p(menu, rng):
    p = rng.random(menu.size)
    sums = {}
    for m, v in zip(menu, p):
        sums[int(m)] = sums.get(int(m), 0.0) + v
    return np.array([v / sums[int(m)] for m, v in zip(menu, p)])


@pytest.mark.parametrize("n", [2, 3, 5, 8, 9])
def test_butterfly_is_kronecker(n):
    rng = np.random.default_rng(n)
    v = rng.standard_normal(n << (n - 1))
    for coefs in [(1.0, -1.0, 0.0, 1.0), (1.0, 1.0, 0.0, 1.0), (1.0, 0.0, 1.0, 1.0), (0.3, -1.7, 2.1, 0.9)]:
        got = O.butterfly(v, n, *coefs)
        want = O.butterfly_dense(v, n, *coefs)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12), coefs
    back = O.butterfly(O.butterfly(v, n, 1.0, -1.0, 0.0, 1.0), n, 1.0, 1.0, 0.0, 1.0)
    assert np.allclose(back, v, atol=1e-12)


def test_butterfly_f32_matches_f64_rounded():
    n = 7
    v = np.random.default_rng(0).standard_normal(n << (n - 1)).astype(np.float32)
    a = O.butterfly(v, n, 1.0, -1.0, 0.0, 1.0)
    b = O.butterfly(v.astype(np.float64), n, 1.0, -1.0, 0.0, 1.0)
    assert np.allclose(a, b, rtol=1e-5, atol=1e-4)


def _menu_members(S, n):
    return [i for i in range(n) if (S >> i) & 1]


@pytest.mark.parametrize("n", [3, 5, 7])
def test_prox_kkt(n):
    """Each menu's output u solves min sum_i c_i/2 (u_i - t_i/c_i)^2 on the simplex, c_i = b + 2 w_i:
    u >= 0, sum u = 1, and c_i u_i - t_i = -mu on the support, >= -mu off it."""
    rng = np.random.default_rng(100 + n)
    menu, alt = _rows_all_menus(n, rng, frac=0.6)
    p = _random_p(menu, rng)
    obs = O.obs_index(n, menu, alt)
    # observe only part of the cells so that mixed menus occur
    keep = rng.random(menu.size) < 0.7
    xhat = O.scatter(n, obs[keep], p[keep])
    wmask = O.scatter(n, obs[keep], np.ones(int(keep.sum())))
    beta = 1.7
    xk = rng.standard_normal(n << (n - 1))
    out = O.prox(n, xk, xhat, wmask, beta)
    Nloc = 1 << (n - 1)
    for S in range(1, 1 << n):
        mem = _menu_members(S, n)
        pos = [i * Nloc + O.compress_bit(S & ~(1 << i), i) for i in mem]
        u = out[pos]
        if len(mem) == 1:
            assert u[0] == 1.0
            continue
        assert np.all(u >= 0) and abs(u.sum() - 1.0) < 1e-12
        c = beta + 2 * wmask[pos]
        t = beta * xk[pos] + 2 * wmask[pos] * xhat[pos]
        g = c * u - t                           # gradient of sum c/2 u^2 - t u
        sup = u > 1e-14
        mu = -g[sup].mean()
        assert np.allclose(g[sup], -mu, atol=1e-10)
        assert np.all(g[~sup] >= -mu - 1e-10)


def test_prox_equal_curvature_is_sorting_projection():
    n = 6
    rng = np.random.default_rng(7)
    m2 = n << (n - 1)
    xk = rng.standard_normal(m2)
    z = np.zeros(m2)
    out = O.prox(n, xk, z, z, 2.5)
    Nloc = 1 << (n - 1)
    for S in range(3, 1 << n):
        mem = _menu_members(S, n)
        if len(mem) < 2:
            continue
        pos = [i * Nloc + O.compress_bit(S & ~(1 << i), i) for i in mem]
        vals = (2.5 * xk[pos]) / 2.5
        assert np.array_equal(out[pos], O.project_simplex(vals)) or np.allclose(out[pos], O.project_simplex(vals),
                                                                                atol=1e-15)


def _mixture_point(n, menu, alt, rng, R=4):
    orders = [rng.permutation(n) for _ in range(R)]
    w = rng.dirichlet(np.ones(R))
    x = np.zeros(menu.size)
    for o, wr in zip(orders, w):
        pos = np.empty(n, dtype=np.int64)
        pos[o] = np.arange(n)
        for j, (m, a) in enumerate(zip(menu, alt)):
            mem = _menu_members(int(m), n)
            top = min(mem, key=lambda i: pos[i])
            x[j] += wr * (top == a)
    return x


@pytest.mark.parametrize("n", [4, 6])
def test_price_mask_brute_force(n):
    rng = np.random.default_rng(n)
    menu, alt = _rows_all_menus(n, rng, frac=0.5)
    scores = rng.standard_normal(menu.size).astype(np.float32)
    perm, masks_x, offsets = O.xmajor(menu, alt, n)
    vmax, rank, V, pred = O.price_mask(n, masks_x, scores[perm], offsets)
    best, _arg = O.brute_pricing(n, menu, alt, scores.astype(np.float64))
    assert abs(vmax - best) < 1e-4
    col = O.ranking_column(n, menu, alt, rank)
    assert abs(float((scores.astype(np.float64) * col).sum()) - best) < 1e-4


def test_price_mask_quantized_exact():
    """With scores k 2^-10 (|k| small) every float32 sum is exact: vmax equals the brute force
    value exactly, and the returned ranking attains it."""
    n = 7
    rng = np.random.default_rng(3)
    menu, alt = _rows_all_menus(n, rng, frac=0.3)
    scores = (rng.integers(-512, 512, size=menu.size) / 1024.0).astype(np.float32)
    perm, masks_x, offsets = O.xmajor(menu, alt, n)
    vmax, rank, _V, _pred = O.price_mask(n, masks_x, scores[perm], offsets)
    best, _ = O.brute_pricing(n, menu, alt, scores.astype(np.float64))
    assert vmax == best
    col = O.ranking_column(n, menu, alt, rank)
    assert float((scores.astype(np.float64) * col).sum()) == best


@pytest.mark.parametrize("n", [6, 9, 12])
def test_stream_f32_equals_mask(n):
    rng = np.random.default_rng(20 + n)
    menu, alt = _rows_all_menus(n, rng, frac=min(1.0, 300.0 / (1 << n)))
    scores = rng.standard_normal(menu.size).astype(np.float32)
    perm, masks_x, offsets = O.xmajor(menu, alt, n)
    vmask, _rank, V, pred = O.price_mask(n, masks_x, scores[perm], offsets)
    vs, margin, pred_top = O.price_stream(n, "f32", masks_x, scores[perm], offsets, k_pred=1)
    assert vs == vmask and margin == 0.0
    # the stored predecessors of the top layers equal the mask sweep's pred at the same subsets
    off = 0
    for k in range(1, n + 1):
        cnt = math.comb(n, k)
        subsets = [m for m in range(1 << n) if bin(m).count("1") == k]      # ascending = colex order
        assert np.array_equal(pred_top[off:off + cnt], pred[subsets])
        off += cnt


@pytest.mark.parametrize("mode", ["u16", "u24"])
def test_stream_quantized_within_margin(mode):
    n = 11
    rng = np.random.default_rng(5)
    menu, alt = _rows_all_menus(n, rng, frac=0.15)
    scores = rng.standard_normal(menu.size).astype(np.float32)
    perm, masks_x, offsets = O.xmajor(menu, alt, n)
    vmask = O.price_mask(n, masks_x, scores[perm], offsets)[0]
    vq, margin, _ = O.price_stream(n, mode, masks_x, scores[perm], offsets)
    assert margin > 0
    assert abs(vq - vmask) <= margin + 1e-6 * abs(vmask)


@pytest.mark.parametrize("n,k_pred", [(7, 1), (7, 4), (7, 7), (10, 1), (10, 3), (12, 2), (12, 6)])
def test_stream_ranking_f32_equals_mask_ranking(n, k_pred):
    """In float32 mode the stream sweep's predecessors are the mask sweep's (lowest member on ties),
    so its ranking must be the mask ranking for every k_pred: k_pred = 1 walks the stored
    predecessors only, k_pred = n is the sub-lattice DP alone (every row inside the full set,
    nothing relabelled), values between split the walk at every depth. A wrong colex rank, layer
    offset or relabelling changes the ranking."""
    rng = np.random.default_rng(60 + n + k_pred)
    menu, alt = _rows_all_menus(n, rng, frac=min(1.0, 400.0 / (1 << n)))
    perm, masks_x, offsets = O.xmajor(menu, alt, n)
    for quant in (False, True):
        scores = rng.standard_normal(menu.size).astype(np.float32)
        if quant:            # multiples of 1/4: ties everywhere, exact sums
            scores = (np.round(scores * 4) / 4).astype(np.float32)
        sx = scores[perm]
        _v, rank_mask, _V, _pred = O.price_mask(n, masks_x, sx, offsets)
        _vs, _m, pred_top = O.price_stream(n, "f32", masks_x, sx, offsets, k_pred=k_pred)
        rank = O.stream_ranking(n, masks_x, sx, offsets, pred_top, k_pred=k_pred)
        assert np.array_equal(rank, rank_mask), (quant, rank, rank_mask)


@pytest.mark.parametrize("mode", ["u16", "u24"])
@pytest.mark.parametrize("n", [7, 11])
def test_stream_ranking_quantized_value(mode, n):
    """u16 / u24: the ranking is a permutation whose exact value (its column against the scores) is
    within two margins of the exact maximum (brute force over the n! rankings at n = 7, the mask
    sweep at n = 11), and at most that maximum."""
    rng = np.random.default_rng(80 + n)
    menu, alt = _rows_all_menus(n, rng, frac=min(1.0, 500.0 / (1 << n)))
    scores = rng.standard_normal(menu.size).astype(np.float32)
    perm, masks_x, offsets = O.xmajor(menu, alt, n)
    sx = scores[perm]
    for k_pred in (1, max(1, n - 4)):
        _vq, margin, pred_top = O.price_stream(n, mode, masks_x, sx, offsets, k_pred=k_pred)
        rank = O.stream_ranking(n, masks_x, sx, offsets, pred_top, k_pred=k_pred)
        assert sorted(rank.tolist()) == list(range(n))
        col = O.ranking_column(n, masks_x, np.repeat(np.arange(n), np.diff(offsets.astype(np.int64))), rank)
        val = float((sx.astype(np.float64) * col).sum())
        if n <= 7:
            best, _ = O.brute_pricing(n, menu, alt, scores.astype(np.float64))
        else:
            best = O.price_mask(n, masks_x, sx, offsets)[0]
        assert val <= best + 1e-5 and val >= best - 2.0 * margin - 1e-5, (val, best, margin)


def test_colex_rank_is_ascending_order():
    n = 9
    for k in range(n + 1):
        subsets = [m for m in range(1 << n) if bin(m).count("1") == k]
        assert [O.colex_rank(m) for m in subsets] == list(range(len(subsets)))


def test_qp_simplex_kkt_and_enumeration():
    rng = np.random.default_rng(11)
    for trial in range(20):
        K = int(rng.integers(2, 8))
        V = rng.random((12, K))
        p = rng.random(12)
        Q, h = V.T @ V, V.T @ p
        lam, sup = O.qp_simplex(Q, h)
        assert abs(lam.sum() - 1) < 1e-12 and lam.min() >= 0
        g = Q @ lam - h
        mu = -g[lam > 1e-9].mean()
        assert np.all(g + mu >= -1e-8)
        # brute force over supports: the best feasible equality solution
        best = math.inf
        for r in range(1, K + 1):
            for A in itertools.combinations(range(K), r):
                A = list(A)
                QA = Q[np.ix_(A, A)]
                try:
                    qh = np.linalg.solve(QA, h[A])
                    qe = np.linalg.solve(QA, np.ones(r))
                except np.linalg.LinAlgError:
                    continue
                mu2 = (qh.sum() - 1) / qe.sum()
                la = qh - mu2 * qe
                if la.min() < -1e-12:
                    continue
                full = np.zeros(K)
                full[A] = la
                best = min(best, 0.5 * full @ Q @ full - h @ full)
        assert 0.5 * lam @ Q @ lam - h @ lam <= best + 1e-9


def test_exact_projection_zero_for_a_mixture():
    n = 5
    rng = np.random.default_rng(2)
    menu, alt = _rows_all_menus(n, rng)
    x = _mixture_point(n, menu, alt, rng)
    d, lb, _ = O.exact_projection(n, menu, alt, x)
    assert d < 1e-7 and lb <= d + 1e-12


@pytest.mark.parametrize("n", [5, 6])
def test_push_lb_and_certified_admm_bracket_exact_distance(n):
    rng = np.random.default_rng(40 + n)
    menu, alt = _rows_all_menus(n, rng, frac=0.7)
    p = _random_p(menu, rng)
    d_exact, lb_exact, x_exact = O.exact_projection(n, menu, alt, p)
    assert abs(d_exact - lb_exact) < 1e-8
    obs = O.obs_index(n, menu, alt)
    # the lower bound at the exact projection is tight
    lb_at_opt, _ = O.fw_lb64(n, obs, x_exact, p)
    assert abs(lb_at_opt - d_exact) < 1e-6
    r = O.bm_certify(n, menu, alt, p, beta=1.0, abs_tol=1e-7, check=50, max_iter=100000)
    assert r["lb"] <= d_exact + 1e-9 <= r["ub"] + 2e-9
    assert r["ub"] - r["lb"] <= 1e-7
    # the pushed point is a RUM point: menu sums 1, Block-Marschak polynomials >= 0
    st = r["state"]
    ub, fit = O.push_ub64(n, st.x2.astype(np.float64), obs, p)
    assert ub >= d_exact - 1e-9
    sums = {}
    for m, v in zip(menu, fit):
        sums[int(m)] = sums.get(int(m), 0.0) + v
    full_menus = {int(m) for m in menu if sum(1 for a in alt[menu == m]) == bin(int(m)).count("1")}
    for m in full_menus:
        assert abs(sums[m] - 1.0) < 1e-9


def test_admm_iterate_stays_in_menu_simplexes():
    n = 6
    rng = np.random.default_rng(9)
    menu, alt = _rows_all_menus(n, rng, frac=0.5)
    p = _random_p(menu, rng)
    st = O.AdmmState(n, menu, alt, p, beta=0.8)
    st.run(37)
    Nloc = 1 << (n - 1)
    for S in range(1, 1 << n):
        mem = _menu_members(S, n)
        pos = [i * Nloc + O.compress_bit(S & ~(1 << i), i) for i in mem]
        u = st.x2[pos]
        assert np.all(u >= 0) and abs(u.sum() - 1) < 1e-12
