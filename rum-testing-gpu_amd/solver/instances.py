"""Seeded instances: the reference's paper generators and the benchmark configurations.

Paper generators (data/gen_instances.py:37-152 of the reference; legacy RandomState streams, so
an instance is byte-identical to the one its certified optimum in data/opt_refs.json was computed
on, which its sha256 fingerprint proves):
    gen_pairwise(n)          every pair {i, j}, both alternatives, p from rand(2) normalized
    gen_medium(n, mult)      uniform random menus of size >= 2 (no repeats) until mp >= mult n (n-1),
                             all members observed, p per menu from rand(k) normalized
    gen_dense(n)             every menu of size >= 2
    get_instance(n, cat)     the category names: sparse, medium, dense, d1pct, m<mult>
    fingerprint(cd)          sha256[:16] of the (M, 3) float64 rows

Benchmark configurations (synthetic, numpy Generator streams; seeds fixed here):
    config1 ... config5      the five workloads the benchmark and the parity tests use (docstrings)
Rows come as (M, 3) float64 arrays (menu, alternative, p) like the reference's choice_data.
"""
import hashlib

import numpy as np


def fingerprint(cd):
    return hashlib.sha256(np.ascontiguousarray(cd, dtype=np.float64).tobytes()).hexdigest()[:16]


def _rows(om, ox, ph):
    return np.column_stack([np.asarray(om, np.float64), np.asarray(ox, np.float64), np.asarray(ph, np.float64)])


def gen_pairwise(n, phat_seed=100):
    """All n(n-1)/2 pairs, both alternatives observed."""
    rng = np.random.RandomState(phat_seed)
    om, ox, ph = [], [], []
    for i in range(n):
        for j in range(i + 1, n):
            g = rng.rand(2)
            g = g / g.sum()
            m = (1 << i) | (1 << j)
            om += [m, m]
            ox += [i, j]
            ph += [float(g[0]), float(g[1])]
    return _rows(om, ox, ph), len(ph)


def gen_medium(n, mult=20, menu_seed=42, phat_seed=100):
    """Uniform random menus of size >= 2, without repeats, until mp >= mult n (n - 1)."""
    target = mult * n * (n - 1)
    rng = np.random.RandomState(menu_seed)
    seen, menus, mp = set(), [], 0
    while mp < target:
        m = int(rng.randint(0, 1 << n))
        if m in seen:
            continue
        k = bin(m).count("1")
        if k < 2:
            continue
        seen.add(m)
        menus.append(m)
        mp += k
    prng = np.random.RandomState(phat_seed)
    om, ox, ph = [], [], []
    for m in menus:
        items = [i for i in range(n) if (m >> i) & 1]
        g = prng.rand(len(items))
        g = g / g.sum()
        for kk, x in enumerate(items):
            om.append(m)
            ox.append(x)
            ph.append(float(g[kk]))
    return _rows(om, ox, ph), len(ph)


def gen_dense(n, phat_seed=100):
    """Every menu of size >= 2: rows item-major, p = rand normalized per menu."""
    N = 1 << n
    allm = np.arange(N, dtype=np.int64)
    pop = np.zeros(N, dtype=np.int64)
    for b in range(n):
        pop += (allm >> b) & 1
    keep = pop >= 2
    om = np.concatenate([allm[((allm >> i) & 1).astype(bool) & keep] for i in range(n)])
    ox = np.concatenate([np.full(int((((allm >> i) & 1).astype(bool) & keep).sum()), i, dtype=np.int64)
                         for i in range(n)])
    g = np.random.RandomState(phat_seed).rand(len(om))
    ph = g / np.bincount(om, weights=g, minlength=N)[om]
    return _rows(om, ox, ph), len(ph)


def get_instance(n, cat):
    """Instance by the reference's category name (data/gen_instances.py:137-152)."""
    if cat == "sparse":
        return gen_pairwise(n)
    if cat == "medium":
        return gen_medium(n, mult=20)
    if cat == "dense":
        return gen_dense(n)
    if cat == "d1pct":
        return gen_medium(n, mult=round(0.01 * (n << (n - 1)) / (n * (n - 1))))
    if cat.startswith("m") and cat[1:].isdigit():
        return gen_medium(n, mult=int(cat[1:]))
    raise ValueError(cat)


def warmup_instance(n=10, seed=0):
    """All binary menus, both alternatives observed (common/instances.py:5-14 of the reference)."""
    rng = np.random.RandomState(seed)
    om, ox, ph = [], [], []
    for i in range(n):
        for j in range(i + 1, n):
            g = rng.rand(2)
            g /= g.sum()
            om += [(1 << i) | (1 << j)] * 2
            ox += [i, j]
            ph += [g[0], g[1]]
    return _rows(om, ox, ph)


# ------------------------------------------------------------------------------------- configs
def mixture_choice(n, menus, orders, weights):
    """q(a | D) = sum_r w_r [a is the best member of D under order r] for the rows of `menus`
    (all members, ascending); orders are permutations listed best first. Returns (om, ox, q)."""
    menus = np.asarray(menus, dtype=np.int64)
    pos = np.empty((len(orders), n), dtype=np.int64)
    for r, o in enumerate(orders):
        pos[r, np.asarray(o)] = np.arange(n)
    memb = ((menus[:, None] >> np.arange(n)[None, :]) & 1).astype(bool)            # (U, n)
    q = np.zeros((menus.size, n))
    for r in range(len(orders)):
        top = np.argmin(np.where(memb, pos[r][None, :], n), axis=1)
        q[np.arange(menus.size), top] += weights[r]
    mi, ai = np.nonzero(memb)
    return menus[mi], ai, q[mi, ai], q


def _random_menus(rng, n, target_cells):
    """Uniform bitmasks, popcount >= 2, no repeats, all members kept, until target cells."""
    seen, menus, cells = set(), [], 0
    while cells < target_cells:
        m = int(rng.integers(0, 1 << n))
        k = bin(m).count("1")
        if k < 2 or m in seen:
            continue
        seen.add(m)
        menus.append(m)
        cells += k
    return menus


def _noisy_rows(rng, n, menus, R, N):
    """Truth from an R-order Dirichlet(1) mixture, then Multinomial(N, q(. | D)) / N per menu."""
    orders = [rng.permutation(n) for _ in range(R)]
    w = rng.dirichlet(np.ones(R))
    menus = list(menus)
    om, ox, _q, qm = mixture_choice(n, menus, orders, w)
    memb = ((np.asarray(menus, dtype=np.int64)[:, None] >> np.arange(n)[None, :]) & 1).astype(bool)
    ph = np.zeros_like(qm)
    for u in range(len(menus)):
        idx = np.flatnonzero(memb[u])
        qq = qm[u, idx]
        qq = qq / qq.sum()
        ph[u, idx] = rng.multinomial(N, qq) / N
    mi, ai = np.nonzero(memb)
    return _rows(np.asarray(menus, dtype=np.int64)[mi], ai, ph[mi, ai]), dict(orders=orders, weights=w, truth=qm)


def config1(seed=8001):
    """n = 8, every menu of size >= 2 (247 menus, 1,016 cells); a 20-order Dirichlet(1) mixture
    gives exact RUM choice probabilities, then 200 replies per menu (multinomial) move p_hat
    slightly off the polytope."""
    rng = np.random.default_rng(seed)
    n = 8
    menus = [m for m in range(1 << n) if bin(m).count("1") >= 2]
    return _noisy_rows(rng, n, menus, 20, 200)[0], n


def config2(seed=18001, n=18, density=0.01, R=200, N=500):
    """n = 18, 1 % of the n 2^(n-1) cells observed (~23.6k cells): uniform menus of popcount >= 2
    with all members; p_hat from a 200-order Dirichlet mixture plus 500-reply multinomial noise."""
    rng = np.random.default_rng(seed)
    menus = _random_menus(rng, n, round(density * (n << (n - 1))))
    return _noisy_rows(rng, n, menus, R, N)[0], n


def config3(seed=22001):
    """n = 22, 0.5 % density (~230k cells), 1,000 orders, 1,000 replies per menu."""
    return config2(seed=seed, n=22, density=0.005, R=1000, N=1000)


def _panel(rng, n, menus, R_resp, per_resp, orders, weights):
    """A panel: each respondent answers per_resp menus drawn uniformly from `menus`, choosing the
    best member under an order drawn from the mixture. Returns dict of columns."""
    resp, men, cho = [], [], []
    menus = np.asarray(menus, dtype=np.int64)
    pos = np.empty((len(orders), n), dtype=np.int64)
    for r, o in enumerate(orders):
        pos[r, np.asarray(o)] = np.arange(n)
    for i in range(R_resp):
        mm = menus[rng.integers(0, menus.size, size=per_resp)]
        oo = rng.choice(len(orders), size=per_resp, p=weights)
        for m, o in zip(mm, oo):
            members = [a for a in range(n) if (int(m) >> a) & 1]
            best = min(members, key=lambda a: pos[o, a])
            resp.append(i)
            men.append(int(m))
            cho.append(int(best))
    return dict(respondent=np.asarray(resp), menu=np.asarray(men, dtype=np.int64), chosen=np.asarray(cho, dtype=np.int64))


def config4_alt0(seed=16002):
    """Batched-bootstrap workload: config5's generator at n = 16 (301 cells of alternative 0)."""
    return config5(seed, 16)


def config4(seed=16001):
    """Batched bootstrap workload: n = 16, menus at 2 % of the cells; a panel of 520 respondents
    each answering 12 menus drawn uniformly from them, choices from a 50-order mixture. Returns
    (panel columns, n); build the rows with inference.cells_from_panel(panel, n)."""
    rng = np.random.default_rng(seed)
    n = 16
    menus = _random_menus(rng, n, round(0.02 * (n << (n - 1))))
    orders = [rng.permutation(n) for _ in range(50)]
    w = rng.dirichlet(np.ones(50))
    return _panel(rng, n, menus, 520, 12, orders, w), n


def config5(seed=25001, n=25):
    """Large-n bootstrap workload: n = 25; 301 menus containing alternative 0 with popcount uniform
    in 2..25 (other members uniform); a panel of 520 respondents x 15 menus from a 100-order
    mixture. Rows: cells_from_panel(panel, 25, alternatives=[0]). The same generator at n = 16
    (seed 16002) is the batched-bootstrap workload whose draws straddle tau (config4's panel, with
    about 5 answers per menu, puts p_hat so far outside the polytope that the screen decides every
    draw)."""
    rng = np.random.default_rng(seed)
    seen, menus = set(), []
    while len(menus) < 301:
        k = int(rng.integers(2, n + 1))
        others = rng.choice(np.arange(1, n), size=k - 1, replace=False)
        m = 1 | int(np.bitwise_or.reduce(np.left_shift(1, others)))
        if m in seen:
            continue
        seen.add(m)
        menus.append(m)
    orders = [rng.permutation(n) for _ in range(100)]
    w = rng.dirichlet(np.ones(100))
    return _panel(rng, n, menus, 520, 15, orders, w), n
