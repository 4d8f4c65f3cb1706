"""CPU oracle for the BM-ADMM / Flow-CG hot path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module, and
only as the checker (or the timed CPU baseline). The product path never touches it.

The arithmetic lives in rum_oracle.cpp (a restatement of the reference's kernels, built by the
Makefile next to it into liboracle.so); this module wraps it with ctypes and adds the pieces that
are only ever run at small sizes:

    brute_pricing        max over all n! rankings (pins pricing and the lower bound at n <= 8)
    qp_simplex           active-set solver for min 1/2 l'Ql - h'l on the simplex: a restatement of
                         common/qp_native.py:37-100 (solve_qp_native) and :20-33 (project_simplex)
    exact_projection     column generation with brute-force pricing: the exact distance at n <= 8
    bm_certify           the reference's certified BM-ADMM loop (rum_solver.py:869-925 with
                         bm_admm/solver.py:148-377) on the C iteration, for n <= 12

Parity status of each function is recorded in the header of tests/test_oracle.py, which pins it.
"""
import ctypes
import itertools
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    """Compile rum_oracle.cpp (idempotent: make decides)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i64, i32, f64, f32 = ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_float
        sig = {
            "orc_butterfly_f64": (None, [P, i32, f64, f64, f64, f64]),
            "orc_butterfly_f32": (None, [P, i32, f32, f32, f32, f32]),
            "orc_obs_index": (None, [i32, i64, P, P, P]),
            "orc_prox_f64": (None, [i32, P, P, P, P, f64]),
            "orc_prox_f32": (None, [i32, P, P, P, P, f32]),
            "orc_inv_gl": (None, [i32, f64, f64, f64, P]),
            "orc_admm_iters_f64": (None, [i32, i64, P, P, P, P, P, P, P, f64, f64, f64, f64, P]),
            "orc_admm_iters_f32": (None, [i32, i64, P, P, P, P, P, P, P, f64, f64, f64, f64, P]),
            "orc_push_ub64": (f64, [i32, P, i64, P, P, P]),
            "orc_fw_lb64": (f64, [i32, i64, P, P, P, P]),
            "orc_price_mask": (f32, [i32, P, P, P, P, P, P]),
            "orc_colex_rank": (ctypes.c_uint64, [ctypes.c_uint64]),
            "orc_price_stream": (f64, [i32, i32, i32, P, P, P, P, P]),
            "orc_ranking_column": (None, [i32, i64, P, P, P, P]),
            "orc_threads": (i32, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ----------------------------------------------------------------------------- slab basics
def compress_bit(m, i):
    return (m & ((1 << i) - 1)) | ((m >> (i + 1)) << i)


def obs_index(n, menu, alt):
    menu = np.ascontiguousarray(menu, dtype=np.uint64)
    alt = np.ascontiguousarray(alt, dtype=np.int32)
    out = np.empty(menu.size, dtype=np.int64)
    lib().orc_obs_index(n, menu.size, _p(menu), _p(alt), _p(out))
    return out


def scatter(n, idx, vals, dtype=np.float64):
    s = np.zeros(n << (n - 1), dtype=dtype)
    s[idx] = vals
    return s


def butterfly(v, n, a, b, c, d):
    """(a, b, c, d) 2x2 butterfly over every rest bit of every item block (returns a copy)."""
    v = np.array(v, copy=True)
    if v.dtype == np.float64:
        lib().orc_butterfly_f64(_p(v), n, a, b, c, d)
    elif v.dtype == np.float32:
        lib().orc_butterfly_f32(_p(v), n, a, b, c, d)
    else:
        raise TypeError(v.dtype)
    return v


def butterfly_dense(v, n, a, b, c, d):
    """The same transform as an explicit Kronecker product (n <= 10): the independent check."""
    Nloc = 1 << (n - 1)
    M2 = np.array([[a, b], [c, d]], dtype=np.float64)
    K = np.ones((1, 1))
    for _ in range(n - 1):
        K = np.kron(M2, K)              # bit 0 is the fastest index: kron(M_{n-2}, ..., M_0)
    V = np.asarray(v, dtype=np.float64).reshape(n, Nloc)
    return (V @ K.T).reshape(-1)


def prox(n, xk, xhat, wmask, beta2):
    xk = np.ascontiguousarray(xk)
    out = np.empty_like(xk)
    xhat = np.ascontiguousarray(xhat, dtype=xk.dtype)
    wmask = np.ascontiguousarray(wmask, dtype=xk.dtype)
    if xk.dtype == np.float64:
        lib().orc_prox_f64(n, _p(xk), _p(out), _p(xhat), _p(wmask), float(beta2))
    else:
        lib().orc_prox_f32(n, _p(xk), _p(out), _p(xhat), _p(wmask), float(np.float32(beta2)))
    return out


def eigen(precond_p, beta1, beta2):
    """numpy eigh of [[1,-1],[-1,1+p^2]] as the reference computes it (bm_admm/solver.py:207-214):
    (e0, e1, gamma, Q row-major)."""
    p = float(precond_p)
    G = np.array([[1.0, -1.0], [-1.0, 1.0 + p * p]])
    evals, Q = np.linalg.eigh(G)
    return float(evals[0]), float(evals[1]), float(beta2 / beta1), np.ascontiguousarray(Q, dtype=np.float64).ravel()


def inv_gl(n, e0, e1, gamma):
    out = np.empty(1 << (n - 1))
    lib().orc_inv_gl(n, e0, e1, gamma, _p(out))
    return out


class AdmmState:
    """The reference's BM-ADMM state for rows (menu, alt, p): x2, s, y, z slabs, dense xhat/wmask."""

    def __init__(self, n, menu, alt, p, beta, alpha=1.95, precond_p=math.sqrt(2.0), dtype=np.float64):
        self.n, self.dt = n, np.dtype(dtype)
        self.beta, self.alpha, self.pp = float(beta), float(alpha), float(precond_p)
        self.obs = obs_index(n, menu, alt)
        self.p = np.asarray(p, dtype=np.float64)
        m2 = n << (n - 1)
        self.xhat = scatter(n, self.obs, self.p.astype(self.dt), self.dt)
        self.wmask = scatter(n, self.obs, np.ones(len(self.obs), dtype=self.dt), self.dt)
        self.e0, self.e1, self.gamma, self.Q = eigen(precond_p, beta, beta)
        self.inv_gl = inv_gl(n, self.e0, self.e1, self.gamma).astype(self.dt)
        self.x2 = prox(n, self.xhat, self.xhat, self.wmask, self.beta)        # cold start (solver.py:264-267)
        self.s = np.zeros(m2, self.dt)
        self.y = np.zeros(m2, self.dt)
        self.z = np.zeros(m2, self.dt)

    def run(self, iters):
        fn = lib().orc_admm_iters_f64 if self.dt == np.float64 else lib().orc_admm_iters_f32
        fn(self.n, int(iters), _p(self.x2), _p(self.s), _p(self.y), _p(self.z), _p(self.xhat), _p(self.wmask),
           _p(self.inv_gl), self.beta, self.beta, self.alpha, self.pp, _p(self.Q))

    def phi(self):
        d = self.x2[self.obs].astype(np.float64) - self.p
        return math.sqrt(math.fsum(d * d))


def push_ub64(n, x, obs, phat):
    x = np.ascontiguousarray(x, dtype=np.float64)
    obs = np.ascontiguousarray(obs, dtype=np.int64)
    phat = np.ascontiguousarray(phat, dtype=np.float64)
    out = np.empty(obs.size)
    ub = lib().orc_push_ub64(n, _p(x), obs.size, _p(obs), _p(phat), _p(out))
    return ub, out


def fw_lb64(n, obs, y, phat):
    obs = np.ascontiguousarray(obs, dtype=np.int64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    phat = np.ascontiguousarray(phat, dtype=np.float64)
    rc = ctypes.c_double()
    lb = lib().orc_fw_lb64(n, obs.size, _p(obs), _p(y), _p(phat), ctypes.byref(rc))
    return lb, rc.value


def bm_certify(n, menu, alt, p, beta, abs_tol, rel_tol=0.0, check=50, max_iter=200000, dtype=np.float64):
    """The certified BM-ADMM loop (rum_solver.py:869-925 driving bm_admm/solver.py:308-352): every
    `check` iterations the float64 push (ub, running min) and the float64 lower bound (running max)
    at the iterate; stop at ub - lb <= abs_tol or (ub - lb)/lb <= rel_tol. Returns dict."""
    st = AdmmState(n, menu, alt, p, beta, dtype=dtype)
    ub, lb, it = math.inf, 0.0, 0
    p_fit = None
    while it < max_iter:
        st.run(check)
        it += check
        x = st.x2.astype(np.float64)
        u, xo = push_ub64(n, x, st.obs, st.p)
        if u < ub:
            ub, p_fit = u, xo
        lb = max(lb, fw_lb64(n, st.obs, x[st.obs], st.p)[0])
        if ub - lb <= abs_tol or (lb > 0 and (ub - lb) / lb <= rel_tol):
            break
    return dict(ub=ub, lb=lb, iters=it, p_fit=p_fit, state=st)


# ----------------------------------------------------------------------------- pricing
def xmajor(menu, alt, n):
    """Rows in the pricer's x-major order (flow_cg/cg_core.py:184-195): stable sort by alternative."""
    alt = np.asarray(alt, dtype=np.int32)
    perm = np.argsort(alt, kind="stable")
    counts = np.bincount(alt, minlength=n)
    offsets = np.concatenate(([0], np.cumsum(counts))).astype(np.uint64)
    return perm, np.ascontiguousarray(np.asarray(menu, dtype=np.uint64)[perm]), offsets


def price_mask(n, masks_x, scores_x, offsets):
    scores_x = np.ascontiguousarray(scores_x, dtype=np.float32)
    masks_x = np.ascontiguousarray(masks_x, dtype=np.uint64)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    V = np.empty(1 << n, dtype=np.float32)
    pred = np.empty(1 << n, dtype=np.uint8)
    rank = np.empty(n, dtype=np.int32)
    vmax = lib().orc_price_mask(n, _p(masks_x), _p(scores_x), _p(offsets), _p(V), _p(pred), _p(rank))
    return float(np.float32(vmax)), rank, V, pred


def price_stream(n, mode, masks_x, scores_x, offsets, k_pred=None):
    """mode: 'f32', 'u16', 'u24'. Returns (vmax, margin, pred_top)."""
    m = {"f32": 0, "u16": 1, "u24": 2}[mode]
    k_pred = max(1, min(n - 10 if k_pred is None else int(k_pred), n))
    tot = sum(math.comb(n, k) for k in range(k_pred, n + 1))
    pred_top = np.zeros(tot, dtype=np.uint8)
    scores_x = np.ascontiguousarray(scores_x, dtype=np.float32)
    masks_x = np.ascontiguousarray(masks_x, dtype=np.uint64)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    margin = ctypes.c_double()
    vmax = lib().orc_price_stream(n, m, k_pred, _p(masks_x), _p(scores_x), _p(offsets), _p(pred_top),
                                  ctypes.byref(margin))
    return vmax, margin.value, pred_top


def stream_ranking(n, masks_x, scores_x, offsets, pred_top, k_pred=None):
    """The stream pricer's ranking (StreamPricer.backtrace / _solve_sublattice,
    flow_cg/pricing_stream.py:348-405): from the full set, take the stored predecessor of each layer
    k = n .. k_pred+1 (pred_top holds layers k_pred..n back to back, each in colex order); the node
    reached has k_pred members, whose order comes from the exact float32 mask DP on its sub-lattice:
    the rows whose menu lies inside it, members relabelled 0..K-1 in ascending order, x-major by the
    new label with the rows' order kept. Returns rank_pos[n] (0 = most preferred)."""
    k_pred = max(1, min(n - 10 if k_pred is None else int(k_pred), n))
    masks_x = np.asarray(masks_x, dtype=np.uint64)
    scores_x = np.asarray(scores_x, dtype=np.float32)
    offsets = np.asarray(offsets, dtype=np.uint64)
    layer_off, acc = {}, 0
    for k in range(k_pred, n + 1):
        layer_off[k] = acc
        acc += math.comb(n, k)
    rank_pos = np.zeros(n, dtype=np.int32)
    D, pos = (1 << n) - 1, 0
    for k in range(n, k_pred, -1):
        x = int(pred_top[layer_off[k] + colex_rank(D)])
        rank_pos[x] = pos
        pos += 1
        D ^= 1 << x
    elems = [b for b in range(n) if (D >> b) & 1]
    K = len(elems)
    assert K == k_pred
    alt_x = np.repeat(np.arange(n), np.diff(offsets.astype(np.int64)))
    inside = (masks_x & np.uint64(D)) == masks_x
    sub_m = np.zeros(int(inside.sum()), dtype=np.uint64)
    for j, e in enumerate(elems):
        sub_m |= ((masks_x[inside] >> np.uint64(e)) & np.uint64(1)) << np.uint64(j)
    relabel = {e: j for j, e in enumerate(elems)}
    sub_a = np.array([relabel[int(a)] for a in alt_x[inside]], dtype=np.int64)
    sub_s = scores_x[inside]
    order = np.argsort(sub_a, kind="stable")
    sub_off = np.concatenate(([0], np.cumsum(np.bincount(sub_a, minlength=K)))).astype(np.uint64)
    _v, sub_rank, _V, _p = price_mask(K, sub_m[order], sub_s[order], sub_off)
    for j, e in enumerate(elems):
        rank_pos[e] = pos + int(sub_rank[j])
    return rank_pos


def threads():
    """host threads the oracle's parallel loops use"""
    return int(lib().orc_threads())


def colex_rank(mask):
    return int(lib().orc_colex_rank(int(mask)))


def ranking_column(n, menu, alt, rank_pos):
    menu = np.ascontiguousarray(menu, dtype=np.uint64)
    alt = np.ascontiguousarray(alt, dtype=np.int32)
    rank_pos = np.ascontiguousarray(rank_pos, dtype=np.int32)
    out = np.empty(menu.size, dtype=np.uint8)
    lib().orc_ranking_column(n, menu.size, _p(menu), _p(alt), _p(rank_pos), _p(out))
    return out


def brute_pricing(n, menu, alt, scores):
    """max over all n! rankings of sum_rows scores * [alt is the top member of menu] (pure Python;
    n <= 8). Returns (value, the first maximizing ranking as rank positions)."""
    menu = np.asarray(menu, dtype=np.int64)
    alt = np.asarray(alt, dtype=np.int64)
    scores = np.asarray(scores, dtype=np.float64)
    best, arg = -math.inf, None
    memb = ((menu[:, None] >> np.arange(n)[None, :]) & 1).astype(bool)
    for perm in itertools.permutations(range(n)):
        pos = np.empty(n, dtype=np.int64)
        pos[list(perm)] = np.arange(n)
        top = np.argmin(np.where(memb, pos[None, :], n), axis=1)
        v = float(scores[top == alt].sum())
        if v > best:
            best, arg = v, pos.copy()
    return best, arg


# ----------------------------------------------------------------------------- master QP
def project_simplex(yv):
    """Euclidean projection onto {x >= 0, sum x = 1} by sorting (qp_native.py:20-33)."""
    k = yv.size
    u = np.sort(yv)[::-1]
    css = np.cumsum(u) - 1.0
    ok = u - css / np.arange(1, k + 1) > 0
    if not ok.any():
        return np.full(k, 1.0 / k)
    rho = int(np.flatnonzero(ok)[-1]) + 1
    return np.maximum(yv - css[rho - 1] / rho, 0.0)


def qp_simplex(Q, h, A_init=None, max_outer=200, tol=1e-10, reg=1e-12):
    """Active set for min 1/2 l'Ql - h'l, l >= 0, sum l = 1 (qp_native.py:37-100): solve the
    equality problem on the working set by Cholesky, drop the most negative weight, else add the
    most violated inactive column, else stop; finish with a simplex projection."""
    K = h.size
    A = [int(i) for i in (A_init or []) if 0 <= i < K] or [int(np.argmax(h))]
    lam_A, A_arr = None, None
    for _ in range(max_outer):
        A_arr = np.asarray(A, dtype=np.int64)
        k = A_arr.size
        QA = Q[np.ix_(A_arr, A_arr)] + reg * np.eye(k)
        rhs = np.column_stack([h[A_arr], np.ones(k)])
        try:
            L = np.linalg.cholesky(QA)
            sol = np.linalg.solve(L.T, np.linalg.solve(L, rhs))
        except np.linalg.LinAlgError:
            sol = np.linalg.lstsq(QA, rhs, rcond=None)[0]
        qh, qe = sol[:, 0], sol[:, 1]
        eQe, eQh = float(qe.sum()), float(qh.sum())
        mu = (eQh - 1.0) / eQe if abs(eQe) > 1e-30 else 0.0
        lam_A = qh - mu * qe
        if k > 0 and float(lam_A.min()) < -tol:
            del A[int(np.argmin(lam_A))]
            continue
        if K > k:
            inact = np.setdiff1d(np.arange(K), A_arr)
            grad = Q[np.ix_(inact, A_arr)] @ lam_A + mu - h[inact]
            i = int(np.argmin(grad))
            if grad[i] < -tol:
                A.append(int(inact[i]))
                continue
        break
    lam = np.zeros(K)
    lam[A_arr] = lam_A
    lam = project_simplex(lam)
    return lam, [int(i) for i in A if lam[i] > 1e-9]


def exact_projection(n, menu, alt, p, tol=1e-12, max_cols=5000):
    """dist(p) to the RUM polytope for n <= 8: column generation whose pricing is exact by
    enumeration of all n! rankings, with the restated active-set master. Returns (dist, lb, x)."""
    menu = np.asarray(menu, dtype=np.int64)
    alt = np.asarray(alt, dtype=np.int64)
    p = np.asarray(p, dtype=np.float64)
    memb = ((menu[:, None] >> np.arange(n)[None, :]) & 1).astype(bool)
    perms = list(itertools.permutations(range(n)))
    cols_all = np.empty((len(perms), menu.size), dtype=np.float64)
    for r, perm in enumerate(perms):
        pos = np.empty(n, dtype=np.int64)
        pos[list(perm)] = np.arange(n)
        top = np.argmin(np.where(memb, pos[None, :], n), axis=1)
        cols_all[r] = (top == alt)
    cols = [0]
    lam = np.array([1.0])
    A = [0]
    lb = 0.0
    for _ in range(max_cols):
        V = cols_all[cols].T
        Q = V.T @ V
        h = V.T @ p
        lam, A = qp_simplex(Q, h, A_init=A)
        x = V @ lam
        r = p - x
        f = float(r @ r)
        vals = cols_all @ r
        j = int(np.argmax(vals))
        gap = float(vals[j] - r @ x)
        lb = max(lb, math.sqrt(max(0.0, f - 2.0 * gap)))
        if gap <= tol or j in cols:
            break
        cols.append(j)
    return math.sqrt(f), lb, x
