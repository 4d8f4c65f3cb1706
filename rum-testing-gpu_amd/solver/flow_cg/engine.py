"""Flow-CG host engine: the column-generation loop of the reference's solve_cgbfqp_cholbitset
(solver/flow_cg/solver.py:28-279) over the GPU pricing of librumgpu (rumgpu_cg_*).

Per column: the master QP on (Q, h) (common/qp_native.py), the mixture p_out updated by the
change in lam, the scalars err^2 = lam'Q lam - 2 h'lam + p'p and inner = h'lam - lam'Q lam, one
pricing sweep on the GPU with scores float32(p - p_out), the certified gap
LB = sqrt(max(err^2 - 2 (Vmax + margin - inner), 0)) (a running max), the stop tests, and the
priced ranking's 0/1 column. Q = V'V and h = V'p are grown from packed column bitsets
(qp_bitset_incremental.py:36-141 of the reference): Q[k, j] = popcount(bits_k & bits_j).
"""
import collections.abc
import ctypes
import math
import time

import numpy as np

from .. import _lib
from ..common.qp_native import QPState, solve_qp_native, solve_qp_numpy

LAYOUTS = {"mask": 0, "slab": 1, "stream": 2}
DTYPES = {"float32": 0, "uint16": 1, "uint24": 2}


def xmajor(menu, alt, n):
    """(perm, masks, alts, offsets): rows stably sorted by alternative (cg_core.py:184-195)."""
    alt = np.asarray(alt, dtype=np.int64)
    perm = np.argsort(alt, kind="stable")
    counts = np.bincount(alt, minlength=n)
    offsets = np.concatenate(([0], np.cumsum(counts))).astype(np.uint64)
    return perm, np.ascontiguousarray(np.asarray(menu, dtype=np.uint64)[perm]), alt[perm], offsets


class Pricer:
    """A rumgpu_cg object: price(scores) -> (vmax, rank_pos, margin)."""

    def __init__(self, n, masks_x, offsets, layout="mask", dtype="float32", k_pred=None, device=0):
        self.ctx = _lib.context(device)
        self.lib = self.ctx.lib
        self.n = int(n)
        self.masks = _lib.c_array(masks_x, np.uint64)
        self.offsets = _lib.c_array(offsets, np.uint64)
        self.M = int(self.masks.size)
        self.layout, self.dtype = layout, dtype
        h = ctypes.c_void_p()
        _lib.check(self.lib.rumgpu_cg_create(self.ctx.h, self.n, _lib.ptr(self.masks), _lib.ptr(self.offsets), self.M,
                                             LAYOUTS[layout], DTYPES[dtype], 0 if k_pred is None else int(k_pred),
                                             ctypes.byref(h)), "cg_create")
        self.h = h
        self.rank = np.empty(self.n, dtype=np.int32)

    def price(self, scores_x):
        """one sweep over scores_x (float32, x-major), or over the scores cols_scores left on the
        device when scores_x is None"""
        s = None if scores_x is None else _lib.c_array(scores_x, np.float32)
        vmax, margin = ctypes.c_double(), ctypes.c_double()
        _lib.check(self.lib.rumgpu_cg_price(self.h, None if s is None else _lib.ptr(s), ctypes.byref(vmax),
                                            _lib.ptr(self.rank), ctypes.byref(margin)), "cg_price")
        return vmax.value, self.rank.copy(), margin.value

    # column store of one run (rumgpu_cg_cols_*)
    def cols_init(self, p_x, cap):
        p = _lib.c_array(p_x, np.float64)
        _lib.check(self.lib.rumgpu_cg_cols_init(self.h, _lib.ptr(p), int(cap)), "cg_cols_init")

    def cols_load(self, bits):
        b = np.ascontiguousarray(bits, dtype=np.uint64)
        _lib.check(self.lib.rumgpu_cg_cols_load(self.h, int(b.shape[0]), _lib.ptr(b) if b.size else None),
                   "cg_cols_load")

    def cols_add(self, rank_pos, k):
        """append the column of a ranking; returns (its inner products with columns 0..k, its inner
        product with p, the packed bitset)"""
        r = _lib.c_array(rank_pos, np.int32)
        gram = np.empty(k + 1)
        hv = ctypes.c_double()
        bits = np.empty((self.M + 63) // 64, dtype=np.uint64)
        _lib.check(self.lib.rumgpu_cg_cols_add(self.h, _lib.ptr(r), _lib.ptr(gram), ctypes.byref(hv),
                                               _lib.ptr(bits)), "cg_cols_add")
        return gram, hv.value, bits

    def cols_scores(self, idx, lam):
        i = _lib.c_array(idx, np.int32)
        v = _lib.c_array(lam, np.float64)
        _lib.check(self.lib.rumgpu_cg_cols_scores(self.h, int(i.size), _lib.ptr(i) if i.size else None,
                                                  _lib.ptr(v) if v.size else None), "cg_cols_scores")

    def cols_pout(self):
        out = np.empty(self.M)
        _lib.check(self.lib.rumgpu_cg_cols_pout(self.h, _lib.ptr(out)), "cg_cols_pout")
        return out

    def price_batch(self, scores):
        """D sweeps at once (mask layout): scores (D, M) x-major -> vmax[D], rank_pos (D, n)"""
        s = np.ascontiguousarray(scores, dtype=np.float32)
        D = s.shape[0]
        vmax = np.empty(D)
        rank = np.empty((D, self.n), dtype=np.int32)
        _lib.check(self.lib.rumgpu_cg_price_batch(self.h, D, _lib.ptr(s), _lib.ptr(vmax), _lib.ptr(rank)),
                   "cg_price_batch")
        return vmax, rank

    def price_batch_lb(self, scores, R):
        """price_batch(scores) and vmax64_batch(R) in one call (the same values): on the window-pass
        sweep the float64 maxima come out of the pricing sweep -> vmax[D], rank_pos (D, n), vmax64[L]"""
        s = np.ascontiguousarray(scores, dtype=np.float32)
        r = np.ascontiguousarray(R, dtype=np.float64)
        D, L = s.shape[0], r.shape[0]
        vmax, v64 = np.empty(D), np.empty(L)
        rank = np.empty((D, self.n), dtype=np.int32)
        _lib.check(self.lib.rumgpu_cg_price_batch_lb(self.h, D, _lib.ptr(s), _lib.ptr(vmax), _lib.ptr(rank), L,
                                                     _lib.ptr(r), _lib.ptr(v64)), "cg_price_batch_lb")
        return vmax, rank, v64

    def vmax64(self, r_x):
        """float64 longest path for residuals r per x-major row (mask or stream layout)"""
        r = _lib.c_array(r_x, np.float64)
        v = ctypes.c_double()
        _lib.check(self.lib.rumgpu_cg_vmax64(self.h, _lib.ptr(r), ctypes.byref(v)), "cg_vmax64")
        return v.value

    def vmax64_batch(self, R):
        """vmax64 for D residual vectors at once (mask layout): R (D, M) x-major -> vmax[D]"""
        r = np.ascontiguousarray(R, dtype=np.float64)
        v = np.empty(r.shape[0])
        _lib.check(self.lib.rumgpu_cg_vmax64_batch(self.h, int(r.shape[0]), _lib.ptr(r), _lib.ptr(v)),
                   "cg_vmax64_batch")
        return v

    def lb64(self, perm, y, p):
        """float64 Frank-Wolfe lower bound at y (input row order), fw_lb_rows_gpu's formula:
        r = p - y, lb = sqrt(max(0, |r|^2 - 2 (vmax(r) - r.y))) with compensated sums"""
        y = np.asarray(y, dtype=np.float64)
        r = np.asarray(p, dtype=np.float64) - y
        vmax = self.vmax64(r[perm])
        f = math.fsum(r * r)
        ry = math.fsum(r * y)
        return math.sqrt(max(0.0, f - 2.0 * (vmax - ry)))

    SWEEPS = {"layers": 0, "auto": 1, "window": 2}
    CLOSURE_MAX_N = 20      # the library refuses the closure sweep above this n
    WINDOW_MAX_N = 26
    WINDOW_MAX_ROWS = 4096  # rows of the one alternative the window-pass sweep takes

    def set_sweep(self, mode):
        """'layers' (the default): the layer-by-layer sweep for any rows (the reference's). When every
        row belongs to one alternative: 'window', the window-pass sweep (V as the subset maximum of
        that alternative's row sums, 128 bytes per subset of the other alternatives, streamed in
        passes of 8 bits; n <= 26), or 'auto', the closure sweep (4 bytes per node; refused above
        n = 20: its first run at n = 25 ended in a GPU fault whose cause has not been shown). Both
        return the layered sweep's values and rankings bit for bit."""
        _lib.check(self.lib.rumgpu_cg_set_sweep(self.h, self.SWEEPS[mode]), "cg_set_sweep")

    def set_slab_tile(self, bits=0):
        """tile bits of the slab layout's layered max (0 = the built-in choice); same results for all"""
        _lib.check(self.lib.rumgpu_cg_set_slab_tile(self.h, int(bits)), "cg_set_slab_tile")

    def one_alternative(self):
        """the alternative the closure sweep runs on, or -1 when the layered sweep runs"""
        a = ctypes.c_int()
        _lib.check(self.lib.rumgpu_cg_sweep_info(self.h, ctypes.byref(a)), "cg_sweep_info")
        return a.value

    def info(self):
        b, ms = ctypes.c_size_t(), ctypes.c_double()
        _lib.check(self.lib.rumgpu_cg_info(self.h, ctypes.byref(b), ctypes.byref(ms)), "cg_info")
        return b.value, ms.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.rumgpu_cg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Columns:
    """Ranking columns as packed bitsets with the Gram matrix Q, h = V'p, and p_out = V lam.

    Storage grows by doubling up to max_K columns. With few rows (M <= DENSE_ROWS) the columns are
    also kept as a dense float64 matrix, so p_out and h of a template are matrix products."""

    DENSE_ROWS = 4096

    def __init__(self, M, p, max_K, start=64):
        self.M, self.W = int(M), (int(M) + 63) // 64
        self.p = np.asarray(p, dtype=np.float64)
        self.max_K = int(max_K)
        self.cap = 0
        self.bits = np.zeros((0, self.W), dtype=np.uint64)
        self.Q = np.zeros((0, 0))
        self.h = np.zeros(0)
        self.C = None
        self.dense = self.M <= self.DENSE_ROWS
        self._grow(min(self.max_K, max(1, int(start))))
        self.pos = []
        self.keys = []
        self.K = 0
        self.p_out = np.zeros(self.M)

    def _grow(self, cap):
        cap = min(self.max_K, cap)
        k = self.cap
        bits = np.zeros((cap, self.W), dtype=np.uint64)
        Q = np.zeros((cap, cap))
        h = np.zeros(cap)
        bits[:k], Q[:k, :k], h[:k] = self.bits[:k], self.Q[:k, :k], self.h[:k]
        self.bits, self.Q, self.h = bits, Q, h
        if self.dense:
            C = np.zeros((cap, self.M))
            if self.C is not None:
                C[:k] = self.C[:k]
            self.C = C
        self.cap = cap

    def add(self, col):
        k = self.K
        if k >= self.max_K:
            raise RuntimeError("max_K = %d columns exceeded" % self.max_K)
        if k >= self.cap:
            self._grow(2 * self.cap)
        padded = np.zeros(self.W * 64, dtype=np.uint8)
        padded[:self.M] = col
        b = np.packbits(padded, bitorder="little").view(np.uint64)
        self.bits[k] = b
        if k:
            cnt = np.bitwise_count(self.bits[:k] & b).sum(axis=1, dtype=np.int64).astype(np.float64)
            self.Q[k, :k] = cnt
            self.Q[:k, k] = cnt
        self.Q[k, k] = float(np.bitwise_count(b).sum(dtype=np.int64))
        nz = np.flatnonzero(col)
        self.h[k] = float(self.p[nz].sum())
        if self.dense:
            self.C[k] = col
        self.pos.append(nz)
        self.keys.append(bytes(np.asarray(col, dtype=np.uint8)))
        self.K += 1

    def load_template(self, tpl):
        """the columns of a template (bits, Gram, positions, keys of the same rows) as the first
        columns; h from this object's p"""
        k = tpl.K
        if k > self.max_K:
            raise RuntimeError("max_K = %d columns exceeded" % self.max_K)
        if k > self.cap:
            self._grow(max(k + 64, 2 * self.cap))
        self.bits[:k] = tpl.bits[:k]
        self.Q[:k, :k] = tpl.Q[:k, :k]
        if self.dense and tpl.dense:
            self.C[:k] = tpl.C[:k]
            self.h[:k] = tpl.C[:k] @ self.p
        else:
            for i in range(k):
                self.h[i] = float(self.p[tpl.pos[i]].sum())
        self.pos = list(tpl.pos)
        self.keys = list(tpl.keys)
        self.K = k

    def apply_delta(self, dlam):
        """p_out += dlam_a * column a for every |dlam_a| > 1e-12, columns ascending: the reference's
        accumulation (qp_bitset_incremental.py:127-141), so p_out and the next scores are its bits"""
        for a in np.flatnonzero(np.abs(dlam) > 1e-12):
            self.p_out[self.pos[int(a)]] += float(dlam[a])

    def view(self):
        return self.Q[:self.K, :self.K], self.h[:self.K]


class PackedColumns(collections.abc.Sequence):
    """0/1 uint8 columns over M rows, kept as packed bitsets (bit r % 64 of word r / 64) and
    unpacked when one is read."""

    def __init__(self, packed, M):
        self._packed, self.M = list(packed), int(M)

    def __len__(self):
        return len(self._packed)

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[j] for j in range(*i.indices(len(self)))]
        return np.unpackbits(self._packed[i].view(np.uint8), bitorder="little")[:self.M]


class DeviceColumns:
    """The ranking columns of one run in the pricer's device column store: the bitsets, p_out = V lam
    and the float32 residual scores of the next sweep live on the GPU (rumgpu_cg_cols_*); the host
    keeps the Gram matrix Q (exact integer counts, as Columns has them), h = V'p (summed on the
    device) and the packed bitsets. set_lambda(lam) recomputes p_out and the scores from lam's
    support."""

    def __init__(self, pricer, M, p, max_K, start=64):
        self.pricer = pricer
        self.M, self.W = int(M), (int(M) + 63) // 64
        self.p = np.asarray(p, dtype=np.float64)
        self.max_K = int(max_K)
        self.cap = 0
        self.Q = np.zeros((0, 0))
        self.h = np.zeros(0)
        self._grow(min(self.max_K, max(1, int(start))))
        self.packed = []
        self.K = 0
        pricer.cols_init(self.p, self.cap)

    def _grow(self, cap):
        cap = min(self.max_K, cap)
        k = self.cap
        Q, h = np.zeros((cap, cap)), np.zeros(cap)
        Q[:k, :k], h[:k] = self.Q[:k, :k], self.h[:k]
        self.Q, self.h, self.cap = Q, h, cap

    def add_ranking(self, rank_pos):
        k = self.K
        if k >= self.max_K:
            raise RuntimeError("max_K = %d columns exceeded" % self.max_K)
        if k >= self.cap:
            self._grow(2 * self.cap)
        gram, hk, bits = self.pricer.cols_add(rank_pos, k)
        self.Q[k, :k] = gram[:k]
        self.Q[:k, k] = gram[:k]
        self.Q[k, k] = gram[k]
        self.h[k] = hk
        self.packed.append(bits)
        self.K += 1

    def load_template(self, tpl):
        k = tpl.K
        if k > self.max_K:
            raise RuntimeError("max_K = %d columns exceeded" % self.max_K)
        if k > self.cap:
            self._grow(max(k + 64, 2 * self.cap))
        self.Q[:k, :k] = tpl.Q[:k, :k]
        for i in range(k):
            self.h[i] = float(self.p[tpl.pos[i]].sum())
        self.pricer.cols_load(tpl.bits[:k])
        self.packed = [tpl.bits[i].copy() for i in range(k)]
        self.K = k

    def set_lambda(self, lam):
        lam = np.asarray(lam, dtype=np.float64)[:self.K]
        idx = np.flatnonzero(lam)
        self.pricer.cols_scores(idx.astype(np.int32), lam[idx])

    def view(self):
        return self.Q[:self.K, :self.K], self.h[:self.K]

    @property
    def p_out(self):
        return self.pricer.cols_pout()

    def columns(self):
        return PackedColumns(self.packed[:self.K], self.M)


RESIDENT = "resident"      # a request to price the scores DeviceColumns.set_lambda left on the device


class NativeRounds:
    """A run's host steps between two sweeps in the library (rumgpu_cgr_*): the ranking columns as
    bitsets with their integer Gram matrix and h, the master QP warm from the last working set and
    factor, p_out over lam's support and the next sweep's float32 scores, one call per round."""

    def __init__(self, n, masks_x, alt_x, p, max_K):
        self.lib = _lib.load()
        self.M, self.n = int(np.asarray(p).size), int(n)
        m = np.ascontiguousarray(masks_x, dtype=np.uint64)
        a = np.ascontiguousarray(alt_x, dtype=np.int32)
        pv = np.ascontiguousarray(p, dtype=np.float64)
        h = ctypes.c_void_p()
        _lib.check(self.lib.rumgpu_cgr_create(int(n), self.M, _lib.ptr(m), _lib.ptr(a), _lib.ptr(pv), int(max_K),
                                              ctypes.byref(h)), "cgr_create")
        self.h = h
        self._out = np.empty(2)
        self._scores = np.empty(self.M, dtype=np.float32)
        self._k = ctypes.c_int()

    def add_ranking(self, rank_pos):
        rp = np.ascontiguousarray(rank_pos, dtype=np.int32)
        _lib.check(self.lib.rumgpu_cgr_add_ranking(self.h, _lib.ptr(rp)), "cgr_add_ranking")

    def add_column(self, col):
        c = np.ascontiguousarray(col, dtype=np.uint8)
        _lib.check(self.lib.rumgpu_cgr_add_column(self.h, _lib.ptr(c)), "cgr_add_column")

    def copy_columns(self, src):
        _lib.check(self.lib.rumgpu_cgr_copy_columns(self.h, src.h), "cgr_copy_columns")

    def set_lambda(self, lam, working_set):
        lv = np.ascontiguousarray(lam, dtype=np.float64)
        A = np.ascontiguousarray(working_set, dtype=np.int32)
        _lib.check(self.lib.rumgpu_cgr_set_lambda(self.h, _lib.ptr(lv), int(lv.size), _lib.ptr(A), int(A.size)),
                   "cgr_set_lambda")

    MASTER = (200, 1e-10, 1e-12)      # active-set passes, KKT tolerance, ridge of the master QP

    def master(self):
        """the master QP, p_out and the next scores: (lam'Q lam, h'lam, scores)"""
        mo, tol, reg = self.MASTER
        _lib.check(self.lib.rumgpu_cgr_master(self.h, mo, tol, reg, _lib.ptr(self._out), _lib.ptr(self._scores)),
                   "cgr_master")
        return float(self._out[0]), float(self._out[1]), self._scores.copy()

    @property
    def K(self):
        _lib.check(self.lib.rumgpu_cgr_get(self.h, ctypes.byref(self._k), None, None, None), "cgr_get")
        return self._k.value

    def result(self):
        """(lam, p_out, columns as 0/1 uint8 arrays)"""
        K = self.K
        lam, p_out = np.empty(K), np.empty(self.M)
        cols = np.empty((K, self.M), dtype=np.uint8)
        _lib.check(self.lib.rumgpu_cgr_get(self.h, None, _lib.ptr(lam), _lib.ptr(p_out), _lib.ptr(cols)), "cgr_get")
        return lam, p_out, list(cols)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.rumgpu_cgr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MasterCall:
    """A request of the native loop: its round's new ranking column, if any (NativeRounds.add_ranking),
    then the master QP, p_out and next scores (NativeRounds.master). The loop yields it instead of
    calling the library itself, so that a driver of many loops can run these steps on host threads
    together (the library call releases the GIL); drive() runs it at once. The reply is master()'s
    (lam'Q lam, h'lam, scores)."""
    __slots__ = ("cols", "rank")

    def __init__(self, cols, rank=None):
        self.cols, self.rank = cols, rank

    def run(self):
        if self.rank is not None:
            self.cols.add_ranking(self.rank)
        return self.cols.master()


def master_batch(calls):
    """MasterCall.run() of runs on the same rows in one library call, on the library's host threads
    (rumgpu_cgr_master_batch; each run's result is the one it gets alone): per call (reply, None),
    or (None, the exception its own call would raise)."""
    if not calls:
        return []
    if len(calls) == 1:
        try:
            return [(calls[0].run(), None)]
        except Exception as e:        # noqa: BLE001 -- handed to the caller's loop
            return [(None, e)]
    lib = calls[0].cols.lib
    M, n = calls[0].cols.M, calls[0].cols.n
    hs = np.array([c.cols.h.value for c in calls], dtype=np.uint64)
    has = np.array([c.rank is not None for c in calls], dtype=np.int32)
    ranks = np.zeros((len(calls), n), dtype=np.int32)
    for i, c in enumerate(calls):
        if c.rank is not None:
            ranks[i] = c.rank
    out = np.empty((len(calls), 2))
    sc = np.empty((len(calls), M), dtype=np.float32)
    st = np.zeros(len(calls), dtype=np.int32)
    mo, tol, reg = NativeRounds.MASTER
    rc = lib.rumgpu_cgr_master_batch(_lib.ptr(hs), len(calls), _lib.ptr(ranks), _lib.ptr(has), mo, tol, reg,
                                     _lib.ptr(out), _lib.ptr(sc), _lib.ptr(st))
    if rc != _lib.OK and not st.any():
        _lib.check(rc, "cgr_master_batch")
    res = []
    for i in range(len(calls)):
        if st[i] != _lib.OK:
            try:
                _lib.check(int(st[i]), "cgr_master")
            except Exception as e:    # noqa: BLE001
                res.append((None, e))
        else:
            res.append(((float(out[i, 0]), float(out[i, 1]), sc[i].copy()), None))
    return res


_NATIVE_TEMPLATES = []


def _native_template(init_columns, n, masks_x, alt_x):
    """A warm-start column list as native rounds' columns (bitsets and integer Gram), built once per
    list; each draw copies it and sums h over its own p."""
    M = int(np.asarray(alt_x).size)
    for lst, tpl in _NATIVE_TEMPLATES:
        if lst is init_columns and tpl.M == M:
            return tpl
    tpl = NativeRounds(n, masks_x, alt_x, np.zeros(M), max(1, len(init_columns)))
    for c in init_columns:
        tpl.add_column(c)
    _NATIVE_TEMPLATES.insert(0, (init_columns, tpl))
    del _NATIVE_TEMPLATES[4:]
    return tpl

_TEMPLATES = []


def _warm_template(init_columns, M):
    """Bitsets and Gram matrix of a warm-start column list, built once per list (a session's kept
    columns start every bootstrap draw): the integer Gram is the same whatever the draw's p."""
    for lst, tpl in _TEMPLATES:
        if lst is init_columns and tpl.M == M:
            return tpl
    tpl = Columns(M, np.zeros(M), len(init_columns))
    for c in init_columns:
        tpl.add(np.asarray(c, dtype=np.uint8))
    _TEMPLATES.insert(0, (init_columns, tpl))
    del _TEMPLATES[4:]
    return tpl


def ranking_column(memb, alt_x, rank_pos, n):
    """0/1 column of a ranking: row = 1 iff its alternative is the menu's top member
    (cg_core.py:63-80 semantics). memb: (M, n) bool membership of the x-major rows."""
    top = np.argmin(np.where(memb, np.asarray(rank_pos)[None, :], n + 1), axis=1)
    return (top == alt_x).astype(np.uint8)


def _cg_steps_native(n, masks_x, alt_x, perm, p_obs, pricer, t_setup, rc_tol, max_iters, time_limit_s, seed, verbose,
                     phi_ub_target, gap_rel_tol, gap_abs_tol, init_columns, init_lambda, early_ub_check, lb_stop):
    """cg_steps with the host columns and the native master (NativeRounds): the same loop, requests,
    stop tests and result; each round's host steps are one library call, which the loop asks its
    driver to make (a MasterCall request before each pricing request)."""
    M = p_obs.size
    p_sq = float(p_obs @ p_obs)
    n_init = len(init_columns) if init_columns is not None else 0
    cols = NativeRounds(n, masks_x, alt_x, p_obs, max_iters + 2 + n_init)
    try:
        if n_init:
            cols.copy_columns(_native_template(init_columns, n, masks_x, alt_x))
            lam = (np.asarray(init_lambda, dtype=np.float64).copy() if init_lambda is not None
                   else np.full(n_init, 1.0 / n_init))
            lam = np.maximum(lam, 0.0)
            lam /= lam.sum()
            cols.set_lambda(lam, [int(i) for i in np.flatnonzero(lam > 1e-12)] or [0])
        else:
            cols.add_ranking(np.random.RandomState(seed).permutation(n))
            cols.set_lambda(np.array([1.0]), [0])
        setup_s = time.perf_counter() - t_setup
        UB, lb_best, gap_rel, rc_final = float("inf"), 0.0, float("inf"), float("nan")
        stop = "max_iters"
        margins, trace = [], []
        qp_s = price_s = 0.0
        t0 = time.perf_counter()
        it = 0
        K = cols.K
        pending = None                  # the last priced ranking, added with the next master
        for it in range(1, max_iters + 1):
            if time_limit_s > 0 and time.perf_counter() - t0 > time_limit_s:
                stop = "time_limit"
                break
            tq = time.perf_counter()
            lQl, hl, req = yield MasterCall(cols, pending)
            pending = None
            err2 = lQl - 2.0 * hl + p_sq
            inner = hl - lQl
            qp_s += time.perf_counter() - tq
            if early_ub_check and phi_ub_target is not None:
                UB = float(np.sqrt(max(0.0, err2)))
                if UB <= phi_ub_target:
                    stop = "phi_ub"
                    break
            tpz = time.perf_counter()
            vmax, rank_pos, margin = yield req
            price_s += time.perf_counter() - tpz
            if margin:
                margins.append(margin)
            fw_gap = (vmax + margin) - inner
            rc_final = fw_gap
            UB = float(np.sqrt(max(0.0, err2)))
            lb_best = max(lb_best, float(np.sqrt(max(0.0, err2 - 2.0 * fw_gap))))
            gap_rel = (UB - lb_best) / lb_best if lb_best > 0 else float("inf")
            trace.append((it, time.perf_counter() - t0, UB, lb_best, K))
            if verbose and (it <= 5 or it % 25 == 0):
                print("[flow_cg %4d] K=%4d dist=%.6f rc=%.3g gap=%.2e" % (it, K, UB, fw_gap, gap_rel), flush=True)
            if phi_ub_target is not None and UB <= phi_ub_target:
                stop = "phi_ub"
                break
            if lb_stop is not None and lb_best >= lb_stop:
                stop = "lb_stop"
                break
            if gap_rel_tol is not None and gap_rel <= gap_rel_tol:
                stop = "gap_rel"
                break
            if gap_abs_tol is not None and UB - lb_best <= gap_abs_tol:
                stop = "gap_abs"
                break
            if fw_gap < rc_tol:
                stop = "rc"
                break
            pending = rank_pos
            K += 1
        if pending is not None:         # a stop before the next master: the column is kept all the same
            cols.add_ranking(pending)
        wall = time.perf_counter() - t0
        lam, p_out, columns = cols.result()
    finally:
        cols.close()
    return dict(wall_s=wall, setup_s=setup_s, trace=trace, margins=margins, iters=it, n_cols=lam.size, lam=lam,
                p_out=p_out, perm=perm, columns=columns, rc_final=rc_final, distance=UB, lb=lb_best, gap_rel=gap_rel,
                stop_reason=stop, qp_s=qp_s, price_s=price_s, device_bytes=pricer.info()[0], layout=pricer.layout,
                dtype=pricer.dtype)


def solve_cg(n, menu, alt, p, pricer=None, layout="mask", dtype="float32", **kw):
    """Column generation to the first stop (solver.py:28-279 semantics). Rows in input order; the
    engine works x-major and returns p_out x-major with `perm` (input row of each x-major row).
    The loop itself is cg_steps; here every pricing request it yields goes to `pricer`."""
    own = pricer is None
    if own:
        perm, masks_x, _alt_x, offsets = xmajor(menu, alt, n)
        pricer = Pricer(n, masks_x, offsets, layout=layout, dtype=dtype)
    try:
        return drive(cg_steps(n, menu, alt, p, pricer, **kw), pricer)
    finally:
        if own:
            pricer.close()


def answer(pricer, req):
    """The reply to one request of a runner: a pricing sweep for float32 scores, the float64
    longest path for ("vmax64", r), the master QP for a MasterCall."""
    if isinstance(req, MasterCall):
        return req.run()
    if isinstance(req, tuple):
        return pricer.vmax64(req[1])
    if isinstance(req, str) and req == RESIDENT:
        return pricer.price(None)
    return pricer.price(req)


def drive(gen, pricer):
    """Run a request generator (cg_steps, or a runner built on it) with one pricer."""
    try:
        req = next(gen)
        while True:
            req = gen.send(answer(pricer, req))
    except StopIteration as e:
        return e.value


def cg_steps(n, menu, alt, p, pricer, rc_tol=0.0, max_iters=4000, time_limit_s=0.0, seed=0, verbose=False,
             phi_ub_target=None, gap_rel_tol=None, gap_abs_tol=None, init_columns=None, init_lambda=None,
             early_ub_check=False, lb_stop=None, qp="native", resident=False):
    """solve_cg's loop as a generator: yields the float32 scores of each pricing sweep, receives
    (vmax, rank_pos, margin), returns the result dict. A driver may price many of these loops in
    one batched sweep (inference/batched.py); each loop's arithmetic is the same either way.
    resident=True (a driver that gives this loop the pricer to itself) keeps the columns of a run
    with more than Columns.DENSE_ROWS rows in the pricer's device column store (DeviceColumns) and
    yields RESIDENT instead of the scores: the column of each ranking, its Gram row, p_out and the
    scores are then computed on the GPU. With the native master on host columns the loop also
    yields a MasterCall before each pricing request (answer() runs it)."""
    t_setup = time.perf_counter()
    perm, masks_x, alt_x, offsets = xmajor(menu, alt, n)
    p32 = np.asarray(p, dtype=np.float64)[perm].astype(np.float32)
    p_obs = p32.astype(np.float64)
    M = p_obs.size
    dev = bool(resident) and M > Columns.DENSE_ROWS
    if not dev and qp == "native":
        return (yield from _cg_steps_native(n, masks_x, alt_x, perm, p_obs, pricer, t_setup, rc_tol, max_iters,
                                            time_limit_s, seed, verbose, phi_ub_target, gap_rel_tol, gap_abs_tol,
                                            init_columns, init_lambda, early_ub_check, lb_stop))
    qp_state = QPState() if qp == "native" else None
    qp_solve = {"native": lambda Q, h, A_init: solve_qp_native(Q, h, A_init=A_init, state=qp_state),
                "numpy": solve_qp_numpy}[qp]
    memb = None if dev else ((masks_x[:, None] >> np.arange(n, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(bool)
    setup_s = time.perf_counter() - t_setup
    p_sq = float(p_obs @ p_obs)
    n_init = len(init_columns) if init_columns is not None else 0
    cols = (DeviceColumns(pricer, M, p_obs, max_iters + 2 + n_init) if dev
            else Columns(M, p_obs, max_iters + 2 + n_init))
    if n_init:
        cols.load_template(_warm_template(init_columns, M))
        lam = (np.asarray(init_lambda, dtype=np.float64).copy() if init_lambda is not None
               else np.full(n_init, 1.0 / n_init))
        lam = np.maximum(lam, 0.0)
        lam /= lam.sum()
        if not dev:
            cols.apply_delta(lam)
        A_warm = [int(i) for i in np.flatnonzero(lam > 1e-12)] or [0]
    else:
        rank0 = np.random.RandomState(seed).permutation(n)
        if dev:
            cols.add_ranking(rank0)
        else:
            cols.add(ranking_column(memb, alt_x, rank0, n))
        lam = np.array([1.0])
        if not dev:
            cols.apply_delta(lam)
        A_warm = [0]
    UB, lb_best, gap_rel, rc_final = float("inf"), 0.0, float("inf"), float("nan")
    stop = "max_iters"
    margins, trace = [], []
    qp_s = price_s = 0.0
    t0 = time.perf_counter()
    it = 0
    for it in range(1, max_iters + 1):
        if time_limit_s > 0 and time.perf_counter() - t0 > time_limit_s:
            stop = "time_limit"
            break
        tq = time.perf_counter()
        Qv, hv = cols.view()
        lam_new, A_warm = qp_solve(Qv, hv, A_init=A_warm)
        if not dev:
            cols.apply_delta(lam_new - lam)
        lam = lam_new
        lQl = float(lam @ (Qv @ lam))
        hl = float(hv @ lam)
        err2 = lQl - 2.0 * hl + p_sq
        inner = hl - lQl
        qp_s += time.perf_counter() - tq
        if early_ub_check and phi_ub_target is not None:
            UB = float(np.sqrt(max(0.0, err2)))
            if UB <= phi_ub_target:
                stop = "phi_ub"
                break
        if dev:
            cols.set_lambda(lam)
            req = RESIDENT
        else:
            req = (p_obs - cols.p_out).astype(np.float32)
        tpz = time.perf_counter()
        vmax, rank_pos, margin = yield req
        price_s += time.perf_counter() - tpz
        if margin:
            margins.append(margin)
        fw_gap = (vmax + margin) - inner
        rc_final = fw_gap
        UB = float(np.sqrt(max(0.0, err2)))
        lb_best = max(lb_best, float(np.sqrt(max(0.0, err2 - 2.0 * fw_gap))))
        gap_rel = (UB - lb_best) / lb_best if lb_best > 0 else float("inf")
        trace.append((it, time.perf_counter() - t0, UB, lb_best, cols.K))
        if verbose and (it <= 5 or it % 25 == 0):
            print("[flow_cg %4d] K=%4d dist=%.6f rc=%.3g gap=%.2e" % (it, cols.K, UB, fw_gap, gap_rel), flush=True)
        if phi_ub_target is not None and UB <= phi_ub_target:
            stop = "phi_ub"
            break
        if lb_stop is not None and lb_best >= lb_stop:
            stop = "lb_stop"
            break
        if gap_rel_tol is not None and gap_rel <= gap_rel_tol:
            stop = "gap_rel"
            break
        if gap_abs_tol is not None and UB - lb_best <= gap_abs_tol:
            stop = "gap_abs"
            break
        if fw_gap < rc_tol:
            stop = "rc"
            break
        if dev:
            cols.add_ranking(rank_pos)
        else:
            cols.add(ranking_column(memb, alt_x, rank_pos, n))
        lam = np.concatenate([lam, [0.0]])
    wall = time.perf_counter() - t0
    dev_bytes = pricer.info()[0]
    if dev:
        cols.set_lambda(lam)
    return dict(wall_s=wall, setup_s=setup_s, trace=trace, margins=margins, iters=it, n_cols=cols.K, lam=lam,
                p_out=cols.p_out.copy(), perm=perm, columns=(cols.columns() if dev else [np.frombuffer(k, dtype=np.uint8).copy() for k in cols.keys]),
                rc_final=rc_final, distance=UB, lb=lb_best, gap_rel=gap_rel, stop_reason=stop, qp_s=qp_s,
                price_s=price_s, device_bytes=dev_bytes, layout=pricer.layout, dtype=pricer.dtype)
