"""Restricted-master QP of Flow-CG: min 1/2 lam'Q lam - h'lam subject to lam >= 0, sum lam = 1.

Same method as the reference (solver/common/qp_native.py:37-100): an active set whose equality
problem is solved by Cholesky,
    lam_A = Q_AA^-1 (h_A - mu 1),   mu = (1'Q_AA^-1 h_A - 1) / (1'Q_AA^-1 1),
dropping the most negative weight, else adding the most violated inactive column (gradient below
-tol), else stopping; regularization reg on the diagonal; a final Euclidean simplex projection;
support = weights above 1e-9. The working set warm-starts from the previous call's support.

solve_qp_native runs it in librumgpu (rumgpu_qp_simplex: the working set's Cholesky factor grows
by one row per entering column instead of being refactored); solve_qp_numpy is the same method in
numpy/scipy, step for step as the reference writes it.
"""
import ctypes

import numpy as np
from scipy.linalg import cho_factor, cho_solve

from .. import _lib


def project_simplex(v):
    """Euclidean projection onto the probability simplex (sort-based threshold)."""
    k = v.size
    u = -np.sort(-v)
    cs = np.cumsum(u) - 1.0
    ok = u - cs / np.arange(1, k + 1) > 0
    if not ok.any():
        return np.full(k, 1.0 / k)
    r = int(np.flatnonzero(ok)[-1])
    return np.maximum(v - cs[r] / (r + 1), 0.0)


class QPState:
    """The last working set of a sequence of master QPs and its Cholesky factor (rumgpu_qp), for a
    column generation whose master only grows."""

    def __init__(self):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self.lib.rumgpu_qp_create(ctypes.byref(h)), "qp_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.rumgpu_qp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_qp_native(Q, h, A_init=None, max_outer=200, tol=1e-10, reg=1e-12, state=None):
    """Active-set solve of the simplex QP in librumgpu; returns (lam, support as an int32 array, in
    working-set order). With a QPState the factor of a working set equal to the last call's is
    reused."""
    K = int(h.size)
    Q = np.asarray(Q, dtype=np.float64)
    if Q.strides[1] != 8 or Q.shape != (K, K):
        Q = np.ascontiguousarray(Q)
    ldq = Q.strides[0] // 8
    h = np.ascontiguousarray(h, dtype=np.float64)
    A = np.ascontiguousarray(A_init if A_init is not None else [], dtype=np.int32)
    lam = np.empty(K)
    sup = np.empty(K, dtype=np.int32)
    ns = ctypes.c_int()
    lib = _lib.load()
    if state is None:
        rc = lib.rumgpu_qp_simplex(K, ctypes.c_void_p(Q.ctypes.data), ldq, _lib.ptr(h), _lib.ptr(A), int(A.size),
                                   int(max_outer), float(tol), float(reg), _lib.ptr(lam), _lib.ptr(sup), ctypes.byref(ns))
    else:
        rc = lib.rumgpu_qp_simplex_warm(state.h, K, ctypes.c_void_p(Q.ctypes.data), ldq, _lib.ptr(h), _lib.ptr(A),
                                        int(A.size), int(max_outer), float(tol), float(reg), _lib.ptr(lam),
                                        _lib.ptr(sup), ctypes.byref(ns))
    _lib.check(rc, "qp_simplex")
    return lam, sup[:ns.value].copy()


def solve_qp_numpy(Q, h, A_init=None, max_outer=200, tol=1e-10, reg=1e-12):
    """Active-set solve of the simplex QP in numpy/scipy; returns (lam, support list)."""
    K = int(h.size)
    work = [int(i) for i in (A_init if A_init is not None else []) if 0 <= int(i) < K]
    if not work:
        work = [int(np.argmax(h))]
    lam_w, idx = None, None
    for _ in range(max_outer):
        idx = np.asarray(work, dtype=np.int64)
        k = idx.size
        Qw = Q[np.ix_(idx, idx)] + reg * np.eye(k)
        ones = np.ones(k)
        try:
            sol = cho_solve(cho_factor(Qw, lower=True, check_finite=False), np.column_stack([h[idx], ones]),
                            check_finite=False)
            qh, qe = sol[:, 0], sol[:, 1]
        except np.linalg.LinAlgError:
            # one least-squares solve per right-hand side, as the reference does (its minimum-norm
            # solutions, bit for bit)
            qh = np.linalg.lstsq(Qw, h[idx], rcond=None)[0]
            qe = np.linalg.lstsq(Qw, ones, rcond=None)[0]
        # 1'x as a dot product with a ones vector (the reference's summation order)
        se = float(ones @ qe)
        mu = (float(ones @ qh) - 1.0) / se if abs(se) > 1e-30 else 0.0
        lam_w = qh - mu * qe
        if k and float(lam_w.min()) < -tol:
            work.pop(int(np.argmin(lam_w)))
            continue
        if K > k:
            free = np.ones(K, dtype=bool)
            free[idx] = False
            cand = np.flatnonzero(free)
            grad = Q[np.ix_(cand, idx)] @ lam_w + mu - h[cand]
            j = int(np.argmin(grad))
            if grad[j] < -tol:
                work.append(int(cand[j]))
                continue
        break
    lam = np.zeros(K)
    lam[idx] = lam_w
    lam = project_simplex(lam)
    return lam, [int(i) for i in work if lam[i] > 1e-9]
