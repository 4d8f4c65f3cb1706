"""Golden vectors of the reference's restricted-master QP and p_out accumulation.

Runs the reference's own functions (numpy/scipy only), loaded by file path so that none of its
package __init__ files (which need CuPy and numba) run:
  solver/common/qp_native.py              solve_qp_native(Q, h, A_init) -> (lam, support)
  solver/flow_cg/qp_bitset_incremental.py BitsetGramMaintainer (Q, h of growing columns) and
                                          PoutMaintainer.apply_delta (p_out += dlam_a * column a)
on masters built like column generation's (0/1 ranking columns over rows, Q = integer Gram,
h = C p), random and degenerate: duplicated columns, more columns than rows, ties in h, warm
working sets (valid, partly out of range, empty). Writes qp_native_vectors.npz next to this file
(flat arrays with offsets; np.load with allow_pickle=False reads it).

Run where the reference's sources are at /root/reference (the tests read only the .npz):
    python3 tests/golden/make_qp_native_vectors.py
"""
import importlib.util
import os

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _load(name, rel):
    spec = importlib.util.spec_from_file_location(name, os.path.join(REF, rel))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _columns(rng, M, K, dens):
    C = (rng.random((K, M)) < dens).astype(np.uint8)
    C[C.sum(axis=1) == 0, 0] = 1
    return C


def _cases(rng):
    """(C, p, A_init) triples: C the K x M 0/1 columns, A_init None or a list"""
    out = []
    for M, K, dens in ((20, 5, 0.4), (20, 15, 0.3), (60, 40, 0.35), (12, 30, 0.5), (40, 80, 0.25)):
        for rep in range(3):
            C = _columns(rng, M, K, dens)
            p = rng.uniform(0.05, 0.95, M)
            out.append((C, p, None))
    # duplicated columns: the same column several times (a singular Q_AA without the ridge)
    for rep in range(3):
        C = _columns(rng, 25, 10, 0.4)
        C = np.concatenate([C, C[[1, 3, 3, 7]]])
        out.append((C, rng.uniform(0.05, 0.95, 25), None))
    # ties in h: p constant on rows, columns of equal size
    for rep in range(2):
        M, K = 16, 12
        C = np.zeros((K, M), dtype=np.uint8)
        for k in range(K):
            C[k, rng.choice(M, 6, replace=False)] = 1
        out.append((C, np.full(M, 0.5), None))
    # p itself a mixture of two columns (distance 0, a face of optimal lam)
    for rep in range(2):
        C = _columns(rng, 30, 20, 0.4)
        p = 0.3 * C[2] + 0.7 * C[11]
        out.append((C.copy(), p.astype(np.float64), None))
    # warm working sets: a valid set, one with out-of-range entries, an empty one
    base = list(out)
    for C, p, _ in base[:6]:
        K = C.shape[0]
        out.append((C, p, [int(i) for i in rng.choice(K, min(3, K), replace=False)]))
        out.append((C, p, [K + 5, 0, -1]))
        out.append((C, p, []))
    return out


def main():
    qpn = _load("ref_qp_native", "solver/common/qp_native.py")
    qbi = _load("ref_qp_bitset_incremental", "solver/flow_cg/qp_bitset_incremental.py")
    rng = np.random.default_rng(20260504)
    cases = _cases(rng)
    Qs, hs, ainit, lams, sups = [], [], [], [], []
    off_q, off_h, off_a, off_s = [0], [0], [0], [0]
    a_none = []
    for C, p, A0 in cases:
        Cf = C.astype(np.float64)
        Q = Cf @ Cf.T
        h = Cf @ p
        lam, sup = qpn.solve_qp_native(Q, h, A_init=A0)
        Qs.append(Q.ravel())
        hs.append(h)
        lams.append(np.asarray(lam, dtype=np.float64))
        sups.append(np.asarray(sup, dtype=np.int64))
        a_none.append(A0 is None)
        ainit.append(np.asarray(A0 if A0 is not None else [], dtype=np.int64))
        off_q.append(off_q[-1] + Q.size)
        off_h.append(off_h[-1] + h.size)
        off_a.append(off_a[-1] + len(ainit[-1]))
        off_s.append(off_s[-1] + len(sup))
    # p_out along a growing master: columns added one at a time to the reference's Gram and h
    # (BitsetGramMaintainer), each master warm from the last support, p_out moved by the change of
    # lam (PoutMaintainer): the reference's column generation bookkeeping
    pc_cols, pc_p, pc_steps, pc_lam = [], [], [], []
    off_pc = [0]
    off_steps = [0]
    for M, K, dens in ((30, 12, 0.35), (50, 20, 0.3)):
        C = _columns(rng, M, K, dens)
        p = rng.uniform(0.05, 0.95, M)
        gm = qbi.BitsetGramMaintainer(M, p, max_K=K)
        pm = qbi.PoutMaintainer(M, gm)
        lam_prev = np.zeros(0)
        sup = None
        for k in range(1, K + 1):
            gm.add_column(C[k - 1])
            pm.add_column(C[k - 1])
            Qv, hv = gm.view()
            lam, sup = qpn.solve_qp_native(Qv, hv, A_init=sup)
            d = lam.copy()
            d[:lam_prev.size] -= lam_prev
            pm.apply_delta(d)
            lam_prev = lam
            pc_steps.append(pm.p_out.copy())
            pc_lam.append(np.asarray(lam, dtype=np.float64))
        pc_cols.append(C.ravel())
        pc_p.append(p)
        off_pc.append(off_pc[-1] + M)
        off_steps.append(off_steps[-1] + K)
    np.savez_compressed(
        os.path.join(HERE, "qp_native_vectors.npz"),
        K=np.array([int(h.size) for h in hs], dtype=np.int64), Q=np.concatenate(Qs), off_q=np.array(off_q),
        h=np.concatenate(hs), off_h=np.array(off_h), a_init=np.concatenate(ainit), off_a=np.array(off_a),
        a_none=np.array(a_none), lam=np.concatenate(lams), support=np.concatenate(sups), off_s=np.array(off_s),
        pout_M=np.array([30, 50]), pout_K=np.array([12, 20]), pout_cols=np.concatenate(pc_cols),
        pout_p=np.concatenate(pc_p), pout_off=np.array(off_pc), pout_steps=np.concatenate(pc_steps),
        pout_off_steps=np.array(off_steps), pout_lam=np.concatenate(pc_lam))
    print("wrote %d masters and %d p_out steps" % (len(hs), len(pc_steps)))


if __name__ == "__main__":
    main()
