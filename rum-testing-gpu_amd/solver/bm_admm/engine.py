"""BM-ADMM host engine: the check loop of the reference's solve_kadmm_v2
(solver/bm_admm/solver.py:148-377) and its hybrid float32 -> float64 driver (:380-425), over the
HIP iteration in librumgpu (rumgpu_bm_*).

Handle    one rumgpu_bm object: the slab state (x2, s, y, z) of one dtype on the device, the
          observation structures, the float64 certificate scratch. set_p() swaps the probabilities
          (a bootstrap draw) without reallocating; keep()/load_kept() are Session keep=/warm=.
solve_kadmm(handle, ...)   `check` iterations at a time (one HIP graph replay when enabled), then
          phi, the pushed upper bound and the lower bound through hooks, and the reference's stop
          tests (ub target, lb target, absolute / relative gap, stall, time, iterations).
solve_mixed(h32, h64, ...) the tuned certificate: float32 iterates while the certified gap keeps
          shrinking, then the same state in float64 until the tolerance is met. Every bound is
          computed in float64 from the iterate, so the bracket is certified in both legs.
"""
import ctypes
import math
import time

import numpy as np

from .. import _lib


def eigen(precond_p, beta1, beta2):
    """(e0, e1, gamma, Q row-major) as the reference builds them (solver.py:206-214): numpy eigh
    of [[1, -1], [-1, 1 + p^2]], gamma = beta2 / beta1."""
    p = float(precond_p)
    G = np.array([[1.0, -1.0], [-1.0, 1.0 + p * p]])
    evals, Q = np.linalg.eigh(G)
    return float(evals[0]), float(evals[1]), float(beta2) / float(beta1), np.ascontiguousarray(Q, dtype=np.float64).ravel()


class Handle:
    """A rumgpu_bm object (see include/rumgpu.h)."""

    def __init__(self, n, menu, alt, p, dtype="float64", beta=1.0, alpha=1.95, precond_p=math.sqrt(2.0),
                 plain=False, device=0, slots=1):
        self.ctx = _lib.context(device)
        self.lib = self.ctx.lib
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.dtype(np.float32), np.dtype(np.float64)):
            raise ValueError("BM-ADMM iterates are float32 or float64")
        self.menu = _lib.c_array(menu, np.uint64)
        self.alt = _lib.c_array(alt, np.int32)
        self.M = int(self.menu.size)
        self.beta, self.alpha, self.precond_p = float(beta), float(alpha), float(precond_p)
        e0, e1, gamma, Q = eigen(precond_p, beta, beta)
        h = ctypes.c_void_p()
        pv = _lib.c_array(p, np.float64)
        self.slots = int(slots)
        if self.slots == 1:
            rc = self.lib.rumgpu_bm_create(self.ctx.h, self.n, _lib.ptr(self.menu), _lib.ptr(self.alt), _lib.ptr(pv),
                                           self.M, self.dtype.itemsize, self.beta, self.alpha, self.precond_p,
                                           e0, e1, gamma, _lib.ptr(Q), 1 if plain else 0, ctypes.byref(h))
        else:
            if plain:
                raise ValueError("the plain schedule runs one slot")
            rc = self.lib.rumgpu_bm_create_batch(self.ctx.h, self.n, _lib.ptr(self.menu), _lib.ptr(self.alt),
                                                 _lib.ptr(pv), self.M, self.dtype.itemsize, self.beta, self.alpha,
                                                 self.precond_p, e0, e1, gamma, _lib.ptr(Q), self.slots, ctypes.byref(h))
        _lib.check(rc, "bm_create")
        self.h = h
        self.plain = bool(plain)
        self.graph_len = 0
        self.m2 = self.n << (self.n - 1)

    # ---------------------------------------------------------------- state
    def set_p(self, p):
        pv = _lib.c_array(p, np.float64)
        if pv.size != self.M:
            raise ValueError("p: expected %d values" % self.M)
        _lib.check(self.lib.rumgpu_bm_set_p(self.h, _lib.ptr(pv)), "bm_set_p")

    def init_cold(self):
        _lib.check(self.lib.rumgpu_bm_init_cold(self.h), "bm_init_cold")

    def set_state(self, st):
        arrs = [_lib.c_array(st[k], np.float64) for k in ("x2", "s", "y", "z")]
        if any(a.size != self.m2 for a in arrs):
            raise ValueError("state slabs must hold n 2^(n-1) = %d values" % self.m2)
        _lib.check(self.lib.rumgpu_bm_set_state(self.h, *[_lib.ptr(a) for a in arrs]), "bm_set_state")

    def get_state(self):
        out = {k: np.empty(self.m2) for k in ("x2", "s", "y", "z")}
        _lib.check(self.lib.rumgpu_bm_get_state(self.h, *[_lib.ptr(out[k]) for k in ("x2", "s", "y", "z")]),
                   "bm_get_state")
        return out

    def copy_state_from(self, other):
        _lib.check(self.lib.rumgpu_bm_copy_state(self.h, other.h), "bm_copy_state")

    def keep(self):
        _lib.check(self.lib.rumgpu_bm_keep(self.h), "bm_keep")

    def load_kept(self):
        _lib.check(self.lib.rumgpu_bm_load_kept(self.h), "bm_load_kept")

    # ---------------------------------------------------------------- iterations and bounds
    def set_graph(self, length):
        _lib.check(self.lib.rumgpu_bm_set_graph(self.h, int(length)), "bm_set_graph")
        self.graph_len = int(length)

    def set_plan(self, lo_bits, lo_rl, hi_rl):
        _lib.check(self.lib.rumgpu_bm_set_plan(self.h, int(lo_bits), int(lo_rl), int(hi_rl)), "bm_set_plan")

    def set_beta(self, beta):
        _lib.check(self.lib.rumgpu_bm_set_beta(self.h, float(beta)), "bm_set_beta")
        self.beta = float(beta)

    def run(self, iters):
        _lib.check(self.lib.rumgpu_bm_run(self.h, int(iters)), "bm_run")

    def finalize(self):
        _lib.check(self.lib.rumgpu_bm_finalize(self.h), "bm_finalize")

    def phi(self):
        v = ctypes.c_double()
        _lib.check(self.lib.rumgpu_bm_phi(self.h, ctypes.byref(v)), "bm_phi")
        return v.value

    def push(self, want_fit=False):
        """float64 pushed upper bound; with want_fit also the fitted probabilities (row order)."""
        v = ctypes.c_double()
        fit = np.empty(self.M) if want_fit else None
        _lib.check(self.lib.rumgpu_bm_push_ub64(self.h, ctypes.byref(v), _lib.ptr(fit)), "bm_push_ub64")
        return (v.value, fit) if want_fit else v.value

    def get_fit(self):
        fit = np.empty(self.M)
        _lib.check(self.lib.rumgpu_bm_get_fit(self.h, _lib.ptr(fit)), "bm_get_fit")
        return fit

    def lb(self):
        v = ctypes.c_double()
        _lib.check(self.lib.rumgpu_bm_lb64(self.h, ctypes.byref(v)), "bm_lb64")
        return v.value

    def x_obs(self):
        out = np.empty(self.M)
        _lib.check(self.lib.rumgpu_bm_x_obs(self.h, _lib.ptr(out)), "bm_x_obs")
        return out

    # ---------------------------------------------------------------- draw slots (batch objects)
    def slot_set_p(self, slot, p):
        pv = _lib.c_array(p, np.float64)
        if pv.size != self.M:
            raise ValueError("p: expected %d values" % self.M)
        _lib.check(self.lib.rumgpu_bm_slot_set_p(self.h, int(slot), _lib.ptr(pv)), "bm_slot_set_p")

    def slot_load(self, slot, src):
        """slot <- the state `src` kept (src.keep())"""
        _lib.check(self.lib.rumgpu_bm_slot_load(self.h, int(slot), src.h), "bm_slot_load")

    def slot_cold(self, slot):
        _lib.check(self.lib.rumgpu_bm_slot_cold(self.h, int(slot)), "bm_slot_cold")

    def slot_add_obs(self, slot, dx2=None, dz=None):
        a = None if dx2 is None else _lib.c_array(dx2, np.float64)
        b = None if dz is None else _lib.c_array(dz, np.float64)
        for v in (a, b):
            if v is not None and v.size != self.M:
                raise ValueError("corrections: expected %d values" % self.M)
        _lib.check(self.lib.rumgpu_bm_slot_add_obs(self.h, int(slot), _lib.ptr(a), _lib.ptr(b)), "bm_slot_add_obs")

    def slot_get_obs(self, slot):
        x, z = np.empty(self.M), np.empty(self.M)
        _lib.check(self.lib.rumgpu_bm_slot_get_obs(self.h, int(slot), _lib.ptr(x), _lib.ptr(z)), "bm_slot_get_obs")
        return x, z

    def certify(self, want_lb=True):
        """float64 (ub, lb, phi) of every slot (lb None unless want_lb)"""
        ub, phi = np.empty(self.slots), np.empty(self.slots)
        lb = np.empty(self.slots) if want_lb else None
        _lib.check(self.lib.rumgpu_bm_certify(self.h, 1 if want_lb else 0, _lib.ptr(ub), _lib.ptr(lb), _lib.ptr(phi)),
                   "bm_certify")
        return ub, lb, phi

    def slot_fit(self, slot):
        fit = np.empty(self.M)
        _lib.check(self.lib.rumgpu_bm_slot_fit(self.h, int(slot), _lib.ptr(fit)), "bm_slot_fit")
        return fit

    # ---------------------------------------------------------------- device draw queue
    def queue_start(self, src, total, capacity, check, lb_every, max_iter, abs_tol, rel_tol, corr=0):
        _lib.check(self.lib.rumgpu_bm_queue_start(self.h, src.h, int(total), int(capacity), int(check),
                                                  int(lb_every), int(max_iter), float(abs_tol), float(rel_tol),
                                                  int(corr)), "bm_queue_start")

    def queue_push(self, P, below, above, dx2=None, dz=None):
        P = _lib.c_array(P, np.float64)
        k = P.shape[0] if P.ndim == 2 else 0
        if P.ndim != 2 or P.shape[1] != self.M:
            raise ValueError("queue_push: p must be k x %d" % self.M)
        below, above = _lib.c_array(below, np.float64), _lib.c_array(above, np.float64)
        cors = [None if c is None else _lib.c_array(c, np.float64) for c in (dx2, dz)]
        for v in [below, above]:
            if v.size != k:
                raise ValueError("queue_push: %d thresholds for %d draws" % (v.size, k))
        for c in cors:
            if c is not None and c.shape != P.shape:
                raise ValueError("queue_push: corrections must be k x %d" % self.M)
        _lib.check(self.lib.rumgpu_bm_queue_push(self.h, k, _lib.ptr(P), _lib.ptr(below), _lib.ptr(above),
                                                 _lib.ptr(cors[0]), _lib.ptr(cors[1])), "bm_queue_push")

    def queue_run(self, cycles):
        """-> (taken, pushed, retired, live, checks)"""
        st = np.zeros(5, dtype=np.int64)
        _lib.check(self.lib.rumgpu_bm_queue_run(self.h, int(cycles), _lib.ptr(st)), "bm_queue_run")
        return tuple(int(v) for v in st)

    def queue_results(self, first, count):
        """-> (lb, ub, iterations, retirement rank, stop) of draws first .. first + count - 1"""
        lb, ub = np.empty(count), np.empty(count)
        it, order = np.empty(count, dtype=np.int64), np.empty(count, dtype=np.int64)
        stop = np.empty(count, dtype=np.int32)
        _lib.check(self.lib.rumgpu_bm_queue_results(self.h, int(first), int(count), _lib.ptr(lb), _lib.ptr(ub),
                                                    _lib.ptr(it), _lib.ptr(order), _lib.ptr(stop)),
                   "bm_queue_results")
        return lb, ub, it, order, stop

    def queue_slots(self):
        draw, it = np.empty(self.slots, dtype=np.int64), np.empty(self.slots, dtype=np.int64)
        _lib.check(self.lib.rumgpu_bm_queue_slots(self.h, _lib.ptr(draw), _lib.ptr(it)), "bm_queue_slots")
        return draw, it

    def queue_expire(self, kill):
        k = np.ascontiguousarray(kill, dtype=np.int32)
        if k.size != self.slots:
            raise ValueError("queue_expire: one flag per slot")
        _lib.check(self.lib.rumgpu_bm_queue_expire(self.h, _lib.ptr(k)), "bm_queue_expire")

    def queue_end(self):
        _lib.check(self.lib.rumgpu_bm_queue_end(self.h), "bm_queue_end")

    def prox_slab(self, xk):
        xk = _lib.c_array(xk, np.float64)
        out = np.empty_like(xk)
        _lib.check(self.lib.rumgpu_prox_slab(self.h, _lib.ptr(xk), _lib.ptr(out)), "prox_slab")
        return out

    def windows(self):
        it = ctypes.c_int64()
        nw = ctypes.c_int()
        bits = (ctypes.c_int * 8)()
        _lib.check(self.lib.rumgpu_bm_info(self.h, ctypes.byref(it), ctypes.byref(nw), bits), "bm_info")
        return [bits[i] for i in range(nw.value)]

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.rumgpu_bm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def slab_bytes_per_entry(itemsize, plain=False):
    """Device bytes per slab entry of one Handle: 6 state/work slabs of the iterate dtype, the
    float64 certificate slab, (plain: 6 more work slabs), plus the per-node arrays (about
    (4 + 4 + 8 + 8 + 4) / (n / 2) bytes per entry, counted as 3)."""
    return (12 if plain else 6) * itemsize + 8 + 3


def solve_kadmm(h, max_iter=10 ** 9, check=50, ub_target=None, ub_eps=1e-2, time_limit=None, push_fn=None,
                verbose=False, stall_checks=0, stall_rtol=1e-3, ub_arm_immediately=False, lb_fn=None,
                lb_target=None, gap_abs_tol=None, gap_rel_tol=None, ub_running_min=False, ub_abs=None,
                start=None, gap_stall=None, beta_adapt=None):
    """The reference's check loop (solver.py:294-377) on a Handle whose state is already set.

    beta_adapt: dict(every, ratio, factor, freeze_on_flip) turns on the adaptive penalty
    (params.BM_BETA_ADAPT): every `every` checks, ub - phi against phi - lb decides whether beta
    grows or shrinks (rumgpu_bm_set_beta keeps the state).

    push_fn(h) -> ub and lb_fn(h) -> lb are the runner's hooks (they may raise to stop). start:
    (ub, lb) carried over from an earlier leg (the hybrid's second leg). gap_stall = (checks,
    factor): stop with 'stall' once the certified gap ub - lb has not fallen below factor times
    its best value for that many checks (the mixed driver's precision-floor test)."""
    t0 = time.perf_counter()
    ub_last = float("nan") if start is None else start[0]
    lb_last = 0.0 if start is None else start[1]
    ub_armed = bool(ub_arm_immediately)
    phi = float("nan")
    stop = "max_iter"
    best_prog, n_flat, stall_was_armed = -math.inf, 0, False
    best_gap, gap_flat = math.inf, 0
    n_check, last_dir, frozen, beta_changes = 0, 0, False, []
    it, n_push, push_s = 0, 0, 0.0
    trace = []
    while it < max_iter:
        k = int(min(check, max_iter - it))
        h.run(k)
        it += k
        if time_limit is not None and time.perf_counter() - t0 > time_limit:
            stop = "time_limit"
            break
        phi = h.phi()
        rec = [it, time.perf_counter() - t0, phi, None]
        trace.append(rec)
        if push_fn is not None:
            if not ub_armed:
                ub_armed = ub_target is None or phi >= ub_target * (1.0 - 3.0 * ub_eps)
            if ub_armed:
                tp = time.perf_counter()
                u = float(push_fn(h))
                ub_last = u if (not ub_running_min or math.isnan(ub_last)) else min(ub_last, u)
                push_s += time.perf_counter() - tp
                n_push += 1
                rec[1] = time.perf_counter() - t0
                rec[3] = ub_last
                if ub_target is not None and ((ub_last - ub_target) / ub_target <= ub_eps if ub_target != 0
                                              else ub_abs is not None and ub_last - ub_target <= ub_abs):
                    stop = "ub_target"
                    break
        if lb_fn is not None and (lb_target is not None or gap_abs_tol is not None or gap_rel_tol is not None):
            lb_last = max(lb_last, float(lb_fn(h)))
            if lb_target is not None and lb_last >= lb_target:
                stop = "lb_target"
                break
            if not math.isnan(ub_last):
                gap = ub_last - lb_last
                if gap_abs_tol is not None and gap <= gap_abs_tol:
                    stop = "gap_abs"
                    break
                if gap_rel_tol is not None and lb_last > 0 and gap / lb_last <= gap_rel_tol:
                    stop = "gap_rel"
                    break
                if gap_stall is not None:
                    if gap < gap_stall[1] * best_gap:
                        best_gap, gap_flat = gap, 0
                    else:
                        gap_flat += 1
                        if gap_flat >= gap_stall[0]:
                            stop = "stall"
                            break
        n_check += 1
        if (beta_adapt is not None and not frozen and lb_fn is not None and not math.isnan(ub_last)
                and n_check % int(beta_adapt["every"]) == 0
                and len(beta_changes) < int(beta_adapt.get("max_changes", 1 << 30))):
            inf_part = max(ub_last - phi, 1e-300)
            opt_part = max(phi - lb_last, 1e-300)
            d = 1 if inf_part > beta_adapt["ratio"] * opt_part else -1 if opt_part > beta_adapt["ratio"] * inf_part else 0
            if d:
                f = float(beta_adapt["factor"])
                if last_dir and d != last_dir and beta_adapt.get("freeze_on_flip", True):
                    f, frozen = math.sqrt(f), True
                h.set_beta(h.beta * f ** d)
                last_dir = d
                beta_changes.append((it, h.beta))
        if stall_checks > 0:
            if ub_armed != stall_was_armed:
                best_prog, n_flat = -math.inf, 0
                stall_was_armed = ub_armed
            prog = phi if not ub_armed else -ub_last
            if prog > best_prog * (1.0 + np.sign(best_prog) * stall_rtol) or best_prog == -math.inf:
                best_prog, n_flat = prog, 0
            else:
                n_flat += 1
                if n_flat >= stall_checks:
                    stop = "stall"
                    break
        if verbose and it % (check * 20) == 0:
            print("   it=%d phi=%.6f ub=%.6f lb=%.6f" % (it, phi, ub_last, lb_last), flush=True)
    wall = time.perf_counter() - t0
    return dict(distance=phi, iters=it, wall_s=wall, per_iter_ms=wall / max(1, it) * 1e3, stop_reason=stop,
                ub=ub_last, lb=lb_last, push_s=push_s, n_push=n_push, trace=trace, dtype=str(h.dtype),
                use_graph=h.graph_len > 0, beta_final=h.beta, beta_changes=beta_changes)


def solve_mixed(h32, h64, gap_stall=(3, 0.9), **kw):
    """Tuned certificate: float32 leg until the certified gap stops shrinking by 10 % per three
    checks (the float32 iterate's precision floor), then the same state in float64 (device
    conversion) until the tolerance. Iterations and traces are summed over both legs. Without a
    lower bound (a race) there is no certified gap: the float32 leg ends when phi, then the pushed
    ub, stops improving by 1e-3 over three checks, the reference hybrid's plateau rule
    (bm_admm/solver.py:380-425, stall_checks=3, stall_rtol=1e-3)."""
    t0 = time.perf_counter()
    if kw.get("lb_fn") is None or all(kw.get(k) is None for k in ("lb_target", "gap_abs_tol", "gap_rel_tol")):
        r32 = solve_kadmm(h32, **dict(dict(stall_checks=3, stall_rtol=1e-3), **kw))
    else:
        r32 = solve_kadmm(h32, gap_stall=gap_stall, **kw)
    if r32["stop_reason"] != "stall":
        r32.update(iters32=r32["iters"], iters64=0, wall32_s=r32["wall_s"], wall64_s=0.0)
        return r32
    h64.copy_state_from(h32)
    if h64.beta != h32.beta:                  # an adaptive penalty carries over
        h64.set_beta(h32.beta)
    kw2 = dict(kw)
    if kw2.get("max_iter") is not None:
        kw2["max_iter"] = max(1, kw2["max_iter"] - r32["iters"])
    if kw2.get("time_limit") is not None:
        kw2["time_limit"] = max(0.0, kw2["time_limit"] - (time.perf_counter() - t0))
    r64 = solve_kadmm(h64, start=(r32["ub"], r32["lb"]), **kw2)
    off_it, off_t = r32["iters"], r32["wall_s"]
    r64["trace"] = r32["trace"] + [[a + off_it, b + off_t, c, d] for a, b, c, d in r64["trace"]]
    r64.update(iters32=r32["iters"], iters64=r64["iters"], wall32_s=r32["wall_s"], wall64_s=r64["wall_s"])
    r64["iters"] = r32["iters"] + r64["iters"]
    r64["wall_s"] = time.perf_counter() - t0
    r64["push_s"] += r32["push_s"]
    r64["n_push"] += r32["n_push"]
    r64["dtype"] = "mixed"
    return r64
