"""Batched bootstrap solves: many draws per GPU launch, retired and refilled one by one.

The reference solves the open draws one after another, each a Session.solve warm from p_hat with the
band's early stops (inference/bootstrap.py:308-335, rum_solver.py:914-925). Here one batch object
(rumgpu_bm_create_batch) holds `slots` draws side by side in every slab: each launch of the fused
iteration advances all of them, `check` iterations are replayed as one HIP graph, and after every
check the float64 certificates of all slots (pushed upper bound, Frank-Wolfe lower bound) come
back in one call. A slot whose draw is decided (ub below tau_lo, lb at or above tau_hi) or whose
bracket meets the tolerance is retired: its bracket goes to the bootstrap and the next open draw
is loaded into that slot (its p uploaded, its state copied from the kept p_hat solve on the device,
optionally corrected by the learned warm start). Every retired bracket is a float64 certificate
of the same kind the per-draw solve gives, so decisions rest on certified bounds alone.

The stop tests per slot are the per-draw solve's (bm_admm/solver.py:294-377 as the session runs
it): running-min ub and running-max lb, ub target = below, lb target = above, absolute and
relative gap on the non-singleton part, a time limit per draw (draw_time) and max_iter.
"""
import math
import time

import numpy as np

from .. import defaults
from ..flow_cg.engine import Pricer as _Pricer

SLOT_CHOICES = (4, 8, 16, 32, 64, 128, 256)
_TUNED = {}


def slot_bytes(n, itemsize):
    """device bytes per draw slot: 6 iterate slabs, the float64 certificate slab, 2 node arrays"""
    m2 = n << (n - 1)
    return 6 * itemsize * m2 + 8 * m2 + 16 * (1 << n)


def autotune_slots(n, menu, alt, p, dtype, beta, alpha, precond_p, check, max_slots, budget_bytes):
    """Draws per launch for this n: time `check` graph-replayed iterations plus one certificate at
    each candidate slot count and keep the smallest count within 5 % of the best draw-iterations
    per second (more slots only add tail waste and memory). Cached per (n, rows, dtype)."""
    from ..bm_admm.engine import Handle
    key = (n, len(menu), str(dtype), int(check))
    if key in _TUNED:
        return _TUNED[key]
    per = slot_bytes(n, np.dtype(dtype).itemsize)
    cands = [d for d in SLOT_CHOICES if d <= max(4, max_slots) and d * per <= budget_bytes]
    if not cands:
        cands = [1]
    rates = {}
    for d in cands:
        h = Handle(n, menu, alt, p, dtype=dtype, beta=beta, alpha=alpha, precond_p=precond_p, slots=d)
        try:
            h.set_graph(check)
            h.run(check)                           # first iteration eager, graph warm
            t0 = time.perf_counter()
            h.run(check)
            h.certify(want_lb=True)
            rates[d] = d * check / (time.perf_counter() - t0)
        finally:
            h.close()
    best = max(rates.values())
    pick = min(d for d, r in rates.items() if r >= 0.95 * best)
    _TUNED[key] = (pick, rates)
    return _TUNED[key]


class BatchedSolver:
    """Solves bootstrap draws in slots; used by Bootstrap(batch=dict(...)).

    slots   draws per launch (None: the autotuner's pick for this n)
    dtype   iterate dtype of the batch ("float64"; "float32" iterates still get float64 bounds)
    model   learned warm start: an object with correction(dp) -> (dx2, dz) on the non-singleton
            rows (dp = P_b - p_hat there), or None; the certificate decides every draw either way
    collect keep (dp, x2_obs - x2_hat_obs, z_obs - z_hat_obs, iterations) of every draw retired
            with a met tolerance (True) or of every retired draw ("all"), as training pairs for the
            warm-start model
    queue   retire and refill on the device (rumgpu_bm_queue_*): the stop tests and refills run in
            the replayed graphs and the host only tops up the pending draws between runs; False
            takes every decision on the host after each check. Both retire every draw at the same
            check with the same bracket. Collecting training pairs and cold slots (no kept p_hat
            state) use the host loop.
    """

    def __init__(self, boot, slots=None, dtype="float64", model=None, collect=False, max_slots=256, verbose=False,
                 queue=True):
        from ..bm_admm.engine import Handle
        sess = boot.session
        if sess.method != "bm_admm":
            raise ValueError("the batched bootstrap runs bm_admm draws (the session's method is %s)" % sess.method)
        rum = boot.rum
        self.boot, self.rum, self.verbose = boot, rum, verbose
        a = sess._args
        self.check, self.lb_every, self.max_iter = int(a["check"]), int(a["lb_every"]), int(a["max_iter"])
        self.beta, self.alpha, self.precond_p = float(a["beta"]), float(a["alpha"]), float(a["precond_p"])
        g = rum._geom
        self.n, self.menu, self.alt = rum.n, g.menu, g.alternative.astype(np.int32)
        self.rest, self.single = rum._rest, rum._single
        self.dtype = np.dtype(dtype)
        self.model, self.collect, self.pairs, self.queue = model, collect, [], bool(queue)
        kept = sess._kept
        self.src = kept["handle"] if kept is not None else None
        if kept is not None and "beta" in kept:
            # the penalty the kept p_hat state was reached with (an adaptive penalty ends elsewhere
            # than it starts); its unscaled multipliers belong to that penalty
            self.beta = float(kept["beta"])
        p0 = boot.p_hat[self.rest]
        self.p_hat_rest = p0
        from .. import machine as _machine
        budget = defaults.AUTO["gpu_memory_fraction"] * _machine.gpu_free_bytes()
        t0 = time.perf_counter()
        self.tune = None
        if slots is None:
            slots, rates = autotune_slots(self.n, self.menu, self.alt, p0, self.dtype, self.beta, self.alpha,
                                          self.precond_p, self.check, max_slots, budget)
            self.tune = rates
        self.D = int(slots)
        if self.D * slot_bytes(self.n, self.dtype.itemsize) > budget:
            raise MemoryError("out of memory: %d slots at n = %d need %.1f GB" %
                              (self.D, self.n, self.D * slot_bytes(self.n, self.dtype.itemsize) / 2 ** 30))
        self.h = Handle(self.n, self.menu, self.alt, p0, dtype=self.dtype, beta=self.beta, alpha=self.alpha,
                        precond_p=self.precond_p, slots=self.D)
        self.h.set_graph(self.check)
        self.x_hat_obs = self.z_hat_obs = None
        self.stats = dict(slots=self.D, autotune=self.tune, setup_s=time.perf_counter() - t0, draws=0, launches=0,
                          checks=0, iterations=0, slot_iterations=0, busy_slot_iterations=0, solve_s=0.0,
                          certify_s=0.0, refill_s=0.0, model_s=0.0)

    # ------------------------------------------------------------------ one draw into a slot
    def _load(self, slot, P_rest):
        t0 = time.perf_counter()
        h = self.h
        h.slot_set_p(slot, P_rest)
        if self.src is not None:
            h.slot_load(slot, self.src)
        else:
            h.slot_cold(slot)
        if self.model is not None and self.src is not None:
            tm = time.perf_counter()
            dx, dz = self.model.correction(P_rest - self.p_hat_rest)
            self.stats["model_s"] += time.perf_counter() - tm
            if dx is not None or dz is not None:
                h.slot_add_obs(slot, dx, dz)
        self.stats["refill_s"] += time.perf_counter() - t0

    def solve(self, idx, P_of, abs_tol, rel_tol, tau_lo, tau_hi, draw_time=None):
        """Yields (b, L, U, exit_flag) per draw of idx (full-distance bracket, Result flags)."""
        idx = [int(b) for b in idx]
        if not idx:
            return
        h = self.h
        kept = self.boot.session._kept               # a refinement round may have re-kept p_hat
        self.src = kept["handle"] if kept is not None else None
        if kept is not None and "beta" in kept and float(kept["beta"]) != h.beta:
            self.beta = float(kept["beta"])
            h.set_beta(self.beta)
        if self.src is not None and (self.model is not None or self.collect):
            h.slot_load(0, self.src)
            self.x_hat_obs, self.z_hat_obs = h.slot_get_obs(0)
        draw_time = float(defaults.BOOTSTRAP["draw_time"] if draw_time is None else draw_time)
        if self.queue and self.src is not None and not self.collect:
            yield from self._solve_on_device(idx, P_of, abs_tol, rel_tol, tau_lo, tau_hi, draw_time)
        else:
            yield from self._solve_on_host(idx, P_of, abs_tol, rel_tol, tau_lo, tau_hi, draw_time)

    def _draw_rows(self, sel, P_of, tau_lo, tau_hi):
        """p on the non-singleton rows, the singleton part's constant and the thresholds on the
        non-singleton part, per draw of sel"""
        P = np.asarray(P_of(sel), dtype=np.float64)
        const = np.empty(len(sel))
        for k in range(len(sel)):
            r_s = 1.0 - P[k, self.single]
            const[k] = float(r_s @ r_s)
        below = np.array([_on_rest(tau_lo, c, False) for c in const])
        above = np.array([_on_rest(tau_hi, c, True) for c in const])
        return np.ascontiguousarray(P[:, self.rest]), const, below, above

    def _solve_on_device(self, idx, P_of, abs_tol, rel_tol, tau_lo, tau_hi, draw_time):
        """The draw queue: draws are pushed in order into a ring, the device retires and refills
        slots after every check, and each run of lb_every cycles ends with the host reading the
        retired draws (in retirement order), topping up the ring and applying time stops."""
        D, h, total = self.D, self.h, len(idx)
        cap = min(total, max(8 * D, 256))
        h.queue_start(self.src, total, cap, self.check, self.lb_every, self.max_iter, abs_tol, rel_tol,
                      corr=2 if self.model is not None else 0)
        const = np.empty(total)
        pushed = 0
        done = np.zeros(total, dtype=bool)
        lo = 0                                     # every draw before lo has been yielded
        t_seen = {}
        taken = 0
        stops = {1: "threshold", 2: "tol", 3: "time", 4: "iters"}
        t_all = time.perf_counter()
        try:
            while True:
                upto = min(total, taken + cap)      # the ring holds cap draws no slot has taken
                if upto > pushed:
                    t0 = time.perf_counter()
                    sel = idx[pushed:upto]
                    Pr, c, below, above = self._draw_rows(sel, P_of, tau_lo, tau_hi)
                    dz = None
                    if self.model is not None:
                        tm = time.perf_counter()
                        _dx, dz = self.model.corrections(Pr - self.p_hat_rest)
                        self.stats["model_s"] += time.perf_counter() - tm
                    h.queue_push(Pr, below, above, None, dz)
                    const[pushed:upto] = c
                    pushed = upto
                    self.stats["refill_s"] += time.perf_counter() - t0
                t0 = time.perf_counter()
                taken, _pushed, retired, _live, _checks = h.queue_run(self.lb_every)
                self.stats["solve_s"] += time.perf_counter() - t0
                self.stats["checks"] = self.stats.get("checks", 0) + self.lb_every
                self.stats["iterations"] += self.check * self.lb_every
                self.stats["slot_iterations"] += self.check * self.lb_every * D
                if taken > lo:
                    lb, ub, it, order, stop = h.queue_results(lo, taken - lo)
                    fin = np.flatnonzero((stop != 0) & ~done[lo:taken])
                    for k in fin[np.argsort(order[fin], kind="stable")]:
                        g = lo + int(k)
                        c = const[g]
                        L = math.sqrt(max(lb[k], 0.0) ** 2 + c)
                        U = math.sqrt(ub[k] ** 2 + c) if math.isfinite(ub[k]) else math.inf
                        done[g] = True
                        self.stats["draws"] += 1
                        self.stats["busy_slot_iterations"] += int(it[k])
                        yield idx[g], L, U, _flag(L, U, abs_tol, rel_tol, stops[int(stop[k])])
                    while lo < taken and done[lo]:
                        lo += 1
                if retired >= total:
                    break
                if math.isfinite(draw_time):
                    draw, _it = h.queue_slots()
                    now = time.perf_counter()
                    kill = np.zeros(D, dtype=np.int32)
                    for s_, g in enumerate(draw):
                        if g >= 0 and now - t_seen.setdefault(int(g), now) > draw_time:
                            kill[s_] = 1
                    if kill.any():
                        h.queue_expire(kill)
        finally:
            h.queue_end()
        self.stats["wall_s"] = self.stats.get("wall_s", 0.0) + time.perf_counter() - t_all
        self.stats["slot_use"] = self.stats["busy_slot_iterations"] / max(1, self.stats["slot_iterations"])

    def _solve_on_host(self, idx, P_of, abs_tol, rel_tol, tau_lo, tau_hi, draw_time):
        D, h = self.D, self.h
        chunk = max(4 * D, 64)
        queue_pos, buf, buf_pos = 0, None, 0
        # per-slot bookkeeping
        b_of = np.full(D, -1, dtype=np.int64)
        below = np.full(D, np.nan)
        above = np.full(D, np.nan)
        const = np.zeros(D)
        ub = np.full(D, np.inf)
        lb = np.zeros(D)
        it = np.zeros(D, dtype=np.int64)
        t_in = np.zeros(D)
        dp_of = [None] * D

        def next_draw():
            nonlocal queue_pos, buf, buf_pos
            if queue_pos >= len(idx):
                return None
            if buf is None or buf_pos >= len(buf):
                sel = idx[queue_pos:queue_pos + chunk]
                buf, buf_pos = np.asarray(P_of(sel), dtype=np.float64), 0
            P = buf[buf_pos]
            b = idx[queue_pos]
            buf_pos += 1
            queue_pos += 1
            return b, P

        def fill(slot):
            nd = next_draw()
            if nd is None:
                b_of[slot] = -1
                return
            b, P = nd
            r_s = 1.0 - P[self.single]
            c = float(r_s @ r_s)
            Pr = np.ascontiguousarray(P[self.rest])
            self._load(slot, Pr)
            b_of[slot], const[slot] = b, c
            below[slot] = _on_rest(tau_lo, c, False)
            above[slot] = _on_rest(tau_hi, c, True)
            ub[slot], lb[slot], it[slot] = np.inf, 0.0, 0
            t_in[slot] = time.perf_counter()
            dp_of[slot] = (Pr - self.p_hat_rest) if self.collect else None

        t_all = time.perf_counter()
        for s in range(D):
            fill(s)
        checks = 0
        while (b_of >= 0).any():
            t0 = time.perf_counter()
            h.run(self.check)
            t1 = time.perf_counter()
            want_lb = (checks % self.lb_every) == 0
            u, l, _phi = h.certify(want_lb=want_lb)
            t2 = time.perf_counter()
            checks += 1
            self.stats["solve_s"] += t1 - t0
            self.stats["certify_s"] += t2 - t1
            self.stats["checks"] += 1
            self.stats["iterations"] += self.check
            self.stats["slot_iterations"] += self.check * D
            live = b_of >= 0
            self.stats["busy_slot_iterations"] += self.check * int(live.sum())
            it[live] += self.check
            ub[live] = np.minimum(ub[live], u[live])
            if want_lb:
                lb[live] = np.maximum(lb[live], l[live])
            now = time.perf_counter()
            for s in np.flatnonzero(live):
                stop = None
                if not math.isnan(below[s]) and below[s] > 0 and ub[s] <= below[s]:
                    stop = "threshold"
                elif not math.isnan(above[s]) and lb[s] >= above[s]:
                    stop = "threshold"
                else:
                    gap = ub[s] - lb[s]
                    if gap <= abs_tol or (lb[s] > 0 and gap / lb[s] <= rel_tol):
                        stop = "tol"
                    elif now - t_in[s] > draw_time:
                        stop = "time"
                    elif it[s] >= self.max_iter:
                        stop = "iters"
                if stop is None:
                    continue
                L = math.sqrt(max(lb[s], 0.0) ** 2 + const[s])
                U = math.sqrt(ub[s] ** 2 + const[s]) if math.isfinite(ub[s]) else math.inf
                flag = _flag(L, U, abs_tol, rel_tol, stop)
                if self.collect and (flag in (1, 2, 3) or self.collect == "all") and self.x_hat_obs is not None:
                    x, z = h.slot_get_obs(s)
                    self.pairs.append((dp_of[s], x - self.x_hat_obs, z - self.z_hat_obs, int(it[s])))
                self.stats["draws"] += 1
                yield int(b_of[s]), L, U, flag
                fill(s)
        self.stats["wall_s"] = self.stats.get("wall_s", 0.0) + time.perf_counter() - t_all
        sl = max(1, self.stats["slot_iterations"])
        self.stats["slot_use"] = self.stats["busy_slot_iterations"] / sl

    def close(self):
        if getattr(self, "h", None) is not None:
            self.h.close()
            self.h = None


def _on_rest(tau, const, above):
    """the threshold on the non-singleton part (rum_solver._on_rest), NaN for none"""
    if tau is None:
        return math.nan
    tau = float(tau)
    if tau * tau <= const:
        return 0.0 if above else math.nan
    return math.sqrt(tau * tau - const)


def _flag(L, U, abs_tol, rel_tol, stop):
    """Result.exit_flag of a bracket (rum_solver.py:42-46 semantics): tolerances first"""
    gap = U - L
    met_abs, met_rel = gap <= abs_tol, L > 0 and gap / L <= rel_tol
    if met_abs and met_rel:
        return 3
    if met_abs:
        return 1
    if met_rel:
        return 2
    return {"threshold": 5, "time": 0, "iters": -1}.get(stop, -2)


class BatchedCG:
    """flow_cg draws priced together: each open draw runs the session's own flow_cg runner (the
    per-draw Session.solve, warm from the kept p_hat columns, with the band's stops), and every
    round the pricing requests of all draws in flight go to the GPU as one batched mask sweep
    (rumgpu_cg_price_batch), whose result for each draw equals its own sweep bit for bit. So every
    draw's columns, weights and bracket are those of its per-draw solve; only the sweeps are shared.

    slots   draws in flight (a multiple of 32 fills the sweeps)
    overlap two halves of the slots alternate, so one half's host steps run while the GPU prices
            the other half
    A half's master QPs (each draw's round of the native loop) run in one library call on its host
    threads; each draw's results are the ones it gets alone.
    """

    def __init__(self, boot, slots=64, overlap=True, verbose=False, **_ignored):
        sess = boot.session
        if sess.method != "flow_cg":
            raise ValueError("BatchedCG runs flow_cg draws (the session's method is %s)" % sess.method)
        if sess._args.get("pricing") != "mask":
            # the batched sweep is a mask sweep: the session's draws then price on the mask layout
            # (the same pricing value; float32 scores only)
            if sess._args.get("dtype", "float32") != "float32":
                raise ValueError("batched sweeps price float32 scores on the mask layout (the plan uses %s %s)"
                                 % (sess._args.get("pricing"), sess._args.get("dtype")))
            sess._args["pricing"] = "mask"
        self.boot, self.sess, self.D, self.verbose = boot, sess, int(slots), verbose
        self.overlap = bool(overlap) and self.D >= 2
        from .. import machine as _machine
        n = boot.rum.n
        one_alt = np.unique(boot.rum._geom.alternative).size == 1
        mode = sess._args.get("_sweep")
        if one_alt and mode is None and n <= _Pricer.WINDOW_MAX_N and boot.rum._geom.M <= _Pricer.WINDOW_MAX_ROWS:
            # rows of one alternative (the bootstrap panels' shape): the window-pass sweep, whose
            # values and rankings are the layered sweep's bit for bit; "_sweep": "layers" keeps the latter
            sess._args["_sweep"] = mode = "window"
        if one_alt and mode == "window" and n <= _Pricer.WINDOW_MAX_N:
            # the window-pass sweep: a 128-byte row per subset of the other alternatives, the scores
            need = 128 * (1 << (n - 1)) + 32 * 4 * boot.rum.M + 64 * boot.rum.M
        elif one_alt and mode == "auto" and n <= _Pricer.CLOSURE_MAX_N:
            # the closure sweep: 4 bytes per subset of the other alternatives, and the scores
            need = 4 * (1 << (n - 1)) + 32 * 4 * boot.rum.M
        else:           # batched V, pred, float64 V (single and 8 draws), scores
            need = (32 * 5 + 8 + 64) * (1 << n) + 32 * 4 * boot.rum.M
        budget = defaults.AUTO["gpu_memory_fraction"] * _machine.gpu_free_bytes()
        if need > budget:
            raise MemoryError("out of memory: batched pricing at n = %d needs %.1f GB, %.1f GB may be used"
                              % (n, need / 2 ** 30, budget / 2 ** 30))
        self.stats = dict(slots=self.D, draws=0, rounds=0, sweeps=0, price_s=0.0, host_s=0.0, wall_s=0.0,
                          draw_rounds=0)

    def solve(self, idx, P_of, abs_tol, rel_tol, tau_lo, tau_hi, draw_time=None):
        """Yields (b, L, U, exit_flag) per draw of idx, as Bootstrap._solve records them."""
        from ..rum_solver import Result, _flow_cg_steps, _cg_pricer, _TimeUp, _Decided, _is_oom
        from ..flow_cg.engine import MasterCall, master_batch
        idx = [int(b) for b in idx]
        if not idx:
            return
        sess, D = self.sess, self.D
        draw_time = float(defaults.BOOTSTRAP["draw_time"] if draw_time is None else draw_time)
        chunk = max(4 * D, 64)
        pos, buf = 0, None
        live = {}                       # slot -> [b, task, gen, request]
        pricer = None
        t_all = time.perf_counter()

        def step(slot, reply, exc=None):
            """advance one draw's runner (exc: raise that in it); returns a finished (b, L, U, flag)
            or None"""
            b, task, gen = live[slot][:3]
            msg, stop = "", None
            try:
                if exc is not None:
                    live[slot][3] = gen.throw(exc)
                else:
                    live[slot][3] = next(gen) if reply is None else gen.send(reply)
                return None
            except StopIteration as e:
                stop = e.value
            except _TimeUp:
                stop = "time"
            except _Decided:
                stop = "threshold"
            except Exception as e:        # noqa: BLE001 -- reported through the exit flag, as Session.solve
                stop = "oom" if _is_oom(e) else "error"
                msg = "%s: %s" % (type(e).__name__, e)
            res = sess._finish(task, stop, msg)
            del live[slot]
            return b, res.lb, res.ub, res.exit_flag

        def start(slot):
            nonlocal pos, buf, pricer
            while pos < len(idx):
                k = pos % chunk
                if k == 0:
                    buf = np.asarray(P_of(idx[pos:pos + chunk]), dtype=np.float64)
                b, P = idx[pos], buf[k]
                pos += 1
                task = sess._task(P, abs_tol, rel_tol, draw_time, tau_lo, tau_hi, True, False, False)
                if isinstance(task, Result):
                    yield b, task.lb, task.ub, task.exit_flag
                    continue
                if pricer is None:
                    pricer = _cg_pricer(task)
                task.t_run = time.perf_counter()
                live[slot] = [b, task, _flow_cg_steps(task, pricer), None]
                done = step(slot, None)
                if done is None:
                    return
                self.stats["draws"] += 1
                yield done

        def masters(slots):
            """the master QPs that the loops of `slots` asked for, in one library call on host threads
            (each loop's own sequence of requests and its results are unchanged), each result sent to
            its loop (an error raised in it); yields the draws that finish, refilling their slots"""
            todo = [q for q in slots if q in live and isinstance(live[q][3], MasterCall)]
            while todo:
                res = master_batch([live[q][3] for q in todo])
                nxt = []
                for q, (r, e) in zip(todo, res):
                    done = step(q, r, e)
                    if done is not None:
                        self.stats["draws"] += 1
                        yield done
                        yield from start(q)
                    if q in live and isinstance(live[q][3], MasterCall):
                        nxt.append(q)
                todo = nxt

        yield from self._rounds(live, step, start, masters, pricer_of=lambda: pricer)
        self.stats["wall_s"] += time.perf_counter() - t_all

    def _rounds(self, live, step, start, masters, pricer_of):
        """the rounds of BatchedCG.solve once its closures exist: every slot started, then batched
        sweeps for the slots' pricing requests and their replies, until no draw is left"""
        D = self.D
        for s in range(D):
            yield from start(s)
        yield from masters(list(live))
        pricer = pricer_of()
        if not self.overlap:
            while live:
                slots = sorted(live)
                t0 = time.perf_counter()
                replies = self._serve(pricer, slots, {q: live[q][3] for q in slots})
                t1 = time.perf_counter()
                self.stats["price_s"] += t1 - t0
                for s in slots:
                    done = step(s, replies[s])
                    if done is not None:
                        self.stats["draws"] += 1
                        yield done
                        yield from start(s)
                yield from masters(slots)
                self.stats["host_s"] += time.perf_counter() - t1
        else:
            # two halves of the slots take turns: while the GPU prices one half (a worker thread;
            # the library call releases the GIL), the host runs the other half's column
            # generation steps. Each draw's own sequence of requests and replies is unchanged.
            from concurrent.futures import ThreadPoolExecutor
            halves = [set(range(0, D, 2)), set(range(1, D, 2))]
            ex = ThreadPoolExecutor(max_workers=1)

            def launch(g):
                slots = sorted(q for q in live if q in halves[g])
                if not slots:
                    return None
                reqs = {q: live[q][3] for q in slots}
                return ex.submit(lambda: (slots, self._serve(pricer, slots, reqs)))

            try:
                futs = [launch(0), launch(1)]
                g = 0
                while futs[0] is not None or futs[1] is not None:
                    if futs[g] is None:
                        g ^= 1
                        continue
                    t0 = time.perf_counter()
                    slots, replies = futs[g].result()
                    futs[g] = None
                    t1 = time.perf_counter()
                    self.stats["price_s"] += t1 - t0          # time the host waited for the GPU
                    for s in slots:
                        done = step(s, replies[s])
                        if done is not None:
                            self.stats["draws"] += 1
                            yield done
                            yield from start(s)
                    yield from masters(slots)
                    futs[g] = launch(g)
                    self.stats["host_s"] += time.perf_counter() - t1
                    g ^= 1
            finally:
                ex.shutdown(wait=True)

    def _serve(self, pricer, slots, reqs):
        """Replies to the requests of `slots` (reqs: slot -> request): one batched sweep for every
        pricing request, one batched float64 sweep for the longest paths of the finishing draws
        (each draw's value equals its own sweep's)."""
        get = reqs.__getitem__
        price = [q for q in slots if not isinstance(get(q), tuple)]
        lbs = [q for q in slots if isinstance(get(q), tuple)]
        out = {}
        if price:
            # up to 8 float64 bounds ride in a sweep of at most 32 draws (the values of separate calls)
            fused = lbs[:8] if (len(price) <= 32 and pricer.layout == "mask") else []
            if fused:
                vmax, rank, v64 = pricer.price_batch_lb(np.stack([get(q) for q in price]),
                                                        np.stack([get(q)[1] for q in fused]))
                for i, q in enumerate(fused):
                    out[q] = float(v64[i])
                lbs = lbs[len(fused):]
                self.stats["lb_fused"] = self.stats.get("lb_fused", 0) + len(fused)
            else:
                vmax, rank = pricer.price_batch(np.stack([get(q) for q in price]))
            self.stats["rounds"] += 1
            self.stats["sweeps"] += -(-len(price) // 32)
            self.stats["draw_rounds"] += len(price)
            for i, q in enumerate(price):
                out[q] = (float(vmax[i]), rank[i].copy(), 0.0)
        if len(lbs) == 1 or (lbs and pricer.layout != "mask"):
            for q in lbs:
                out[q] = pricer.vmax64(get(q)[1])
        elif lbs:
            v = pricer.vmax64_batch(np.stack([get(q)[1] for q in lbs]))
            for i, q in enumerate(lbs):
                out[q] = float(v[i])
        self.stats["lb_sweeps"] = self.stats.get("lb_sweeps", 0) + len(lbs)
        return out

    def close(self):
        pass
