"""Hong and Li's bootstrap test of RUM and its intervals, on certified distance brackets.

    boot = rum.bootstrap(panel=panel)          # n = respondents (solver/defaults.py BOOTSTRAP)
    boot.p_value(); boot.confidence_interval(); boot.distance_interval(); boot.save(path)

Contract (reference: solver/inference/bootstrap.py): the arguments, the three results and their
fields, the record / checkpoint JSON (keys, fingerprints), the messages, and which draws get solved
to which tolerance (that decides every bracket).

How it runs here. Readiness: p_hat is certified once by a Session that keeps its state and fitted
point; then one pass over the draws screens them, since ||P_b - p_fit|| bounds dist(P_b) from
above. Every later step is a list of draw indexes handed to a solver strategy with a tolerance and a
stop rule: _PerDraw (one Session.solve per draw, warm from p_hat, the reference's loop) or _Batched
(inference/batched.py: many draws per GPU launch). Solved brackets land in a _Ledger, which also
reads and writes the draws' part of the record.
"""
import hashlib
import json
import math
import os
import time
from dataclasses import asdict, dataclass

import numpy as np
import scipy.sparse as sp

from .. import defaults
from . import hong_li
from .draws import Draws, MenuDraws

VERSION = 1
NEVER = 127                 # exit flag of a draw that was never solved
CHUNK_BYTES = 1 << 23       # perturbed draws are made this many bytes at a time for the per-draw loop


def _table(rows):
    width = max(len(k) for k, _ in rows)
    return "\n".join("%s : %s" % (k.rjust(width), v) for k, v in rows)


@dataclass
class PValue:
    p: float
    p_lo: float
    p_hi: float
    alpha: float
    reject: bool
    certain: bool
    n_below: int
    n_above: int
    n_undecided: int
    n_screened: int
    n_solved: int
    B: int
    tau: tuple
    dist_hat: tuple
    T: tuple
    stopped_early: bool
    time: float

    def __repr__(self):
        verdict = "reject H0" if self.reject else "do not reject H0"
        return _table([
            ("p", "%.6g  (bracket [%.6g, %.6g])" % (self.p, self.p_lo, self.p_hi)),
            ("decision", "%s at alpha = %g%s" % (verdict, self.alpha, "" if self.certain else " (not certain)")),
            ("draws", "%d: %d below, %d above, %d undecided; %d screened, %d solved"
             % (self.B, self.n_below, self.n_above, self.n_undecided, self.n_screened, self.n_solved)),
            ("dist(p_hat)", "[%.10g, %.10g]" % self.dist_hat),
            ("tau", "[%.10g, %.10g]" % self.tau),
            ("T", "[%.10g, %.10g]" % self.T),
            ("time", "%.2f s%s" % (self.time, ", stopped once the decision was locked" if self.stopped_early else ""))])


@dataclass
class Interval:
    lower: float
    lower_bracket: tuple
    lower_clipped: float
    upper: float
    alpha: float
    q: tuple
    reject: bool
    n_solved: int
    time: float

    def __repr__(self):
        return _table([
            ("interval", "[%.6g, inf)  at %g %%" % (self.lower, 100 * (1 - self.alpha))),
            ("lower end", "in [%.6g, %.6g]; max(0, lower) = %.6g" % (self.lower_bracket + (self.lower_clipped,))),
            ("q", "in [%.6g, %.6g]" % self.q),
            ("H0: dist = 0", "rejected (0 is outside)" if self.reject else "not rejected"),
            ("time", "%.2f s, %d draws solved to tolerance" % (self.time, self.n_solved))])


@dataclass
class DistanceInterval:
    lower: float
    upper: float
    alpha: float
    half_width: float
    dist_hat: tuple
    reject: bool
    time: float

    def __repr__(self):
        return _table([
            ("interval", "[%.6g, %.6g]  at %g %%, two-sided" % (self.lower, self.upper, 100 * (1 - self.alpha))),
            ("half-width", "%.6g, the (1 - alpha)-quantile of ||p_hat_b - p_hat||" % self.half_width),
            ("dist(p_hat)", "[%.10g, %.10g]" % self.dist_hat),
            ("H0: dist = 0", "rejected (0 is outside)" if self.reject else "not rejected"),
            ("time", "%.2f s, no draw solved" % self.time)])


class _Ledger:
    """Per-draw results: the bracket [L, U] (U starts at the screen's bound), the tightest tolerance
    a solve reached with band stops (tol_stop) and without (tol_full), and the last exit flag."""

    FIELDS = (("lb", "L"), ("ub", "U"), ("tol_stop", "tol_stop"), ("tol_full", "tol_full"), ("exit_flag", "flag"))

    def __init__(self, B):
        self.L, self.U = np.zeros(B), np.full(B, np.inf)
        self.tol_stop, self.tol_full = np.full(B, np.inf), np.full(B, np.inf)
        self.flag = np.full(B, NEVER, dtype=np.int8)

    def note(self, b, lb, ub, flag, tol, full):
        self.L[b] = max(self.L[b], lb)
        self.U[b] = min(self.U[b], ub)
        self.tol_stop[b] = min(self.tol_stop[b], tol)
        if full:
            self.tol_full[b] = min(self.tol_full[b], tol)
        self.flag[b] = flag

    def n_solved(self):
        return int((self.flag != NEVER).sum())

    def as_record(self):
        return {key: getattr(self, attr).tolist() for key, attr in self.FIELDS}

    def load(self, rec):
        for key, attr in self.FIELDS:
            setattr(self, attr, np.array(rec[key], dtype=np.int8 if attr == "flag" else np.float64))


class _PerDraw:
    """One Session.solve per draw, warm from the kept p_hat state (the reference's loop)."""

    def __init__(self, boot):
        self.boot = boot

    def run(self, idx, abs_tol, rel_tol, stops, until):
        bt, s = self.boot, self.boot.session
        below, above = bt.tau if stops else (None, None)
        per_chunk = max(1, CHUNK_BYTES // (8 * bt.draws.M))
        t0, every = time.perf_counter(), max(1, len(idx) // 10)
        done = 0
        for lo in range(0, len(idx), per_chunk):
            part = idx[lo:lo + per_chunk]
            P = hong_li.perturb(bt.p_hat, bt.draws.rows(part), bt.r_n, bt.eps_n)
            for b, p in zip(part, P):
                r = s.solve(p, abs_tol=abs_tol, rel_tol=rel_tol, max_time=bt.draw_time, warm=True, below=below,
                            above=above)
                if r.exit_flag == -4:
                    bt._save_checkpoint(force=True)
                    raise KeyboardInterrupt("interrupted at draw %d; the checkpoint keeps every bracket" % b)
                bt.ledger.note(b, r.lb, r.ub, r.exit_flag, abs_tol, not stops)
                done += 1
                if bt.verbose and done % every == 0:
                    bt._say("  %d / %d draws at abs_tol %g (%.0f s)" % (done, len(idx), abs_tol,
                                                                       time.perf_counter() - t0))
                bt._save_checkpoint()
                if until is not None and done % 25 == 0 and until():
                    return


class _Batched:
    """The open draws many at a time (inference/batched.py): BatchedCG for a flow_cg session,
    BatchedSolver (draw slots of one BM-ADMM batch object) otherwise."""

    def __init__(self, boot, options):
        from .batched import BatchedCG, BatchedSolver
        kind = BatchedCG if boot.session.method == "flow_cg" else BatchedSolver
        self.boot, self.engine = boot, kind(boot, **(options or {}))

    def run(self, idx, abs_tol, rel_tol, stops, until):
        bt = self.boot
        below, above = bt.tau if stops else (None, None)
        make = lambda sel: hong_li.perturb(bt.p_hat, bt.draws.rows(sel), bt.r_n, bt.eps_n)  # noqa: E731
        t0, every = time.perf_counter(), max(1, len(idx) // 10)
        for done, (b, lb, ub, flag) in enumerate(self.engine.solve(idx, make, abs_tol, rel_tol, below, above,
                                                                   bt.draw_time), 1):
            bt.ledger.note(b, lb, ub, flag, abs_tol, not stops)
            if bt.verbose and done % every == 0:
                bt._say("  %d / %d draws at abs_tol %g (%.0f s)" % (done, len(idx), abs_tol, time.perf_counter() - t0))
            bt._save_checkpoint()
        bt.batch_stats = self.engine.stats

    def close(self):
        self.engine.close()


def _resolve(value, key):
    return defaults.BOOTSTRAP[key] if value is None else value


class Bootstrap:
    def __init__(self, rum, n=None, panel=None, replies=None, B=None, weights=None, seed=None, r_n=None, eps_n=None,
                 method=None, method_args=None, abs_tol=None, rel_tol=None, refine=None, max_time=None,
                 draw_time=None, vcov=None, machine=None, checkpoint=None, verbose=False, batch=None):
        self.rum, self.p_hat = rum, rum.choice_probability
        self.B = int(_resolve(B, "B"))
        self.abs_tol, self.rel_tol = float(_resolve(abs_tol, "abs_tol")), float(_resolve(rel_tol, "rel_tol"))
        self.refine, self.max_time = int(_resolve(refine, "refine")), float(_resolve(max_time, "max_time"))
        self.draw_time = float(_resolve(draw_time, "draw_time"))
        if min(self.abs_tol, self.rel_tol) < 0 or self.abs_tol + self.rel_tol == 0 or self.refine < 0:
            raise ValueError("abs_tol and rel_tol must be >= 0 and not both 0, and refine >= 0")
        self.method, self.method_args, self.vcov, self.machine = method, method_args, vcov, machine
        self.verbose, self.batch = verbose, batch
        self.draws = self._make_draws(panel, replies, weights, seed)
        n = self.draws.default_n if n is None else n
        self.r_n, self.eps_n, self.s = hong_li.scaling(n, r_n, eps_n)
        self.n = int(n)
        self.draws.check_n(self.n)
        self.settings = dict(n=self.n, input=self.draws.input, B=self.B, weights=self.draws.weights,
                             seed=self.draws.seed, r_n=self.r_n, eps_n=self.eps_n, method=method,
                             method_args=method_args, abs_tol=self.abs_tol, rel_tol=self.rel_tol, refine=self.refine,
                             max_time=self.max_time, draw_time=self.draw_time, weighted=vcov is not None)
        rows = np.column_stack([rum.menu.astype(np.float64), rum.alternative.astype(np.float64), self.p_hat])
        self.fingerprints = {"data": _sha(rows), self.draws.input: self.draws.input_fingerprint, "draws": None}
        self.ledger = _Ledger(self.B)
        self.L0 = self.U0 = None
        self.certificates = []
        self.results = {"p_value": {}, "confidence_interval": {}, "distance_interval": {}}
        self.dev = None
        self.screened, self.n_screened = False, 0
        self.time = dict(certificate=0.0, screen=0.0, solve=0.0)
        self.checkpoint, self._saved, self._resumed_draws = checkpoint, time.perf_counter(), None
        self._session, self._runner, self.batch_stats = None, None, {}
        if checkpoint and os.path.exists(checkpoint):
            self._resume(checkpoint)

    def _make_draws(self, panel, replies, weights, seed):
        if (panel is None) == (replies is None):
            raise ValueError("give the choices behind p_hat once: panel= (respondent, menu, chosen; respondents are "
                             "resampled) or replies= (menu, chosen; each menu's replies are resampled)")
        rum = self.rum
        if panel is not None:
            return Draws(panel, rum.menu, rum.alternative, self.p_hat, self.B, _resolve(weights, "weights"),
                         _resolve(seed, "seed"))
        if weights not in (None, MenuDraws.weights):
            raise ValueError("weights = %r is for panel=; replies= are resampled menu by menu" % (weights,))
        return MenuDraws(replies, rum.menu, rum.alternative, self.p_hat, self.B, _resolve(seed, "seed"))

    # the per-draw arrays live in the ledger
    L = property(lambda self: self.ledger.L)
    U = property(lambda self: self.ledger.U)
    tol_stop = property(lambda self: self.ledger.tol_stop)
    tol_full = property(lambda self: self.ledger.tol_full)
    flag = property(lambda self: self.ledger.flag)

    @property
    def tau(self):
        return (self.s * self.L0, self.s * self.U0)

    @property
    def session(self):
        if self._session is None:
            self._session = self.rum.session(method=self.method, method_args=self.method_args, vcov=self.vcov,
                                             machine=self.machine)
        return self._session

    # ------------------------------------------------------------------------ the three results
    def p_value(self, alpha=0.05, stop_when_decided=False):
        alpha, t0 = hong_li.check_alpha(alpha), time.perf_counter()
        self._start()
        stopped = False
        if self.U0 > hong_li.ZERO:
            locked = (lambda: hong_li.locked(*self._above_open(), self.B, alpha)) if stop_when_decided else None
            for tol_a, tol_r in self._rounds(lambda: self._undecided().any() and not (locked and locked())):
                todo = np.flatnonzero(self._undecided() & (self.tol_stop > tol_a))
                self._run(todo, tol_a, tol_r, stops=True, until=locked)
            stopped = bool(locked and locked() and self._undecided().any())
        dec = self._decisions()
        counts = {k: int((dec == v).sum()) for k, v in (("n_below", hong_li.BELOW), ("n_above", hong_li.ABOVE),
                                                         ("n_undecided", hong_li.UNDECIDED))}
        p_lo, p_hi = hong_li.p_bracket(dec)
        res = PValue(p=p_hi, p_lo=p_lo, p_hi=p_hi, alpha=alpha, reject=p_hi <= alpha,
                     certain=p_hi <= alpha or p_lo > alpha, n_screened=self.n_screened,
                     n_solved=self.ledger.n_solved(), B=self.B, tau=self.tau, dist_hat=(self.L0, self.U0),
                     T=(self.r_n * self.L0, self.r_n * self.U0), stopped_early=stopped, time=time.perf_counter() - t0,
                     **counts)
        return self._keep("p_value", alpha, res)

    def confidence_interval(self, alpha=0.05):
        alpha, t0 = hong_li.check_alpha(alpha), time.perf_counter()
        self._start()
        full_before = int(np.isfinite(self.tol_full).sum())
        per_batch = max(1, math.ceil(alpha * self.B / 4))

        def wide_open():
            lo, hi, _q = hong_li.ci_lower(self.L0, self.U0, self.L, self.U, alpha, self.r_n, self.eps_n)
            return hi - lo > 2 * self.abs_tol

        for tol_a, tol_r in self._rounds(wide_open):
            # the draws whose bracket meets the quantile band, largest U first, a batch at a time
            while True:
                q_l, q_u = hong_li.quantile_band(self.L, self.U, alpha)
                hot = np.flatnonzero((self.U >= q_l) & (self.L <= q_u) & (self.tol_full > tol_a))
                if hot.size == 0:
                    break
                first = hot[np.argsort(-self.U[hot], kind="stable")[:per_batch]]
                self._run(np.sort(first), tol_a, tol_r, stops=False)
        low, high, q = hong_li.ci_lower(self.L0, self.U0, self.L, self.U, alpha, self.r_n, self.eps_n)
        res = Interval(lower=low, lower_bracket=(low, high), lower_clipped=max(0.0, low), upper=math.inf, alpha=alpha,
                       q=q, reject=low > 0, n_solved=int(np.isfinite(self.tol_full).sum()) - full_before,
                       time=time.perf_counter() - t0)
        return self._keep("confidence_interval", alpha, res)

    def distance_interval(self, alpha=0.05):
        alpha, t0 = hong_li.check_alpha(alpha), time.perf_counter()
        self._start()
        if self.dev is None:
            self.dev = np.concatenate([self._norm(block - self.p_hat) for _a, _b, block in self.draws.blocks()])
        lower, upper, c = hong_li.distance_interval(self.L0, self.U0, self.dev, alpha)
        res = DistanceInterval(lower=lower, upper=upper, alpha=alpha, half_width=c, dist_hat=(self.L0, self.U0),
                               reject=lower > 0, time=time.perf_counter() - t0)
        return self._keep("distance_interval", alpha, res)

    def _keep(self, kind, alpha, res):
        self.results.setdefault(kind, {})[repr(alpha)] = asdict(res)
        self._save_checkpoint(force=True)
        return res

    def _rounds(self, still_open):
        """(abs_tol, rel_tol) of the first round and of up to `refine` more, each ten times tighter,
        while still_open(); p_hat is certified again whenever its bracket is wider than the new
        tolerance."""
        tol_a, tol_r = self.abs_tol, self.rel_tol
        yield tol_a, tol_r
        for _ in range(self.refine):
            if not still_open():
                return
            tol_a, tol_r = tol_a / 10, tol_r / 10
            if self.U0 - self.L0 > tol_a:
                self._certify(tol_a, tol_r)
            yield tol_a, tol_r

    # ------------------------------------------------------------------------ the steps
    def _start(self):
        if self.session.fit is None:
            self._certify(self.abs_tol, self.rel_tol)
        if not self.screened:
            self._screen()

    def _certify(self, abs_tol, rel_tol):
        t0 = time.perf_counter()
        r = self.session.solve(self.p_hat, abs_tol=abs_tol, rel_tol=rel_tol, max_time=self.max_time, keep=True,
                               verbose=self.verbose)
        if not math.isfinite(r.ub):
            raise RuntimeError("p_hat could not be certified: %s (%s)" % (r.status, r.message))
        self.L0 = r.lb if self.L0 is None else max(self.L0, r.lb)
        self.U0 = r.ub if self.U0 is None else min(self.U0, r.ub)
        self.certificates.append(dict(abs_tol=abs_tol, rel_tol=rel_tol, lb=r.lb, ub=r.ub, exit_flag=r.exit_flag,
                                      status=r.status, message=r.message, iterations=r.iterations, time=r.time,
                                      method=r.method, method_args=r.method_args, machine=r.machine))
        self.time["certificate"] += time.perf_counter() - t0
        self._say("p_hat: dist in [%.10g, %.10g], %s (%s, %.1f s)" % (r.lb, r.ub, r.status, r.method, r.time))

    def _screen(self):
        """||P_b - p_fit|| >= dist(P_b): a draw whose bound is below tau_lo needs no solve."""
        t0, fit = time.perf_counter(), self.session.fit
        self.dev = np.empty(self.B)
        for a, b, block in self.draws.blocks():
            self.dev[a:b] = self._norm(block - self.p_hat)
            if fit is not None:
                P = hong_li.perturb(self.p_hat, block, self.r_n, self.eps_n)
                self.U[a:b] = np.minimum(self.U[a:b], self._norm(P - fit))
        self.fingerprints["draws"] = self.draws.fingerprint
        if self._resumed_draws not in (None, self.draws.fingerprint):
            raise ValueError("the draws differ from the checkpoint's (fingerprint %s, checkpoint %s)"
                             % (self.draws.fingerprint, self._resumed_draws))
        self.screened, self.n_screened = True, int((self.U < self.tau[0]).sum())
        self.time["screen"] += time.perf_counter() - t0
        self._say("screen: %d of %d draws below tau_lo = %.10g with no solve (%.1f s)"
                  % (self.n_screened, self.B, self.tau[0], time.perf_counter() - t0))

    def _run(self, idx, abs_tol, rel_tol, stops, until=None):
        """Solve the draws idx (in this order) to the tolerance, with the band's stops or without."""
        if len(idx) == 0:
            return
        t0 = time.perf_counter()
        try:
            if self.batch is not None:
                try:
                    if self._runner is None:
                        self._runner = _Batched(self, self.batch)
                    return self._runner.run(idx, abs_tol, rel_tol, stops, until)
                except MemoryError as e:            # the batch does not fit: one draw at a time
                    self._say("batched solves do not fit (%s); solving one draw at a time" % e)
                    self.batch = None
                    self._drop_runner()
            _PerDraw(self).run(idx, abs_tol, rel_tol, stops, until)
        finally:
            self.time["solve"] += time.perf_counter() - t0

    def _drop_runner(self):
        if self._runner is not None:
            self._runner.close()
            self._runner = None

    def _norm(self, D):
        if self.vcov is None:
            return np.sqrt(np.einsum("ij,ij->i", D, D))
        DV = (self.vcov @ D.T).T if sp.issparse(self.vcov) else D @ np.asarray(self.vcov)
        return np.sqrt(np.maximum(np.einsum("ij,ij->i", np.asarray(DV), D), 0.0))

    def _decisions(self):
        if self.U0 <= hong_li.ZERO:
            return np.full(self.B, hong_li.ABOVE, dtype=np.int8)
        return hong_li.decide(self.L, self.U, *self.tau)

    def _undecided(self):
        return self._decisions() == hong_li.UNDECIDED

    def _above_open(self):
        dec = self._decisions()
        return int((dec == hong_li.ABOVE).sum()), int((dec == hong_li.UNDECIDED).sum())

    # ------------------------------------------------------------------------ record and checkpoint
    def record(self):
        s = self._session
        plan = None if s is None else dict(method=s.method, method_args=s.method_args, auto_reason=s.auto_reason)
        rec = dict(version=VERSION, settings=self.settings, fingerprints=self.fingerprints, plan=plan,
                   p_hat=dict(lb=self.L0, ub=self.U0, certificates=self.certificates),
                   tau=None if self.L0 is None else list(self.tau),
                   scaling=dict(r_n=self.r_n, eps_n=self.eps_n, s=self.s),
                   screen=dict(done=self.screened, n_below=self.n_screened))
        rec[self.draws.input] = self.draws.summary()
        rec.update(draws=self.ledger.as_record(), results=self.results, time=self.time,
                   when=time.strftime("%Y-%m-%d %H:%M:%S"))
        return rec

    def save(self, path):
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.record(), f, indent=1, default=_plain)
        os.replace(tmp, path)

    def _save_checkpoint(self, force=False):
        due = force or time.perf_counter() - self._saved >= defaults.BOOTSTRAP["checkpoint_every"]
        if self.checkpoint and due:
            self.save(self.checkpoint)
            self._saved = time.perf_counter()

    def _resume(self, path):
        with open(path) as f:
            rec = json.load(f)
        fp = rec.get("fingerprints", {})
        same = {"version": rec.get("version") == VERSION,
                "settings": json.loads(json.dumps(self.settings, default=_plain)) == rec.get("settings"),
                "data": fp.get("data") == self.fingerprints["data"],
                self.draws.input: fp.get(self.draws.input) == self.fingerprints[self.draws.input]}
        if not all(same.values()):
            raise ValueError("checkpoint %s belongs to another run (differs in: %s); delete it or pass another path"
                             % (path, ", ".join(k for k, v in same.items() if not v)))
        self.ledger.load(rec["draws"])
        self.L0, self.U0 = rec["p_hat"]["lb"], rec["p_hat"]["ub"]
        self.certificates, self.results, self.time = rec["p_hat"]["certificates"], rec["results"], rec["time"]
        self._resumed_draws = fp["draws"]
        self._say("resumed %s: %d draws solved before" % (path, self.ledger.n_solved()))

    # ------------------------------------------------------------------------ the rest
    def close(self):
        self._drop_runner()
        if self._session is not None:
            self._session.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _say(self, msg):
        if self.verbose:
            print("[bootstrap] " + msg, flush=True)

    def __repr__(self):
        rows = [("bootstrap", "n = %d, B = %d %s (seed %s); r_n = %.6g, eps_n = %.6g, s = %.6g"
                 % (self.n, self.B, self.draws.describe, self.draws.seed, self.r_n, self.eps_n, self.s))]
        if self._session is not None:
            s = self._session
            rows.append(("plan", "%s %s%s" % (s.method, s.method_args.get("dtype", ""),
                                              "  [%s]" % s.auto_reason if s.auto_reason else "")))
        if self.L0 is not None:
            rows.append(("dist(p_hat)", "[%.10g, %.10g]; tau = [%.10g, %.10g]" % ((self.L0, self.U0) + self.tau)))
            dec = self._decisions()
            rows.append(("draws", "%d screened below; %d solved; %d below, %d above, %d undecided"
                         % (self.n_screened, self.ledger.n_solved(), int((dec == 0).sum()), int((dec == 1).sum()),
                            int((dec == 2).sum()))))
        for a, r in self.results["p_value"].items():
            rows.append(("p_value", "%.6g at alpha %s (%s)" % (r["p"], a, "reject" if r["reject"] else "do not reject")))
        for a, r in self.results["confidence_interval"].items():
            rows.append(("interval", "[%.6g, inf) at alpha %s" % (r["lower"], a)))
        for a, r in self.results.get("distance_interval", {}).items():
            rows.append(("two-sided", "[%.6g, %.6g] at alpha %s" % (r["lower"], r["upper"], a)))
        rows.append(("time", ", ".join("%s %.1f s" % kv for kv in self.time.items())))
        return _table(rows)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def _plain(x):
    if isinstance(x, np.generic):
        return x.item()
    if isinstance(x, np.ndarray):
        return x.tolist()
    raise TypeError("not JSON serializable: %r" % type(x))
