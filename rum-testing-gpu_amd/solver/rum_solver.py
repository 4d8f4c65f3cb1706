"""The RUM class: the distance from observed choice probabilities to the random utility model,
computed by the MI355X engines of this package (BM-ADMM and Flow-CG, HIP kernels behind
librumgpu). The surface is the reference's (solver/rum_solver.py:179-658):

    from solver.rum_solver import RUM
    rum = RUM(menu, alternative, choice_probability)          # one row per observed cell
    rum = RUM.from_rows(cd, n)                                # the same from (M, 3) rows
    res = rum.calculate_distance()                            # defaults: solver/defaults.py
    res = rum.calculate_distance(abs_tol=1e-6, rel_tol=0, method="bm_admm",
                                 method_args={"dtype": "mixed"}, max_time=600)
    with rum.session() as s:                                  # one plan, many points (bootstrap)
        r0 = s.solve(rum.choice_probability, keep=True)
        rb = s.solve(p_b, below=tau_lo, above=tau_hi, warm=True)

Rows of singleton menus are taken out (their probability must be 1) and added back as a constant.
Every result brackets the distance: lb <= dist <= ub, distance = ub = the distance of p_fit.

    exit_flag   1 abs_tol met    2 rel_tol met    3 both met    4 opt target reached (race)
                5 threshold reached (Session.solve: ub < below or lb >= above)
                0 time limit    -1 iteration limit    -2 stalled    -3 out of memory
               -4 interrupted   -5 solver error (message says which)

Methods: bm_admm and flow_cg run here (GPU). The reference's Gurobi and Flow-ADMM methods
(flow_admm, ip_cg, ip_cg_heur, bm_qp, flow_qp) are outside this package and raise when named.
"""
import math
import time
import types
import warnings
from dataclasses import dataclass, field, fields

import numpy as np
import scipy.sparse as sp

from . import ALIASES, defaults, params
from . import machine as _machine

METHODS = ("bm_admm", "flow_cg", "flow_admm", "ip_cg", "ip_cg_heur", "bm_qp", "flow_qp")
BUILT = ("bm_admm", "flow_cg")
GPU_METHODS = ("bm_admm", "flow_cg", "flow_admm")
RACE_METHODS = ("bm_admm", "flow_cg")
STATUS = {5: "threshold reached", 4: "opt target reached", 3: "abs_tol and rel_tol met", 1: "abs_tol met",
          2: "rel_tol met", 0: "time limit", -1: "iteration limit", -2: "stalled", -3: "out of memory",
          -4: "interrupted", -5: "solver error"}
_FLAG = {"threshold": 5, "target": 4, "time": 0, "iters": -1, "stalled": -2, "oom": -3, "interrupted": -4,
         "error": -5}
_UNIT = {"bm_admm": "ADMM iterations", "flow_cg": "columns"}


def _table(rows):
    """Right-aligned 'label : text' lines."""
    width = max(len(k) for k, _ in rows)
    return "\n".join("%s : %s" % (k.rjust(width), v) for k, v in rows)


@dataclass
class Result:
    exit_flag: int
    status: str
    message: str
    distance: float
    lb: float
    ub: float
    abs_gap: float
    rel_gap: float
    abs_tol: float
    rel_tol: float
    met_abs: bool
    met_rel: bool
    method: str
    auto_reason: str
    method_args: dict
    iterations: int
    iteration_unit: str
    time: float
    time_setup: float
    time_solve: float
    seed: object
    p_fit: np.ndarray = None
    trace: list = field(default_factory=list)      # (iteration, seconds since the call, ub, lb)
    machine: str = ""
    info: dict = field(default_factory=dict)
    mode: str = "certificate"
    opt: float = None
    target: float = None
    abs_to_opt: float = None
    rel_to_opt: float = None

    def _rows(self):
        """(label, text) pairs of the printed result, in order."""
        yield "exit_flag", "%d (%s)" % (self.exit_flag, self.status)
        yield "distance", "%.10g" % self.distance
        if self.mode == "race":
            yield "mode", "race to opt (no lower bound)"
            yield "opt, target", "%.10g, %.10g" % (self.opt, self.target)
            yield "ub - opt", "-" if self.abs_to_opt is None else "%.3g" % self.abs_to_opt
        else:
            yield "lb, ub", "%.10g, %.10g" % (self.lb, self.ub)
            for name, gap, tol, met in (("abs", self.abs_gap, self.abs_tol, self.met_abs),
                                        ("rel", self.rel_gap, self.rel_tol, self.met_rel)):
                yield name + "_gap", "%.3g (%s_tol %.3g, %s)" % (gap, name, tol, "met" if met else "not met")
        yield "method", self.method + ("  [%s]" % self.auto_reason if self.auto_reason else "")
        yield "iterations", "%d %s" % (self.iterations, self.iteration_unit)
        yield "time", "%.3f s (setup %.3f s, solve %.3f s)" % (self.time, self.time_setup, self.time_solve)
        yield "seed", str(self.seed)
        if self.message:
            yield "message", self.message
        yield "method_args", str(dict(self.method_args))
        yield "p_fit", "None" if self.p_fit is None else "array of %d" % len(self.p_fit)
        yield "trace", "%d entries" % len(self.trace)
        yield "machine", self.machine
        if self.info:
            yield "info", str(self.info)

    def __repr__(self):
        return _table(list(self._rows()))

    def to_dict(self, p_fit=False):
        d = {f.name: getattr(self, f.name) for f in fields(self) if f.name != "p_fit"}
        d["trace"] = [list(row) for row in self.trace]
        if p_fit:
            d["p_fit"] = None if self.p_fit is None else self.p_fit.tolist()
        return d


class _TimeUp(Exception):
    pass


class _Decided(Exception):
    pass


class _Geometry:
    """The non-singleton rows (menu, alternative) and n."""

    def __init__(self, menu, alternative, n):
        self.menu = np.asarray(menu, dtype=np.uint64)
        self.alternative = np.asarray(alternative, dtype=np.int64)
        self.num_alternatives = int(n)
        self.M = int(self.menu.size)

    def choice_data(self, p):
        return np.column_stack([self.menu.astype(np.float64), self.alternative.astype(np.float64),
                                np.asarray(p, dtype=np.float64)])


def _tolerances(abs_tol, rel_tol, max_time):
    """The stop rule of one call, defaults filled in and checked."""
    abs_tol = float(defaults.ABS_TOL if abs_tol is None else abs_tol)
    rel_tol = float(defaults.REL_TOL if rel_tol is None else rel_tol)
    max_time = float(defaults.MAX_TIME if max_time is None else max_time)
    if abs_tol < 0 or rel_tol < 0 or not max_time > 0:
        raise ValueError("abs_tol and rel_tol must be >= 0 and max_time > 0")
    return abs_tol, rel_tol, max_time


def _task(rum, p, W, args, abs_tol, rel_tol, deadline, seed, verbose, t0, const, ram, **extra):
    """What a runner reads and writes: the non-singleton rows and their p, the plan, the stop rule,
    and the running state (bounds, fitted point, iterations, trace, info)."""
    state = dict(lb=0.0, ub=math.inf, p_fit=None, iterations=0, trace=[], info={}, setup=0.0)
    base = dict(geom=rum._geom, p=p[rum._rest], W=W, args=args, abs_tol=abs_tol, rel_tol=rel_tol, deadline=deadline,
                seed=seed, verbose=verbose, t0=t0, state=state, const=const, ram=ram, engines=None)
    base.update(extra)
    return types.SimpleNamespace(**base)


def _check_rows(menu, alternative, num_alternatives):
    """(menu, alternative, n) as int64 arrays and the number of alternatives, after the row checks
    of the reference's RUM constructor (same messages), each run only once the ones before it pass."""
    menu, alternative = _as_int(menu, "menu"), _as_int(alternative, "alternative")
    if menu.ndim != 1 or alternative.shape != menu.shape:
        raise ValueError("menu and alternative must be 1-D arrays of the same length")
    if menu.size == 0:
        raise ValueError("no observations")
    for bad, msg in ((menu <= 0, "row %d: empty menu"), (alternative < 0, "row %d: negative alternative")):
        i = _first_true(bad)
        if i is not None:
            raise ValueError(msg % i)
    need = max(int(menu.max()).bit_length(), int(alternative.max()) + 1)
    n = need if num_alternatives is None else int(num_alternatives)
    if n < need:
        raise ValueError("num_alternatives = %d, but the data use alternatives up to %d" % (n, need - 1))
    if n > 62:
        raise ValueError("menus are 64-bit masks: at most 62 alternatives")
    i = _first_true(((menu >> alternative) & 1) == 0)
    if i is not None:
        raise ValueError("row %d: alternative %d is not in its menu (mask %d)" % (i, alternative[i], menu[i]))
    # a repeated (menu, alternative) pair: the smallest such pair in (menu, alternative) order, at
    # its first row (a stable sort keeps equal pairs in row order)
    order = np.lexsort((alternative, menu))
    sm, sa = menu[order], alternative[order]
    j = _first_true((sm[1:] == sm[:-1]) & (sa[1:] == sa[:-1]))
    if j is not None:
        i = int(order[j])
        raise ValueError("row %d: (menu %d, alternative %d) appears more than once" % (i, menu[i], alternative[i]))
    return menu, alternative, n


def _first_true(mask):
    hit = np.flatnonzero(mask)
    return int(hit[0]) if hit.size else None


class RUM:
    def __init__(self, menu, alternative, choice_probability, num_alternatives=None):
        menu, alternative, n = _check_rows(menu, alternative, num_alternatives)
        self.n, self.M, self.menu, self.alternative = n, int(menu.size), menu, alternative
        self.singleton = (menu & (menu - 1)) == 0
        self._rest = np.flatnonzero(~self.singleton)
        self._single = np.flatnonzero(self.singleton)
        self._geom = _Geometry(menu[self._rest], alternative[self._rest], n)
        self._menu_of_row = np.unique(menu, return_inverse=True)[1]
        self.choice_probability = self._check_p(choice_probability)
        self._w_key = self._w_val = None

    @classmethod
    def from_rows(cls, rows, num_alternatives=None):
        rows = np.asarray(rows)
        if rows.ndim != 2 or rows.shape[1] != 3:
            raise ValueError("rows must be an (M, 3) array: menu, alternative, choice probability")
        return cls(rows[:, 0], rows[:, 1], rows[:, 2], num_alternatives=num_alternatives)

    def __repr__(self):
        return "RUM(n=%d, %d rows, %d menus, %d singleton rows)" % (self.n, self.M, np.unique(self.menu).size,
                                                                    self._single.size)

    # ------------------------------------------------------------------ distance
    def calculate_distance(self, abs_tol=None, rel_tol=None, method=None, method_args=None, max_time=None, p=None,
                           vcov=None, seed=0, verbose=False, machine=None, opt=None, opt_rel_tol=None, opt_floor=None):
        """dist(p) with a certified bracket (the reference's rum_solver.py:233-296), or with opt the
        race to ub <= max((1 + opt_rel_tol) opt, opt_floor)."""
        if opt is not None:
            return self._race(opt, opt_rel_tol, opt_floor, abs_tol, rel_tol, method, method_args, max_time, p,
                              vcov, seed, verbose, machine)
        if opt_rel_tol is not None or opt_floor is not None:
            raise ValueError("opt_rel_tol and opt_floor set a race target and need opt")
        t0 = time.perf_counter()
        abs_tol, rel_tol, max_time = _tolerances(abs_tol, rel_tol, max_time)
        if seed is None:
            seed = int(np.random.SeedSequence().entropy % (2 ** 31))
        p = self.choice_probability if p is None else self._check_p(p)
        W, const = self._weights(vcov, p)
        prof = _machine.load() if machine is None else machine
        if self._geom.M == 0:
            return self._singletons_only(t0, abs_tol, rel_tol, seed, const, prof)
        name, reason, args = self._plan(method, method_args, W is not None, prof)
        task = _task(self, p, W, args, abs_tol, rel_tol, t0 + max_time, seed, verbose, t0, const,
                     prof.get("ram_bytes"))
        t_run = time.perf_counter()
        stop, message = _execute(name, task, own_engines=True)
        return self._result(t0, t_run, name, reason, args, abs_tol, rel_tol, seed, task.state, stop, message, const,
                            prof)

    def _singletons_only(self, t0, abs_tol, rel_tol, seed, const, prof):
        """Every row is a singleton menu: the distance is the constant part, exactly."""
        d = math.sqrt(const)
        return self._result(t0, t0, "none", "only singleton menus", {}, abs_tol, rel_tol, seed,
                            dict(lb=d, ub=d, p_fit=np.zeros(0), iterations=0, trace=[], info={}, setup=0.0),
                            "tol", "", 0.0, prof)

    # ------------------------------------------------------------------ session / bootstrap
    def session(self, method=None, method_args=None, vcov=None, machine=None, seed=0):
        return Session(self, method, method_args, vcov, machine, seed)

    def bootstrap(self, n=None, panel=None, replies=None, B=None, weights=None, seed=None, r_n=None, eps_n=None,
                  method=None, method_args=None, abs_tol=None, rel_tol=None, refine=None, max_time=None,
                  draw_time=None, vcov=None, machine=None, checkpoint=None, verbose=False, batch=None):
        """Hong and Li's bootstrap p-value and confidence interval (solver/inference), the reference's
        rum_solver.py:304-331. batch: None or a dict of options of the batched draw solver
        (solver/inference/batched.py); None solves the draws one at a time as the reference does."""
        from .inference.bootstrap import Bootstrap
        return Bootstrap(self, n, panel=panel, replies=replies, B=B, weights=weights, seed=seed, r_n=r_n, eps_n=eps_n,
                         method=method, method_args=method_args, abs_tol=abs_tol, rel_tol=rel_tol, refine=refine,
                         max_time=max_time, draw_time=draw_time, vcov=vcov, machine=machine, checkpoint=checkpoint,
                         verbose=verbose, batch=batch)

    # ------------------------------------------------------------------ race
    def _race(self, opt, opt_rel_tol, opt_floor, abs_tol, rel_tol, method, method_args, max_time, p, vcov, seed,
              verbose, machine):
        """The race to a known optimum (rum_solver.py:333-453 of the reference): run until the
        pushed (bm_admm) or master (flow_cg) upper bound reaches the target; no lower bound."""
        t0 = time.perf_counter()
        if abs_tol is not None or rel_tol is not None:
            raise ValueError("with opt the call is a race to ub <= target and computes no lower bound, so abs_tol "
                             "and rel_tol do not apply; opt_rel_tol and opt_floor set the target")
        opt = float(opt)
        r = float(defaults.OPT_REL_TOL if opt_rel_tol is None else opt_rel_tol)
        floor = float(defaults.OPT_FLOOR if opt_floor is None else opt_floor)
        max_time = float(defaults.MAX_TIME if max_time is None else max_time)
        if not (math.isfinite(opt) and opt >= 0 and r >= 0 and floor >= 0 and max_time > 0):
            raise ValueError("opt, opt_rel_tol and opt_floor must be finite and >= 0, and max_time > 0")
        target = max((1.0 + r) * opt, floor)
        if not target > 0:
            raise ValueError("opt = 0 makes the target (1 + opt_rel_tol) * opt = 0, which a race never reaches: "
                             "pass opt_floor > 0")
        p = self.choice_probability if p is None else self._check_p(p)
        W, const = self._weights(vcov, p)
        if W is not None:
            raise ValueError("a race computes the plain L2 distance (vcov None or the identity)")
        if self._geom.M == 0 or target ** 2 <= const:
            raise ValueError("the singleton-menu rows alone are at distance %.10g, so a target of %.10g leaves "
                             "nothing to race" % (math.sqrt(const), target))
        prof = _machine.load() if machine is None else machine
        name, reason, args = self._plan(method, method_args, False, prof)
        rest = math.sqrt(target ** 2 - const) if const else target
        task = _task(self, p, None, args, 0.0, 0.0, t0 + max_time, seed, verbose, t0, const, prof.get("ram_bytes"),
                     below=rest, above=None, race=True)
        st = task.state
        t_run = time.perf_counter()
        stop, message = _execute(name, task, own_engines=True)
        t_end = time.perf_counter()
        ub = math.sqrt(st["ub"] ** 2 + const) if math.isfinite(st["ub"]) else math.inf
        flag = 4 if (stop == "threshold" or ub <= target) else _FLAG.get(stop, -2)
        nan = math.nan
        return Result(exit_flag=flag, status=STATUS[flag], message=message, distance=ub, lb=nan, ub=ub, abs_gap=nan,
                      rel_gap=nan, abs_tol=nan, rel_tol=nan, met_abs=False, met_rel=False, method=name,
                      auto_reason=reason, method_args={k: v for k, v in args.items() if not k.startswith("_")},
                      iterations=int(st["iterations"]), iteration_unit=_UNIT[name], time=t_end - t0,
                      time_setup=(t_run - t0) + st["setup"], time_solve=t_end - t_run - st["setup"], seed=seed,
                      p_fit=None, trace=[(i, s, math.sqrt(u * u + const) if math.isfinite(u) else u, nan)
                                         for i, s, u, _l in st["trace"]],
                      machine=_machine_line(prof), info=st["info"], mode="race", opt=opt, target=target,
                      abs_to_opt=ub - opt if math.isfinite(ub) else None,
                      rel_to_opt=(ub - opt) / opt if math.isfinite(ub) and opt > 0 else None)

    # ------------------------------------------------------------------ helpers
    def _check_p(self, p):
        p = np.asarray(p, dtype=np.float64)
        if p.shape != (self.M,):
            raise ValueError("choice probabilities: expected %d values, got shape %s" % (self.M, p.shape))
        if not np.isfinite(p).all():
            raise ValueError("choice probabilities must be finite")
        if (p < -1e-12).any() or (p > 1 + 1e-12).any():
            i = int(np.flatnonzero((p < -1e-12) | (p > 1 + 1e-12))[0])
            raise ValueError("row %d: probability %.6g is outside [0, 1]" % (i, p[i]))
        sums = np.bincount(self._menu_of_row, weights=p)
        if (sums > 1 + 1e-9).any():
            warnings.warn("%d menus have observed probabilities summing to more than 1 (largest %.6g)"
                          % (int((sums > 1 + 1e-9).sum()), sums.max()))
        return p

    def _finite_p(self, p):
        p = np.asarray(p, dtype=np.float64)
        if p.shape != (self.M,):
            raise ValueError("p: expected %d values, got shape %s" % (self.M, p.shape))
        if not np.isfinite(p).all():
            raise ValueError("p must be finite")
        return p

    def _weights(self, vcov, p):
        """(W on the non-singleton rows or None for L2, the singleton part of dist^2)."""
        r_s = 1.0 - p[self._single]
        if vcov is None:
            return None, float(r_s @ r_s)
        V = vcov.tocsr() if sp.issparse(vcov) else np.asarray(vcov, dtype=np.float64)
        if V.shape != (self.M, self.M):
            raise ValueError("vcov must be (%d, %d); got %s" % (self.M, self.M, V.shape))
        if _is_identity(V):
            return None, float(r_s @ r_s)
        return V, float(r_s @ r_s)

    def _plan(self, method, method_args, weighted, prof):
        n, mp = self.n, self._geom.M
        if method is None:
            if method_args:
                raise ValueError("method_args needs a named method")
            name, reason = choose_method(n, mp, weighted, prof)
        else:
            name, reason = _method_name(method), ""
        if name not in BUILT:
            raise RuntimeError("%s is not part of this package (the MI355X engines are bm_admm and flow_cg; the "
                               "reference's Gurobi and Flow-ADMM methods were left out)" % name)
        args = dict(defaults.METHOD_ARGS[name])
        bad = set(method_args or {}) - set(args)
        if bad:
            raise ValueError("unknown method_args for %s: %s; allowed: %s" % (name, sorted(bad), sorted(args)))
        args.update(method_args or {})
        if name == "bm_admm" and args.get("beta") == "adaptive" and "check" not in (method_args or {}):
            args["check"] = params.BM_BETA_ADAPT_CHECK
        return name, reason, _PLAN[name](n, mp, weighted, prof, args)

    def _result(self, t0, t_run, name, reason, args, abs_tol, rel_tol, seed, st, stop, message, const, prof):
        lb = math.sqrt(max(st["lb"], 0.0) ** 2 + const)
        ub = math.sqrt(st["ub"] ** 2 + const) if math.isfinite(st["ub"]) else math.inf
        abs_gap = ub - lb
        rel_gap = abs_gap / lb if lb > 0 else math.inf
        met_abs, met_rel = abs_gap <= abs_tol, lb > 0 and rel_gap <= rel_tol
        flag = 3 if (met_abs and met_rel) else 1 if met_abs else 2 if met_rel else _FLAG.get(stop, -2)
        message = message or st.get("message", "")
        if flag == -2 and not message:
            message = ("the method stopped on its own test, but the bracket does not meet the tolerances"
                       if stop == "tol" else "the method stopped without meeting the tolerances")
        elif flag == 0 and not message:
            message = "max_time reached"
        elif flag == -1 and not message:
            message = "the method's iteration limit was reached (see method_args)"
        info, met = st["info"], flag in (1, 2, 3)
        notes = []
        if name == "bm_admm" and args.get("dtype") == "float32" and not met:
            notes.append("a float32 state's pushed upper bound usually stalls above the tolerances; dtype='mixed' "
                         "or 'float64' certifies where it fits")
        if info.get("lb_rigorous") is False:
            notes.append("lb is the %s pricing bound, since no float64 bound fits the GPU at n = %d, so it does "
                         "not cover the float32 rounding inside the sweep" % (info.get("lb_dtype"), self.n))
        if notes:
            message = "; ".join(([message] if message else []) + notes)
        p_fit = None
        if st["p_fit"] is not None:
            p_fit = np.ones(self.M)
            p_fit[self._rest] = st["p_fit"]
        t_end = time.perf_counter()
        setup = (t_run - t0) + st["setup"]
        return Result(exit_flag=flag, status=STATUS[flag], message=message, distance=ub, lb=lb, ub=ub,
                      abs_gap=abs_gap, rel_gap=rel_gap, abs_tol=abs_tol, rel_tol=rel_tol, met_abs=bool(met_abs),
                      met_rel=bool(met_rel), method=name, auto_reason=reason,
                      method_args={k: v for k, v in args.items() if not k.startswith("_")},
                      iterations=int(st["iterations"]), iteration_unit=_UNIT.get(name, ""), time=t_end - t0,
                      time_setup=setup, time_solve=(t_end - t0) - setup, seed=seed, p_fit=p_fit, trace=st["trace"],
                      machine=_machine_line(prof), info=st["info"])


# ---------------------------------------------------------------------- session
class Session:
    """One plan, many solves (the bootstrap's solver; rum_solver.py:570-658 of the reference).

    The method, its storage format and its GPU budget are planned once; the session keeps its GPU
    engines (BM-ADMM slabs or the Flow-CG pricer) between solves, so a new p costs an upload of p,
    not an allocation. keep= stores the solve's state (BM-ADMM: the iterate on the device;
    Flow-CG: the columns and weights) and its fitted point; warm= starts from it.
    """

    def __init__(self, rum, method=None, method_args=None, vcov=None, machine=None, seed=0):
        t0 = time.perf_counter()
        self.rum, self.vcov, self.seed = rum, vcov, seed
        self.prof = _machine.load() if machine is None else machine
        weighted = rum._weights(vcov, rum.choice_probability)[0] is not None
        self.method, self.auto_reason, self._args = rum._plan(method, method_args, weighted, self.prof)
        self.method_args = {k: v for k, v in self._args.items() if not k.startswith("_")}
        self.fit = None
        self._kept = None
        self._engines = {}
        self.time_plan = time.perf_counter() - t0

    def solve(self, p, abs_tol=None, rel_tol=None, max_time=None, below=None, above=None, warm=False, keep=False,
              verbose=False):
        task = self._task(p, abs_tol, rel_tol, max_time, below, above, warm, keep, verbose)
        if isinstance(task, Result):
            return task
        task.t_run = time.perf_counter()
        stop, message = _execute(self.method, task, own_engines=False)
        return self._finish(task, stop, message)

    def _task(self, p, abs_tol, rel_tol, max_time, below, above, warm, keep, verbose):
        """The runner's task for one solve (or the Result itself when only singleton menus remain)."""
        rum, t0 = self.rum, time.perf_counter()
        abs_tol, rel_tol, max_time = _tolerances(abs_tol, rel_tol, max_time)
        if warm and self._kept is None:
            raise ValueError("warm=True starts from a kept solve: call solve(..., keep=True) first")
        p = rum._finite_p(p)
        W, const = rum._weights(self.vcov, p)
        if rum._geom.M == 0:
            return rum._singletons_only(t0, abs_tol, rel_tol, self.seed, const, self.prof)
        return _task(rum, p, W, self._args, abs_tol, rel_tol, t0 + max_time, self.seed, verbose, t0, const,
                     self.prof.get("ram_bytes"), below=_on_rest(below, const, False),
                     above=_on_rest(above, const, True), warm=self._kept if warm else None, keep=bool(keep),
                     engines=self._engines, t_run=t0)

    def _finish(self, task, stop, message):
        st = task.state
        kept = st.pop("keep", None)
        res = self.rum._result(task.t0, task.t_run, self.method, self.auto_reason, self._args, task.abs_tol,
                               task.rel_tol, self.seed, st, stop, message, task.const, self.prof)
        if task.keep:
            self._kept, self.fit = kept, res.p_fit
        return res

    def close(self):
        self._kept = self.fit = None
        for e in list(self._engines.values()):
            try:
                e.close()
            except Exception:
                pass
        self._engines.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __repr__(self):
        return "Session(%s%s, %s; kept state: %s)" % (self.method, "  [%s]" % self.auto_reason if self.auto_reason
                                                      else "", self.method_args, "no" if self._kept is None else "yes")


def _on_rest(tau, const, above):
    """A threshold on the whole distance moved onto the non-singleton part: dist^2 = rest^2 + const.
    A threshold the singleton rows alone reach is met at once (above: from 0 on) or never (below)."""
    if tau is None:
        return None
    tau = float(tau)
    if not math.isfinite(tau):
        raise ValueError("below and above must be finite")
    room = tau * tau - const
    if room > 0:
        return math.sqrt(room)
    return 0.0 if above else None


# ---------------------------------------------------------------------- choosing a method
def choose_method(n, mp, weighted, prof):
    """(method, reason) when none is named (the rule of the reference's rum_solver.py:675-711): BM-ADMM
    from the density defaults.AUTO sets for this n, Flow-CG below that density, for small n, and
    whenever BM-ADMM does not fit. The reason lists every step that decided it."""
    density = mp / (n << (n - 1))
    why = ["n = %d, %d rows, density %.3g%%" % (n, mp, 100 * density)]
    blocked = _why_no_gpu_method(weighted, prof)
    candidates = []
    if blocked:
        why.append(blocked)
    else:
        d_min = bm_admm_min_density(n)
        if d_min is None:
            why.append("n <= %d -> flow_cg at any density" % (min(defaults.AUTO["bm_admm_min_density"]) - 1))
        elif density < d_min:
            why.append("density < %.2g%% at n = %d -> flow_cg" % (100 * d_min, n))
        else:
            candidates.append(("bm_admm", "density >= %.2g%% at n = %d -> bm_admm" % (100 * d_min, n)))
        candidates.append(("flow_cg", None))
    for name, because in candidates:
        try:
            _PLAN[name](n, mp, False, prof, dict(defaults.METHOD_ARGS[name]))
        except RuntimeError as e:           # does not fit: the next candidate
            why.append(str(e))
            continue
        return name, "; ".join(why + ([because] if because else []))
    raise RuntimeError("no method can run on this machine: " + "; ".join(why))


def _why_no_gpu_method(weighted, prof):
    """Why neither GPU method can take this problem, or None."""
    if weighted:
        return "weighted norm: only the Gurobi methods take vcov, and they are not part of this package"
    if not _machine.can_use_gpu(prof):
        g = prof.get("gpu")
        return "no usable GPU (%s)" % ((g or {}).get("reason") or "none found" if prof.get("use_gpu", True)
                                      else "use_gpu = False")
    return None


def bm_admm_min_density(n):
    steps = [d for n0, d in sorted(defaults.AUTO["bm_admm_min_density"].items()) if n >= n0]
    return steps[-1] if steps else None


def _need_gpu(name, weighted, prof):
    if weighted:
        raise RuntimeError("%s computes the unweighted distance only (vcov None or identity)" % name)
    if not _machine.can_use_gpu(prof):
        g = prof.get("gpu")
        raise RuntimeError("%s needs the MI355X engines (librumgpu on a gfx950 GPU): %s" % (
            name, "use_gpu = False in the machine profile" if g and not prof.get("use_gpu", True)
            else (g or {}).get("reason") or "no GPU found"))


# Device bytes per slab entry of one BM-ADMM solve on the MI355X engines (solver/bm_admm/engine.py
# slab_bytes_per_entry): float64 iterates 59, float32 35, mixed (both handles) 94; plus 10 %.
_BM_BYTES_PER_ENTRY = {"float64": 59, "float32": 35, "mixed": 94}


def _bm_admm_bytes(n, dtype):
    return int(1.1 * _BM_BYTES_PER_ENTRY[dtype] * (n << (n - 1)))


def _plan_bm_admm(n, mp, weighted, prof, a):
    _need_gpu("bm_admm", weighted, prof)
    if not 2 <= n <= 26:
        raise RuntimeError("bm_admm handles 2 <= n <= 26 here")
    for k in ("check", "lb_every", "max_iter"):
        if int(a[k]) < 1:
            raise ValueError("bm_admm %s must be >= 1" % k)
    if a["dtype"] not in ("auto", "float64", "float32", "mixed"):
        raise ValueError("bm_admm dtype must be 'auto', 'float64', 'float32' or 'mixed'")
    if a["use_graph"] not in ("auto", True, False):
        raise ValueError("bm_admm use_graph must be 'auto', True or False")
    budget = defaults.AUTO["gpu_memory_fraction"] * _machine.gpu_free_bytes()
    dt = "float64" if a["dtype"] == "auto" else a["dtype"]
    need = _bm_admm_bytes(n, dt)
    if need > budget:
        raise RuntimeError("bm_admm at n = %d in %s needs %.1f GB of GPU memory; %.1f GB may be used. Use flow_cg"
                           % (n, dt, need / 2 ** 30, budget / 2 ** 30))
    g = (n <= 22) if a["use_graph"] == "auto" else bool(a["use_graph"])
    a.update(dtype=dt, use_graph=g, _gpu_bytes=need)
    if a["beta"] is None:
        a["beta"] = params.bm_beta_default(mp)
    elif a["beta"] == "cert":
        a["beta"] = params.bm_beta_cert(mp)
    elif a["beta"] == "adaptive":
        ad = dict(params.BM_BETA_ADAPT)
        ad["every"] = max(1, int(round(ad.pop("every_iters") / int(a["check"]))))     # in checks
        a["beta"], a["_beta_adapt"] = params.bm_beta_adaptive_start(mp), ad
    else:
        a["beta"] = float(a["beta"])
        if not a["beta"] > 0:
            raise ValueError("bm_admm beta must be > 0, None (the race rule), 'cert' or 'adaptive'")
    return a


def _flow_cg_bytes(n, layout, dtype, mp):
    N, cap = 1 << n, math.comb(n, n // 2)
    if layout == "mask":
        return 9 * N + 12 * mp
    if layout == "slab":
        return 9 * N + 4 * (n << (n - 1)) + 20 * mp
    k_pred = max(1, n - 10)
    item = {"float32": 4, "uint24": 3, "uint16": 2}[dtype]
    return (2 * cap * item + sum(math.comb(n, k) for k in range(k_pred, n + 1)) + 9 * (1 << k_pred)
            + 2 * 8 * (cap // 1024 + 1) + 24 * mp)


_CG_ENGINE = (("mask", "float32"), ("stream", "float32"), ("stream", "uint24"), ("stream", "uint16"))
_CG_OLD = {"stream_f32": ("stream", "float32"), "stream_u24": ("stream", "uint24"), "stream_u16": ("stream", "uint16")}


def _cg_storage(a):
    """The requested (layout, dtype) of flow_cg's pricing, old names resolved and checked."""
    layout, dt = a["pricing"], a["dtype"]
    if layout in _CG_OLD:
        if dt not in ("auto", _CG_OLD[layout][1]):
            raise ValueError("pricing=%r is %s storage; it contradicts dtype=%r" % (layout, _CG_OLD[layout][1], dt))
        layout, dt = _CG_OLD[layout]
    if layout not in ("auto", "mask", "slab", "stream"):
        raise ValueError("flow_cg pricing must be 'auto', 'mask', 'slab' or 'stream' (old names: stream_f32, "
                         "stream_u24, stream_u16)")
    if dt not in ("auto", "float32", "uint24", "uint16"):
        raise ValueError("flow_cg dtype must be 'auto', 'float32', 'uint24' or 'uint16'")
    if layout in ("mask", "slab") and dt not in ("auto", "float32"):
        raise ValueError("%s pricing stores float32 only; dtype=%r needs pricing='stream' or 'auto'" % (layout, dt))
    return layout, dt


def _cg_candidates(layout, dt, n, mp):
    """Pricing engines in order of preference for the request. Many rows per alternative put the
    slab form first: the row-order sweeps test every row at every subset (mp 2^(n-1) tests), the
    slab form reads each arc weight once from a subset-sum butterfly."""
    pool = [("slab", "float32")] if layout == "slab" else list(_CG_ENGINE)
    if layout == "auto" and dt in ("auto", "float32") and mp > defaults.AUTO["cg_slab_rows_per_alt"] * n:
        pool.insert(0, ("slab", "float32"))
    for lay, d in pool:
        if layout in ("auto", lay) and dt in ("auto", d) and not (lay in ("mask", "slab") and n > 31):
            yield lay, d


def _plan_flow_cg(n, mp, weighted, prof, a):
    _need_gpu("flow_cg", weighted, prof)
    k = int(a["max_iters"]) + 2
    host = 8 * k * k + 8 * k * ((mp + 63) // 64)
    if prof.get("ram_bytes") and host > 0.5 * prof["ram_bytes"]:
        raise RuntimeError("flow_cg max_iters = %d needs %.1f GB of host memory for its Gram matrix"
                           % (a["max_iters"], host / 2 ** 30))
    layout, dt = _cg_storage(a)
    budget = defaults.AUTO["gpu_memory_fraction"] * _machine.gpu_free_bytes()
    # a single run over 4,096 rows keeps its column bitsets, p and p_out on the device
    # (flow_cg.engine.DeviceColumns) when they fit beside the pricer, else on the host
    cols = 8 * k * ((mp + 63) // 64) + 16 * mp + 20 * k if mp > 4096 else 0
    tried = []
    for lay, d in _cg_candidates(layout, dt, n, mp):
        need = _flow_cg_bytes(n, lay, d, mp)
        if need > budget:
            tried.append("%s %s %.1f GB" % (lay, d, need / 2 ** 30))
            continue
        a.update(pricing=lay, dtype=d, _gpu_bytes=need)
        if cols and need + cols > budget:
            a.setdefault("_cols", "host")
        elif cols:
            a["_gpu_bytes"] = need + cols
        return a
    raise RuntimeError("flow_cg at n = %d does not fit the GPU: %.1f GB may be used (%s)"
                       % (n, budget / 2 ** 30, "; ".join(tried) or "mask needs n <= 31"))


_PLAN = {"bm_admm": _plan_bm_admm, "flow_cg": _plan_flow_cg}


# ---------------------------------------------------------------------- running a method
def _engines(t):
    if t.engines is None:
        t.engines = {}
    return t.engines


def _close_engines(t):
    if t.engines:
        for e in list(t.engines.values()):
            try:
                e.close()
            except Exception:
                pass
        t.engines.clear()


class _EngineSet:
    """GPU objects of one plan (BM-ADMM handles or the Flow-CG pricer), closed together."""

    def __init__(self, key, items):
        self.key, self.items = key, items

    def close(self):
        for h in self.items.values():
            h.close()
        self.items = {}


def _bm_handles(t):
    """The BM-ADMM handle(s) of this task's plan, created once per Session (or per call)."""
    from .bm_admm.engine import Handle
    g, a = t.geom, t.args
    n = g.num_alternatives
    eng = _engines(t)
    key = (a["dtype"], float(a["beta"]), float(a["alpha"]), float(a["precond_p"]), bool(a["use_graph"]),
           int(a["check"]))
    dts = {"float64": ("float64",), "float32": ("float32",), "mixed": ("float32", "float64")}[a["dtype"]]
    if "bm" not in eng or eng["bm"].key != key:
        if "bm" in eng:
            eng.pop("bm").close()
        items = {}
        tuned_s = None
        for d in dts:
            h = Handle(n, g.menu, g.alternative.astype(np.int32), t.p, dtype=d, beta=a["beta"], alpha=a["alpha"],
                       precond_p=a["precond_p"])
            if defaults.AUTO["bm_autotune"] and n >= defaults.AUTO["bm_autotune_min_n"]:
                from .bm_admm.autotune import tune
                t_tune = time.perf_counter()
                plan, _us = tune(h, device=_machine.gpu_name())
                tuned_s = (tuned_s or 0.0) + time.perf_counter() - t_tune
                h.window_plan = list(plan)
                h.init_cold()               # the timing runs moved the state
            if a["use_graph"]:
                h.set_graph(int(a["check"]))
            items[d] = h
        eng["bm"] = _EngineSet(key, items)
        if tuned_s is not None:             # plan lookup (or timing, on a cache miss) inside this call
            t.state["info"]["autotune_s"] = round(tuned_s, 3)
    hs = [eng["bm"].items[d] for d in dts]
    for h in hs:
        h.set_p(t.p)
    plan = getattr(hs[-1], "window_plan", None)
    if plan is not None:                    # every solve on these handles reports their tuned plan
        t.state["info"]["window_plan"] = list(plan)
    return hs


def _run_bm_admm(t):
    """rum_solver.py:869-925 of the reference on the HIP engine: float64 push (running min, with its
    fitted point) and float64 lower bound every lb_every checks, gap stops, Session thresholds."""
    from .bm_admm.engine import solve_kadmm, solve_mixed
    a, st = t.args, t.state
    t_call = time.perf_counter()
    hs = _bm_handles(t)
    checks = [0]
    race = getattr(t, "race", False)

    def push(h):
        if time.perf_counter() >= t.deadline:
            raise _TimeUp
        if not checks[0]:
            st["setup"] = time.perf_counter() - t_call
        ub = h.push()
        checks[0] += 1
        st["iterations"] = checks[0] * int(a["check"])
        if ub < st["ub"]:
            st["ub"] = float(ub)
            st["p_fit"] = h.get_fit()
        if race:
            st["trace"].append((st["iterations"], time.perf_counter() - t.t0, st["ub"], math.nan))
        return ub

    def lower(h):
        if (checks[0] - 1) % int(a["lb_every"]) == 0:
            st["lb"] = max(st["lb"], h.lb())
        st["trace"].append((st["iterations"], time.perf_counter() - t.t0, st["ub"], st["lb"]))
        return st["lb"]

    below, above, warm = getattr(t, "below", None), getattr(t, "above", None), getattr(t, "warm", None)
    # start: warm from the kept state (held by the handle that kept it) or cold
    first = hs[0]
    for h in hs:                               # an adaptive penalty restarts from its start value
        if h.beta != float(a["beta"]):
            h.set_beta(float(a["beta"]))
    if warm is not None:
        src = warm["handle"]
        if src is first:
            first.load_kept()
        else:
            src.load_kept()
            first.copy_state_from(src)
    else:
        first.init_cold()
    kw = dict(max_iter=int(a["max_iter"]), check=int(a["check"]), push_fn=push, ub_arm_immediately=True,
              lb_fn=None if race else lower, ub_running_min=True,
              gap_abs_tol=None if race else t.abs_tol, gap_rel_tol=None if race else t.rel_tol,
              time_limit=max(t.deadline - time.perf_counter(), 0.0), verbose=t.verbose,
              beta_adapt=None if race else a.get("_beta_adapt"))
    if below is not None and below > 0:
        kw.update(ub_target=below, ub_eps=0.0)
    if above is not None:
        kw["lb_target"] = above
    st["info"].update(gpu_bytes_planned=a["_gpu_bytes"], lb_dtype="float64", lb_layout="slab", lb_rigorous=True,
                      iterate_dtype=a["dtype"], windows=first.windows())
    if a["dtype"] == "mixed":
        r = solve_mixed(hs[0], hs[1], **kw)
        final = hs[1] if r.get("iters64") else hs[0]
        st["info"].update(iters32=r["iters32"], iters64=r["iters64"])
    else:
        r = solve_kadmm(first, **kw)
        final = first
    st["iterations"] = int(r["iters"])
    st["info"].update(per_iter_ms=round(r["per_iter_ms"], 4), push_s=round(r["push_s"], 3), beta_final=r["beta_final"],
                      beta_changes=len(r["beta_changes"]))
    if getattr(t, "keep", False):
        final.keep()
        st["keep"] = dict(handle=final, beta=float(final.beta))
    return {"gap_abs": "tol", "gap_rel": "tol", "time_limit": "time", "max_iter": "iters", "ub_target": "threshold",
            "lb_target": "threshold"}.get(r["stop_reason"], "stalled")


def _flow_cg_lb(n, g, p_fit, p, engine_lb, delta, dtype, pricer=None, perm=None):
    """flow_cg's final lower bound in the most precise form that fits (rum_solver.py:928-938 of the
    reference): the float64 Frank-Wolfe bound, its longest path taken on the mask pricer's lattice
    when the pricer has one (no slab), else through a float64 slab, else the engine's bound
    lowered by delta."""
    from . import _lib
    import ctypes
    if pricer is not None and pricer.layout in ("mask", "stream"):
        return pricer.lb64(perm, p_fit, p), "float64", pricer.layout
    need = 16 * (n << (n - 1)) + 16 * (1 << n)
    if need <= defaults.AUTO["gpu_memory_fraction"] * _machine.gpu_free_bytes():
        ctx = _lib.context(0)
        lb = ctypes.c_double()
        menu = _lib.c_array(g.menu, np.uint64)
        alt = _lib.c_array(g.alternative, np.int32)
        y = _lib.c_array(p_fit, np.float64)
        pv = _lib.c_array(p, np.float64)
        _lib.check(ctx.lib.rumgpu_lb64(ctx.h, n, _lib.ptr(menu), _lib.ptr(alt), g.M, _lib.ptr(y), _lib.ptr(pv),
                                       ctypes.byref(lb)), "lb64")
        return lb.value, "float64", "slab"
    return engine_lb - delta, dtype, "pricing sweep"


def _cg_stops(below, above, margin):
    kw = {}
    if below is not None:
        kw.update(phi_ub_target=max(below - margin, 0.0), early_ub_check=True)
    if above is not None:
        kw["lb_stop"] = above + margin
    return kw


def _cg_pricer(t):
    """The session's pricer for this plan, created once."""
    from .flow_cg.engine import Pricer, xmajor
    g, a = t.geom, t.args
    eng = _engines(t)
    key = (a["pricing"], a["dtype"])
    if "cg" not in eng or eng["cg"].key != key:
        if "cg" in eng:
            eng.pop("cg").close()
        perm, masks_x, _alts, offsets = xmajor(g.menu, g.alternative, g.num_alternatives)
        eng["cg"] = _EngineSet(key, {"pricer": Pricer(g.num_alternatives, masks_x, offsets, layout=a["pricing"],
                                                      dtype=a["dtype"])})
    pricer = eng["cg"].items["pricer"]
    if pricer.layout == "mask":
        # "layers" (the default) is the reference's layer-by-layer sweep; for rows of one alternative
        # "window" takes the window-pass sweep and "auto" the closure sweep (n <= 20)
        mode = a.get("_sweep", "layers")
        if mode == "auto" and g.num_alternatives > Pricer.CLOSURE_MAX_N:
            mode = "layers"
        pricer.set_sweep(mode)
    return pricer


def _run_flow_cg(t):
    """rum_solver.py:941-1004 of the reference: the engine solves for the float32-rounded p; the
    bracket is recomputed for the true p (exact ub of the fitted mixture, float64 lower bound), with
    one resume from the engine's columns when the float64 bracket misses the tolerance."""
    from .flow_cg.engine import drive
    pricer = _cg_pricer(t)
    # "_cols": "host" keeps the columns on the host as the reference does (qp_bitset_incremental.py)
    return drive(_flow_cg_steps(t, pricer, resident=t.args.get("_cols", "device") == "device"), pricer)


def _flow_cg_steps(t, pricer, resident=False):
    """_run_flow_cg's body as a generator of pricing requests (flow_cg.engine.cg_steps); returns the
    stop reason. The batched bootstrap drives many of these with one batched sweep per round.
    resident: the pricer is this run's alone, so large runs keep their columns on the device."""
    from .flow_cg.engine import cg_steps
    g, a, st = t.geom, t.args, t.state
    n = g.num_alternatives
    delta = float(np.linalg.norm(t.p - t.p.astype(np.float32).astype(np.float64)))
    rel, abs_e = t.rel_tol, max(t.abs_tol - 2 * delta, 0.0)
    race = getattr(t, "race", False)
    below, above, kept = getattr(t, "below", None), getattr(t, "above", None), getattr(t, "warm", None)
    warm = {} if kept is None else dict(init_columns=kept["columns"], init_lambda=kept["lam"])
    stops = _cg_stops(below, above, 2 * delta)
    r = None
    for attempt in range(2):
        remaining = t.deadline - time.perf_counter()
        if remaining <= 0:
            return "time"
        t_call = time.perf_counter()
        kw = dict(max_iters=int(a["max_iters"]), time_limit_s=remaining, seed=t.seed, verbose=t.verbose, **warm,
                  **stops, qp=a.get("_qp", "native"), resident=resident)
        if not race:
            kw.update(gap_rel_tol=rel, gap_abs_tol=abs_e if abs_e else None)
        r = yield from cg_steps(n, g.menu, g.alternative, t.p, pricer, **kw)
        p_fit = np.empty(g.M)
        p_fit[r["perm"]] = r["p_out"]
        st["p_fit"] = p_fit
        st["ub"] = float(np.linalg.norm(p_fit - t.p))
        lb_prev = st["lb"]
        if race or (below is not None and st["ub"] < below):
            lb, lb_dt, lb_lay = lb_prev, None, "not computed (below the threshold)"
        else:
            # the float64 bound on the pricer's mask lattice, or (a["_lb_form"] == "slab") through a
            # float64 slab as the reference computes it
            if a.get("_lb_form", "mask") == "mask" and pricer.layout in ("mask", "stream"):
                # the longest path of the float64 bound is a request like a pricing sweep, so a
                # batched driver can run it beside other draws' sweeps
                rr = t.p - p_fit
                vmax = yield ("vmax64", np.ascontiguousarray(rr[r["perm"]]))
                l2 = math.fsum(rr * rr) - 2.0 * (vmax - math.fsum(rr * p_fit))
                lb, lb_dt, lb_lay = math.sqrt(max(0.0, l2)), "float64", pricer.layout
            else:
                lb, lb_dt, lb_lay = _flow_cg_lb(n, g, p_fit, t.p, float(r["lb"]), delta, a["dtype"], None, r["perm"])
        st["lb"] = max(lb_prev, lb)
        st["iterations"] = int(r["n_cols"])
        st["setup"] += float(r["setup_s"])
        off = (t_call - t.t0) + float(r["setup_s"])
        st["trace"] += [(k, off + w, ub, max(lb_prev, l_ - delta)) for _it, w, ub, l_, k in r["trace"]]
        if st["trace"]:
            st["trace"][-1] = st["trace"][-1][:2] + (st["ub"], st["lb"])
        st["info"].update(pricing=r["layout"], pricing_dtype=r["dtype"], gpu_bytes=r["device_bytes"],
                          p_float32_shift=delta, pricing_margin=max(r["margins"]) if r["margins"] else 0.0,
                          restarts=attempt, lb_dtype=lb_dt, lb_layout=lb_lay, lb_rigorous=lb_dt in ("float64", None),
                          qp_s=round(r["qp_s"], 3), price_s=round(r["price_s"], 3))
        if getattr(t, "keep", False):
            st["keep"] = dict(columns=r["columns"], lam=r["lam"])
        if race:
            return "threshold" if st["ub"] <= below else {"time_limit": "time", "max_iters": "iters"}.get(
                r["stop_reason"], "stalled")
        if (below is not None and st["ub"] < below) or (above is not None and st["lb"] >= above):
            return "threshold"
        gap = st["ub"] - st["lb"]
        if (attempt or r["stop_reason"] not in ("gap_rel", "gap_abs", "phi_ub", "lb_stop") or gap <= t.abs_tol
                or gap <= t.rel_tol * st["lb"]):
            break
        excess = 1.5 * max(gap - (float(r["distance"]) - float(r["lb"])), 2 * delta)
        rel = max(t.rel_tol - excess / st["lb"], 0.0) if st["lb"] > 0 else 0.0
        abs_e = max(t.abs_tol - 2 * delta - excess, 0.0)
        if rel <= 0 and abs_e <= 0 and not stops:
            break
        stops = _cg_stops(below, above, 2 * delta + excess)
        warm = dict(init_columns=r["columns"], init_lambda=r["lam"])
    return {"gap_rel": "tol", "gap_abs": "tol", "rc": "tol", "time_limit": "time",
            "max_iters": "iters"}.get(r["stop_reason"], "stalled")


_RUN = {"bm_admm": _run_bm_admm, "flow_cg": _run_flow_cg}


def _execute(name, task, own_engines):
    """Run the method's runner on the task: (stop reason, message). A time limit, a decided
    threshold, Ctrl-C, running out of memory and any other failure end the run with a reason (the
    exit flag) instead of an exception. A one-shot call owns its engines and frees them here; a
    session keeps them for its next solve."""
    try:
        return _RUN[name](task), ""
    except _TimeUp:
        return "time", ""
    except _Decided:
        return "threshold", ""
    except KeyboardInterrupt:
        return "interrupted", "interrupted by the user"
    except Exception as e:                # noqa: BLE001 -- reported through the exit flag
        return ("oom" if _is_oom(e) else "error"), "%s: %s" % (type(e).__name__, e)
    finally:
        if own_engines:
            _close_engines(task)


# ---------------------------------------------------------------------- small helpers
def _method_name(method):
    """The canonical method name: case, surrounding blanks and '-' vs '_' do not matter, and the
    reference's aliases (cgbfqp, cgbf, ...) map to their methods."""
    key = str(method).strip().lower().replace("-", "_")
    name = ALIASES.get(key, key)
    if name in METHODS:
        return name
    raise ValueError("unknown method %r; choose from %s" % (method, METHODS))


def _as_int(x, what):
    a = np.asarray(x)
    if a.dtype.kind == "f":
        if not np.all(np.isfinite(a)) or not np.all(a == np.floor(a)):
            raise ValueError("%s must hold integers" % what)
    elif a.dtype.kind not in "iu":
        raise ValueError("%s must hold integers" % what)
    return a.astype(np.int64)


def _is_identity(V):
    n = V.shape[0]
    if sp.issparse(V):
        return (V - sp.identity(n, dtype=V.dtype, format="csr")).count_nonzero() == 0
    return np.array_equal(V, np.eye(n))


def _is_oom(e):
    return isinstance(e, MemoryError) or "OutOfMemory" in type(e).__name__ or "out of memory" in str(e).lower()


def _machine_line(prof):
    g = prof.get("gpu")
    gpu = ("%s (%s)" % (g.get("name"), "used" if _machine.can_use_gpu(prof) else "not used")) if g else "no GPU"
    return "%s; %s cores; no Gurobi" % (gpu, prof.get("cpu_count"))
