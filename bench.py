#!/usr/bin/env python3
"""Benchmark: certified BM-ADMM solves per second at an absolute gap of 1e-5 on one MI355X.

Workload (the metric's configuration): n = 18 alternatives, 1 % of the 18 * 2^17 (menu,
alternative) cells observed (config2 of solver/instances.py, seed 18001: ~23.6k cells, p_hat from
a 200-order Dirichlet mixture plus 500-reply multinomial noise). A step is one certified solve
through the unchanged surface (Session.solve with abs_tol = 1e-5, rel_tol = 0) from a cold start:
the bracket [lb, ub] with ub - lb <= 1e-5 is float64 (push and Frank-Wolfe bound) every check.

    python bench.py [--steps K] [--warmup W] [--dtype mixed|float64] [--plain-steps P] [--no-cpu]

Reported beside the value (one JSON line on stdout):
  plain      the same certified solve with the reference's own schedule in float64 (separate
             elementwise kernels, four full butterflies per iteration) on this GPU: tuned vs plain
  roofline   the dominant fused pass: algorithmic bytes per launch / its mean device time, from
             a start/stop HIP event pair recorded in each pass launch's own dispatch
             (hipExtLaunchKernelGGL) on the engine's stream, right after the timed region
             (rumgpu_bm_profile)
  cpu_baseline  the CPU oracle (oracle/rum_oracle.cpp, OpenMP on every host thread) on a bounded
             sample of the same solve: a few float64 iterations plus one push and one lower bound,
             scaled to a certified solve by the GPU run's iteration, check and lower-bound counts
  dsr_bootstrap  the reference's stored DSR bootstrap (n = 25, B = 10,000) run whole through the
             batched path: its time beside the stored run's, and the draws whose decision differs
Inputs are resident on the GPU before the timed region (the engines are built in warmup); each
solve call returns after its last stream synchronization, so the host clock brackets the device
work exactly.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "rum-testing-gpu_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "certified solves/s at abs gap 1e-5 (n=18); bootstrap draws/s, tuned vs plain"
HBM_PEAK_GBS = 8000.0


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dtype", default="float64", choices=["mixed", "float64", "float32"])
    ap.add_argument("--beta", default="adaptive",
                    help="'cert' (certificate rule), 'adaptive' (balanced on the certified gap), 'ref' (race rule) "
                         "or a number")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--abs-tol", type=float, default=1e-5)
    ap.add_argument("--max-time", type=float, default=120.0)
    ap.add_argument("--profile-iters", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=60)
    ap.add_argument("--no-graph", action="store_true", help="eager launches instead of HIP graph replays")
    ap.add_argument("--boot-draws", type=int, default=512,
                    help="open bootstrap draws of the n=25 panel solved by the batched (tuned) leg; 0 skips it")
    ap.add_argument("--boot-plain-draws", type=int, default=128,
                    help="of those, solved one by one as the reference does (the plain leg)")
    ap.add_argument("--boot-slots", type=int, default=64)
    ap.add_argument("--no-dsr", action="store_true",
                    help="skip the whole DSR alternative-0 bootstrap (n=25, B=10,000) checked against the stored run")
    ap.add_argument("--dsr-all-alternatives", action="store_true",
                    help="only the DSR bootstrap on all 901 cells (every alternative), against the published run")
    ap.add_argument("--reference-instances", action="store_true",
                    help="only the instances with the reference's certified optima (tests/golden/opt_refs_subset.json):"
                         " each solved to rel 1e-3 by the method of its stored record, timed beside the record")
    ap.add_argument("--only", default="",
                    help="with --reference-instances: a comma-separated list of instance names to run")
    ap.add_argument("--bootstrap-only", action="store_true",
                    help="only the bootstrap leg (prints its JSON object): the workload of its rocprof summary")
    ap.add_argument("--traffic-run", action="store_true",
                    help="only the workload for HBM counters: --profile-iters fused iterations on the n=18 "
                         "instance (or config3 with --traffic-n 22) with the autotuned window plan, then a device "
                         "copy of the 4 state slabs (a copy kernel of known size)")
    ap.add_argument("--traffic-n", type=int, default=18, choices=[18, 22])
    ap.add_argument("--config3-only", action="store_true",
                    help="only the n=22 (config3) iteration roofline (prints its JSON object)")
    ap.add_argument("--no-config3", action="store_true", help="skip the n=22 iteration roofline")
    ap.add_argument("--no-slots", action="store_true",
                    help="skip the batched BM-ADMM bootstrap (n=16 alternative-0 panel: draw slots with the device "
                         "queue, host-driven slots, per-draw solves, learned warm start)")
    ap.add_argument("--slots-only", action="store_true", help="only the batched BM-ADMM bootstrap")
    ap.add_argument("--no-ai", action="store_true", help="skip the reference's stored AI run (n=18, 73,492 cells)")
    ap.add_argument("--pricing-sweeps", default="",
                    help="only pricing sweeps on one instance (n20_d1pct, n25_sparse, config2 or config3), "
                         "--sweeps of them: the workload of the pricing profiles")
    ap.add_argument("--sweeps", type=int, default=20)
    ap.add_argument("--pricing-layout", default="slab",
                    help="with --pricing-sweeps: slab, mask, or stream_f32 / stream_u24 / stream_u16")
    ap.add_argument("--slab-tile", type=int, default=0,
                    help="with --pricing-sweeps: low elements per layered-max tile (2..10; 0 = the built-in 7)")
    ap.add_argument("--summarize-stats", default="",
                    help="print the per-kernel table of a rocprofv3 --kernel-trace --stats database (rocpd .db)")
    ap.add_argument("--summarize-traffic", nargs=3, metavar=("FETCH_DB", "WRITE_DB", "COPY_BYTES"),
                    help="print FETCH_SIZE / WRITE_SIZE per dispatch of two rocprofv3 --pmc databases, "
                         "calibrated on the ew_copy kernel of --traffic-run (COPY_BYTES read and written)")
    ap.add_argument("--traffic-json", default="", help="with --summarize-traffic: write bytes per dispatch by kernel")
    return ap.parse_args()


def _pass_bytes(kind, m2, isz, n, M):
    """Algorithmic HBM bytes of one launch of a fused pass (slab transfers x m2 x itemsize, plus
    the per-menu observation arrays of the prox and inv_gl of the D pass)."""
    transfers = {0: 6, 1: 4, 2: 2, 3: 5, 4: 3, 5: 2, 6: 9}[kind]
    extra = 0
    if kind == 4:
        extra = 8 * (1 << n) + 8 * M
    if kind == 2:
        extra = isz * (m2 // n)
    return transfers * m2 * isz + extra


KIND_NAMES = {0: "bm_pass SY (A window 1, s/y update, A' window 1)", 1: "bm_pass RHS (A', rhs, Q^T on window 2)",
              2: "bm_pass D (Q^T, inv_gl, Q on window 1)", 3: "bm_pass XK (Q, xk, A on window 2)",
              4: "bm_prox (per-menu projection, z update)", 5: "bm_pass T (middle window)",
              6: "bm_pass FULL (one window)"}


def main():
    a = _parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: the engines run on a single MI355X and have no multi-GPU path "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from solver import instances, _lib
    from solver.rum_solver import RUM
    from solver.bm_admm.engine import Handle, solve_kadmm

    if a.summarize_stats:
        print(_summarize_stats(a.summarize_stats))
        return
    if a.summarize_traffic:
        text, per_kernel = _summarize_traffic(*a.summarize_traffic)
        print(text)
        if a.traffic_json:
            with open(a.traffic_json, "w") as f:
                json.dump(per_kernel, f, indent=1)
        return
    if a.pricing_sweeps:
        _pricing_sweeps(a.pricing_sweeps, a.sweeps, a.slab_tile, a.pricing_layout)
        return
    if a.config3_only:
        print(json.dumps(_config3_bench(a)), flush=True)
        return
    if a.bootstrap_only:
        print(json.dumps(_bootstrap_bench(a)), flush=True)
        return
    if a.slots_only:
        print(json.dumps(_slots_bench()), flush=True)
        return
    if a.dsr_all_alternatives:
        print(json.dumps(_dsr_all_alternatives()), flush=True)
        return
    if a.reference_instances:
        print(json.dumps(_reference_instances([k for k in a.only.split(",") if k])), flush=True)
        return

    cd, n = instances.config2() if a.traffic_n == 18 or not a.traffic_run else instances.config3()
    rum = RUM.from_rows(cd, n)
    if a.traffic_run:
        from solver import params, machine
        from solver.bm_admm.autotune import tune
        g = rum._geom
        h = Handle(n, g.menu, g.alternative.astype(np.int32), rum.choice_probability[rum._rest], dtype=a.dtype,
                   beta=params.bm_beta_cert(g.M))
        plan, _t = tune(h, device=machine.gpu_name())
        h.init_cold()
        h2 = Handle(n, g.menu, g.alternative.astype(np.int32), rum.choice_probability[rum._rest], dtype=a.dtype,
                    beta=params.bm_beta_cert(g.M))
        h.run(a.profile_iters)
        h2.copy_state_from(h)           # 4 x ew_copy: each reads and writes n 2^(n-1) values
        print("[bench] traffic run: n=%d, window plan %s, %d iterations, state copy of 4 x %d bytes"
              % (n, list(plan), a.profile_iters, (n << (n - 1)) * h.dtype.itemsize), file=sys.stderr, flush=True)
        return
    p = rum.choice_probability
    mp = rum._geom.M
    m2 = n << (n - 1)
    print("[bench] config2: n=%d, %d cells (%.3f %% of %d), fingerprint %s" % (n, mp, 100.0 * mp / m2, m2,
                                                                             instances.fingerprint(cd)),
          file=sys.stderr, flush=True)

    margs = {"dtype": a.dtype}
    if a.no_graph:
        margs["use_graph"] = False
    if a.beta != "ref":
        margs["beta"] = a.beta if a.beta in ("cert", "adaptive") else float(a.beta)
    sess = rum.session(method="bm_admm", method_args=margs)
    results = []
    for i in range(a.warmup):
        r = sess.solve(p, abs_tol=a.abs_tol, rel_tol=0.0, max_time=a.max_time)
        print("[bench] warmup %d: flag %d, [%.10f, %.10f], %d it, %.3f s" % (i, r.exit_flag, r.lb, r.ub, r.iterations,
                                                                            r.time), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for i in range(a.steps):
        r = sess.solve(p, abs_tol=a.abs_tol, rel_tol=0.0, max_time=a.max_time)
        results.append(r)
        print("[bench] step %d: flag %d, [%.10f, %.10f] gap %.2e, %d it (%s), %.3f s"
              % (i, r.exit_flag, r.lb, r.ub, r.ub - r.lb, r.iterations,
                 {k: r.info.get(k) for k in ("iters32", "iters64")}, r.time), file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    certified = all(r.exit_flag in (1, 3) for r in results)
    value = a.steps / elapsed if certified else 0.0
    r_last = results[-1]

    # roofline of the dominant fused pass (profiled on the session's handles, same state)
    eng = sess._engines["bm"].items
    sess_plan = r_last.info.get("window_plan")
    prof = {}
    import ctypes
    for dname, h in eng.items():
        ms = (ctypes.c_double * 7)()
        cnt = (ctypes.c_int64 * 7)()
        _lib.check(h.lib.rumgpu_bm_profile(h.h, a.profile_iters, ms, cnt), "bm_profile")
        prof[dname] = {k: (ms[k], cnt[k]) for k in range(7) if cnt[k]}
    # the handle that ran the most iterations in the timed solves
    it32 = sum(int(r.info.get("iters32") or 0) for r in results)
    it64 = sum(int(r.info.get("iters64") or 0) for r in results)
    dom_dtype = ("float32" if it32 >= it64 else "float64") if a.dtype == "mixed" else a.dtype
    isz = 4 if dom_dtype == "float32" else 8
    kinds = prof[dom_dtype]
    dom = max(kinds, key=lambda k: kinds[k][0])
    avg_ms = kinds[dom][0] / kinds[dom][1]
    alg = _pass_bytes(dom, m2, isz, n, mp)
    achieved = alg / (avg_ms * 1e-3) / 1e9
    iter_ms = sum(v[0] for v in kinds.values()) / a.profile_iters
    iter_bytes = sum(_pass_bytes(k, m2, isz, n, mp) * v[1] for k, v in kinds.items()) / a.profile_iters
    passes = {KIND_NAMES[k]: dict(avg_us=round(1e3 * v[0] / v[1], 3), launches=v[1],
                                  gbs=round(_pass_bytes(k, m2, isz, n, mp) / (v[0] / v[1] * 1e-3) / 1e9, 1))
              for k, v in kinds.items()}
    # HBM bytes per launch of the timed kernel, from the committed rocprofv3 FETCH_SIZE + WRITE_SIZE
    # table of the same window plan (keyed by the kernel's template name); None when no table has it
    traffic, dom_kernel = None, _kernel_name(dom, dom_dtype, sess_plan, n)
    tf = os.path.join(ROOT, "profiles", "traffic_%s_n18.json" % dom_dtype)
    if os.path.exists(tf) and dom_kernel:
        try:
            table = json.load(open(tf))
            hits = [v for k, v in table.items() if _same_kernel(k, dom_kernel)]
            traffic = hits[0] if len(hits) == 1 else None
        except Exception:
            traffic = None
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, kernel=KIND_NAMES[dom],
                    kernel_template=dom_kernel, window_plan=list(sess_plan) if sess_plan else None,
                    working_set="%d float64 slabs of %.1f MB (%.0f MB) fit the 256 MB Infinity Cache: this fraction "
                                "is cache-served; config3's (n=22) is the HBM-bound one"
                                % (6, m2 * isz / 1e6, 6 * m2 * isz / 1e6),
                    dtype=dom_dtype, alg_bytes_per_launch=alg, avg_launch_us=round(avg_ms * 1e3, 3),
                    iteration=dict(us=round(iter_ms * 1e3, 2), alg_bytes=int(iter_bytes),
                                   gbs=round(iter_bytes / (iter_ms * 1e-3) / 1e9, 1)),
                    passes=passes)
    sess.close()

    # plain: the reference schedule in float64, same hooks and stops
    plain = None
    if a.plain_steps > 0:
        from solver import params
        h = Handle(n, rum._geom.menu, rum._geom.alternative.astype(np.int32), p[rum._rest], dtype="float64",
                   beta=params.bm_beta_default(mp), plain=True)
        st = dict(ub=math.inf, lb=0.0)

        def push(hh):
            u = hh.push()
            st["ub"] = min(st["ub"], u)
            return u

        def lower(hh):
            st["lb"] = max(st["lb"], hh.lb())
            return st["lb"]

        times, iters = [], []
        for i in range(a.plain_steps):
            st.update(ub=math.inf, lb=0.0)
            h.init_cold()
            tp = time.perf_counter()
            rp = solve_kadmm(h, max_iter=10 ** 7, check=50, push_fn=push, lb_fn=lower, ub_arm_immediately=True,
                             ub_running_min=True, gap_abs_tol=a.abs_tol, gap_rel_tol=0.0,
                             time_limit=a.max_time * 3)
            times.append(time.perf_counter() - tp)
            iters.append(rp["iters"])
            print("[bench] plain %d: %s, [%.10f, %.10f], %d it, %.3f s" % (i, rp["stop_reason"], rp["lb"], rp["ub"],
                                                                         rp["iters"], times[-1]),
                  file=sys.stderr, flush=True)
        h.close()
        plain = dict(value=round(len(times) / sum(times), 5), unit="certified solves/s", dtype="f64",
                     schedule="reference (plain kernels)", beta=params.bm_beta_default(mp),
                     iterations=iters, seconds=[round(t, 3) for t in times],
                     certified=bool(rp["stop_reason"] == "gap_abs"))

    cpu = None
    if not a.no_cpu:
        cpu = _cpu_baseline(n, rum, p, r_last, a)
    boot = _bootstrap_bench(a) if a.boot_draws > 0 else None
    dsr = None if a.no_dsr else _dsr_bench()
    cfg3 = None if a.no_config3 else _config3_bench(a)
    ai = None if a.no_ai else _ai_bench()
    slots = None if a.no_slots else _slots_bench()

    line = dict(metric=METRIC, value=round(value, 5), unit="certified solves/s", n_gpus=1, steps=a.steps,
                warmup=a.warmup, ms_per_step=round(1e3 * elapsed / a.steps, 2), higher_is_better=True,
                scaling="weak", vs_baseline=None, dtype={"mixed": "f32+f64", "float64": "f64", "float32": "f32"}[a.dtype],
                data="synthetic",
                config=dict(workload="certified BM-ADMM solve (abs gap 1e-5, rel 0), n=18, 1%% density "
                                     "(config2 seed 18001, %d cells)" % mp, n=n, cells=mp, abs_tol=a.abs_tol,
                            beta_rule=a.beta, beta=r_last.method_args["beta"],
                            beta_final=r_last.info.get("beta_final"), check=r_last.method_args["check"],
                            iterations=[r.iterations for r in results],
                            iters32=[r.info.get("iters32") for r in results],
                            iters64=[r.info.get("iters64") for r in results],
                            bracket=[r_last.lb, r_last.ub], certified=certified),
                roofline=roofline, cpu_baseline=cpu, plain=plain, bootstrap=boot, dsr_bootstrap=dsr,
                config3_iteration=cfg3, ai_stored_run=ai, bm_admm_slots=slots,
                tuned_vs_plain=round(value / plain["value"], 3) if plain and plain["value"] > 0 else None)
    print(json.dumps(line), flush=True)


def _bootstrap_bench(a):
    """Bootstrap draws/s on the n = 25 panel (config5: 301 cells of alternative 0, 520 respondents x
    15 menus, B = 10,000, multinomial_reject seed 0), flow_cg (the method the reference runs on
    such panels). p_hat is certified and the draws screened first (untimed); then the open draws
    are solved with the band's stops, warm from p_hat:
      plain  one Session.solve per draw with the float64 lower bound through a slab and the
             numpy/scipy master QP, as the reference's runner computes them (rum_solver.py:928-938,
             common/qp_native.py, bootstrap.py:308-335)
      tuned  the batched driver: every draw in flight priced by one mask sweep per round, the native
             master QP, the final float64 bound on the pricer's lattice
    Both legs run the same column generation per draw; the decisions (below / above / undecided)
    must agree."""
    import numpy as np
    import pandas as pd
    from solver import instances
    from solver.inference import cells_from_panel, hong_li
    from solver.inference.batched import BatchedCG
    from solver.rum_solver import RUM
    panel, n = instances.config5()
    panel = pd.DataFrame(panel)
    menu, alt, p = cells_from_panel(panel, n, alternatives=[0])
    rum = RUM(menu, alt, p, num_alternatives=n)
    boot = rum.bootstrap(panel=panel, B=10000, method="flow_cg", method_args={"pricing": "mask"})
    t0 = time.perf_counter()
    boot._start()
    t_start = time.perf_counter() - t0
    und = np.flatnonzero(boot._undecided())
    K = min(a.boot_draws, und.size)
    Kp = min(a.boot_plain_draws, K)
    idx = und[:K]
    tau = boot.tau
    print("[bench] bootstrap n=25: p_hat [%.8f, %.8f] certified and %d draws screened in %.2f s; %d open, "
          "timing %d (plain %d)" % (boot.L0, boot.U0, boot.n_screened, t_start, und.size, K, Kp),
          file=sys.stderr, flush=True)
    P = hong_li.perturb(boot.p_hat, boot.draws.rows(idx[:Kp]), boot.r_n, boot.eps_n)
    sess = boot.session
    sess._args["_lb_form"] = "slab"       # the reference's float64 slab bound ...
    sess._args["_qp"] = "numpy"           # ... and its numpy/scipy master QP
    plain = {}
    t0 = time.perf_counter()
    for j, b in enumerate(idx[:Kp]):
        r = sess.solve(P[j], abs_tol=boot.abs_tol, rel_tol=boot.rel_tol, max_time=boot.draw_time, warm=True,
                       below=tau[0], above=tau[1])
        plain[int(b)] = (r.lb, r.ub)
    t_plain = time.perf_counter() - t0
    sess._args["_lb_form"] = "mask"
    sess._args["_qp"] = "native"
    print("[bench] bootstrap plain: %d draws %.3f s" % (Kp, t_plain), file=sys.stderr, flush=True)
    P_of = lambda sel: hong_li.perturb(boot.p_hat, boot.draws.rows(sel), boot.r_n, boot.eps_n)  # noqa: E731
    bs = BatchedCG(boot, slots=a.boot_slots)
    t0 = time.perf_counter()
    tuned = {b: (L, U) for b, L, U, _f in bs.solve(idx, P_of, boot.abs_tol, boot.rel_tol, tau[0], tau[1],
                                                   boot.draw_time)}
    t_tuned = time.perf_counter() - t0
    print("[bench] bootstrap tuned: %d draws %.3f s" % (K, t_tuned), file=sys.stderr, flush=True)
    common = [int(b) for b in idx[:Kp]]
    d_plain = hong_li.decide([plain[b][0] for b in common], [plain[b][1] for b in common], *tau)
    d_tuned = hong_li.decide([tuned[b][0] for b in common], [tuned[b][1] for b in common], *tau)
    dec_all = hong_li.decide([tuned[int(b)][0] for b in idx], [tuned[int(b)][1] for b in idx], *tau)
    v_plain, v_tuned = Kp / t_plain, K / t_tuned
    st = bs.stats
    boot.close()
    return dict(workload="bootstrap draws of the n=25 panel (config5 seed 25001: 301 cells of alternative 0, "
                         "520 respondents x 15 menus), B=10,000, flow_cg mask pricing, abs 5e-4; open draws "
                         "after the screen, band stops, warm from p_hat",
                unit="draws/s", tuned=round(v_tuned, 3), plain=round(v_plain, 3),
                tuned_vs_plain=round(v_tuned / v_plain, 3), draws_tuned=K, draws_plain=Kp, open_draws=int(und.size),
                screened=int(boot.n_screened), decisions_identical=bool(np.array_equal(d_plain, d_tuned)),
                max_bracket_diff=float(max(max(abs(plain[b][0] - tuned[b][0]), abs(plain[b][1] - tuned[b][1]))
                                           for b in common)),
                decided=dict(below=int((dec_all == hong_li.BELOW).sum()), above=int((dec_all == hong_li.ABOVE).sum()),
                             undecided=int((dec_all == hong_li.UNDECIDED).sum())),
                projected_open_draws_s=dict(plain=round(und.size / v_plain, 1), tuned=round(und.size / v_tuned, 1)),
                slots=a.boot_slots, rounds=st["rounds"], price_s=round(st["price_s"], 3), host_s=round(st["host_s"], 3),
                start_s=round(t_start, 3),
                learned="not applied to flow_cg draws; the MLP warm-starts BM-ADMM slots (tests/test_gpu_learned.py)")


def _reference_instances(only=()):
    """The 41 instances whose certified optimum brackets the reference stores (data/opt_refs.json,
    subset in tests/golden; generated by solver/instances.py with the stored cd_hash fingerprints),
    smallest first. Each is solved cold through calculate_distance to rel 1e-3 with the method of
    its stored record (CB-GPU -> flow_cg, K-ADMM -> bm_admm), the record's stop rule; the time is
    the call's wall clock (the HIP context exists beforehand) beside the record's wall_s, and the
    bracket must overlap [lb, OPT]."""
    from solver import _lib, instances
    from solver.rum_solver import RUM
    with open(os.path.join(ROOT, "tests", "golden", "opt_refs_subset.json")) as f:
        refs = json.load(f)
    _lib.context(0)         # the HIP runtime starts before the first timed solve
    rows = []
    for key in sorted(refs, key=lambda k: refs[k]["mp"]):
        if only and key not in only:
            continue
        ref = refs[key]
        n, cat = int(key.split("_")[0][1:]), key.split("_", 1)[1]
        cd, mp = instances.get_instance(n, cat)
        method = "bm_admm" if ref["method"].startswith("K-ADMM") else "flow_cg"
        rum = RUM.from_rows(cd, n)
        t0 = time.perf_counter()
        r = rum.calculate_distance(abs_tol=0.0, rel_tol=1e-3, method=method, max_time=300)
        t = time.perf_counter() - t0
        ok = bool(r.exit_flag in (2, 3) and r.lb <= ref["opt"] + 1e-9 and r.ub >= ref["lb"] - 1e-9)
        tune_s = r.info.get("autotune_s")     # window-plan timing of the first bm_admm solve at this n
        rows.append(dict(instance=key, cells=mp, method=method, exit_flag=r.exit_flag, lb=r.lb, ub=r.ub,
                         seconds=round(t, 3), stored_seconds=round(ref["wall_s"], 3),
                         speedup=round(ref["wall_s"] / t, 2), certified_overlap=ok, iterations=r.iterations,
                         autotune_seconds=tune_s))
        print("[bench] %-11s %9d cells %-7s %7.2f s (stored %7.2f s) flag %d overlap %s%s"
              % (key, mp, method, t, ref["wall_s"], r.exit_flag, ok,
                 "" if tune_s is None else " (plan tuning %.2f s)" % tune_s), file=sys.stderr, flush=True)
    sp = [r["speedup"] for r in rows]
    return dict(workload="the reference's certified-optimum instances (n=18-20 and n25_sparse, 41), rel 1e-3, cold",
                instances=rows, all_certified_overlap=all(r["certified_overlap"] for r in rows),
                geomean_speedup=round(math.exp(sum(math.log(x) for x in sp) / len(sp)), 2),
                seconds=round(sum(r["seconds"] for r in rows), 1),
                stored_seconds=round(sum(r["stored_seconds"] for r in rows), 1))


def _slots_bench(B=1024, per_draw=96, train=128, test=128, slots=32):
    """Batched BM-ADMM bootstrap on the n = 16 alternative-0 panel (the n = 25 workload's generator at
    n = 16: 301 cells of alternative 0, 520 respondents x 15 menus, seed 16002; B = 1,024
    multinomial_reject draws, seed 0, abs 5e-4, float64 iterates and certificates). p_hat is
    certified and the draws screened (untimed); then the open draws are solved with the band's
    stops, warm from p_hat:
      per_draw  one Session.solve per draw, the reference's loop (bootstrap.py:308-335), on the
                first `per_draw` open draws
      queue     draw slots of one batch object with retirement and refill on the device
      host      the same slots with the host taking every decision after each check
      learned   the MLP warm start trained on the GPU on the certified states of the first `train`
                open draws, then the next `test` open draws solved by the queue with and without it
    Decisions must agree everywhere; queue and host must give the same brackets bit for bit."""
    import numpy as np
    import pandas as pd
    from solver import instances
    from solver.inference import cells_from_panel, hong_li
    from solver.inference.batched import BatchedSolver
    from solver.learned import WarmStart
    from solver.rum_solver import RUM
    panel, n = instances.config4_alt0()
    panel = pd.DataFrame(panel)
    menu, alt, p = cells_from_panel(panel, n, alternatives=[0])
    rum = RUM(menu, alt, p, num_alternatives=n)
    out = dict(workload="bootstrap draws of the n=16 alternative-0 panel (seed 16002: 301 cells, 520 respondents "
                        "x 15 menus), B=%d, bm_admm float64, abs 5e-4, band stops, warm from p_hat" % B,
               unit="draws/s", slots=slots)
    with rum.bootstrap(panel=panel, B=B, method="bm_admm", method_args={"dtype": "float64"}) as boot:
        t0 = time.perf_counter()
        boot._start()
        und = np.flatnonzero(boot._undecided())
        out.update(p_hat_s=round(time.perf_counter() - t0, 2), open_draws=int(und.size), screened=boot.n_screened)
        print("[bench] slots: p_hat in %.1f s, %d open draws" % (out["p_hat_s"], und.size), file=sys.stderr,
              flush=True)
        tau = boot.tau
        P_of = lambda sel: hong_li.perturb(boot.p_hat, boot.draws.rows(sel), boot.r_n, boot.eps_n)  # noqa: E731

        def run(idx, **kw):
            bs = BatchedSolver(boot, slots=slots, **kw)
            t = time.perf_counter()
            res = {b: (L, U, f) for b, L, U, f in bs.solve(idx, P_of, boot.abs_tol, 0.0, tau[0], tau[1], 120)}
            t = time.perf_counter() - t
            st, pairs = dict(bs.stats), bs.pairs
            bs.close()
            return res, t, st, pairs

        legs = {}
        for name, kw in (("queue", {}), ("host", {"queue": False})):
            res, t, st, _ = run(und, **kw)
            legs[name] = res
            out[name] = round(und.size / t, 2)
            out[name + "_slot_use"] = round(st["slot_use"], 3)
        same_bits = all(legs["queue"][b] == legs["host"][b] for b in und)
        # one Session.solve per draw on the first per_draw open draws
        sess = boot.session
        sample = und[:per_draw]
        t = time.perf_counter()
        one = {}
        for b, P in zip(sample, P_of(sample)):
            r = sess.solve(P, abs_tol=boot.abs_tol, rel_tol=0.0, max_time=120, warm=True, below=tau[0], above=tau[1])
            one[b] = (r.lb, r.ub)
        out["per_draw"] = round(sample.size / (time.perf_counter() - t), 2)
        dq = hong_li.decide([legs["queue"][b][0] for b in sample], [legs["queue"][b][1] for b in sample], *tau)
        d1 = hong_li.decide([one[b][0] for b in sample], [one[b][1] for b in sample], *tau)
        out.update(queue_vs_per_draw=round(out["queue"] / out["per_draw"], 2), queue_equals_host_bitwise=same_bits,
                   decisions_identical=bool(np.array_equal(dq, d1)))
        # learned warm start: trained on the first `train` open draws, tested on the next `test`
        if und.size >= train + test:
            tr, te = und[:train], und[train:train + test]
            _r, _t, _s, pairs = run(tr, collect="all")
            ws = WarmStart(rum.M, 512, seed=0)
            tt = time.perf_counter()
            hist = ws.train(pairs, epochs=300, lr=1e-3, batch=64)
            train_s = time.perf_counter() - tt
            rp, tp, sp, _ = run(te)
            rl, tl, sl, _ = run(te, model=ws)
            ws.close()
            dp = hong_li.decide([rp[b][0] for b in te], [rp[b][1] for b in te], *tau)
            dl = hong_li.decide([rl[b][0] for b in te], [rl[b][1] for b in te], *tau)
            ip, il = sp["busy_slot_iterations"] / te.size, sl["busy_slot_iterations"] / te.size
            out["learned"] = dict(train_draws=int(tr.size), test_draws=int(te.size), train_s=round(train_s, 3),
                                  loss=[round(hist[0], 5), round(hist[-1], 5)],
                                  plain=round(te.size / tp, 2), learned=round(te.size / tl, 2),
                                  iterations_per_draw=[round(ip, 1), round(il, 1)],
                                  iterations_saved=round(1 - il / ip, 4),
                                  decisions_identical=bool(np.array_equal(dp, dl)))
    print("[bench] slots: per draw %.1f, queue %.1f, host %.1f draws/s; learned %s" % (
        out["per_draw"], out["queue"], out["host"], out.get("learned")), file=sys.stderr, flush=True)
    return out


def _dsr_bench():
    """The whole bootstrap of the reference's stored DSR run (tests/golden: the n = 25 overload panel,
    alternative 0, 301 cells, 520 respondents; B = 10,000, multinomial_reject seed 0, abs 5e-4,
    the default method), end to end through rum.bootstrap(..., batch={"slots": 64}).p_value(). The
    stored record holds that run's per-draw brackets and its time on an RTX 5090; every draw's
    decision here is compared with the stored one."""
    import numpy as np
    import pandas as pd
    from solver.inference import cells_from_panel, hong_li
    from solver.rum_solver import RUM
    gold = os.path.join(ROOT, "tests", "golden")
    panel = pd.read_csv(os.path.join(gold, "dsr_overload_data_25_long.csv"))
    with open(os.path.join(gold, "dsr_alt0_record.json")) as f:
        rec = json.load(f)
    menu, alt, p = cells_from_panel(panel, 25, alternatives=[0])
    rum = RUM(menu, alt, p, num_alternatives=25)
    t0 = time.perf_counter()
    with rum.bootstrap(panel=panel, B=10000, batch={"slots": 64}) as boot:
        pv = boot.p_value()
        t = time.perf_counter() - t0
        dec = hong_li.decide(boot.L, boot.U, *boot.tau)
        L, U = boot.L.copy(), boot.U.copy()
        same_draws = boot.fingerprints["draws"] == rec["fingerprints"]["draws"]
        ci = boot.confidence_interval()
    dr = rec["draws"]
    rL, rU = np.asarray(dr["lb"], dtype=np.float64), np.asarray(dr["ub"], dtype=np.float64)
    rdec = hong_li.decide(rL, rU, *rec["tau"])
    both = np.isfinite(U) & np.isfinite(rU)
    overlap = int(((L[both] <= rU[both] + 1e-9) & (rL[both] <= U[both] + 1e-9)).sum())
    rpv = rec["results"]["p_value"]["0.05"]
    print("[bench] DSR bootstrap: p %.4f in %.2f s (stored run: p %.4f in %.1f s); %d draws decided differently"
          % (pv.p, t, rpv["p"], rpv["time"], int((dec != rdec).sum())), file=sys.stderr, flush=True)
    return dict(workload="the reference's stored DSR run: n=25 overload panel, alternative 0 (301 cells, 520 "
                         "respondents), B=10,000, multinomial_reject seed 0, abs 5e-4, whole p_value() end to end",
                seconds=round(t, 2), draws_per_s=round(10000 / t, 1), p=pv.p,
                below=pv.n_below, above=pv.n_above, undecided=pv.n_undecided, screened=pv.n_screened,
                stored_run=dict(seconds=round(rpv["time"], 1), gpu="RTX 5090", p=rpv["p"], below=rpv["n_below"],
                                above=rpv["n_above"], undecided=rpv["n_undecided"]),
                speedup_vs_stored_run=round(rpv["time"] / t, 2), draws_fingerprint_equal=bool(same_draws),
                decisions_differing=int((dec != rdec).sum()), brackets_overlapping=overlap,
                brackets_compared=int(both.sum()),
                confidence_interval=dict(lower=ci.lower, lower_bracket=list(ci.lower_bracket), seconds=round(ci.time, 2),
                                         draws_solved=ci.n_solved,
                                         published=dict(lower=-0.0476, seconds=237, draws_solved=592,
                                                        source="the reference's solver/DESIGN_NOTES.md:422")))


# the reference's published run of the same bootstrap on all 901 cells (solver/DESIGN_NOTES.md:423)
_DSR_ALL_PUBLISHED = dict(p=0.1392, below=8608, above=1392, undecided=0, seconds=1499)


def _dsr_all_alternatives():
    """The DSR bootstrap on every alternative's cells (n = 25, 901 cells, 520 respondents; B = 10,000,
    seed 0, abs 5e-4, the default method), p_value() end to end through the batched driver, beside
    the reference's published p, counts and time for the same run."""
    import pandas as pd
    from solver import _lib
    from solver.inference import cells_from_panel
    from solver.rum_solver import RUM
    panel = pd.read_csv(os.path.join(ROOT, "tests", "golden", "dsr_overload_data_25_long.csv"))
    menu, alt, p = cells_from_panel(panel, 25)
    rum = RUM(menu, alt, p, num_alternatives=25)
    _lib.context(0)
    t0 = time.perf_counter()
    with rum.bootstrap(panel=panel, B=10000, batch={"slots": 64}, verbose=True) as boot:
        pv = boot.p_value()
    t = time.perf_counter() - t0
    pub = _DSR_ALL_PUBLISHED
    return dict(workload="DSR bootstrap on all 901 cells (n=25, 520 respondents), B=10,000, seed 0, abs 5e-4",
                seconds=round(t, 1), p=pv.p, below=pv.n_below, above=pv.n_above, undecided=pv.n_undecided,
                screened=pv.n_screened, published=pub,
                same_as_published=bool((pv.p, pv.n_below, pv.n_above, pv.n_undecided)
                                       == (pub["p"], pub["below"], pub["above"], pub["undecided"])),
                speedup_vs_published=round(pub["seconds"] / t, 2))


def _kernel_name(kind, dtype, plan, n):
    """rocprof's name (prefix) of the kernel that runs pass `kind` (0 SY, 1 RHS, 2 D, 3 XK, 4 prox) for
    window plan (lo_bits, lo_rl, hi_rl): bm_pass<T, RL, KIND, B, C, NT> with the geometry fixed at
    compile time where the launcher has that variant, else bm_pass<T, RL, KIND, 0, -1, NT>; NT is the
    non-temporal variant of float64 objects from n = 20 on."""
    if not plan:
        return None
    T = "double" if dtype == "float64" else "float"
    nt = "true" if (T == "double" and n >= 20) else "false"
    lo, lrl, hrl = (int(v) for v in plan)
    if kind == 4:
        return "void bm_prox_iter<%s, %d>" % (T, n)
    if kind in (0, 2):
        K = {0: 1, 2: 4}[kind]
        if lrl == 3 and 8 <= lo <= 12:
            return "void bm_pass<%s, 3, %d, %d, 0, %s>(PassArgs<%s>)" % (T, K, lo, nt, T)
        return "void bm_pass<%s, %d, %d, 0, -1, %s>(PassArgs<%s>)" % (T, lrl, K, nt, T)
    K, B, CW = {1: 3, 3: 5}[kind], n - 1 - lo, (4 if T == "double" else 5)
    if hrl in (3, 4) and 4 <= B <= 8:
        return "void bm_pass<%s, %d, %d, %d, %d, %s>(PassArgs<%s>)" % (T, hrl, K, B, CW, nt, T)
    return "void bm_pass<%s, %d, %d, 0, -1, %s>(PassArgs<%s>)" % (T, hrl, K, nt, T)


def _same_kernel(a, b):
    """a starts with b once the NT argument (absent from tables taken before it existed; the non-NT
    kernels are the same code) is left out of both names"""
    strip = lambda k: k.replace(", false>", ">").replace(", true>", ">")      # noqa: E731
    return strip(a).startswith(strip(b))


def _config3_bench(a):
    """Fused BM-ADMM iterations at n = 22, 0.5 % (config3, 230,691 cells): one float64 slab is 369 MB,
    so the working set (6 slabs) streams from HBM; no cache serves it. Per dtype: the autotuned window
    plan, the time per iteration (HIP-graph replays, host clock around 100 iterations after a
    stream synchronization), the per-pass device times from start/stop events recorded in each
    pass's own dispatch (rumgpu_bm_profile), the algorithmic bytes per iteration (20 n 2^(n-1) w) and
    the bytes the passes actually move (28 slab transfers with three windows)."""
    import ctypes
    import numpy as np
    from solver import _lib, instances, machine, params
    from solver.bm_admm.autotune import tune
    from solver.bm_admm.engine import Handle
    from solver.rum_solver import RUM
    cd, n = instances.config3()
    rum = RUM.from_rows(cd, n)
    g = rum._geom
    p = rum.choice_probability[rum._rest]
    m2 = n << (n - 1)
    out = dict(workload="fused BM-ADMM iterations, n=22, 0.5%% density (config3 seed 22001, %d cells)" % g.M,
               fingerprint=instances.fingerprint(cd))
    for dt in ("float64", "float32"):
        isz = 8 if dt == "float64" else 4
        h = Handle(n, g.menu, g.alternative.astype(np.int32), p, dtype=dt, beta=params.bm_beta_cert(g.M))
        plan, _t = tune(h, device=machine.gpu_name())
        h.init_cold()
        if not a.no_graph:
            h.set_graph(10)
        h.run(20)
        h.phi()
        t0 = time.perf_counter()
        h.run(100)
        h.phi()
        ms_it = (time.perf_counter() - t0) * 10.0
        ms = (ctypes.c_double * 7)()
        cnt = (ctypes.c_int64 * 7)()
        iters = 50
        _lib.check(h.lib.rumgpu_bm_profile(h.h, iters, ms, cnt), "bm_profile")
        kinds = {k: (ms[k], cnt[k]) for k in range(7) if cnt[k]}
        dev_ms = sum(v[0] for v in kinds.values()) / iters
        moved = sum(_pass_bytes(k, m2, isz, n, g.M) * v[1] for k, v in kinds.items()) / iters
        alg = 20 * m2 * isz
        passes = {KIND_NAMES[k]: dict(avg_us=round(1e3 * v[0] / v[1], 1), launches_per_iteration=v[1] / iters,
                                      gbs=round(_pass_bytes(k, m2, isz, n, g.M) / (v[0] / v[1] * 1e-3) / 1e9, 1),
                                      frac=round(_pass_bytes(k, m2, isz, n, g.M) / (v[0] / v[1] * 1e-3) / 1e9
                                                 / HBM_PEAK_GBS, 3))
                  for k, v in kinds.items()}
        out[dt] = dict(window_plan=list(plan), windows=h.windows(), ms_per_iteration=round(ms_it, 4),
                       device_ms_per_iteration=round(dev_ms, 4), alg_bytes_per_iteration=alg,
                       alg_gbs=round(alg / (ms_it * 1e-3) / 1e9, 1),
                       alg_frac=round(alg / (ms_it * 1e-3) / 1e9 / HBM_PEAK_GBS, 3),
                       moved_bytes_per_iteration=int(moved),
                       moved_gbs=round(moved / (dev_ms * 1e-3) / 1e9, 1),
                       moved_frac=round(moved / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3), passes=passes)
        print("[bench] config3 %s: %.3f ms per iteration, %.0f GB/s algorithmic (%.2f of peak), %.0f GB/s moved"
              % (dt, ms_it, out[dt]["alg_gbs"], out[dt]["alg_frac"], out[dt]["moved_gbs"]), file=sys.stderr, flush=True)
        h.close()
    return out


def _ai_bench():
    """The reference's stored AI run (gpt-5.6-luna, n = 18, 73,492 cells, 100 replicates): its
    float64 BM-ADMM certificate of dist(p_hat) with the record's settings (beta 5, check 25, lower
    bound every 4th check, abs 5e-4), the same certificate with the adaptive penalty, and the whole
    bootstrap (B = 10,000 panel draws, p_value()), each beside the stored time on an RTX 5090."""
    import numpy as np
    import pandas as pd
    from solver import _lib
    from solver.inference import cells_from_panel
    from solver.rum_solver import RUM
    gold = os.path.join(ROOT, "tests", "golden")
    z = np.load(os.path.join(gold, "ai_luna_replies.npz"))
    table = pd.DataFrame({c: z[c].astype(np.int64) for c in ("respondent", "menu", "chosen")})
    rec = json.load(open(os.path.join(gold, "ai_luna_record.json")))
    menu, alt, p = cells_from_panel(table, 18)
    rum = RUM(menu, alt, p, num_alternatives=18)
    cert = rec["p_hat"]["certificate"]
    margs = {k: cert["method_args"][k] for k in ("beta", "check", "lb_every", "dtype")}
    _lib.context(0)
    out = dict(workload="AI gpt-5.6-luna: n=18, 73,492 cells, 100 replicates as respondents", stored_gpu="RTX 5090")
    for name, args in (("certificate_reference_settings", margs),
                       ("certificate_adaptive", {"beta": "adaptive", "dtype": "float64"})):
        rum.calculate_distance(abs_tol=cert["abs_tol"], rel_tol=0.0, method="bm_admm", method_args=args,
                               max_time=120)                       # window plan tuned, kernels loaded
        t0 = time.perf_counter()
        r = rum.calculate_distance(abs_tol=cert["abs_tol"], rel_tol=0.0, method="bm_admm", method_args=args,
                                   max_time=120)
        t = time.perf_counter() - t0
        out[name] = dict(seconds=round(t, 3), iterations=r.iterations, bracket=[r.lb, r.ub], exit_flag=r.exit_flag,
                         stored_seconds=round(cert["time"], 2), stored_iterations=cert["iterations"],
                         stored_bracket=[cert["lb"], cert["ub"]], speedup=round(cert["time"] / t, 2))
    s = rec["panel"]["settings"]
    t0 = time.perf_counter()
    with rum.bootstrap(panel=table, B=s["B"], seed=s["seed"], method=s["method"], method_args=s["method_args"],
                       abs_tol=s["abs_tol"], rel_tol=s["rel_tol"], refine=s["refine"]) as boot:
        pv = boot.p_value()
        tm = dict(boot.time)
    t = time.perf_counter() - t0
    want = rec["panel"]["p_value"]
    out["bootstrap"] = dict(seconds=round(t, 2), p=pv.p, below=pv.n_below, screened=pv.n_screened,
                            certificate_s=round(tm["certificate"], 2), screen_s=round(tm["screen"], 2),
                            stored_seconds=round(want["time"], 2), stored_p=want["p"],
                            stored_certificate_s=round(rec["panel"]["time"]["certificate"], 2),
                            stored_screen_s=round(rec["panel"]["time"]["screen"], 2),
                            speedup=round(want["time"] / t, 2))
    print("[bench] AI luna: certificate %.2f s (stored %.2f s), bootstrap %.2f s (stored %.2f s)"
          % (out["certificate_reference_settings"]["seconds"], cert["time"], t, want["time"]), file=sys.stderr,
          flush=True)
    return out


def _pricing_sweeps(key, sweeps, tile_bits=0, layout="slab"):
    """Pricing sweeps on one instance (random float32 scores) in one layout, for a kernel profile."""
    import numpy as np
    from solver import instances
    from solver.flow_cg.engine import Pricer, xmajor
    lay, sdt = {"slab": ("slab", "float32"), "mask": ("mask", "float32"), "stream_f32": ("stream", "float32"),
               "stream_u24": ("stream", "uint24"), "stream_u16": ("stream", "uint16")}[layout]
    if key in ("config2", "config3"):
        cd, n = getattr(instances, key)()
    else:
        n, cat = int(key.split("_")[0][1:]), key.split("_", 1)[1]
        cd, _mp = instances.get_instance(n, cat)
    menu, alt = cd[:, 0].astype(np.uint64), cd[:, 1].astype(np.int32)
    perm, masks_x, _a, offsets = xmajor(menu, alt, n)
    rng = np.random.default_rng(1)
    pr = Pricer(n, masks_x, offsets, layout=lay, dtype=sdt)
    if lay == "slab":
        pr.set_slab_tile(tile_bits)
    S = rng.standard_normal((2, menu.size)).astype(np.float32)
    pr.price(S[0])
    t0 = time.perf_counter()
    for k in range(sweeps):
        pr.price(S[k % 2])
    dt = (time.perf_counter() - t0) / sweeps
    pr.close()
    print("[bench] %s pricing %s (n=%d, %d rows%s): %.3f ms per sweep over %d sweeps"
          % (layout, key, n, menu.size, ", tile bits %s" % (tile_bits or "7") if lay == "slab" else "",
             dt * 1e3, sweeps), flush=True)


def _summarize_stats(db):
    """Per-kernel calls, total and mean time, share: the top_kernels view of a rocpd database."""
    import sqlite3
    con = sqlite3.connect(db)
    rows = con.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
    out = ["| kernel | calls | total ms | mean us | share % |", "|---|---|---|---|---|"]
    for name, calls, tot, avg, pct in rows[:24]:
        out.append("| `%s` | %d | %.1f | %.3f | %.1f |" % (name[:100], calls, tot / 1e3, avg, pct))
    return "\n".join(out)


def _counters(db, cname):
    import collections
    import sqlite3
    con = sqlite3.connect(db)
    d = collections.defaultdict(list)
    for k, v in con.execute("select kernel_name, value from counters_collection where counter_name = ?", (cname,)):
        d[k].append(float(v))
    return d


def _summarize_traffic(fetch_db, write_db, copy_bytes):
    """FETCH_SIZE and WRITE_SIZE (KiB per dispatch) per kernel, each scaled so that the ew_copy kernel
    (which reads and writes copy_bytes per dispatch) reads what it reads: HBM bytes per dispatch."""
    copy_bytes = int(copy_bytes)
    f, w = _counters(fetch_db, "FETCH_SIZE"), _counters(write_db, "WRITE_SIZE")
    copy = [k for k in f if k.startswith("void ew_copy<")]
    if not copy:
        raise SystemExit("no ew_copy dispatch to calibrate on")
    cf = copy_bytes / (1024.0 * sum(f[copy[0]]) / len(f[copy[0]]))
    cw = copy_bytes / (1024.0 * sum(w[copy[0]]) / len(w[copy[0]]))
    out = ["calibration on ew_copy (reads and writes %d bytes per dispatch): bytes = FETCH_SIZE KiB x 1024 x %.4f, "
           "WRITE_SIZE KiB x 1024 x %.4f" % (copy_bytes, cf, cw), "",
           "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM read MB | HBM write MB |",
           "|---|---|---|---|---|---|"]
    per_kernel = {}
    for k in sorted(f, key=lambda k: -sum(f[k]) / len(f[k])):
        fv = sum(f[k]) / len(f[k])
        wv = sum(w.get(k, [0.0])) / max(1, len(w.get(k, [])))
        rb, wb = fv * 1024 * cf, wv * 1024 * cw
        out.append("| `%s` | %d | %.1f | %.1f | %.2f | %.2f |" % (k[:90], len(f[k]), fv, wv, rb / 1e6, wb / 1e6))
        per_kernel[k] = int(rb + wb)
    return "\n".join(out), per_kernel


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_baseline(n, rum, p, r_last, a):
    """The oracle (oracle/rum_oracle.cpp, OpenMP over every host thread it may use) on a bounded
    sample of the same certified solve: cpu_iters float64 iterations of the reference schedule, one
    push and one lower bound, scaled to the GPU solve's iteration count, its checks (iterations /
    check) and its lower bounds (every lb_every-th check)."""
    import numpy as np
    from oracle import oracle as O
    g = rum._geom
    pr = p[rum._rest]
    st = O.AdmmState(n, g.menu, g.alternative.astype(np.int32), pr, r_last.method_args["beta"])
    t0 = time.perf_counter()
    st.run(a.cpu_iters)
    t_it = (time.perf_counter() - t0) / a.cpu_iters
    t1 = time.perf_counter()
    O.push_ub64(n, st.x2, st.obs, pr)
    t_push = time.perf_counter() - t1
    t1 = time.perf_counter()
    O.fw_lb64(n, st.obs, st.x2[st.obs], pr)
    t_lb = time.perf_counter() - t1
    iters = r_last.iterations
    check, lb_every = int(r_last.method_args["check"]), int(r_last.method_args["lb_every"])
    checks = max(1, iters // check)
    lbs = -(-checks // lb_every)
    t_solve = iters * t_it + checks * t_push + lbs * t_lb
    threads = O.threads()
    return dict(value=round(1.0 / t_solve, 8), unit="certified solves/s", cores=threads, kind="port",
                sample="%d float64 iterations + 1 push + 1 lower bound of the n=18 1%% instance (%.1f s on %d "
                       "threads of %s), scaled to %d iterations, %d checks (every %d) and %d lower bounds"
                       % (a.cpu_iters, a.cpu_iters * t_it + t_push + t_lb, threads, _cpu_model(), iters, checks,
                          check, lbs),
                seconds_per_iteration=round(t_it, 5), seconds_per_push=round(t_push, 5),
                seconds_per_lb=round(t_lb, 5), host_cpus=os.cpu_count(), cpu_model=_cpu_model())


if __name__ == "__main__":
    main()
