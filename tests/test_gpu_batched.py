"""Draw slots of the batch object, through the C ABI, against single-draw objects.

Every kernel of an iteration works item by item (the butterflies within an item block, the prox
within a menu of one slot), and every certificate reduction of a slot runs the same blocks in the
same order, so slot d of a batch object must equal a single object on the same p and start bit for
bit: states, pushed upper bounds, lower bounds, phi and fitted points. The batched bootstrap must
reach the per-draw bootstrap's decision on every draw.
"""
import numpy as np
import pytest

from solver import _lib, instances
from solver.bm_admm.engine import Handle
from solver.inference import cells_from_panel
from solver.rum_solver import RUM

pytestmark = pytest.mark.gpu


def _medium(n, cells, rng):
    seen, om, ox = set(), [], []
    while len(om) < cells:
        m = int(rng.integers(0, 1 << n))
        if bin(m).count("1") < 2 or m in seen:
            continue
        seen.add(m)
        for a in range(n):
            if (m >> a) & 1:
                om.append(m)
                ox.append(a)
    menu, alt = np.array(om, dtype=np.uint64), np.array(ox, dtype=np.int32)
    p = rng.random(menu.size)
    _, inv = np.unique(menu, return_inverse=True)
    return menu, alt, p / np.bincount(inv, weights=p)[inv]


def _perturb(menu, p, rng, eps=0.05):
    q = np.clip(p + eps * rng.standard_normal(p.size), 1e-3, None)
    _, inv = np.unique(menu, return_inverse=True)
    return q / np.bincount(inv, weights=q)[inv]


def _same(a, b):
    assert np.array_equal(a, b), np.abs(np.asarray(a) - np.asarray(b)).max()


@pytest.mark.parametrize("n,dtype,graph", [(9, "float64", 0), (12, "float64", 5), (18, "float64", 5),
                                           (11, "float32", 4)])
def test_slots_equal_single_objects(n, dtype, graph):
    rng = np.random.default_rng(300 + n)
    menu, alt, p = _medium(n, min(6000, n << (n - 3)), rng)
    D = 3
    P = [_perturb(menu, p, rng) for _ in range(D)]
    beta = 0.8
    bat = Handle(n, menu, alt, p, dtype=dtype, beta=beta, slots=D)
    if graph:
        bat.set_graph(graph)
    for d in range(D):
        bat.slot_set_p(d, P[d])
        bat.slot_cold(d)
    singles = []
    for d in range(D):
        h = Handle(n, menu, alt, P[d], dtype=dtype, beta=beta)
        if graph:
            h.set_graph(graph)
        singles.append(h)
    # a warm source: a single object's state after 12 iterations on p, kept
    src = Handle(n, menu, alt, p, dtype="float64", beta=beta)
    src.run(12)
    src.keep()
    for rnd in range(2):
        bat.run(11)
        for h in singles:
            h.run(11)
        ub, lb, phi = bat.certify(want_lb=True)
        for d, h in enumerate(singles):
            u1, fit1 = h.push(want_fit=True)
            assert ub[d] == u1, (rnd, d, ub[d], u1)
            assert lb[d] == h.lb(), (rnd, d)
            assert phi[d] == h.phi(), (rnd, d)
            _same(bat.slot_fit(d), fit1)
            x, z = bat.slot_get_obs(d)
            _same(x, h.x_obs())
        if rnd == 0:
            # refill slot 1 with a new draw warm from the kept state; its single twin copies the
            # same state (src's live state equals what it kept)
            Pn = _perturb(menu, p, rng)
            bat.slot_set_p(1, Pn)
            bat.slot_load(1, src)
            singles[1].close()
            h = Handle(n, menu, alt, Pn, dtype=dtype, beta=beta)
            if graph:
                h.set_graph(graph)
            h.copy_state_from(src)
            singles[1] = h
    for h in singles + [bat, src]:
        h.close()


def test_slot_obs_corrections():
    n = 10
    rng = np.random.default_rng(7)
    menu, alt, p = _medium(n, 900, rng)
    bat = Handle(n, menu, alt, p, dtype="float64", beta=1.0, slots=2)
    bat.run(5)
    x0, z0 = bat.slot_get_obs(1)
    dx, dz = 1e-3 * rng.standard_normal(menu.size), 1e-3 * rng.standard_normal(menu.size)
    bat.slot_add_obs(1, dx, dz)
    x1, z1 = bat.slot_get_obs(1)
    _same(x1, x0 + dx)
    _same(z1, z0 + dz)
    xa, _ = bat.slot_get_obs(0)
    bat.slot_add_obs(0, None, dz)
    xb, zb = bat.slot_get_obs(0)
    _same(xa, xb)
    with pytest.raises(ValueError):
        bat.slot_set_p(2, p)
    bat.close()


def _small_panel(seed=1, n=8):
    """40 menus, 600 respondents x 8 menus from a 6-order mixture: p_hat near the polytope, so the
    draws' distances straddle tau (9 of the first 96 above it, by exact enumeration)"""
    rng = np.random.default_rng(seed)
    menus, seen = [], set()
    while len(menus) < 40:
        m = int(rng.integers(0, 1 << n))
        if bin(m).count("1") < 2 or m in seen:
            continue
        seen.add(m)
        menus.append(m)
    orders = [rng.permutation(n) for _ in range(6)]
    w = rng.dirichlet(np.ones(6))
    return instances._panel(rng, n, menus, 600, 8, orders, w), n


def test_batched_bootstrap_same_decisions():
    import pandas as pd
    panel, n = _small_panel()
    panel = pd.DataFrame(panel)
    menu, alt, p = cells_from_panel(panel, n)
    rum = RUM(menu, alt, p, num_alternatives=n)
    B = 96
    args = dict(method="bm_admm", method_args={"beta": "cert"}, abs_tol=1e-4, refine=1)
    with rum.bootstrap(panel=panel, B=B, **args) as one:
        pv1 = one.p_value()
    with rum.bootstrap(panel=panel, B=B, batch=dict(slots=16), **args) as bat:
        pv2 = bat.p_value()
        st = bat.batch_stats
    with rum.bootstrap(panel=panel, B=B, batch=dict(slots=16, queue=False), **args) as host:
        host.p_value()
    # retire and refill on the device or on the host: the same brackets bit for bit
    for a, b in ((bat.L, host.L), (bat.U, host.U), (bat.flag, host.flag), (bat.tol_stop, host.tol_stop)):
        assert np.array_equal(a, b), np.flatnonzero(a != b)
    print(pv1)
    print(pv2)
    print(st)
    d1, d2 = one._decisions(), bat._decisions()
    assert np.array_equal(d1, d2), np.flatnonzero(d1 != d2)
    assert pv1.p == pv2.p and pv1.n_below == pv2.n_below and pv1.n_above == pv2.n_above
    # both brackets of every solved draw hold the same distance
    solved = (one.flag != 127) & (bat.flag != 127)
    assert (bat.L[solved] <= one.U[solved] + 1e-9).all() and (one.L[solved] <= bat.U[solved] + 1e-9).all()
    assert pv1.n_solved >= 20 and st["draws"] >= pv2.n_solved


@pytest.mark.parametrize("n,D,alt0", [(9, 5, False), (14, 16, False), (18, 21, False), (12, 40, True),
                                       (17, 32, True)])
def test_batched_pricing_equals_single(n, D, alt0):
    """alt0: rows of alternative 0 only (the bootstrap panels' shape), where the batched sweep skips
    the subsets without a row-carrying member"""
    from solver.flow_cg.engine import Pricer, xmajor
    rng = np.random.default_rng(400 + n)
    menu, alt, p = _medium(n, min(8000, n << (n - 3)), rng)
    if alt0:
        keep = alt == 0
        menu, alt, p = menu[keep], alt[keep], p[keep]
    perm, masks_x, _alts, offsets = xmajor(menu, alt, n)
    pr = Pricer(n, masks_x, offsets, layout="mask")
    S = rng.standard_normal((D, menu.size)).astype(np.float32)
    vb, rb = pr.price_batch(S)
    for d in range(D):
        v1, r1, _m = pr.price(S[d])
        assert vb[d] == v1 and np.array_equal(rb[d], r1), d
    pr.close()


@pytest.mark.parametrize("n,a,D,quant", [(2, 1, 3, False), (5, 0, 7, True), (9, 4, 32, True), (9, 8, 40, False),
                                          (13, 0, 32, True), (14, 6, 5, False), (16, 15, 33, True),
                                          (20, 0, 32, False)])
def test_one_alternative_sweep_equals_layers(n, a, D, quant):
    """Rows of one alternative: the closure sweep (opt-in) returns the layered sweep's vmax and
    ranking bit for bit, single and batched, and the same float64 longest path; quantized scores
    (multiples of 1/4) make ties common, which exercises the lowest-member rule. Sizes stay at
    n <= 20: the closure sweep's first run inside the n = 25 bootstrap test ended with a GPU fault
    whose cause has not been shown, so it is not run above n = 20."""
    from oracle import oracle as orc
    from solver.flow_cg.engine import Pricer, xmajor
    rng = np.random.default_rng(900 + n)
    menu, alt, _p = _medium(n, min(6000, max(2, (n << n) >> 3)), rng)
    keep = alt == a
    menu, alt = menu[keep], alt[keep]
    perm, masks_x, _alts, offsets = xmajor(menu, alt, n)
    pr = Pricer(n, masks_x, offsets, layout="mask")
    assert pr.one_alternative() == -1
    pr.set_sweep("auto")
    assert pr.one_alternative() == a
    S = rng.standard_normal((D, menu.size)).astype(np.float32)
    if quant:
        S = (np.round(S * 4) / 4).astype(np.float32)
    r64 = rng.standard_normal(menu.size)
    vb, rb = pr.price_batch(S)
    v1 = [pr.price(S[d])[:2] for d in range(D)]
    l1 = pr.vmax64(r64)
    pr.set_sweep("layers")
    assert pr.one_alternative() == -1
    vb0, rb0 = pr.price_batch(S)
    l0 = pr.vmax64(r64)
    for d in range(D):
        v0, r0, _m = pr.price(S[d])
        assert vb[d] == v0 and np.array_equal(rb[d], r0), d
        assert v1[d][0] == v0 and np.array_equal(v1[d][1], r0), d
        assert vb0[d] == v0 and np.array_equal(rb0[d], r0), d
        if n <= 14 and d < 4:
            vo, ro, _V, _pred = orc.price_mask(n, masks_x, S[d], offsets)
            assert vo == v0 and np.array_equal(ro, r0), d
    assert l1 == l0
    pr.close()


@pytest.mark.parametrize("n,a,D,quant", [(2, 1, 3, False), (5, 0, 7, True), (9, 4, 32, True), (9, 8, 40, False),
                                          (13, 0, 33, True), (14, 6, 5, False), (16, 15, 32, True),
                                          (20, 0, 32, False), (22, 3, 35, True), (25, 0, 32, True)])
def test_window_sweep_equals_layers(n, a, D, quant):
    """Rows of one alternative: the window-pass sweep (V as the subset maximum of a's row sums,
    formed in passes of 8 bits) returns the layered sweep's vmax and ranking bit for bit, batched
    (32 draws per table, padded) and single, and the same float64 longest path, single and for
    up to 8 draws at once; quantized scores (multiples of 1/4) make ties common."""
    from oracle import oracle as orc
    from solver.flow_cg.engine import Pricer, xmajor
    rng = np.random.default_rng(950 + n)
    menu, alt, _p = _medium(n, min(6000, max(2, (n << n) >> 3)), rng)
    keep = alt == a
    menu, alt = menu[keep], alt[keep]
    perm, masks_x, _alts, offsets = xmajor(menu, alt, n)
    pr = Pricer(n, masks_x, offsets, layout="mask")
    pr.set_sweep("window")
    assert pr.one_alternative() == a
    S = rng.standard_normal((D, menu.size)).astype(np.float32)
    if quant:
        S = (np.round(S * 4) / 4).astype(np.float32)
    R64 = rng.standard_normal((5, menu.size))
    vb, rb = pr.price_batch(S)
    v1 = [pr.price(S[d])[:2] for d in range(min(D, 6))]
    l1 = pr.vmax64(R64[0])
    lb1 = pr.vmax64_batch(R64)
    pr.set_sweep("layers")
    assert pr.one_alternative() == -1
    vb0, rb0 = pr.price_batch(S)
    l0 = pr.vmax64(R64[0])
    lb0 = pr.vmax64_batch(R64)
    for d in range(D):
        assert vb[d] == vb0[d] and np.array_equal(rb[d], rb0[d]), d
        if d < len(v1):
            v0, r0, _m = pr.price(S[d])
            assert v1[d][0] == v0 and np.array_equal(v1[d][1], r0), d
        if n <= 14 and d < 3:
            vo, ro, _V, _pred = orc.price_mask(n, masks_x, S[d], offsets)
            assert vo == vb[d] and np.array_equal(ro, rb[d]), d
    assert l1 == l0 and np.array_equal(lb1, lb0)
    pr.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n,a,D,L", [(2, 1, 3, 1), (9, 4, 32, 8), (12, 0, 5, 3), (16, 15, 32, 2), (20, 0, 17, 8),
                                     (25, 0, 32, 5)])
def test_fused_float64_bounds_equal_separate_calls(n, a, D, L):
    """price_batch_lb (the float64 maxima of L draws computed in the window-pass sweep that prices D
    others, on the same subset tests) returns price_batch's vmax and rankings and vmax64_batch's
    values bit for bit, with tied and untied scores; on the layered sweep it is the two calls"""
    from solver.flow_cg.engine import Pricer, xmajor
    rng = np.random.default_rng(970 + n)
    menu, alt, _p = _medium(n, min(6000, max(2, (n << n) >> 3)), rng)
    keep = alt == a
    menu, alt = menu[keep], alt[keep]
    perm, masks_x, _alts, offsets = xmajor(menu, alt, n)
    pr = Pricer(n, masks_x, offsets, layout="mask")
    for mode in ("window", "layers"):
        pr.set_sweep(mode)
        for quant in (False, True):
            S = rng.standard_normal((D, menu.size)).astype(np.float32)
            R64 = rng.standard_normal((L, menu.size))
            if quant:
                S = (np.round(S * 4) / 4).astype(np.float32)
                R64 = np.round(R64 * 8) / 8
            vf, rf, lf = pr.price_batch_lb(S, R64)
            vb, rb = pr.price_batch(S)
            lb = pr.vmax64_batch(R64)
            assert np.array_equal(vf, vb) and np.array_equal(rf, rb) and np.array_equal(lf, lb), (mode, quant)
            assert np.array_equal(lf, [pr.vmax64(R64[i]) for i in range(L)]), (mode, quant)
    pr.close()


def test_one_alternative_sweep_refused_above_n20():
    """Above n = 20 the library refuses the closure sweep (status 4) and the layered sweep stays on;
    nothing is launched: the refusal happens on the host."""
    from solver.flow_cg.engine import Pricer, xmajor
    n = 21
    rng = np.random.default_rng(21)
    menu = np.array([int(m) | 1 for m in rng.integers(2, 1 << n, size=40)], dtype=np.uint64)
    alt = np.zeros(menu.size, dtype=np.int32)
    perm, masks_x, _alts, offsets = xmajor(menu, alt, n)
    pr = Pricer(n, masks_x, offsets, layout="mask")
    with pytest.raises(_lib.GPUError, match="not run above n = 20"):
        pr.set_sweep("auto")
    assert pr.one_alternative() == -1
    pr.close()


def test_batched_flow_cg_bootstrap_identical_brackets():
    import pandas as pd
    panel, n = _small_panel()
    panel = pd.DataFrame(panel)
    menu, alt, p = cells_from_panel(panel, n)
    rum = RUM(menu, alt, p, num_alternatives=n)
    B = 96
    args = dict(method="flow_cg", method_args={"pricing": "mask"}, abs_tol=1e-4, refine=1)
    with rum.bootstrap(panel=panel, B=B, **args) as one:
        pv1 = one.p_value()
    with rum.bootstrap(panel=panel, B=B, batch=dict(slots=24), **args) as bat:
        pv2 = bat.p_value()
        st = bat.batch_stats
    print(pv1)
    print(pv2)
    print(st)
    # the same runner per draw, the same columns: identical brackets and flags
    assert np.array_equal(one.L, bat.L) and np.array_equal(one.U, bat.U)
    assert np.array_equal(one.flag, bat.flag)
    assert pv1.p == pv2.p and pv1.n_solved >= 20


def test_batched_flow_cg_bootstrap_from_a_slab_plan():
    """A flow_cg plan on the slab layout (pricing="slab" here; the default plan takes it above 16
    rows per alternative) runs its batched draws on the mask layout: the session switches to mask,
    and the decisions equal the per-draw solves on the mask layout."""
    import pandas as pd
    panel, n = _small_panel()
    panel = pd.DataFrame(panel)
    menu, alt, p = cells_from_panel(panel, n)
    rum = RUM(menu, alt, p, num_alternatives=n)
    B = 64
    with rum.bootstrap(panel=panel, B=B, method="flow_cg", method_args={"pricing": "mask"}, abs_tol=1e-4,
                       refine=1) as one:
        pv1 = one.p_value()
    with rum.bootstrap(panel=panel, B=B, method="flow_cg", method_args={"pricing": "slab"}, abs_tol=1e-4,
                       refine=1, batch=dict(slots=16)) as bat:
        pv2 = bat.p_value()
        assert bat.session._args["pricing"] == "mask"
    print(pv1)
    print(pv2)
    assert (pv1.n_below, pv1.n_above, pv1.n_undecided) == (pv2.n_below, pv2.n_above, pv2.n_undecided)


def test_batched_confidence_interval_equals_per_draw():
    """Bootstrap.confidence_interval solves upper-tail draws to tolerance with no band stops
    (bootstrap.py:212-240); batched flow_cg gives the per-draw brackets, so the same interval"""
    import pandas as pd
    panel, n = _small_panel()
    panel = pd.DataFrame(panel)
    menu, alt, p = cells_from_panel(panel, n)
    rum = RUM(menu, alt, p, num_alternatives=n)
    args = dict(method="flow_cg", method_args={"pricing": "mask"}, abs_tol=1e-4, refine=1)
    with rum.bootstrap(panel=panel, B=96, **args) as one:
        ci1 = one.confidence_interval()
    with rum.bootstrap(panel=panel, B=96, batch=dict(slots=16), **args) as bat:
        ci2 = bat.confidence_interval()
    print(ci1)
    print(ci2)
    assert ci1.n_solved >= 1
    assert ci1.lower == ci2.lower and ci1.lower_bracket == ci2.lower_bracket and ci1.n_solved == ci2.n_solved


def test_config4_alt0_slots_reach_per_draw_decisions():
    """Batched bootstrap at n = 16: the panel generator of the n = 25 workload at n = 16 (301 cells of
    alternative 0, 520 respondents x 15 menus, seed 16002; config4 as specified is decided by the
    screen alone, see the module notes). p_hat is certified once; its open draws are then solved
    twice from the same screened brackets: one float64 Session.solve per draw (the reference's
    loop) and the draw slots of one float64 batch object (graph-replayed iterations, retire and
    refill on the device, float64 certificates per slot). Every draw's decision must agree, and
    every pair of brackets must overlap. A third pass with the host taking the retire and refill
    decisions must give the device queue's brackets bit for bit."""
    import time
    import pandas as pd
    from solver.inference.bootstrap import _Batched, _PerDraw
    panel, n = instances.config4_alt0()
    panel = pd.DataFrame(panel)
    menu, alt, p = cells_from_panel(panel, n, alternatives=[0])
    rum = RUM(menu, alt, p, num_alternatives=n)
    with rum.bootstrap(panel=panel, B=210, method="bm_admm", method_args={"dtype": "float64"}) as boot:
        t0 = time.perf_counter()
        boot._start()
        open_ = np.flatnonzero(boot._undecided())
        print("[case] config4_alt0: p_hat [%.8f, %.8f] in %.1f s, %d screened, %d open"
              % (boot.L0, boot.U0, time.perf_counter() - t0, boot.n_screened, open_.size), flush=True)
        assert open_.size >= 64
        screened = (boot.L.copy(), boot.U.copy())
        t0 = time.perf_counter()
        _PerDraw(boot).run(open_, boot.abs_tol, boot.rel_tol, stops=True, until=None)
        t_one = time.perf_counter() - t0
        d_one, L1, U1 = boot._decisions().copy(), boot.L.copy(), boot.U.copy()
        boot.L[:], boot.U[:] = screened
        runner = _Batched(boot, {"slots": 32})
        t0 = time.perf_counter()
        runner.run(open_, boot.abs_tol, boot.rel_tol, stops=True, until=None)
        t_bat = time.perf_counter() - t0
        runner.close()
        d_bat = boot._decisions()
        print("per draw %.1f s, slots %.1f s for %d draws; %s" % (t_one, t_bat, open_.size, boot.batch_stats))
        assert np.array_equal(d_one, d_bat), np.flatnonzero(d_one != d_bat)
        assert (boot.L[open_] <= U1[open_] + 1e-9).all() and (L1[open_] <= boot.U[open_] + 1e-9).all()
        # the slots retired and refilled by the host after every check: the same draws retire at the
        # same checks, so every bracket, flag and tolerance equals the device queue's bit for bit
        queue = (boot.L.copy(), boot.U.copy(), boot.flag.copy())
        boot.L[:], boot.U[:] = screened
        runner = _Batched(boot, {"slots": 32, "queue": False})
        t0 = time.perf_counter()
        runner.run(open_, boot.abs_tol, boot.rel_tol, stops=True, until=None)
        t_host = time.perf_counter() - t0
        runner.close()
        print("host-driven slots %.1f s; %s" % (t_host, boot.batch_stats))
        for a, b in zip(queue, (boot.L, boot.U, boot.flag)):
            assert np.array_equal(a, b), np.flatnonzero(a != b)
