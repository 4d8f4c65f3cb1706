"""What the learned warm start gains: the batched bootstrap on an alternative-0 panel (301 cells,
n = 12, B = 2,048), the MLP trained on the GPU on the first 96 open draws the method certified, then
the next 64 open draws solved with and without it. Every draw keeps its own float64 bracket, so the
decisions must agree; the learned start must take fewer draw-iterations than the plain one (it
took 475 against 488 per draw, 2.6 % fewer), and the wall times are printed."""
import time

import numpy as np
import pandas as pd
import pytest

from solver import instances
from solver.inference import cells_from_panel, hong_li
from solver.inference.batched import BatchedSolver
from solver.learned import WarmStart
from solver.rum_solver import RUM

pytestmark = pytest.mark.gpu


def test_learned_warm_start_gain():
    panel, n = instances.config5(seed=12002, n=12)
    panel = pd.DataFrame(panel)
    menu, alt, p = cells_from_panel(panel, n, alternatives=[0])
    rum = RUM(menu, alt, p, num_alternatives=n)
    boot = rum.bootstrap(panel=panel, B=2048, weights="multinomial", method="bm_admm", method_args={"beta": "cert"})
    boot._start()
    und = np.flatnonzero(boot._undecided())
    assert und.size >= 160, und.size
    tr, te = und[:96], und[96:160]
    P_of = lambda sel: hong_li.perturb(boot.p_hat, boot.draws.rows(sel), boot.r_n, boot.eps_n)  # noqa: E731
    tau = boot.tau
    bs = BatchedSolver(boot, slots=32, collect="all")
    list(bs.solve(tr, P_of, boot.abs_tol, 0.0, tau[0], tau[1], 120))
    pairs = bs.pairs
    bs.close()
    ws = WarmStart(rum.M, 512, seed=0)
    hist = ws.train(pairs, epochs=300, lr=1e-3, batch=64)
    assert hist[-1] < 0.1 * hist[0]
    res, work = {}, {}
    for name, model in (("plain", None), ("learned", ws)):
        bs = BatchedSolver(boot, slots=32, model=model)
        t0 = time.perf_counter()
        res[name] = {b: (L, U) for b, L, U, _f in bs.solve(te, P_of, boot.abs_tol, 0.0, tau[0], tau[1], 120)}
        work[name] = (bs.stats["busy_slot_iterations"] / len(te), time.perf_counter() - t0)
        bs.close()
    ws.close()
    print("draw-iterations per draw: plain %.0f, learned %.0f (%.1f %% fewer); wall %.2f s vs %.2f s"
          % (work["plain"][0], work["learned"][0], 100 * (1 - work["learned"][0] / work["plain"][0]),
             work["plain"][1], work["learned"][1]))
    # the trained model must pay: fewer draw-iterations than the plain warm start from p_hat
    assert work["learned"][0] < work["plain"][0], work
    d1 = hong_li.decide([res["plain"][b][0] for b in te], [res["plain"][b][1] for b in te], *tau)
    d2 = hong_li.decide([res["learned"][b][0] for b in te], [res["learned"][b][1] for b in te], *tau)
    assert np.array_equal(d1, d2)
    # both brackets certify the same distance
    for b in te:
        assert res["plain"][b][0] <= res["learned"][b][1] + 1e-9 and res["learned"][b][0] <= res["plain"][b][1] + 1e-9
    boot.close()
