"""Solver constants and the one instance-dependent rule: the BM-ADMM penalty.

Values follow the reference (solver/params.py:152-186): alpha = 1.95, p = sqrt(2), check = 50,
beta = clip(10^-1.48 mp^0.508, 0.3, 100) with mp the number of observed (menu, alternative) cells.
"""
import math

BM_ALPHA = 1.95
BM_PRECOND_P = math.sqrt(2.0)
BM_CHECK_EVERY = 50
BM_BETA_COEF, BM_BETA_EXP = 10 ** -1.48, 0.508
BM_BETA_MIN, BM_BETA_MAX = 0.3, 100.0
CB_DEFAULTS = dict(rc_tol=0.0, max_iters=4000)


def bm_beta_default(mp):
    """Penalty from the number of observed cells (params.py:160-162 of the reference)."""
    return float(min(BM_BETA_MAX, max(BM_BETA_MIN, BM_BETA_COEF * float(mp) ** BM_BETA_EXP)))


# Certificate penalty. The rule above was fitted to the paper's race (ub within 1 % of a known
# optimum). Certifying to an absolute gap of 1e-5 prefers a smaller penalty: measured on this
# build (n = 14-20, 0.5-3 % density) the iteration count is U-shaped in beta with its minimum
# between 0.15 and 0.3 times the race rule, e.g. n = 18 at 1 % (23.6k cells): 82,700 iterations at
# the race rule, 15,300-17,400 at 0.15-0.2 of it. beta="cert" selects this rule.
BM_BETA_CERT_FACTOR = 0.2


def bm_beta_cert(mp):
    """Penalty for certified solves (beta="cert")."""
    return float(max(BM_BETA_MIN * BM_BETA_CERT_FACTOR, BM_BETA_CERT_FACTOR * bm_beta_default(mp)))


# Adaptive penalty (beta="adaptive"): residual balancing on the certified gap. Every `every_iters`
# iterations (a whole number of checks) the gap ub - lb splits at phi = ||x2[obs] - p||: ub - phi is what the push to a feasible
# point costs (primal infeasibility), phi - lb what the iterate still lacks in optimality (the
# dual side). When one part exceeds `ratio` times the other, beta is multiplied (infeasibility) or
# divided (optimality) by `factor`; the first reversal of direction takes a half step (sqrt of
# factor) and freezes beta, and at most `max_changes` changes are made. Start: `start` times the
# race rule. Measured on MI355X (abs gap 1e-5, float64, ten n = 18-20 instances of 0.1-3 %
# density): 37 % fewer iterations than beta="cert" in total, from -47 % (n18_dense, n18_m231,
# n19_m102) to +2 % (n18_m8); without the cap on changes n18_sparse drifted to 4.5 times the race
# rule and needed 12 times the iterations.
BM_BETA_ADAPT = dict(start=0.05, every_iters=500, ratio=10.0, factor=2.0, freeze_on_flip=True, max_changes=3)
# Certificate cadence of beta="adaptive" when method_args names no `check`: a push and a lower bound
# every 100 iterations instead of 50. The rule above acts on iterations, so its schedule is the
# same; over the same ten instances the iterations stay within 0.1 % and the time drops 5.5 %.
BM_BETA_ADAPT_CHECK = 100


def bm_beta_adaptive_start(mp):
    """Initial penalty of beta="adaptive"."""
    return float(max(BM_BETA_MIN * BM_BETA_ADAPT["start"], BM_BETA_ADAPT["start"] * bm_beta_default(mp)))

