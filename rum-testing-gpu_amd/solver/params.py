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
