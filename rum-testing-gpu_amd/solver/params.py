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
