"""Hong and Li's numerical delta method on the RUM distance, from certified brackets.

The formulas of the reference (solver/inference/hong_li.py:1-106):

    r_n = sqrt(n), eps_n = n^(-1/3), s = 1 + eps_n r_n     (defaults; both can be set)
    P_b = p_hat + eps_n r_n (p_hat_b - p_hat)              the b-th perturbed point
    p   = #{b : dist(P_b) >= tau} / B,  tau = s dist(p_hat)
    CI  = [dist(p_hat) - q / r_n, inf),  q = (1 - alpha)-quantile of (dist(P_b) - dist(p_hat)) / eps_n
    two-sided: [max(0, dist(p_hat) - c), dist(p_hat) + c],  c = (1 - alpha)-quantile of ||p_hat_b - p_hat||

Every distance is a bracket ([L0, U0] for p_hat, [L_b, U_b] per draw); a draw is below if
U_b < tau_lo = s L0 (counts 0), above if L_b >= tau_hi = s U0 (counts 1), else undecided (counts 1),
so p is an upper bound on the exact p-value.
"""
import math

import numpy as np

BELOW, ABOVE, UNDECIDED = 0, 1, 2
ZERO = 1e-9


def scaling(n, r_n=None, eps_n=None):
    if int(n) != n or n < 1:
        raise ValueError("n, the sample size, must be a positive integer; got %r" % (n,))
    n = int(n)
    r_n = math.sqrt(n) if r_n is None else float(r_n)
    eps_n = n ** (-1.0 / 3.0) if eps_n is None else float(eps_n)
    if not (math.isfinite(r_n) and r_n > 0 and math.isfinite(eps_n) and eps_n > 0):
        raise ValueError("r_n and eps_n must be finite and > 0")
    return r_n, eps_n, 1.0 + eps_n * r_n


def perturb(p_hat, p_b, r_n, eps_n):
    return p_hat + (eps_n * r_n) * (p_b - p_hat)


def check_alpha(alpha):
    alpha = float(alpha)
    if not 0.0 < alpha < 1.0:
        raise ValueError("alpha must lie in (0, 1); got %r" % alpha)
    return alpha


def decide(L, U, tau_lo, tau_hi):
    L, U = np.asarray(L, dtype=np.float64), np.asarray(U, dtype=np.float64)
    d = np.full(U.shape, UNDECIDED, dtype=np.int8)
    d[L >= tau_hi] = ABOVE
    d[U < tau_lo] = BELOW
    return d


def p_bracket(decision):
    decision = np.asarray(decision)
    above = int((decision == ABOVE).sum())
    undecided = int((decision == UNDECIDED).sum())
    return above / decision.size, (above + undecided) / decision.size


def locked(n_above, n_open, B, alpha):
    return (n_above + n_open) / B <= alpha or n_above / B > alpha


def ci_lower(L0, U0, L, U, alpha, r_n, eps_n):
    q_hi = float(np.quantile((np.asarray(U, dtype=np.float64) - L0) / eps_n, 1.0 - alpha))
    q_lo = float(np.quantile((np.asarray(L, dtype=np.float64) - U0) / eps_n, 1.0 - alpha))
    return L0 - q_hi / r_n, U0 - q_lo / r_n, (q_lo, q_hi)


def distance_interval(L0, U0, dev, alpha):
    c = float(np.quantile(np.asarray(dev, dtype=np.float64), 1.0 - alpha))
    return max(0.0, L0 - c), U0 + c, c


def quantile_band(L, U, alpha):
    return (float(np.quantile(np.asarray(L, dtype=np.float64), 1.0 - alpha)),
            float(np.quantile(np.asarray(U, dtype=np.float64), 1.0 - alpha)))
