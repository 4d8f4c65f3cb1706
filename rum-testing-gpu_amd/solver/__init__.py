"""RUM projection solvers on AMD MI355X: BM-ADMM and Flow-CG as hand-written gfx950 HIP kernels.

A certified distance comes from the RUM class (solver/rum_solver.py):

    from solver.rum_solver import RUM
    res = RUM.from_rows(cd, n).calculate_distance()

cd: (mp, 3) rows (menu bitmask, chosen alternative 0-based, p_hat), one per observed cell. The
method names and aliases are the reference's (solver/__init__.py:187-233).
"""
from . import params  # noqa: F401

METHODS = ("bm_admm", "flow_cg")
ALIASES = {"cgbfqp": "flow_cg", "cgbf": "flow_cg", "cgip": "ip_cg", "cgip_heur": "ip_cg_heur",
           "cgipheuristic_rand": "ip_cg_heur"}
