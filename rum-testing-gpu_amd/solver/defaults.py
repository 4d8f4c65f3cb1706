"""Default settings of the RUM class (rum_solver.py); override per call.

The names and values are the reference's (solver/defaults.py:19-108), so a user's method_args
keep their meaning. A call stops once UB - LB <= abs_tol or (UB - LB)/LB <= rel_tol, or when
max_time runs out (setup included). method_args may only use the names listed for the method.

Storage: bm_admm dtype "auto" is float64; "float32" runs float32 iterates with float64
certificates, and "mixed" runs the float32 sprint then the float64 finish from the same state
(the tuned path; the certificate is float64 either way). flow_cg dtype float32 / uint24 / uint16
is the pricing storage. Bounds are float64 wherever a float64 form fits the GPU.
"""
import math

ABS_TOL = 1e-5        # absolute gap on the distance
REL_TOL = 1e-4        # relative gap (UB - LB) / LB
MAX_TIME = 3600.0     # hard stop in seconds, setup included
SEED = 0              # seed of the random first column (flow_cg)
OPT_REL_TOL = 1e-2    # race: stop at UB <= max((1 + OPT_REL_TOL) opt, OPT_FLOOR)
OPT_FLOOR = 0.0

METHOD_ARGS = {
    "bm_admm": dict(
        beta=None,              # penalty; None = 10^-1.48 mp^0.508 clipped to [0.3, 100] (params.bm_beta_default);
                                # "cert" = 0.2 of it (params.bm_beta_cert), tuned for certificates;
                                # "adaptive" = balanced on the certified gap (params.BM_BETA_ADAPT)
        alpha=1.95,             # over-relaxation
        precond_p=math.sqrt(2.0),
        check=50,               # iterations between upper bounds (feasibility push)
        lb_every=1,             # upper-bound checks between lower bounds
        dtype="auto",           # "auto" = float64; "float32"; "mixed" = float32 sprint, float64 finish (the
                                # float32 leg helps only where float32's Mobius rounding is small, n <~ 14)
        use_graph="auto",       # replay `check` iterations as one HIP graph; auto = n <= 22
        max_iter=10 ** 7,
    ),
    "flow_cg": dict(
        dtype="auto",           # pricing values: "float32", "uint24" or "uint16"; auto = the first that fits
        pricing="auto",         # "mask" (one value per subset), "slab" (butterfly arc weights) or "stream"
                                # (two live layers); old names stream_f32/_u24/_u16 work; auto = slab for many
                                # rows per alternative (AUTO["cg_slab_rows_per_alt"]), else mask where it fits
        layers="auto",          # accepted for compatibility: the engine always stores the layer lists
        max_iters=4000,         # columns; also sizes the host Gram matrix
    ),
}

# Automatic choice when method=None (rum_solver.choose_method): bm_admm from this density at this n,
# flow_cg below it and for n <= 10 (the reference's measured crossovers, defaults.py:82-94).
AUTO = dict(
    bm_admm_min_density={11: 0.02, 14: 0.01, 18: 0.005},
    exact_qp_max_n=10,
    gpu_memory_fraction=0.9,    # share of free GPU memory a method may plan to use
    cg_slab_rows_per_alt=16,    # flow_cg pricing "auto": the slab form above this many rows per alternative
                                # (n = 16: 0.19 vs 0.50 ms; n = 22 at 0.5 %: 0.78 ms vs 1.06 s for stream u24)
    bm_autotune=True,           # time the fused passes' window plans once per (device, n, dtype)
    bm_autotune_min_n=12,       # (solver/bm_admm/autotune.py; cached in ~/.cache/rum-testing-gpu_amd)
)

# rum.bootstrap(...): Hong and Li's p-value and confidence interval (solver/inference).
BOOTSTRAP = dict(
    B=10000,
    weights="multinomial_reject",
    seed=0,
    abs_tol=5e-4,
    rel_tol=0.0,
    refine=2,
    max_time=3600.0,
    draw_time=120.0,
    checkpoint_every=60.0,
)
