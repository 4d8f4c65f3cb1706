"""Writes race_records.json: the reference's stored race records that its own GPU test replays
(tests/test_gpu.py:52-57 there) and the n18_d1pct rows of its results table, reduced to what a
replay compares: the instance, the method, its settings, the stop, the iteration or column counts,
the upper bound and the wall time on the reference's GPU.

    python tests/golden/make_race_records.py /path/to/reference
"""
import csv
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
RECORDS = ("replication/density_grid/results/cgbfqp_race_n18_sparse.json",
           "replication/density_grid/results/bm_admm_race_n18_m771.json",
           "replication/feasibility_in_n/results/race_cgbfqp_n25_sparse.json",
           "replication/feasibility_in_density/results/race_bm_admm_n20_d1pct.json")
KEEP = ("key", "n", "mp", "solver", "criterion", "opt", "ub", "iters", "iters32", "iters64", "stop", "mode",
        "beta", "wall_s", "layers_mode", "race_floor_applied", "gpu")


def main(ref):
    out = []
    for rel in RECORDS:
        rec = json.load(open(os.path.join(ref, rel)))
        row = {k: rec[k] for k in KEEP if k in rec}
        row["source"] = rel
        out.append(row)
    table = os.path.join(ref, "final_results", "Numerical1", "results_n18.csv")
    for i, r in enumerate(csv.DictReader(open(table)), start=2):
        if r["instance"] == "n18_d1pct" and r["solver"] in ("BM-ADMM", "CGBFQP-GPU"):
            out.append(dict(key="n18_d1pct", n=18, mp=int(r["m_p"]), solver="bm_admm" if r["solver"] == "BM-ADMM"
                            else "cgbfqp", criterion="race", opt=float(r["opt"]), ub=float(r["ub_at_stop"]),
                            iters=int(r["count"]), stop=r["stop"], wall_s=float(r["wall_s"]),
                            pricing=r["pricing_backend"] or None,
                            source="final_results/Numerical1/results_n18.csv:%d" % i))
    with open(os.path.join(HERE, "race_records.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
