"""Summaries of rocprofv3 output (the rocpd SQLite database it writes) as short text tables.

    python -m solver.rocpd_summary stats   <run_results.db>            per-kernel calls / mean / share
    python -m solver.rocpd_summary traffic <fetch.db> <write.db>       FETCH_SIZE / WRITE_SIZE per dispatch,
                                                                       calibrated on the ew_copy kernel
    (run from rum-testing-gpu_amd/, or with it on sys.path)

The traffic table calibrates both counters on the state-copy kernel of bench.py --traffic-run
(ew_copy: reads and writes exactly n 2^(n-1) values), so `bytes` is HBM-side traffic in bytes per
dispatch. With --json it also writes {pass kind: bytes per launch} for bench.py's roofline.traffic.
"""
import collections
import json
import sqlite3
import sys

# fused pass kernels of BM-ADMM by template arguments: bm_pass<T, RL, KIND[, B, C]>, KIND 1 SY, 3 RHS, 4 D, 5 XK
PASS_KIND = {1: 0, 3: 1, 4: 2, 5: 3}


def kind_of(name):
    if name.startswith("void bm_prox_iter<"):
        return 4
    if name.startswith("void bm_pass<"):
        k = int(name.split(",")[2].split(">")[0])
        return PASS_KIND.get(k)
    return None


def stats(db):
    con = sqlite3.connect(db)
    rows = con.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
    out = ["| kernel | calls | total ms | mean us | share % |", "|---|---|---|---|---|"]
    for name, calls, tot, avg, pct in rows[:20]:
        out.append("| `%s` | %d | %.1f | %.3f | %.1f |" % (name[:90], calls, tot / 1e3, avg, pct))
    return "\n".join(out)


def _counter(db, cname):
    con = sqlite3.connect(db)
    d = collections.defaultdict(list)
    for k, v in con.execute("select kernel_name, value from counters_collection where counter_name = ?", (cname,)):
        d[k].append(float(v))
    return d


def traffic(fetch_db, write_db, copy_bytes):
    f, w = _counter(fetch_db, "FETCH_SIZE"), _counter(write_db, "WRITE_SIZE")
    copy = [k for k in f if k.startswith("void ew_copy<")]
    if not copy:
        raise SystemExit("no ew_copy dispatch to calibrate on")
    cf = copy_bytes / (1024.0 * sum(f[copy[0]]) / len(f[copy[0]]))
    cw = copy_bytes / (1024.0 * sum(w[copy[0]]) / len(w[copy[0]]))
    out = ["calibration on ew_copy (reads and writes %d bytes per dispatch): bytes = FETCH_SIZE KiB x 1024 x %.4f, "
           "WRITE_SIZE KiB x 1024 x %.4f" % (copy_bytes, cf, cw), "",
           "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM read MB | HBM write MB |", "|---|---|---|---|---|---|"]
    per_kind = {}
    for k in sorted(f, key=lambda k: -sum(f[k]) / len(f[k])):
        fv = sum(f[k]) / len(f[k])
        wv = sum(w.get(k, [0.0])) / max(1, len(w.get(k, [])))
        rb, wb = fv * 1024 * cf, wv * 1024 * cw
        out.append("| `%s` | %d | %.1f | %.1f | %.2f | %.2f |" % (k[:80], len(f[k]), fv, wv, rb / 1e6, wb / 1e6))
        kd = kind_of(k)
        if kd is not None:
            per_kind[str(kd)] = int(rb + wb)
    return "\n".join(out), per_kind


def main(argv):
    if argv[0] == "stats":
        print(stats(argv[1]))
    elif argv[0] == "traffic":
        copy_bytes = int(argv[3]) if len(argv) > 3 and not argv[3].startswith("--") else 18874368
        text, per_kind = traffic(argv[1], argv[2], copy_bytes)
        print(text)
        if "--json" in argv:
            with open(argv[argv.index("--json") + 1], "w") as fh:
                json.dump(per_kind, fh, indent=1)
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
