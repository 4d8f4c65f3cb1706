"""Writes the AI (gpt-5.6-luna, n = 18) fixtures from the reference's own files.

    python tests/golden/make_ai_luna_fixture.py /path/to/reference

ai_luna_replies.npz   the reply table final_data/Application_AI/raw_data/gpt5.6-luna.npz as the
                      reference ships it (respondent int16, menu int32, chosen int8, dropped), copied
                      byte for byte: 880,700 replies, 100 per menu, 8,807 menus
ai_luna_record.json   what the reference's stored bootstrap records hold for it
                      (replication/empirical_ai_n18/results/bootstrap_ai_luna.json, panel draws, and
                      bootstrap_ai_luna_replies.json, draws menu by menu): fingerprints, the certified
                      dist(p_hat) bracket with its certificate's settings, iterations and time, tau,
                      the screen's count, the p-value and the two-sided distance interval
"""
import json
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    shutil.copyfile(os.path.join(ref, "final_data", "Application_AI", "raw_data", "gpt5.6-luna.npz"),
                    os.path.join(HERE, "ai_luna_replies.npz"))
    res = os.path.join(ref, "replication", "empirical_ai_n18", "results")
    panel = json.load(open(os.path.join(res, "bootstrap_ai_luna.json")))
    replies = json.load(open(os.path.join(res, "bootstrap_ai_luna_replies.json")))
    cert = panel["p_hat"]["certificates"][0]
    out = dict(
        panel=dict(settings=panel["settings"], fingerprints=panel["fingerprints"], tau=panel["tau"],
                   screen=panel["screen"], p_value=panel["results"]["p_value"]["0.05"], time=panel["time"],
                   panel=panel["panel"]),
        replies=dict(settings=replies["settings"], fingerprints=replies["fingerprints"],
                     distance_interval=replies["results"]["distance_interval"]["0.05"],
                     p_value=replies["results"]["p_value"]["0.05"]),
        p_hat=dict(lb=panel["p_hat"]["lb"], ub=panel["p_hat"]["ub"],
                   certificate={k: cert[k] for k in ("abs_tol", "rel_tol", "lb", "ub", "exit_flag", "iterations",
                                                     "time", "method", "method_args", "machine")}),
    )
    with open(os.path.join(HERE, "ai_luna_record.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
