"""Writes the DSR alternative-0 record fixture from the reference's own file.

    python tests/golden/make_dsr_alt0_record.py /path/to/reference

dsr_alt0_record.json  the reference's stored record of its alternative-0 bootstrap
                      (replication/empirical_dsr_n25/results/bootstrap_dsr_alt0.json: n = 25 overload
                      panel, B = 10,000, seed 0, abs 5e-4), copied byte for byte. It is the whole
                      record Bootstrap.record() wrote (version, settings, fingerprints, plan, the
                      p_hat certificate, tau, scaling, screen, panel summary, every draw's lb, ub,
                      tol_stop, tol_full and exit_flag, the p-value and interval results, times), so
                      it serves both as known answers (decisions, p-value, interval) and as a
                      checkpoint written by the reference for the resume test.
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    shutil.copyfile(os.path.join(ref, "replication", "empirical_dsr_n25", "results", "bootstrap_dsr_alt0.json"),
                    os.path.join(HERE, "dsr_alt0_record.json"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
