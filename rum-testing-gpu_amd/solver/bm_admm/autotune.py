"""Window-plan autotuner of the fused BM-ADMM passes, per n and dtype.

A plan (lo_bits, lo_rl, hi_rl) splits the n-1 rest bits into the window at bit 0 (up to 13 bits) and
windows of at most 8 bits above it; each window is one fused pass per Kronecker transform, and the register
bits set how many rounds (LDS exchanges) a window takes. Every plan computes the same iteration up
to rounding order. The tuner times graph-replayed iterations of each candidate near the built-in
plan on the caller's own object, a short run each and a full run for the fastest few, and keeps
the fastest; results are cached in memory and in
~/.cache/rum-testing-gpu_amd/plans.json, keyed by device, n and dtype. The reference plans its
butterfly passes with a calibrated DP instead (common/butterfly.py:125-195).
"""
import json
import os
import time

CACHE = os.path.join(os.path.expanduser("~"), ".cache", "rum-testing-gpu_amd", "plans.json")
_MEM = {}


def default_lo(nb):
    """the built-in plan's window at bit 0 (csrc/bfly.hpp plan_windows)"""
    k = 1 if nb <= 12 else 1 + (nb - 12 + 7) // 8
    lo = (nb + k - 1) // k
    lo = max(lo, nb - 8 * (k - 1))
    return min(lo, 12)


def _windows(nb, lo):
    return 1 + -(-(nb - lo) // 8)


def candidates(nb):
    lo0 = default_lo(nb)
    los = sorted({lo for lo in (lo0 - 1, lo0, lo0 + 1, lo0 + 2) if 1 <= lo <= min(12, nb) and nb - lo <= 16})
    out = [(lo, lrl, hrl) for lo in los for lrl in (2, 3) for hrl in (3, 4)]
    # 4 register bits at bit 0 for windows of 10-12 bits (compiled with fixed geometry)
    out += [(lo, 4, hrl) for lo in los if 10 <= lo <= 12 for hrl in (3, 4)]
    # a 13-bit window at bit 0 (4 or 5 register bits) where it saves a window over 12 bits (nb = 21: two
    # windows instead of three); its float64 tile needs more than 64 KiB of LDS, which the library
    # accepts only where the device has it
    if nb > 13 and _windows(nb, 13) < _windows(nb, 12):
        out += [(13, lrl, hrl) for lrl in (4, 5) for hrl in (3, 4)]
    return out


def _key(device, n, dtype):
    return "%s/n%d/%s" % (device, n, dtype)


def _load():
    try:
        with open(CACHE) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _save(d):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        tmp = CACHE + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(d, fh, indent=1)
        os.replace(tmp, CACHE)
    except OSError:
        pass


def _time(h, plan, warm, iters):
    """microseconds per graph-replayed iteration of plan on h, None where h cannot take the plan"""
    try:
        h.set_plan(*plan)
    except ValueError:               # the plan exceeds the pass's thread or LDS limits
        return None
    h.run(warm)
    t0 = time.perf_counter()
    h.run(iters)
    return (time.perf_counter() - t0) / iters * 1e6


def tune(h, device="gpu", iters=200, warm=50, screen=(25, 50), keep=3):
    """The fastest plan for handle h (its n and dtype), applied to h; returns (plan, timings).
    Every candidate is screened on a short run (screen = warm, timed iterations); the `keep`
    fastest are timed again on the full run (warm, iters), which picks the plan."""
    key = _key(device, h.n, str(h.dtype))
    if key in _MEM:
        plan = tuple(_MEM[key]["plan"])
        h.set_plan(*plan)
        return plan, _MEM[key]["us"]
    disk = _load()
    if key in disk:
        _MEM[key] = disk[key]
        plan = tuple(disk[key]["plan"])
        h.set_plan(*plan)
        return plan, disk[key]["us"]
    graph = h.graph_len
    h.set_graph(screen[0])
    times = {}
    for plan in candidates(h.n - 1):
        us = _time(h, plan, *screen)
        if us is not None:
            times[plan] = us
    h.set_graph(warm)
    finalists = sorted(times, key=times.get)[:keep]
    for plan in finalists:
        times[plan] = _time(h, plan, warm, iters)
    best = min(finalists, key=times.get)
    h.set_plan(*best)
    h.set_graph(graph)
    rec = dict(plan=list(best), us={"%d,%d,%d" % k: round(v, 2) for k, v in times.items()})
    _MEM[key] = rec
    disk = _load()
    disk[key] = rec
    _save(disk)
    return best, rec["us"]
