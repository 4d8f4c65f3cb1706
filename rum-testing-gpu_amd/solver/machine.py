"""What this computer offers the solvers: the AMD GPU (through librumgpu), CPU and memory.

The reference detects an NVIDIA card with nvidia-smi and CuPy (solver/machine.py:25-61) and keeps
a profile dict; this module keeps the same profile keys (use_gpu, gpu{name, memory_total,
supported, reason}, cpu_count, ram_bytes, gurobi) and fills them from HIP through the C ABI.
Gurobi is not part of this package (gurobi = None), so the CPU-only methods are unavailable.
"""
import os
import platform

from . import _lib

_PROFILE = None


def _ram_bytes():
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal:"):
                    return int(line.split()[1]) * 1024
    except OSError:
        pass
    return None


def _gpu():
    try:
        if _lib.device_count() < 1:
            return None
        pr = _lib.device_props(0)
    except Exception as e:                    # noqa: BLE001 -- reported in the profile
        return dict(name=None, memory_total=0, supported=False, reason="librumgpu: %s" % e)
    ok = pr["arch"].startswith("gfx950")
    reason = ("%s (%s), %.0f GB, %d CUs" % (pr["name"], pr["arch"], pr["memory_total"] / 2 ** 30, pr["cu_count"])
              if ok else "librumgpu is built for gfx950; this device is %s" % pr["arch"])
    return dict(name=pr["name"], arch=pr["arch"], memory_total=pr["memory_total"], cu_count=pr["cu_count"],
                supported=ok, reason=reason)


def detect(use_gpu=True):
    """A fresh profile of this machine."""
    return dict(use_gpu=bool(use_gpu), gpu=_gpu(), cpu_count=os.cpu_count(), ram_bytes=_ram_bytes(),
                cpu=platform.processor() or platform.machine(), gurobi=None)


def load():
    """The profile of this process (detected once)."""
    global _PROFILE
    if _PROFILE is None:
        _PROFILE = detect()
    return _PROFILE


def setup(use_gpu=True):
    """Detect again (optionally forbidding the GPU) and make it this process's profile."""
    global _PROFILE
    _PROFILE = detect(use_gpu)
    return _PROFILE


def can_use_gpu(prof):
    g = prof.get("gpu")
    return bool(prof.get("use_gpu", True) and g and g.get("supported"))


def can_use_gurobi(prof):
    return False


def gpu_free_bytes(device=0):
    """Free device memory now (hipMemGetInfo)."""
    return _lib.mem_info(device)[0]


def gpu_name(device=0):
    """'<name> <arch>' of the device (the autotuner's cache key), or 'gpu' when unknown"""
    try:
        pr = _lib.device_props(device)
        return "%s %s" % (pr["name"], pr["arch"])
    except Exception:                         # noqa: BLE001
        return "gpu"
