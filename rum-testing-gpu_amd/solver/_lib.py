"""ctypes binding of librumgpu.so (the HIP engines) — the only way the Python host reaches the GPU.

The library is built in-tree next to this package (rum-testing-gpu_amd/librumgpu.so). Loading it
does not touch the GPU; the first context does. If the library is missing the engines raise at
once: there is no CPU fallback.
"""
import ctypes
import os
import threading

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "librumgpu.so")

OK, EINVAL, ENOMEM, EHIP, EUNSUP = 0, 1, 2, 3, 4

# every symbol include/rumgpu.h declares: (restype, argtypes)
_P, _I, _I64, _D, _F, _SZ = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_float,
                             ctypes.c_size_t)
SIGNATURES = {
    "rumgpu_ctx_create": [_I, _P],
    "rumgpu_ctx_destroy": [_P],
    "rumgpu_mem_info": [_I, _P, _P],
    "rumgpu_device_props": [_I, _P, _I, _P, _I, _P, _P],
    "rumgpu_device_count": [_P],
    "rumgpu_last_error": [_P, _I],
    "rumgpu_version": [],
    "rumgpu_bm_create": [_P, _I, _P, _P, _P, _I64, _I, _D, _D, _D, _D, _D, _D, _P, _I, _P],
    "rumgpu_bm_destroy": [_P],
    "rumgpu_bm_set_p": [_P, _P],
    "rumgpu_bm_init_cold": [_P],
    "rumgpu_bm_set_state": [_P, _P, _P, _P, _P],
    "rumgpu_bm_get_state": [_P, _P, _P, _P, _P],
    "rumgpu_bm_copy_state": [_P, _P],
    "rumgpu_bm_keep": [_P],
    "rumgpu_bm_load_kept": [_P],
    "rumgpu_bm_run": [_P, _I64],
    "rumgpu_bm_set_graph": [_P, _I64],
    "rumgpu_bm_finalize": [_P],
    "rumgpu_bm_phi": [_P, _P],
    "rumgpu_bm_push_ub64": [_P, _P, _P],
    "rumgpu_bm_get_fit": [_P, _P],
    "rumgpu_bm_lb64": [_P, _P],
    "rumgpu_bm_x_obs": [_P, _P],
    "rumgpu_bm_info": [_P, _P, _P, _P],
    "rumgpu_bm_profile": [_P, _I64, _P, _P],
    "rumgpu_bm_set_beta": [_P, _D],
    "rumgpu_bm_set_plan": [_P, _I, _I, _I],
    "rumgpu_bm_create_batch": [_P, _I, _P, _P, _P, _I64, _I, _D, _D, _D, _D, _D, _D, _P, _I, _P],
    "rumgpu_bm_slot_set_p": [_P, _I, _P],
    "rumgpu_bm_slot_load": [_P, _I, _P],
    "rumgpu_bm_slot_cold": [_P, _I],
    "rumgpu_bm_slot_add_obs": [_P, _I, _P, _P],
    "rumgpu_bm_slot_get_obs": [_P, _I, _P, _P],
    "rumgpu_bm_certify": [_P, _I, _P, _P, _P],
    "rumgpu_bm_slot_fit": [_P, _I, _P],
    "rumgpu_cgr_create": [_I, _I64, _P, _P, _P, _I, _P],
    "rumgpu_cgr_destroy": [_P],
    "rumgpu_cgr_add_ranking": [_P, _P],
    "rumgpu_cgr_add_column": [_P, _P],
    "rumgpu_cgr_copy_columns": [_P, _P],
    "rumgpu_cgr_set_lambda": [_P, _P, _I, _P, _I],
    "rumgpu_cgr_master": [_P, _I, _D, _D, _P, _P],
    "rumgpu_cgr_master_batch": [_P, _I, _P, _P, _I, _D, _D, _P, _P, _P],
    "rumgpu_cgr_get": [_P, _P, _P, _P, _P],
    "rumgpu_bm_queue_start": [_P, _P, _I64, _I64, _I, _I, _I64, _D, _D, _I],
    "rumgpu_bm_queue_push": [_P, _I64, _P, _P, _P, _P, _P],
    "rumgpu_bm_queue_run": [_P, _I64, _P],
    "rumgpu_bm_queue_results": [_P, _I64, _I64, _P, _P, _P, _P, _P],
    "rumgpu_bm_queue_slots": [_P, _P, _P],
    "rumgpu_bm_queue_expire": [_P, _P],
    "rumgpu_bm_queue_end": [_P],
    "rumgpu_butterfly64": [_P, _I, _P, _D, _D, _D, _D],
    "rumgpu_butterfly32": [_P, _I, _P, _F, _F, _F, _F],
    "rumgpu_prox_slab": [_P, _P, _P],
    "rumgpu_lb64": [_P, _I, _P, _P, _I64, _P, _P, _P],
    "rumgpu_lb64_noalloc": [_P, _I, _P, _P, _I64, _P, _P, _P],
    "rumgpu_cg_create": [_P, _I, _P, _P, _I64, _I, _I, _I, _P],
    "rumgpu_cg_destroy": [_P],
    "rumgpu_cg_price_batch": [_P, _I, _P, _P, _P],
    "rumgpu_cg_vmax64": [_P, _P, _P],
    "rumgpu_cg_vmax64_batch": [_P, _I, _P, _P],
    "rumgpu_cg_price_batch_lb": [_P, _I, _P, _P, _P, _I, _P, _P],
    "rumgpu_cg_set_sweep": [_P, _I],
    "rumgpu_cg_sweep_info": [_P, _P],
    "rumgpu_cg_set_slab_tile": [_P, _I],
    "rumgpu_cg_cols_init": [_P, _P, _I64],
    "rumgpu_cg_cols_load": [_P, _I64, _P],
    "rumgpu_cg_cols_add": [_P, _P, _P, _P, _P],
    "rumgpu_cg_cols_scores": [_P, _I, _P, _P],
    "rumgpu_cg_cols_pout": [_P, _P],
    "rumgpu_qp_simplex": [_I, _P, _I64, _P, _P, _I, _I, _D, _D, _P, _P, _P],
    "rumgpu_qp_create": [_P],
    "rumgpu_qp_destroy": [_P],
    "rumgpu_qp_simplex_warm": [_P, _I, _P, _I64, _P, _P, _I, _I, _D, _D, _P, _P, _P],
    "rumgpu_mlp_create": [_P, _I, _I, _I, _I, _P],
    "rumgpu_mlp_destroy": [_P],
    "rumgpu_mlp_param_count": [_P, _P],
    "rumgpu_mlp_set_params": [_P, _P],
    "rumgpu_mlp_get_params": [_P, _P],
    "rumgpu_mlp_forward": [_P, _I, _P, _P],
    "rumgpu_mlp_backward": [_P, _I, _P, _P, _P, _P],
    "rumgpu_mlp_adam": [_P, _F, _F, _F, _F],
    "rumgpu_mlp_gemm_test": [_P, _I, _I, _I, _P, _P, _P],
    "rumgpu_cg_price": [_P, _P, _P, _P, _P],
    "rumgpu_cg_info": [_P, _P, _P],
}

_lib = None
_lock = threading.Lock()


class GPUError(RuntimeError):
    """A librumgpu call failed (HIP error or unsupported size)."""


def load():
    """The library handle, with argument types set; raises if the library has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise GPUError("librumgpu.so is not built (%s); run __graft_entry__.build() or make -C "
                               "rum-testing-gpu_amd/csrc" % LIB_PATH)
            L = ctypes.CDLL(LIB_PATH)
            for name, args in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = ctypes.c_int
                fn.argtypes = args
            _lib = L
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().rumgpu_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(status, what=""):
    """Map a status code to an exception: out of memory -> MemoryError (exit flag -3 upstream)."""
    if status == OK:
        return
    msg = last_error() or what
    if status == ENOMEM:
        raise MemoryError(msg if "out of memory" in msg else "out of memory: " + msg)
    if status == EINVAL:
        raise ValueError(msg)
    raise GPUError(msg)


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def c_array(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Context:
    """One device and one HIP stream (rumgpu_ctx)."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.rumgpu_ctx_create(int(device), ctypes.byref(h)), "ctx_create")
        self.h = h
        self.device = int(device)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.rumgpu_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_CTX = {}


def context(device=0):
    """The process's shared context for a device (created on first use)."""
    with _lock:
        c = _CTX.get(device)
    if c is None:
        c = Context(device)
        with _lock:
            _CTX[device] = c
    return c


def device_count():
    n = ctypes.c_int(0)
    st = load().rumgpu_device_count(ctypes.byref(n))
    return n.value if st == OK else 0


def mem_info(device=0):
    fr, tot = ctypes.c_size_t(), ctypes.c_size_t()
    check(load().rumgpu_mem_info(int(device), ctypes.byref(fr), ctypes.byref(tot)), "mem_info")
    return fr.value, tot.value


def device_props(device=0):
    name = ctypes.create_string_buffer(256)
    arch = ctypes.create_string_buffer(256)
    tot, cus = ctypes.c_size_t(), ctypes.c_int()
    check(load().rumgpu_device_props(int(device), name, 256, arch, 256, ctypes.byref(tot), ctypes.byref(cus)))
    return dict(name=name.value.decode(), arch=arch.value.decode(), memory_total=tot.value, cu_count=cus.value)
