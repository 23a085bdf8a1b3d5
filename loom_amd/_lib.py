"""ctypes binding of libloomgpu.so (declarations: include/loomgpu.h).

The library is built in-tree (``python -m loom_amd.build``); importing this
module without it raises immediately -- there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libloomgpu.so")

u32p = ctypes.POINTER(ctypes.c_uint32)
u32pp = ctypes.POINTER(u32p)
sizep = ctypes.POINTER(ctypes.c_size_t)

EXPORTS = {
    "lg_last_error": (ctypes.c_char_p, []),
    "lg_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "lg_ctx_destroy": (None, [ctypes.c_void_p]),
    "lg_plan_create": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "lg_plan_destroy": (None, [ctypes.c_void_p]),
    "lg_plan_check": (ctypes.c_int, [u32p, ctypes.c_size_t, sizep]),
    "lg_plan_set_setup": (ctypes.c_int, [ctypes.c_void_p, u32pp, ctypes.c_size_t, u32p]),
    "lg_plan_num_inputs": (ctypes.c_int, [ctypes.c_void_p, sizep, sizep]),
    "lg_prove": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, u32pp, u32pp, u32pp, sizep, sizep, u32pp, sizep, sizep]),
    "lg_prove_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, u32pp, u32pp, u32pp, sizep, sizep, u32pp, sizep,
                                       sizep]),
    "lg_plan_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, u32pp, u32pp, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)]),
    "lg_plan_set_options": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "lg_plan_set_grinding": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "lg_prove_chained": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, u32pp, u32pp, ctypes.c_int, u32p, u32pp,
                                        sizep, sizep, u32pp, sizep, sizep]),
    "lg_verify_chained": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, u32pp, sizep, u32p, ctypes.c_size_t, u32pp,
                                         ctypes.c_int, u32p, ctypes.POINTER(ctypes.c_int32)]),
    "lg_plan_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]),
    "lg_ctx_set_graphs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "lg_plan_graph_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
    "lg_verify": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, u32pp, sizep, u32p, ctypes.c_size_t, u32pp, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_int32)]),
    "lg_plan_verify_message": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t, sizep]),
    "lg_ctx_set_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "lg_ctx_timing_report": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t, sizep]),
    "lg_plan_last_phases": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "lg_hash_leaf_pairs": (ctypes.c_int, [ctypes.c_void_p, u32pp, ctypes.c_int, u32pp, ctypes.c_int, ctypes.c_size_t, u32p]),
    "lg_merkle_build": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_size_t, u32pp, sizep, ctypes.c_int]),
    "lg_grand_product": (ctypes.c_int, [ctypes.c_void_p, u32p, u32p, u32p, ctypes.c_size_t]),
    "lg_probe_issue_rate": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "lg_grand_product_noalloc": (ctypes.c_int, [ctypes.c_void_p, u32p, u32p, u32p, ctypes.c_size_t]),
    "lg_rs_encode": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_size_t, ctypes.c_int, u32p]),
    "lg_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "lg_server_destroy": (None, [ctypes.c_void_p]),
    "lg_server_submit": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_size_t, ctypes.c_int, u32pp, u32pp,
                                        ctypes.POINTER(ctypes.c_uint64)]),
    "lg_server_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, u32pp, sizep, sizep]),
    "lg_server_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t]),
    "lg_plan_bytes_per_instance": (ctypes.c_int, [u32p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint64)]),
}

_lib = None


class LoomGpuError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LoomGpuError("libloomgpu.so is not built (run `python -m loom_amd.build`); "
                               "the prove path has no CPU fallback")
        # PyTorch ships its own HIP runtime (same soname as /opt/rocm's).  It
        # must be the one in the process before the library loads, so that
        # the library binds to it instead of pulling in a second runtime.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        msg = lib().lg_last_error().decode(errors="replace")
        raise LoomGpuError("%s failed (%d): %s" % (what, rc, msg))
