"""ctypes binding of libemsa.so (the C ABI declared in include/emsa.h).

The library must be built in-tree (``python -m cuda_emicp_softassign_amd.build``
or ``__graft_entry__.build()``). There is no fallback: a missing library
raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libemsa.so")

EMSA_OK = 0
EMSA_EINVAL = -1
EMSA_ENOMEM = -2
EMSA_EHIP = -3
EMSA_ENONFINITE = -4
EMSA_ENODEV = -5
SOFTASSIGN_MATERIALISED = 0
SOFTASSIGN_LOGDOMAIN = 1
TRACE_W = 32


class emsa_params(C.Structure):
    """Mirror of struct emsa_params (include/emsa.h)."""
    _fields_ = [
        ("JMAX", C.c_int), ("I0", C.c_int), ("I1", C.c_int),
        ("alpha", C.c_float), ("T_0", C.c_float), ("TFACTOR", C.c_float),
        ("moutlier", C.c_float),
        ("sigma_p2", C.c_float), ("sigma_inf", C.c_float), ("sigma_factor", C.c_float),
        ("d_02", C.c_float),
        ("maxIteration", C.c_int), ("noviewer", C.c_int), ("nostop", C.c_int),
        ("notimer", C.c_int),
    ]


class emsa_pair(C.Structure):
    """Mirror of struct emsa_pair (include/emsa.h)."""
    _fields_ = [("X", C.c_void_p), ("N", C.c_int), ("Y", C.c_void_p), ("M", C.c_int)]


class emsa_pair_grad(C.Structure):
    """Mirror of struct emsa_pair_grad (include/emsa.h)."""
    _fields_ = [("gX", C.c_void_p), ("gY", C.c_void_p)]


# Every function the header declares, with its ctypes signature.
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_pp = C.POINTER(emsa_params)
_ip = C.POINTER(C.c_int)
# void (*)(int algorithm, int iteration, const float *R, const float *t, void *user)
emsa_iteration_fn = C.CFUNCTYPE(None, C.c_int, C.c_int, _fp, _fp, C.c_void_p)
SIGNATURES = {
    "emsa_default_params": (None, [_pp]),
    "emsa_set_iteration_callback": (None, [emsa_iteration_fn, C.c_void_p]),
    "emsa_last_error": (C.c_char_p, []),
    "emsa_emicp_iterations": (C.c_int, [_pp]),
    "emsa_emicp": (C.c_int, [_fp, C.c_int, _fp, C.c_int, _fp, _fp, _pp]),
    "emsa_emicp_traced": (C.c_int, [_fp, C.c_int, _fp, C.c_int, _fp, _fp, _pp, _dp, C.c_int]),
    "emsa_emicp_step": (C.c_int, [_fp, C.c_int, _fp, C.c_int, _fp, _fp, C.c_float, C.c_float,
                                  _dp, _dp, _dp]),
    "emsa_emicp_cpu": (C.c_int, [_fp, C.c_int, _fp, C.c_int, _fp, _fp, _pp, _dp, C.c_int]),
    "emsa_softassign": (C.c_int, [_fp, C.c_int, _fp, C.c_int, _fp, _fp, _pp, C.c_int]),
    "emsa_softassign_traced": (C.c_int, [_fp, C.c_int, _fp, C.c_int, _fp, _fp, _pp, C.c_int,
                                         _dp, C.c_int]),
    "emsa_softassign_step": (C.c_int, [_fp, C.c_int, _fp, C.c_int, _fp, _fp, C.c_float,
                                       C.c_float, C.c_int, _fp, _fp, C.c_int, _dp, _dp, _dp,
                                       _dp]),
    "emsa_find_rt_from_s": (C.c_int, [_fp, _fp, _fp, _fp, _fp]),
    "emsa_icp": (C.c_int, [_fp, C.c_int, _fp, C.c_int, _fp, _fp, _pp]),
    "emsa_icp_traced": (C.c_int, [_fp, C.c_int, _fp, C.c_int, _fp, _fp, _pp, _dp, C.c_int]),
    "emsa_icp_nearest": (C.c_int, [_fp, C.c_int, _fp, C.c_int, _fp, _fp, _ip, _fp]),
    "emsa_emicp_dev": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, _pp,
                                 C.c_void_p]),
    "emsa_emicp_dev_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                         _pp, C.c_void_p]),
    "emsa_probe_rates": (C.c_int, [_dp, _dp]),
    "emsa_probe_issue_cycles": (C.c_int, [_dp, _dp, _dp]),
    "emsa_emicp_batched": (C.c_int, [C.POINTER(emsa_pair), C.c_int, _fp, C.c_int, _pp, _fp,
                                     _dp]),
    "emsa_emicp_batched_dev": (C.c_int, [C.POINTER(emsa_pair), C.c_int, C.c_void_p, C.c_int,
                                         _pp, C.c_void_p, C.c_void_p, C.c_void_p]),
    "emsa_autotune_emicp": (C.c_int, [C.POINTER(emsa_pair), C.c_int, C.c_int, _pp, C.c_void_p,
                                      C.POINTER(C.c_int)]),
    "emsa_graph_cache_clear": (C.c_int, []),
    "emsa_set_tiling": (C.c_int, [C.c_int, C.c_int]),
    "emsa_probe_emicp_affinity": (C.c_int, [C.POINTER(emsa_pair), C.c_int, C.c_int, C.c_float,
                                            C.c_int, _dp, _dp]),
    "emsa_probe_set_affinity_path": (C.c_int, [C.c_int]),
    "emsa_probe_set_affinity_cull": (C.c_int, [C.c_int]),
    "emsa_probe_set_row_fuse": (C.c_int, [C.c_int]),
    "emsa_probe_set_affinity_balance": (C.c_int, [C.c_int]),
    "emsa_probe_set_affinity_expanded": (C.c_int, [C.c_float]),
    "emsa_probe_hilbert_order": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "emsa_probe_emicp_sequence": (C.c_int, [C.POINTER(emsa_pair), C.c_int, C.c_int,
                                            C.POINTER(emsa_params), _dp, _dp,
                                            C.POINTER(C.c_int), _dp, C.c_int]),
    "emsa_emicp_unrolled_ckpt_size": (C.c_longlong, [C.POINTER(emsa_pair), C.c_int, C.c_int]),
    "emsa_emicp_unrolled_fwd": (C.c_int, [C.POINTER(emsa_pair), C.c_int, C.c_void_p, C.c_float,
                                          C.c_float, C.c_float, C.c_int, C.c_void_p, C.c_void_p,
                                          C.c_void_p]),
    "emsa_emicp_unrolled_bwd": (C.c_int, [C.POINTER(emsa_pair), C.c_int, C.c_void_p, C.c_float,
                                          C.c_float, C.c_float, C.c_int, C.c_void_p, C.c_void_p,
                                          C.POINTER(emsa_pair_grad), C.c_void_p, C.c_void_p,
                                          C.c_void_p]),
    "emsa_softassign_unrolled_ckpt_bytes": (C.c_longlong, [C.c_int] * 6),
    "emsa_softassign_unrolled_fwd": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                               C.c_void_p, C.c_float, C.c_float, C.c_float,
                                               C.c_float, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                               C.c_void_p, C.c_void_p]),
    "emsa_softassign_unrolled_bwd": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                               C.c_void_p, C.c_float, C.c_float, C.c_float,
                                               C.c_float, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_void_p]),
}

_lib = None


class EmsaError(RuntimeError):
    """A libemsa entry point returned a negative status."""

    def __init__(self, fn, code, msg):
        super().__init__("%s failed (%d): %s" % (fn, code, msg))
        self.code = code


def load():
    global _lib
    if _lib is None:
        # torch bundles its own libamdhip64.so.7. Import it first so that this
        # library binds to the same HIP runtime (same soname) instead of
        # loading a second one from /opt/rocm into the process.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libemsa.so is not built (%s); run python -m cuda_emicp_softassign_amd.build"
                % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(fn, rc):
    if rc < 0:
        msg = load().emsa_last_error()
        raise EmsaError(fn, rc, msg.decode() if msg else "")
    return rc
