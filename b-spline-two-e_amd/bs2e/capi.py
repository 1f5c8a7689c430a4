"""ctypes binding of include/bs2e.h (libbs2e.so, the HIP library).

There is no CPU fallback: if the library is missing or no HIP device is
present, calls raise Bs2eError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libbs2e.so")

i64 = C.c_int64
dbl = C.c_double
vp = C.c_void_p
P = C.POINTER

BS2E_OK = 0
BS2E_EINVAL = -1
BS2E_EDEVICE = -2
BS2E_ENOMEM = -3
BS2E_ERANGE = -4
BS2E_ESTATE = -5
PREC_FP64 = 0
PREC_MIXED = 1

# (name, restype, argtypes) of every symbol include/bs2e.h declares
SIGNATURES = [
    ("bs2e_create", i64, [P(vp), i64]),
    ("bs2e_destroy", i64, [vp]),
    ("bs2e_last_error", C.c_char_p, [vp]),
    ("bs2e_set_stream", i64, [vp, vp]),
    ("bs2e_generate_grid", i64, [i64, i64, i64, dbl, dbl, P(dbl), i64]),
    ("bs2e_count_nnz", i64, [i64, i64, P(dbl), P(i64), P(i64), P(i64)]),
    ("bs2e_slater_primitives", i64, [vp, i64, i64, P(dbl), i64, i64, i64, P(dbl), P(dbl), P(i64), P(i64), P(i64),
                                     i64, P(dbl), P(i64), P(i64), P(i64), P(i64), P(i64)]),
    ("bs2e_rk_create", i64, [vp, i64, i64, P(dbl), i64, i64, i64, P(vp)]),
    ("bs2e_rk_from_primitives", i64, [vp, i64, i64, P(dbl), i64, i64, P(dbl), P(dbl), P(i64), P(i64), P(i64),
                                      i64, P(dbl), P(i64), P(i64), P(i64), P(i64), P(i64), P(vp)]),
    ("bs2e_rk_compute", i64, [vp]),
    ("bs2e_rk_synchronize", i64, [vp]),
    ("bs2e_rk_info", i64, [vp, P(i64), i64]),
    ("bs2e_rk_device_ptr", i64, [vp, P(vp), P(i64)]),
    ("bs2e_rk_pair_index", i64, [vp, i64, i64]),
    ("bs2e_rk_copy_to_host", i64, [vp, P(dbl), i64]),
    ("bs2e_rk_copy_rows", i64, [vp, i64, i64, P(dbl)]),
    ("bs2e_rk_get_val", i64, [vp, P(i64), P(dbl), i64]),
    ("bs2e_rk_set_profiling", i64, [vp, i64]),
    ("bs2e_rk_stage_times", i64, [vp, P(dbl), i64]),
    ("bs2e_rk_release", i64, [vp]),
    ("bs2e_ang_k_LS", dbl, [i64] * 6),
    ("bs2e_three_j", dbl, [i64] * 6),
    ("bs2e_six_j", dbl, [i64] * 6),
    ("bs2e_count_configs", i64, [i64] * 8 + [P(i64), i64]),
    ("bs2e_block_create", i64, [vp, i64, i64, P(i64), i64, P(vp)]),
    ("bs2e_block_sizes", i64, [vp, P(i64), i64]),
    ("bs2e_block_fill", i64, [vp, i64, P(dbl), P(dbl), P(i64), P(i64), P(dbl), P(i64), P(i64), P(dbl)]),
    ("bs2e_block_release", i64, [vp]),
    ("bs2e_block_times", i64, [vp, P(dbl), i64]),
    ("bs2e_rk_set_graph", i64, [vp, i64]),
    ("bs2e_setup_S", i64, [vp, i64, i64, P(dbl), i64, P(dbl)]),
    ("bs2e_setup_H_one_particle", i64, [vp, i64, i64, P(dbl), i64, i64, i64, i64, i64, dbl, dbl, dbl, P(dbl)]),
    ("bs2e_rk_autotune", i64, [vp, i64]),
    ("bs2e_rk_tuning", i64, [vp, P(i64), i64]),
    ("bs2e_rk_set_tuning", i64, [vp, i64, i64]),
    ("bs2e_store_ceiling", i64, [vp, vp, i64, i64, P(dbl)]),
    ("bs2e_rk_build_static", i64, [i64, i64, P(dbl), i64, i64, i64, vp, i64, vp]),
    ("bs2e_blockset_create", i64, [vp, i64, P(i64), P(i64), P(i64), i64, P(vp)]),
    ("bs2e_blockset_sizes", i64, [vp, P(i64), i64]),
    ("bs2e_blockset_fill", i64, [vp, i64, P(dbl), P(dbl)]),
    ("bs2e_blockset_get", i64, [vp, i64, P(i64), P(i64), P(dbl), P(i64), P(i64), P(dbl)]),
    ("bs2e_blockset_get_rows", i64, [vp, i64, i64, i64, P(i64), P(i64), P(dbl), P(i64), P(i64), P(dbl)]),
    ("bs2e_blockset_fill_rows", i64, [vp, i64, P(dbl), P(dbl), i64, i64]),
    ("bs2e_blockset_row_counts", i64, [vp, i64, i64, P(i64), P(i64)]),
    ("bs2e_blockset_reserve", i64, [vp, i64, i64]),
    ("bs2e_blockset_refill", i64, [vp]),
    ("bs2e_blockset_set_graph", i64, [vp, i64]),
    ("bs2e_blockset_replay", i64, [vp]),
    ("bs2e_blockset_synchronize", i64, [vp]),
    ("bs2e_blockset_times", i64, [vp, P(dbl), i64]),
    ("bs2e_blockset_release", i64, [vp]),
    ("bs2e_setup_r_k_int", i64, [vp, i64, i64, P(dbl), i64, i64, dbl, P(dbl)]),
    ("bs2e_setup_radial_dip", i64, [vp, i64, i64, P(dbl), i64, i64, P(dbl), P(dbl)]),
    ("bs2e_dip_block_create", i64, [vp, i64, i64, i64, i64, i64, i64, P(i64), i64, i64, i64, i64, P(i64), i64, i64,
                                    P(vp)]),
    ("bs2e_dip_block_sizes", i64, [vp, P(i64), i64]),
    ("bs2e_dip_block_fill", i64, [vp, i64, P(dbl), P(dbl), P(dbl), P(i64), P(i64), P(dbl)]),
    ("bs2e_dip_block_times", i64, [vp, P(dbl), i64]),
    ("bs2e_dip_block_release", i64, [vp]),
    ("bs2e_server_create", i64, [i64, i64, i64, P(vp)]),
    ("bs2e_server_destroy", i64, [vp]),
    ("bs2e_server_last_error", C.c_char_p, [vp]),
    ("bs2e_footprint", i64, [i64, i64, P(dbl), i64, i64, P(i64)]),
    ("bs2e_server_submit", i64, [vp, i64, i64, P(dbl), i64, i64]),
    ("bs2e_server_submit_job", i64, [vp, vp]),
    ("bs2e_server_blocks", i64, [vp, i64, P(vp)]),
    ("bs2e_server_cache", i64, [vp, P(i64), i64]),
    ("bs2e_server_progress", i64, [vp]),
    ("bs2e_server_status", i64, [vp, i64, P(i64), P(dbl)]),
    ("bs2e_server_wait", i64, [vp, i64]),
    ("bs2e_server_tensor", i64, [vp, i64, P(vp)]),
    ("bs2e_server_retire", i64, [vp, i64]),
    ("bs2e_server_finished", i64, [vp, P(i64), P(dbl), i64]),
    ("bs2e_server_memory", i64, [vp, P(i64)]),
]


class Bs2eError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bs2e error {code}: {msg}")
        self.code = code


class Job(C.Structure):
    """struct bs2e_job (include/bs2e.h)."""
    _fields_ = [("k", i64), ("n_knots", i64), ("knots", P(dbl)), ("k_gl", i64), ("max_k", i64),
                ("precision", i64), ("n_sym", i64), ("L", P(i64)), ("n_config", P(i64)), ("configs", P(i64)),
                ("full", i64), ("Z", i64), ("max_l_1p", i64), ("cap_order", i64), ("cap_r0", dbl),
                ("cap_eta_re", dbl), ("cap_eta_im", dbl), ("sym_configs", P(P(i64)))]


_lib = None


def lib():
    """Loads libbs2e.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise Bs2eError(BS2E_ESTATE, f"{LIB_PATH} is missing: build it with "
                                         "`python b-spline-two-e_amd/build.py` (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, ctx=None):
    if rc < 0:
        msg = lib().bs2e_last_error(ctx)
        raise Bs2eError(int(rc), msg.decode() if msg else "")
    return rc


def ptr(a, ct=dbl):
    return a.ctypes.data_as(P(ct))
