"""ctypes binding of libima_gpu.so (the C ABI in include/ima_gpu.h).

Loading fails loudly when the built library is missing: the engine has no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libima_gpu.so")

IMA_OK = 0
ERR_ARG, ERR_HIP, ERR_STATE, ERR_SIZE, ERR_NODEV = 1, 2, 3, 4, 5
MODEL_HKY, MODEL_IS, MODEL_SW = 0, 1, 2
PREC_FP64, PREC_FP32 = 0, 1
STATUS_OK, STATUS_IS_REJECT, STATUS_SW_ZERO, STATUS_IS_NOSITE, STATUS_PRIOR_BAD = 0, 1, 2, 3, 4
LMODE_REF_COMPAT, LMODE_EXACT = 0, 1
POPMIG_CALC, POPMIG_MARGIN = 0, 1
REJECT_IS = -1000000000.0
MAX_GENES = 256
K_HKY64, K_HKY32, K_IS, K_SW, K_LMODE, K_SWAP, K_MC3, K_ACCEPT, K_PRIOR = range(9)
MC3_GRAPH, MC3_SCALE, MC3_TIMED, MC3_RY1 = 1, 2, 4, 8
MC3_UPDATE_U, MC3_U_SERIAL, MC3_TOPOLOGY, MC3_NW = 16, 32, 64, 128


class McfLocus(C.Structure):
    _fields_ = [("model", C.c_int32), ("n", C.c_int32), ("nsites", C.c_int32)]


class McfState(C.Structure):
    _fields_ = [(k, C.POINTER(C.c_double)) for k in
                ("tvals", "u", "kappa", "pi", "time", "dlikeA", "mig_time")] + \
               [(k, C.POINTER(C.c_int32)) for k in
                ("up0", "up1", "down", "mut", "pop", "A", "mig_count", "mig_to")] + \
               [("mig_capacity", C.c_int64)]


class Item(C.Structure):
    _fields_ = [("slot", C.c_int32), ("flags", C.c_int32), ("u", C.c_double),
                ("kappa", C.c_double), ("tr", C.c_double)]


class SWMove(C.Structure):
    """ima_sw_move."""
    _fields_ = [(f, C.c_int32) for f in ("slot", "edge", "downedge", "sisedge", "newsisedge",
                                          "copyA0", "copyA1", "down_was_root", "hold_down_A",
                                          "minA", "maxA", "flip")] + \
               [("copy_dlikeA", C.c_double * 3), ("newsis_dlikeA", C.c_double),
                ("u", C.c_double), ("uni", C.c_double)]


i32p = C.POINTER(C.c_int32)
i8p = C.POINTER(C.c_int8)
dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)
vp = C.c_void_p

# name -> (restype, argtypes)
PROTOS = {
    "ima_abi_version": (C.c_int, []),
    "ima_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "ima_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "ima_ctx_destroy": (None, [vp]),
    "ima_ctx_last_error": (C.c_char_p, [vp]),
    "ima_locus_add_hky": (C.c_int, [vp, C.c_int, C.c_int, i8p, i32p, C.c_int, dp,
                                    C.POINTER(C.c_int)]),
    "ima_locus_add_is": (C.c_int, [vp, C.c_int, C.c_int, i8p, C.POINTER(C.c_int)]),
    "ima_locus_add_sw": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int)]),
    "ima_locus_set_sumlogk": (C.c_int, [vp, C.c_int, C.c_double]),
    "ima_locus_sumlogk_from_slot": (C.c_int, [vp, C.c_int, C.c_int, dp]),
    "ima_slots_alloc": (C.c_int, [vp, C.c_int, i32p, C.POINTER(C.c_int)]),
    "ima_slots_write": (C.c_int, [vp, C.c_int, C.c_int, i32p, i32p, i32p, dp, i32p, i32p]),
    "ima_slots_read": (C.c_int, [vp, C.c_int, C.c_int, i32p, i32p, i32p, dp, i32p, i32p]),
    "ima_slots_count": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "ima_slot_info": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ima_score": (C.c_int, [vp, C.POINTER(Item), C.c_int, C.c_int, dp, i32p]),
    "ima_model_add": (C.c_int, [vp, C.c_int, i32p, i32p, i32p, i32p, i32p, C.c_int, i32p, i32p,
                                C.c_int, i32p, i32p, dp, dp, C.POINTER(C.c_int)]),
    "ima_locus_set_model": (C.c_int, [vp, C.c_int, C.c_int, i32p, C.c_double]),
    "ima_slots_write_pops": (C.c_int, [vp, C.c_int, C.c_int, i32p, i32p, i32p, dp, i32p]),
    "ima_treeweight": (C.c_int, [vp, i32p, C.c_int, dp, i32p, dp, dp, i32p, dp, dp]),
    "ima_slots_reserve_migs": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
    "ima_t_nw_update": (C.c_int, [vp, i32p, C.c_int, dp, C.c_int, C.c_double, C.c_uint64,
                                  C.c_uint64, dp, i32p]),
    "ima_prior": (C.c_int, [vp, i32p, i32p, C.c_int, dp, dp, dp, dp]),
    "ima_mc3_set_prior": (C.c_int, [vp, dp, C.c_double]),
    "ima_mc3_read_prior": (C.c_int, [vp, dp, dp]),
    "ima_slots_read_mig": (C.c_int, [vp, C.c_int, dp]),
    "ima_sw_update_A": (C.c_int, [vp, vp, C.c_int, C.c_int, dp, dp, i32p, dp]),
    "ima_mc3_set_prior_group": (C.c_int, [vp, C.c_int]),
    "ima_mc3_set_analyses": (C.c_int, [vp, C.c_int]),
    "ima_mc3_prior_group": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "ima_mc3_tune": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_int), dp]),
    "ima_device_memory": (C.c_int, [C.c_int, dp, dp]),
    "ima_server_create": (C.c_int, [C.c_int, C.c_double, C.c_int, C.POINTER(vp)]),
    "ima_server_destroy": (None, [vp]),
    "ima_server_last_error": (C.c_char_p, [vp]),
    "ima_server_table": (C.c_int, [vp, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.POINTER(C.c_int)]),
    "ima_server_submit_lmode": (C.c_int, [vp, C.c_int, dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_int, C.POINTER(C.c_int64)]),
    "ima_server_submit_analysis": (C.c_int, [vp, vp, C.POINTER(C.c_int64)]),
    "ima_server_step": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "ima_server_status": (C.c_int, [vp, C.c_int64, C.POINTER(C.c_int), dp, dp]),
    "ima_server_lmode_result": (C.c_int, [vp, C.c_int64, dp, dp]),
    "ima_server_analysis_result": (C.c_int, [vp, C.c_int64, dp, dp, dp, C.POINTER(C.c_uint64)]),
    "ima_server_release": (C.c_int, [vp, C.c_int64]),
    "ima_server_latency": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), dp, dp, dp]),
    "ima_server_usage": (C.c_int, [vp, dp, dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ima_score_repeat": (C.c_int, [vp, C.POINTER(Item), C.c_int, C.c_int, C.c_int, dp, i32p]),
    "ima_is_site_nodes": (C.c_int, [vp, C.c_int, i32p]),
    "ima_sw_dlike": (C.c_int, [vp, C.POINTER(Item), dp]),
    "ima_likelihood_hky": (C.c_int, [vp, C.c_int, i32p, i32p, i32p, dp, C.c_int, C.c_double,
                                     C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, dp]),
    "ima_swap": (C.c_int, [vp, C.c_int, dp, dp, C.c_int, C.c_uint64, C.c_uint64,
                           C.POINTER(C.c_int), i32p]),
    "ima_lmode_table": (C.c_int, [vp, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.POINTER(C.c_int)]),
    "ima_lmode_table_free": (C.c_int, [vp, C.c_int]),
    "ima_lmode_jointp": (C.c_int, [vp, C.c_int, dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int, dp, dp]),
    "ima_lmode_jointp_noalloc": (C.c_int, [vp, C.c_int, vp, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_int, vp, vp]),
    "ima_lmode_joint_peak": (C.c_int, [vp, C.c_int, C.c_int, dp, dp, dp, C.c_int, C.c_uint64,
                                       C.c_int, dp, dp, dp, C.POINTER(C.c_int64)]),
    "ima_lmode_marginal_peak": (C.c_int, [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                          C.c_double, dp]),
    "ima_lmode_popmig": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_int, dp]),
    "ima_lmode_popmig_peak": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_int, C.c_int, C.c_int, dp, dp,
                                        C.POINTER(C.c_int64)]),
    "ima_lmode_greater_than": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp, dp]),
    "ima_slots_read_mig_events": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int32), i32p, dp, i32p]),
    "ima_set_site_block": (C.c_int, [vp, C.c_int]),
    "ima_mc3_site_block": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "ima_mcf_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, vp, vp]),
    "ima_mcf_read": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, vp, dp, vp,
                               C.POINTER(C.c_int64)]),
    "ima_lmode_margincalc": (C.c_int, [vp, C.c_int, C.c_int, dp, C.c_int, C.c_int,
                                       C.c_double, C.c_int, C.c_double, dp]),
    "ima_mc3_setup": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp, dp, dp, C.c_uint64, C.c_int,
                                C.c_int, C.c_double]),
    "ima_mc3_run": (C.c_int, [vp, C.c_int, C.c_int]),
    "ima_mc3_read": (C.c_int, [vp, dp, dp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "ima_mc3_read_rates": (C.c_int, [vp, dp, dp]),
    "ima_mc3_decide": (C.c_int, [vp, C.c_int, dp, dp, C.c_double, i32p]),
    "ima_slots_read_pops": (C.c_int, [vp, C.c_int, C.c_int, i32p]),
    "ima_mc3_record": (C.c_int, [vp, C.c_int]),
    "ima_ti_write": (C.c_int, [C.c_char_p, C.c_char_p, fp, C.c_int, C.c_int, C.c_int]),
    "ima_ti_count": (C.c_int, [C.c_char_p, C.POINTER(C.c_int)]),
    "ima_ti_load": (C.c_int, [C.POINTER(C.c_char_p), C.c_int, C.c_int, C.c_int, fp,
                              C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ima_mc3_read_weights": (C.c_int, [vp, dp, dp]),
    "ima_lmode_table_reserve": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.POINTER(C.c_int)]),
    "ima_lmode_table_info": (C.c_int, [vp, C.c_int] + [C.POINTER(C.c_int)] * 5),
    "ima_lmode_table_read": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, fp]),
    "ima_integrate_tree_prob": (C.c_int, [vp, C.c_int, dp, C.c_int, dp, dp, dp]),
    "ima_mc3_prior_plan": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
    "ima_timing_enable": (C.c_int, [vp, C.c_int]),
    "ima_timing_get": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int64), dp]),
    "ima_timing_reset": (C.c_int, [vp]),
}

_lib = None


def load():
    """Load libima_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m ima2p_amd.build` "
                           "(the engine has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in PROTOS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def declared_symbols(header=None):
    """Entry points declared in include/ima_gpu.h."""
    import re
    header = header or os.path.join(os.path.dirname(HERE), "include", "ima_gpu.h")
    txt = open(header).read()
    return sorted(set(re.findall(r"\b(ima_[A-Za-z0-9_]+)\s*\(", txt)))
