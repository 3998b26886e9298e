"""ctypes view of include/dem_capi.h (the engine's C ABI).

The structures mirror the C layouts field for field; ``libgpucode_amd.so``
must have been built (``python -m gpucode_amd.build``).  There is no
fallback: if the library is missing every call raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpucode_amd.so")

MAX_LEVELS = 8
WALLS = 3

DEM_OK = 0
ERRORS = {-1: "DEM_ERR_ARG", -2: "DEM_ERR_HIP", -3: "DEM_ERR_OVERFLOW",
          -4: "DEM_ERR_LEVEL", -5: "DEM_ERR_NOMEM", -6: "DEM_ERR_STATE"}
PREC_FP64 = 0
PREC_MIXED = 1


class DemInput(C.Structure):
    _fields_ = [
        ("NP", C.c_int), ("nLevels", C.c_int), ("loadStage", C.c_int), ("forceCap", C.c_int),
        ("tPacking", C.c_double), ("realSize", C.c_double), ("cutGap", C.c_double),
        ("expandSize", C.c_double),
        ("cirX", C.c_double), ("cirY", C.c_double), ("Bw", C.c_double), ("R", C.c_double),
        ("H", C.c_double), ("feedLmin", C.c_double), ("feedLmax", C.c_double),
        ("mat", (C.c_double * 7) * 2),
        ("stension", C.c_double), ("vol", C.c_double), ("gapmn", C.c_double),
        ("dtFactor", C.c_double), ("nOutPacking", C.c_int), ("sizeLevelInput", C.c_int),
        ("levelDia", C.c_double * MAX_LEVELS),
        ("gravity", C.c_double), ("dtOverride", C.c_double), ("theta", C.c_double),
        ("ruptureMode", C.c_int), ("gridClampQuirk", C.c_int),
    ]


class DemCalpara(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("emod", "pois", "sfrc", "ystress", "ydsp_t", "yforce_t", "ystiff_t")]


class DemSetup(C.Structure):
    _fields_ = (
        [(n, C.c_double) for n in ("rlunit", "rmunit", "rfunit", "rsunit", "rtunit", "rvunit",
                                   "reunit", "dt", "searchGap")]
        + [("origin", C.c_double * 3), ("endPoint", C.c_double * 3)]
        + [(n, C.c_double) for n in ("cirX", "cirY", "R", "Tw", "Bw", "topv", "btmv",
                                     "feedLmin", "feedLmax", "tPacking")]
        + [(n, C.c_double * 2) for n in ("density", "emod", "ymod", "pois", "sfrc", "rfrc",
                                         "dmpn", "yldp")]
        + [(n, C.c_double) for n in ("stension", "vol", "gapmn", "layer", "brknmx", "theta")]
        + [("pp", DemCalpara), ("pw", DemCalpara)]
        + [(n, C.c_int) for n in ("nLevels", "forceCap", "loadStage", "ruptureMode",
                                  "gridClampQuirk", "nPPCnt")]
        + [("levelSize", C.c_double * MAX_LEVELS), ("levelNNebors", C.c_int * MAX_LEVELS),
           ("levelNPars", C.c_int * MAX_LEVELS), ("levelNebSId", C.c_int * MAX_LEVELS),
           ("cellSize", C.c_double * MAX_LEVELS), ("gridSize", (C.c_int * 3) * MAX_LEVELS),
           ("numCells", C.c_int * MAX_LEVELS), ("ncellIdx", C.c_int * MAX_LEVELS),
           ("sysNCells", C.c_int), ("radMin", C.c_double), ("radMax", C.c_double)]
    )


class DemStats(C.Structure):
    _fields_ = [(n, C.c_longlong) for n in ("n_particles", "n_nebor", "n_contact",
                                            "n_wall_slots", "n_wall_contact", "n_bridge")]


class DemTuneResult(C.Structure):
    _fields_ = [("cutGap", C.c_double), ("kernel", C.c_int), ("morton", C.c_int),
                ("ms_per_step", C.c_double), ("ms_default", C.c_double), ("n_tried", C.c_int)]


class DemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_LIB = None

P = C.c_void_p
D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)
F = C.POINTER(C.c_float)

_SIGS = {
    "dem_last_error": (C.c_char_p, []),
    "dem_derive": (C.c_int, [C.POINTER(DemInput), D, C.POINTER(DemSetup), D, I, I]),
    "dem_create": (C.c_int, [C.POINTER(DemInput), D, C.c_int, C.POINTER(P)]),
    "dem_destroy": (C.c_int, [P]),
    "dem_get_setup": (C.c_int, [P, C.POINTER(DemSetup)]),
    "dem_set_material": (C.c_int, [P, D]),
    "dem_set_liquid": (C.c_int, [P, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_int, C.c_int]),
    "dem_set_parameters": (C.c_int, [P, C.c_double, C.c_double]),
    "dem_set_particles": (C.c_int, [P, C.c_int, D, D, D]),
    "dem_feed": (C.c_int, [P, C.c_int, D]),
    "dem_total": (C.c_int, [P]),
    "dem_get_particles": (C.c_int, [P, D, D, D]),
    "dem_get_forces": (C.c_int, [P, D, D, D]),
    "dem_set_store_forces": (C.c_int, [P, C.c_int]),
    "dem_get_pmove": (C.c_int, [P, D]),
    "dem_get_lqdcnt": (C.c_int, [P, I]),
    "dem_get_nebor": (C.c_int, [P, I, I, C.c_int]),
    "dem_get_hist_pp": (C.c_int, [P, C.c_int, D, D, D, D, D, I, I]),
    "dem_get_walls": (C.c_int, [P, I, I]),
    "dem_get_hist_pw": (C.c_int, [P, D, D, D, D, D, I, I]),
    "dem_set_contacts": (C.c_int, [P, C.c_int, I, I, D, D, D, D, D, I, I, I, I, D, D, D, D, D,
                                   I, I]),
    "dem_get_hash": (C.c_int, [P, I, I]),
    "dem_get_cells": (C.c_int, [P, I, I]),
    "dem_cnt_hstry_save": (C.c_int, [P]),
    "dem_bld_nebor_list": (C.c_int, [P]),
    "dem_bld_cir_wall": (C.c_int, [P]),
    "dem_hstry_copy": (C.c_int, [P]),
    "dem_var_init": (C.c_int, [P]),
    "dem_pp_inter_calc": (C.c_int, [P]),
    "dem_cir_pw_calc": (C.c_int, [P]),
    "dem_update_system": (C.c_int, [P, I]),
    "dem_rebuild": (C.c_int, [P]),
    "dem_flush_rebuild": (C.c_int, [P]),
    "dem_step_unfused": (C.c_int, [P]),
    "dem_step": (C.c_int, [P, C.c_int]),
    "dem_time_steps": (C.c_int, [P, C.c_int, D, I]),
    "dem_sync": (C.c_int, [P]),
    "dem_set_kernel": (C.c_int, [P, C.c_int]),
    "dem_set_order": (C.c_int, [P, C.c_int]),
    "dem_set_relayout": (C.c_int, [P, C.c_int]),
    "dem_n_relayout": (C.c_int, [P]),
    "dem_set_wave_split": (C.c_int, [P, C.c_int]),
    "dem_set_pair_records": (C.c_int, [P, C.c_int]),
    "dem_pair_records_used": (C.c_int, [P]),
    "dem_iter": (C.c_int, [P]),
    "dem_ctime": (C.c_double, [P]),
    "dem_n_rebuild": (C.c_int, [P]),
    "dem_contact_stats": (C.c_int, [P, C.POINTER(DemStats)]),
    "dem_get_timing": (C.c_int, [P, D, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), D, I]),
    "dem_device_bytes": (C.c_longlong, [P]),
    "dem_step_kernel_name": (C.c_char_p, [P]),
    "dem_stream_copy": (C.c_int, [C.c_longlong, C.c_int, D]),
    "dem_read_deck": (C.c_int, [C.c_char_p, C.POINTER(DemInput)]),
    "dem_particle_feed": (C.c_int, [C.POINTER(DemSetup), D, D, C.c_int, C.c_int,
                                    C.POINTER(C.c_ulonglong)]),
    "dem_write_tec": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.c_int, D, D, D, D, D]),
    "dem_load_para": (C.c_int, [C.c_char_p, I, I, I]),
    "dem_profile_rebuild": (C.c_int, [P, C.c_int, D]),
    "dem_log_header": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "dem_log_output": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                 C.c_int, C.c_int]),
    "dem_log_total": (C.c_int, [C.c_char_p, C.c_double]),
    "dem_key_bits": (C.c_int, [P]),
    "dem_set_neighbor_search": (C.c_int, [P, C.c_int]),
    "dem_neighbor_search_used": (C.c_int, [P]),
    "dem_scale_velocities": (C.c_int, [P, C.c_double]),
    "dem_dump_write": (C.c_int, [P, C.c_char_p, C.c_int]),
    "dem_dump_write_binary": (C.c_int, [P, C.c_char_p, C.c_int]),
    "dem_dump_read": (C.c_int, [P, C.c_char_p, I]),
    "dem_set_lqdcnt": (C.c_int, [P, I]),
    "dem_set_clock": (C.c_int, [P, C.c_int, C.c_double, C.c_double]),
    # dem_batch.h
    "dem_batch_create": (C.c_int, [C.c_int, C.POINTER(DemInput), C.POINTER(D), C.c_int,
                                   C.POINTER(P)]),
    "dem_batch_destroy": (C.c_int, [P]),
    "dem_batch_size": (C.c_int, [P]),
    "dem_batch_instance": (P, [P, C.c_int]),
    "dem_batch_step": (C.c_int, [P, C.c_int]),
    "dem_batch_set_schedule": (C.c_int, [P, C.c_int, C.c_int]),
    "dem_batch_flush_rebuilds": (C.c_int, [P]),
    "dem_batch_schedule_stats": (C.c_int, [P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "dem_batch_timing": (C.c_int, [P, D, C.POINTER(C.c_longlong), D, C.POINTER(C.c_longlong)]),
    "dem_batch_analyze": (C.c_int, [P, D, D, D]),
    "dem_analyze": (C.c_int, [P, D, D, D]),
    # dem_tune.h
    "dem_autotune": (C.c_int, [C.POINTER(DemInput), D, C.c_int, D, D, C.c_int, D, C.c_int, C.c_int,
                               C.POINTER(DemTuneResult), D]),
    # dem_surrogate.h
    "sur_nparams": (C.c_int, []),
    "sur_create": (C.c_int, [C.c_ulonglong, C.POINTER(P)]),
    "sur_destroy": (C.c_int, [P]),
    "sur_get_params": (C.c_int, [P, F]),
    "sur_set_params": (C.c_int, [P, F]),
    "sur_forward": (C.c_int, [P, C.c_int, F, F]),
    "sur_forward_noalloc": (C.c_int, [P, C.c_int, C.c_void_p, C.c_void_p]),
    "sur_loss_grad": (C.c_int, [P, C.c_int, F, F, F, F]),
    "sur_train": (C.c_int, [P, C.c_int, F, F, C.c_int, C.c_int, C.c_float, C.c_ulonglong, F]),
}


def declared_symbols():
    return list(_SIGS)


def lib():
    """Loads the engine library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m gpucode_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != DEM_OK:
        msg = lib().dem_last_error()
        raise DemError(rc, msg.decode() if msg else "")
    return rc


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(D)


def fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(F)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(I)
