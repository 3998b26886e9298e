"""ctypes wrapper of the FP64 CPU oracle (oracle/libdem_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / CPU baseline.  The product
(gpucode_amd) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdem_oracle.so")
ML = 8
WALLS = 3


class OrcInput(C.Structure):
    _fields_ = [
        ("NP", C.c_int), ("nLevels", C.c_int), ("loadStage", C.c_int), ("forceCap", C.c_int),
        ("tPacking", C.c_double), ("realSize", C.c_double), ("cutGap", C.c_double),
        ("expandSize", C.c_double),
        ("cirX", C.c_double), ("cirY", C.c_double), ("Bw", C.c_double), ("R", C.c_double),
        ("H", C.c_double), ("feedLmin", C.c_double), ("feedLmax", C.c_double),
        ("mat", (C.c_double * 7) * 2),
        ("stension", C.c_double), ("vol", C.c_double), ("gapmn", C.c_double),
        ("dtFactor", C.c_double), ("nOutPacking", C.c_int), ("sizeLevelInput", C.c_int),
        ("levelDia", C.c_double * ML),
        ("gravity", C.c_double), ("dtOverride", C.c_double), ("theta", C.c_double),
        ("ruptureMode", C.c_int), ("gridClampQuirk", C.c_int),
    ]


class OrcCalpara(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("emod", "pois", "sfrc", "ystress", "ydsp_t", "yforce_t", "ystiff_t")]


class OrcSetup(C.Structure):
    _fields_ = (
        [(n, C.c_double) for n in ("rlunit", "rmunit", "rfunit", "rsunit", "rtunit", "rvunit",
                                   "reunit", "dt", "searchGap")]
        + [("origin", C.c_double * 3), ("endPoint", C.c_double * 3)]
        + [(n, C.c_double) for n in ("cirX", "cirY", "R", "Tw", "Bw", "topv", "btmv",
                                     "feedLmin", "feedLmax", "tPacking")]
        + [(n, C.c_double * 2) for n in ("density", "emod", "ymod", "pois", "sfrc", "rfrc",
                                         "dmpn", "yldp")]
        + [(n, C.c_double) for n in ("stension", "vol", "gapmn", "layer", "brknmx", "theta")]
        + [("pp", OrcCalpara), ("pw", OrcCalpara)]
        + [(n, C.c_int) for n in ("nLevels", "forceCap", "loadStage", "ruptureMode",
                                  "gridClampQuirk", "nPPCnt")]
        + [("levelSize", C.c_double * ML), ("levelNNebors", C.c_int * ML),
           ("levelNPars", C.c_int * ML), ("levelNebSId", C.c_int * ML),
           ("cellSize", C.c_double * ML), ("gridSize", (C.c_int * 3) * ML),
           ("numCells", C.c_int * ML), ("ncellIdx", C.c_int * ML),
           ("sysNCells", C.c_int), ("radMin", C.c_double), ("radMax", C.c_double)]
    )


class OrcHist1(C.Structure):
    _fields_ = [("dispt", C.c_double * 3), ("nDsp", C.c_double), ("nDsp_mx", C.c_double),
                ("plstRad", C.c_double), ("plstDfm", C.c_double), ("stage", C.c_int),
                ("lqdBrg", C.c_int)]


D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)
P = C.c_void_p
_SIGS = {
    "orc_derive": (C.c_int, [C.POINTER(OrcInput), D, C.POINTER(OrcSetup), D, I, I]),
    "orc_contact_pp": (None, [C.POINTER(OrcSetup), C.POINTER(OrcHist1), C.c_double, D,
                              C.c_double, D, D, C.c_double, D, D, C.c_int, D, D, D]),
    "orc_contact_pw": (None, [C.POINTER(OrcSetup), C.POINTER(OrcHist1), C.c_double, D,
                              C.c_double, D, D, D, C.c_int, D, D, D]),
    "orc_fcap": (C.c_double, [C.POINTER(OrcSetup), C.c_double, C.c_double, C.c_double]),
    "orc_fcap_coef": (C.c_double, [D, C.POINTER(OrcSetup), C.c_double, C.c_double,
                                   C.c_double]),
    "orc_rupture": (C.c_double, [C.POINTER(OrcSetup), C.c_double]),
    "orc_exp": (C.c_double, [C.c_double]),
    "orc_log": (C.c_double, [C.c_double]),
    "orc_pow": (C.c_double, [C.c_double, C.c_double]),
    "orc_create": (P, [C.POINTER(OrcInput), D]),
    "orc_free": (None, [P]),
    "orc_get_setup": (C.POINTER(OrcSetup), [P]),
    "orc_num_threads": (C.c_int, [C.c_int]),
    "orc_set_particles": (None, [P, C.c_int, D, D, D]),
    "orc_feed": (None, [P, C.c_int, D]),
    "orc_get_particles": (None, [P, D, D, D]),
    "orc_get_forces": (None, [P, D, D, D]),
    "orc_get_pmove": (None, [P, D]),
    "orc_get_lqdcnt": (None, [P, I]),
    "orc_total": (C.c_int, [P]),
    "orc_iter": (C.c_int, [P]),
    "orc_ctime": (C.c_double, [P]),
    "orc_tw": (C.c_double, [P]),
    "orc_n_rebuild": (C.c_int, [P]),
    "orc_overflow": (C.c_int, [P]),
    "orc_get_nebor": (C.c_int, [P, I, I, C.c_int]),
    "orc_get_hist_pp": (C.c_int, [P, C.c_int, D, D, D, D, D, I, I]),
    "orc_get_walls": (None, [P, I, I]),
    "orc_get_hist_pw": (None, [P, D, D, D, D, D, I, I]),
    "orc_get_hash": (None, [P, I, I]),
    "orc_get_cells": (None, [P, I, I]),
    "orc_set_contacts": (C.c_int, [P, C.c_int, I, I, D, D, D, D, D, I, I, I, I, D, D, D, D, D,
                                   I, I]),
    "orc_set_lqdcnt": (None, [P, I]),
    "orc_rebuild": (C.c_int, [P]),
    "orc_step": (C.c_int, [P]),
    "orc_run": (C.c_int, [P, C.c_int]),
    "orc_force_rebuild": (None, [P]),
    "orc_time_force_steps": (C.c_double, [P, C.c_int]),
    "orc_var_init": (None, [P]),
    "orc_pp_force": (None, [P]),
    "orc_pw_force": (None, [P]),
    "orc_update": (C.c_int, [P]),
}
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for n, (r, a) in _SIGS.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _LIB = L
    return _LIB


def _d(a):
    return None if a is None else a.ctypes.data_as(D)


def _i(a):
    return None if a is None else a.ctypes.data_as(I)


def _contact_args(num, lst, hp, wnum, wlst, hw):
    f = lambda a, t: np.ascontiguousarray(a, t)  # noqa: E731
    i32, f64 = np.int32, np.float64
    return [f(num, i32), f(lst, i32), f(hp["dispt"], f64), f(hp["nDsp"], f64),
            f(hp["nDsp_mx"], f64), f(hp["plstRad"], f64), f(hp["plstDfm"], f64),
            f(hp["stage"], i32), f(hp["lqdBrg"], i32), f(wnum, i32), f(wlst, i32),
            f(hw["dispt"], f64), f(hw["nDsp"], f64), f(hw["nDsp_mx"], f64),
            f(hw["plstRad"], f64), f(hw["plstDfm"], f64), f(hw["stage"], i32),
            f(hw["lqdBrg"], i32)]


def as_input(inp):
    """Copies any ctypes structure with the dem_input layout."""
    return OrcInput.from_buffer_copy(bytes(inp))


def derive(inp, dia=None):
    inp = as_input(inp)
    S = OrcSetup()
    n = inp.NP
    rad, lvl, idx = np.zeros(n), np.zeros(n, np.int32), np.zeros(n, np.int32)
    d = None if dia is None else np.ascontiguousarray(dia, np.float64)
    rc = lib().orc_derive(C.byref(inp), _d(d), C.byref(S), _d(rad), _i(lvl), _i(idx))
    if rc != 0:
        raise ValueError(f"orc_derive failed ({rc})")
    return S, rad, lvl, idx


def hist1(**kw):
    h = OrcHist1()
    for k, v in kw.items():
        if k == "dispt":
            h.dispt[:] = list(v)
        else:
            setattr(h, k, v)
    return h


def contact_pp(S, h, nrmDsp, n, ri, vi, wi, rj, vj, wj, stage=1):
    F, M, Mr = np.zeros(3), np.zeros(3), C.c_double()
    a = [np.ascontiguousarray(x, np.float64) for x in (n, vi, wi, vj, wj)]
    lib().orc_contact_pp(C.byref(S), C.byref(h), nrmDsp, _d(a[0]), ri, _d(a[1]), _d(a[2]), rj,
                         _d(a[3]), _d(a[4]), stage, _d(F), _d(M), C.byref(Mr))
    return F, M, Mr.value


def contact_pw(S, h, nrmDsp, n, ri, vi, wi, vj, stage=1):
    F, M, Mr = np.zeros(3), np.zeros(3), C.c_double()
    a = [np.ascontiguousarray(x, np.float64) for x in (n, vi, wi, vj)]
    lib().orc_contact_pw(C.byref(S), C.byref(h), nrmDsp, _d(a[0]), ri, _d(a[1]), _d(a[2]),
                         _d(a[3]), stage, _d(F), _d(M), C.byref(Mr))
    return F, M, Mr.value


class Oracle:
    """Whole-state FP64 oracle of one packing."""

    def __init__(self, inp, dia=None):
        self.inp = as_input(inp)
        d = None if dia is None else np.ascontiguousarray(dia, np.float64)
        self.st = lib().orc_create(C.byref(self.inp), _d(d))
        if not self.st:
            raise ValueError("orc_create failed")
        self.S = lib().orc_get_setup(self.st).contents
        self.NP = self.inp.NP
        self.K = max(self.S.levelNNebors[l] for l in range(self.S.nLevels))

    def close(self):
        if self.st:
            lib().orc_free(self.st)
            self.st = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def total(self):
        return lib().orc_total(self.st)

    @property
    def iter(self):
        return lib().orc_iter(self.st)

    @property
    def ctime(self):
        return lib().orc_ctime(self.st)

    @property
    def n_rebuild(self):
        return lib().orc_n_rebuild(self.st)

    def set_particles(self, pos, vel=None, angv=None):
        pos = np.ascontiguousarray(pos, np.float64)
        v = None if vel is None else np.ascontiguousarray(vel, np.float64)
        w = None if angv is None else np.ascontiguousarray(angv, np.float64)
        lib().orc_set_particles(self.st, len(pos), _d(pos), _d(v), _d(w))

    def feed(self, pos):
        pos = np.ascontiguousarray(pos, np.float64)
        lib().orc_feed(self.st, len(pos), _d(pos))

    def particles(self):
        n = self.total
        p, v, w = (np.zeros((n, 3)) for _ in range(3))
        lib().orc_get_particles(self.st, _d(p), _d(v), _d(w))
        return p, v, w

    def forces(self):
        n = self.total
        F, M, Mr = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
        lib().orc_get_forces(self.st, _d(F), _d(M), _d(Mr))
        return F, M, Mr

    def pmove(self):
        a = np.zeros(self.total)
        lib().orc_get_pmove(self.st, _d(a))
        return a

    def lqdcnt(self):
        a = np.zeros(self.total, np.int32)
        lib().orc_get_lqdcnt(self.st, _i(a))
        return a

    def nebor(self, maxK=None):
        maxK = maxK or self.K
        n = self.total
        num, lst = np.zeros(n, np.int32), np.zeros((n, maxK), np.int32)
        if lib().orc_get_nebor(self.st, _i(num), _i(lst), maxK) != 0:
            raise ValueError("maxK too small")
        return num, lst

    def hist_pp(self, maxK=None):
        maxK = maxK or self.K
        n = self.total
        d = np.zeros((n, maxK, 3))
        f = [np.zeros((n, maxK)) for _ in range(4)]
        s, b = np.zeros((n, maxK), np.int32), np.zeros((n, maxK), np.int32)
        lib().orc_get_hist_pp(self.st, maxK, _d(d), *[_d(x) for x in f], _i(s), _i(b))
        return dict(dispt=d, nDsp=f[0], nDsp_mx=f[1], plstRad=f[2], plstDfm=f[3], stage=s,
                    lqdBrg=b)

    def walls(self):
        n = self.total
        num, lst = np.zeros(n, np.int32), np.zeros((n, WALLS), np.int32)
        lib().orc_get_walls(self.st, _i(num), _i(lst))
        return num, lst

    def hist_pw(self):
        n = self.total
        d = np.zeros((n, WALLS, 3))
        f = [np.zeros((n, WALLS)) for _ in range(4)]
        s, b = np.zeros((n, WALLS), np.int32), np.zeros((n, WALLS), np.int32)
        lib().orc_get_hist_pw(self.st, _d(d), *[_d(x) for x in f], _i(s), _i(b))
        return dict(dispt=d, nDsp=f[0], nDsp_mx=f[1], plstRad=f[2], plstDfm=f[3], stage=s,
                    lqdBrg=b)

    def set_contacts(self, num, lst, hp, wnum, wlst, hw):
        """Restart state (lists, history) as the getters return it."""
        c = _contact_args(num, lst, hp, wnum, wlst, hw)
        if lib().orc_set_contacts(self.st, lst.shape[1], *[_i(a) if a.dtype == np.int32 else _d(a)
                                                           for a in c]) != 0:
            raise ValueError("a list is longer than its size level allows")

    def set_lqdcnt(self, a):
        a = np.ascontiguousarray(a, np.int32)
        lib().orc_set_lqdcnt(self.st, _i(a))

    def hash_index(self):
        n = self.total
        h, i = np.zeros(n, np.int32), np.zeros(n, np.int32)
        lib().orc_get_hash(self.st, _i(h), _i(i))
        return h, i

    def cells(self):
        n = self.S.sysNCells
        s, e = np.zeros(n, np.int32), np.zeros(n, np.int32)
        lib().orc_get_cells(self.st, _i(s), _i(e))
        return s, e

    def rebuild(self):
        if lib().orc_rebuild(self.st) != 0:
            raise ValueError("neighbour list overflow")

    def step(self, n=1):
        if lib().orc_run(self.st, int(n)) != 0:
            raise ValueError("neighbour list overflow")

    def force_rebuild(self):
        lib().orc_force_rebuild(self.st)

    def time_force_steps(self, n):
        return lib().orc_time_force_steps(self.st, int(n))


def threads(n=0):
    return lib().orc_num_threads(int(n))
