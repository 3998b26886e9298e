"""Python host wrapper of the DEM engine's C ABI.

Mirrors the reference's host-side flow (main.cpp:10-404 setup, 425-623 time
loop) on top of include/dem_capi.h.  Every method calls into
libgpucode_amd.so; nothing here computes physics.
"""
import ctypes as C

import numpy as np

from . import _capi as A
from ._capi import check, dptr, iptr


def derive(inp, dia=None):
    """Host-side derived setup: returns (DemSetup, rad, pLevel, neborIdx)."""
    n = inp.NP
    S = A.DemSetup()
    rad = np.zeros(n)
    lvl = np.zeros(n, np.int32)
    idx = np.zeros(n, np.int32)
    d = None if dia is None else np.ascontiguousarray(dia, np.float64)
    check(A.lib().dem_derive(C.byref(inp), dptr(d), C.byref(S), dptr(rad), iptr(lvl), iptr(idx)))
    return S, rad, lvl, idx


def read_deck(path):
    inp = A.DemInput()
    check(A.lib().dem_read_deck(str(path).encode(), C.byref(inp)))
    return inp


def particle_feed(setup, rad, pos, n_placed, seed):
    """ParticleFeed: fills pos[n_placed:] in place; returns (count, seed)."""
    s = C.c_ulonglong(seed)
    r = A.lib().dem_particle_feed(C.byref(setup), dptr(rad), dptr(pos), int(n_placed),
                                  int(len(rad)), C.byref(s))
    if r < 0:
        check(A.DEM_ERR_ARG if r == -1 else r)
    return r, s.value


def write_tec(path, outs, time_s, pos, rad, vel, angv, force):
    n = len(pos)
    arrs = [np.ascontiguousarray(a, np.float64) for a in (pos, rad, vel, angv, force)]
    check(A.lib().dem_write_tec(str(path).encode(), int(outs), float(time_s), n,
                                *[dptr(a) for a in arrs]))


def stream_copy(nbytes=1 << 30, iters=20):
    """Attainable HBM bandwidth [GB/s] of a plain device copy (read + write)."""
    g = C.c_double()
    check(A.lib().dem_stream_copy(int(nbytes), int(iters), C.byref(g)))
    return g.value


def autotune(inp, pos, dia=None, vel=None, steps=200, cut_gaps=None, precision=A.PREC_FP64):
    """Times skins x step kernels x storage orders on this packing
    (dem_autotune); returns (best as a dict, table [skin][LOOP/COMPACT/WAVE]
    [row-major/Morton] of ms per step, -1 where skipped)."""
    pos = np.ascontiguousarray(pos, np.float64)
    v = None if vel is None else np.ascontiguousarray(vel, np.float64)
    d = None if dia is None else np.ascontiguousarray(dia, np.float64)
    cg = None if cut_gaps is None else np.ascontiguousarray(cut_gaps, np.float64)
    ncut = 5 if cg is None else len(cg)
    table = np.zeros((ncut, 3, 2))
    r = A.DemTuneResult()
    check(A.lib().dem_autotune(C.byref(inp), dptr(d), len(pos), dptr(pos), dptr(v), int(steps),
                               dptr(cg), ncut, int(precision), C.byref(r), dptr(table)))
    return {f: getattr(r, f) for f, _ in A.DemTuneResult._fields_}, table


def tune_instances(instances, precision=A.PREC_MIXED, steps=200, cut_gaps=(0.1, 0.2, 0.3, 0.5),
                   log=None, kernels=("loop", "wave")):
    """dem_autotune on each instance's own packing: every skin of cut_gaps
    (which sets the cell size and the list length) x block shape (the LOOP
    kernel, one thread per particle, or the WAVE kernel, a wave per chunk of
    64 with contact compaction) x storage order, timed for `steps` steps
    from the instance's start.  Returns one dict per instance: cutGap,
    kernel, morton, ms_per_step (the choice), ms_default (the instance's
    own configuration).  kernels: the block shapes that may be chosen (a
    batch whose instances all take one shape steps in one launch per step,
    not two)."""
    from .workloads import min_cut_gap

    kq_ok = tuple(q for q, name in ((0, "loop"), (2, "wave")) if name in kernels)

    out = []
    for k, w in enumerate(instances):
        # a wet packing's skins start at the smallest that reaches its
        # bridges' rupture distance
        g0 = min_cut_gap(w["inp"], w.get("dia"))
        gaps = sorted({float(g) for g in cut_gaps if g >= g0} | ({g0, 1.5 * g0} if g0 > 0 else set()))
        best, table = autotune(w["inp"], w["pos"], w.get("dia"), w.get("vel"), steps=steps,
                               cut_gaps=gaps, precision=precision)
        # a batch runs the LOOP and WAVE kernels; the COMPACT kernel (block
        # compaction) is a single-instance form, so it is not a candidate
        cand = [(table[s, kq, mo], s, kq, mo) for s in range(len(gaps)) for kq in kq_ok
                for mo in (0, 1) if table[s, kq, mo] >= 0]
        ms, s, kq, mo = min(cand)
        out.append(dict(cutGap=gaps[s], kernel=(Engine.KERNEL_LOOP, None, Engine.KERNEL_WAVE)[kq],
                        morton=mo, ms_per_step=float(ms), ms_default=float(best["ms_default"])))
        if log and (k + 1) % 32 == 0:
            log(f"tuned {k + 1} of {len(instances)}")
    return out


def load_para(path):
    v = [C.c_int(), C.c_int(), C.c_int()]
    check(A.lib().dem_load_para(str(path).encode(), *[C.byref(x) for x in v]))
    return tuple(x.value for x in v)


class Engine:
    """One packing on the GPU (a dem_ctx)."""

    def __init__(self, inp, dia=None, precision=A.PREC_FP64, _ctx=None):
        self._owned = _ctx is None
        if _ctx is None:
            self._ctx = C.c_void_p()
            d = None if dia is None else np.ascontiguousarray(dia, np.float64)
            check(A.lib().dem_create(C.byref(inp), dptr(d), int(precision), C.byref(self._ctx)))
        else:
            self._ctx = _ctx
        self.inp = inp
        self.NP = inp.NP
        self.setup = A.DemSetup()
        check(A.lib().dem_get_setup(self._ctx, C.byref(self.setup)))
        self.K = max(self.setup.levelNNebors[l] for l in range(self.setup.nLevels))

    def close(self):
        if self._ctx and self._owned:
            A.lib().dem_destroy(self._ctx)
        self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- state ------------------------------------------------------------
    @property
    def total(self):
        return A.lib().dem_total(self._ctx)

    def set_particles(self, pos, vel=None, angv=None):
        pos = np.ascontiguousarray(pos, np.float64)
        v = None if vel is None else np.ascontiguousarray(vel, np.float64)
        w = None if angv is None else np.ascontiguousarray(angv, np.float64)
        check(A.lib().dem_set_particles(self._ctx, len(pos), dptr(pos), dptr(v), dptr(w)))

    def feed(self, pos):
        pos = np.ascontiguousarray(pos, np.float64)
        check(A.lib().dem_feed(self._ctx, len(pos), dptr(pos)))

    def particles(self):
        n = self.total
        pos, vel, w = (np.zeros((n, 3)) for _ in range(3))
        check(A.lib().dem_get_particles(self._ctx, dptr(pos), dptr(vel), dptr(w)))
        return pos, vel, w

    def forces(self):
        n = self.total
        F, M = np.zeros((n, 3)), np.zeros((n, 3))
        Mr = np.zeros(n)
        check(A.lib().dem_get_forces(self._ctx, dptr(F), dptr(M), dptr(Mr)))
        return F, M, Mr

    def store_forces(self, on=True):
        check(A.lib().dem_set_store_forces(self._ctx, int(bool(on))))

    def pmove(self):
        a = np.zeros(self.total)
        check(A.lib().dem_get_pmove(self._ctx, dptr(a)))
        return a

    def lqdcnt(self):
        a = np.zeros(self.total, np.int32)
        check(A.lib().dem_get_lqdcnt(self._ctx, iptr(a)))
        return a

    def nebor(self, maxK=None):
        maxK = maxK or self.K
        n = self.total
        num = np.zeros(n, np.int32)
        lst = np.zeros((n, maxK), np.int32)
        check(A.lib().dem_get_nebor(self._ctx, iptr(num), iptr(lst), maxK))
        return num, lst

    def hist_pp(self, maxK=None):
        maxK = maxK or self.K
        n = self.total
        d = np.zeros((n, maxK, 3))
        f = [np.zeros((n, maxK)) for _ in range(4)]
        s, b = np.zeros((n, maxK), np.int32), np.zeros((n, maxK), np.int32)
        check(A.lib().dem_get_hist_pp(self._ctx, maxK, dptr(d), *[dptr(x) for x in f],
                                      iptr(s), iptr(b)))
        return dict(dispt=d, nDsp=f[0], nDsp_mx=f[1], plstRad=f[2], plstDfm=f[3], stage=s,
                    lqdBrg=b)

    def walls(self):
        n = self.total
        num = np.zeros(n, np.int32)
        lst = np.zeros((n, A.WALLS), np.int32)
        check(A.lib().dem_get_walls(self._ctx, iptr(num), iptr(lst)))
        return num, lst

    def hist_pw(self):
        n = self.total
        d = np.zeros((n, A.WALLS, 3))
        f = [np.zeros((n, A.WALLS)) for _ in range(4)]
        s, b = np.zeros((n, A.WALLS), np.int32), np.zeros((n, A.WALLS), np.int32)
        check(A.lib().dem_get_hist_pw(self._ctx, dptr(d), *[dptr(x) for x in f], iptr(s),
                                      iptr(b)))
        return dict(dispt=d, nDsp=f[0], nDsp_mx=f[1], plstRad=f[2], plstDfm=f[3], stage=s,
                    lqdBrg=b)

    def set_contacts(self, num, lst, hp, wnum, wlst, hw):
        """dem_set_contacts: restart lists and history, as nebor(), hist_pp(),
        walls() and hist_pw() return them."""
        i32, f64 = np.int32, np.float64
        a = [np.ascontiguousarray(x, t) for x, t in (
            (num, i32), (lst, i32), (hp["dispt"], f64), (hp["nDsp"], f64), (hp["nDsp_mx"], f64),
            (hp["plstRad"], f64), (hp["plstDfm"], f64), (hp["stage"], i32), (hp["lqdBrg"], i32),
            (wnum, i32), (wlst, i32), (hw["dispt"], f64), (hw["nDsp"], f64), (hw["nDsp_mx"], f64),
            (hw["plstRad"], f64), (hw["plstDfm"], f64), (hw["stage"], i32), (hw["lqdBrg"], i32))]
        check(A.lib().dem_set_contacts(self._ctx, a[1].shape[1],
                                       *[iptr(x) if x.dtype == i32 else dptr(x) for x in a]))

    def set_lqdcnt(self, a):
        a = np.ascontiguousarray(a, np.int32)
        check(A.lib().dem_set_lqdcnt(self._ctx, iptr(a)))

    def hash_index(self):
        n = self.total
        h, i = np.zeros(n, np.int32), np.zeros(n, np.int32)
        check(A.lib().dem_get_hash(self._ctx, iptr(h), iptr(i)))
        return h, i

    def cells(self):
        n = self.setup.sysNCells
        s, e = np.zeros(n, np.int32), np.zeros(n, np.int32)
        check(A.lib().dem_get_cells(self._ctx, iptr(s), iptr(e)))
        return s, e

    # --- stepping -----------------------------------------------------------
    def rebuild(self):
        check(A.lib().dem_rebuild(self._ctx))

    def flush_rebuild(self):
        """The rebuild the next step would start with, done now."""
        check(A.lib().dem_flush_rebuild(self._ctx))

    def step(self, n=1):
        check(A.lib().dem_step(self._ctx, int(n)))

    def step_unfused(self, n=1):
        for _ in range(n):
            check(A.lib().dem_step_unfused(self._ctx))

    def var_init(self):
        check(A.lib().dem_var_init(self._ctx))

    def pp_inter_calc(self):
        check(A.lib().dem_pp_inter_calc(self._ctx))

    def cir_pw_calc(self):
        check(A.lib().dem_cir_pw_calc(self._ctx))

    def update_system(self):
        c = C.c_int()
        check(A.lib().dem_update_system(self._ctx, C.byref(c)))
        return c.value

    def time_steps(self, n):
        ms, rb = C.c_double(), C.c_int()
        check(A.lib().dem_time_steps(self._ctx, int(n), C.byref(ms), C.byref(rb)))
        return ms.value, bool(rb.value)

    def sync(self):
        check(A.lib().dem_sync(self._ctx))

    KERNEL_AUTO, KERNEL_LOOP, KERNEL_COMPACT, KERNEL_WAVE = 0, 1, 2, 3

    def set_kernel(self, variant):
        check(A.lib().dem_set_kernel(self._ctx, int(variant)))

    def set_order(self, morton):
        """Storage order between rebuilds: Morton (True) or row-major."""
        check(A.lib().dem_set_order(self._ctx, int(bool(morton))))

    def set_relayout(self, every):
        """Steps between history layout refreshes in step() (0: never)."""
        check(A.lib().dem_set_relayout(self._ctx, int(every)))

    @property
    def n_relayout(self):
        return A.lib().dem_n_relayout(self._ctx)

    def set_wave_split(self, parts):
        """Wave kernel block shape: waves per chunk of 64 particles (1, 2, 4;
        0: one)."""
        check(A.lib().dem_set_wave_split(self._ctx, int(parts)))

    def set_pair_records(self, on):
        """Shared pair records for the wave kernel (default on): one history
        record per pair of equal radii instead of two mirrored ones."""
        check(A.lib().dem_set_pair_records(self._ctx, int(bool(on))))

    @property
    def pair_records_used(self):
        return bool(A.lib().dem_pair_records_used(self._ctx))

    @property
    def iter(self):
        return A.lib().dem_iter(self._ctx)

    @property
    def ctime(self):
        return A.lib().dem_ctime(self._ctx)

    @property
    def n_rebuild(self):
        return A.lib().dem_n_rebuild(self._ctx)

    def stats(self):
        s = A.DemStats()
        check(A.lib().dem_contact_stats(self._ctx, C.byref(s)))
        return {f: getattr(s, f) for f, _ in A.DemStats._fields_}

    def timing(self):
        ms, st, la, rms, nr = (C.c_double(), C.c_longlong(), C.c_longlong(), C.c_double(),
                               C.c_int())
        check(A.lib().dem_get_timing(self._ctx, C.byref(ms), C.byref(st), C.byref(la),
                                     C.byref(rms), C.byref(nr)))
        return dict(fused_ms=ms.value, fused_steps=st.value, fused_launches=la.value,
                    rebuild_ms=rms.value, n_rebuild=nr.value)

    def device_bytes(self):
        return A.lib().dem_device_bytes(self._ctx)

    def analyze(self):
        """(phi, Z, top) of the current packing (GPU analysis kernel)."""
        v = [C.c_double(), C.c_double(), C.c_double()]
        check(A.lib().dem_analyze(self._ctx, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def step_kernel_name(self):
        """Name of the step kernel dem_step launches for this packing."""
        return A.lib().dem_step_kernel_name(self._ctx).decode()

    # --- restart --------------------------------------------------------------
    REBUILD_STAGES = ("save", "hash", "sort", "permute", "cell_range", "neighbor_find",
                      "active_slots", "wall_list", "hist_remap")

    def scale_velocities(self, factor):
        check(A.lib().dem_scale_velocities(self._ctx, float(factor)))

    NSEARCH_AUTO, NSEARCH_PLAIN, NSEARCH_TILED = 0, 1, 2

    def set_neighbor_search(self, mode):
        check(A.lib().dem_set_neighbor_search(self._ctx, int(mode)))

    def neighbor_search_used(self):
        return A.lib().dem_neighbor_search_used(self._ctx)

    def key_bits(self):
        return A.lib().dem_key_bits(self._ctx)

    def profile_rebuild(self, reps=3):
        """Mean ms per rebuild stage (dem_profile_rebuild), by stage name."""
        ms = np.zeros(len(self.REBUILD_STAGES))
        check(A.lib().dem_profile_rebuild(self._ctx, int(reps), dptr(ms)))
        return dict(zip(self.REBUILD_STAGES, ms.tolist()))

    def dump(self, path, nBatch=1, binary=False):
        f = A.lib().dem_dump_write_binary if binary else A.lib().dem_dump_write
        check(f(self._ctx, str(path).encode(), int(nBatch)))

    def load(self, path):
        nb = C.c_int()
        check(A.lib().dem_dump_read(self._ctx, str(path).encode(), C.byref(nb)))
        return nb.value


class Batch:
    """Many packings stepped together (a dem_batch); instance k is an Engine
    view owned by the batch."""

    def __init__(self, inputs, dias=None, precision=A.PREC_FP64):
        n = len(inputs)
        arr = (A.DemInput * n)(*inputs)
        self._keep = []
        dptrs = None
        if dias is not None:
            dptrs = (A.D * n)()
            for k, d in enumerate(dias):
                if d is not None:
                    a = np.ascontiguousarray(d, np.float64)
                    self._keep.append(a)
                    dptrs[k] = dptr(a)
        self._b = C.c_void_p()
        check(A.lib().dem_batch_create(n, arr, dptrs, int(precision), C.byref(self._b)))
        self.inst = [Engine(inputs[k], None, precision, _ctx=C.c_void_p(
            A.lib().dem_batch_instance(self._b, k))) for k in range(n)]

    @classmethod
    def from_instances(cls, instances, choices=None, precision=A.PREC_FP64):
        """A batch of workload instances (dicts with inp, dia, pos and
        optionally vel), each with its tuned skin, step kernel and storage
        order when choices (from tune_instances) are given, particles set."""
        inputs = []
        for k, w in enumerate(instances):
            inp = A.DemInput.from_buffer_copy(w["inp"])
            if choices is not None:
                inp.cutGap = choices[k]["cutGap"]
            inputs.append(inp)
        b = cls(inputs, [w.get("dia") for w in instances], precision)
        for k, w in enumerate(instances):
            if choices is not None:
                b[k].set_kernel(choices[k]["kernel"])
                b[k].set_order(choices[k]["morton"])
            b[k].set_particles(w["pos"], w.get("vel"))
        return b

    def __len__(self):
        return len(self.inst)

    def __getitem__(self, k):
        return self.inst[k]

    def close(self):
        if self._b:
            for e in self.inst:
                e._ctx = C.c_void_p()
            A.lib().dem_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def flush_rebuilds(self):
        check(A.lib().dem_batch_flush_rebuilds(self._b))

    def set_schedule(self, device=True, slot_every=4):
        check(A.lib().dem_batch_set_schedule(self._b, int(bool(device)), int(slot_every)))

    def schedule_stats(self):
        a, b = C.c_longlong(), C.c_longlong()
        check(A.lib().dem_batch_schedule_stats(self._b, C.byref(a), C.byref(b)))
        return {"slots": a.value, "host_syncs": b.value}

    def step(self, n):
        check(A.lib().dem_batch_step(self._b, int(n)))

    def analyze(self):
        n = len(self.inst)
        phi, Z, top = np.zeros(n), np.zeros(n), np.zeros(n)
        check(A.lib().dem_batch_analyze(self._b, dptr(phi), dptr(Z), dptr(top)))
        return phi, Z, top

    def timing(self):
        ms, la, rms, nr = C.c_double(), C.c_longlong(), C.c_double(), C.c_longlong()
        check(A.lib().dem_batch_timing(self._b, C.byref(ms), C.byref(la), C.byref(rms),
                                       C.byref(nr)))
        return dict(step_ms=ms.value, launches=la.value, rebuild_ms=rms.value,
                    n_rebuild=nr.value)
