"""GPU parity: the HIP engine (through the C ABI) against the FP64 oracle on
identical seeded inputs.

Bars (also stated per test):
  * integer work (cell hash, sorted index, cell ranges, neighbour and wall
    lists, history slots) -- bit-exact, including slot order;
  * FP64 path, one step -- per-particle force error / (sum_j |f_ij| + m g)
    <= 1e-12 (the GPU uses x*sqrt(x) and sqrt where the reference calls pow,
    a few ulp apart);
  * FP64 path, N steps -- positions within 1e-9 d, velocities within 1e-7 of
    the largest speed (differences grow only from those ulps).
"""
import math

import numpy as np
import pytest

import gpucode_amd as G
from gpucode_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def pair(w, precision=G.PREC_FP64, vel=None, angv=None):
    e = G.Engine(w["inp"], w["dia"], precision)
    o = O.Oracle(w["inp"], w["dia"])
    e.set_particles(w["pos"], vel, angv)
    o.set_particles(w["pos"], vel, angv)
    return e, o


def force_scale(o, e_forces=None):
    """sum_j |f_ij| + m g per particle from the oracle's contact magnitudes."""
    F, M, Mr = o.forces()
    S = o.S
    pos = o.particles()[0]
    num, lst = o.nebor()
    h = o.hist_pp()
    n = len(pos)
    scale = np.zeros(n)
    rad = None
    # bound |f_ij| by the normal force magnitude implied by the overlap
    p = pos
    for i in range(n):
        for k in range(num[i]):
            j = lst[i, k]
            d = np.linalg.norm(p[i] - p[j])
            ov = 1.0 - d if rad is None else rad[i] + rad[j] - d
            if ov > 0:
                scale[i] += 4 / 3 * S.pp.emod * math.sqrt(0.25) * ov ** 1.5 * (1 + S.pp.sfrc)
    return scale + 1.0


def test_hash_sort_cells_bit_exact():
    w = W.small_bed(n=3000)
    e, o = pair(w)
    e.rebuild()
    o.rebuild()
    h1, i1 = e.hash_index()
    h2, i2 = o.hash_index()
    assert np.array_equal(h1, h2) and np.array_equal(i1, i2)
    s1, t1 = e.cells()
    s2, t2 = o.cells()
    assert np.array_equal(s1, s2) and np.array_equal(t1, t2)


def test_hash_sort_polydisperse_levels():
    w = W.wet_poly(n=4000, R=10e-3, H=60e-3, nLevels=3)
    e, o = pair(w)
    e.rebuild()
    o.rebuild()
    assert np.array_equal(e.hash_index()[0], o.hash_index()[0])
    assert np.array_equal(e.hash_index()[1], o.hash_index()[1])
    for a, b in zip(e.cells(), o.cells()):
        assert np.array_equal(a, b)
    n1, l1 = e.nebor(o.K)
    n2, l2 = o.nebor()
    assert np.array_equal(n1, n2) and np.array_equal(l1, l2)


def test_neighbour_and_wall_lists_bit_exact():
    w = W.small_bed(n=3000)
    e, o = pair(w)
    e.rebuild()
    o.rebuild()
    n1, l1 = e.nebor()
    n2, l2 = o.nebor()
    assert np.array_equal(n1, n2)
    assert np.array_equal(l1, l2)  # same slot order, as original ids
    w1, wl1 = e.walls()
    w2, wl2 = o.walls()
    assert np.array_equal(w1, w2) and np.array_equal(wl1, wl2)


@pytest.mark.parametrize("make", [
    lambda: W.small_bed(n=3000),
    lambda: W.pour_2M(n=20000, R=10e-3, H=30e-3),
], ids=["dense_bed", "sparse_pour"])
def test_tiled_and_plain_neighbor_search_equal(make):
    # the LDS-staged tile search and the one-thread search give the same
    # lists, hence the same steps, rebuilds included
    w = make()
    out = {}
    for mode in (G.Engine.NSEARCH_PLAIN, G.Engine.NSEARCH_TILED):
        e = G.Engine(w["inp"])
        e.set_neighbor_search(mode)
        e.set_particles(w["pos"], w.get("vel"))
        e.step(1000)
        assert e.neighbor_search_used() == mode
        out[mode] = (e.nebor(), e.particles(), e.n_rebuild)
        e.close()
    a, b = out[G.Engine.NSEARCH_PLAIN], out[G.Engine.NSEARCH_TILED]
    assert a[2] == b[2] and a[2] >= (2 if w["name"] == "pour_2M" else 1)
    for u, v in zip(a[0] + a[1], b[0] + b[1]):
        assert np.array_equal(u, v)


def test_tiled_search_crowded_tiles_fall_back():
    # 2,000 spheres crowded into a cube of 6.6 d (about ten per cell): a
    # tile's rim then holds more candidates than its LDS stage, those tiles
    # scan global memory, and the lists overflow their 35 slots: reported,
    # as with the one-thread search
    w = W.small_bed(n=2000)
    rng = np.random.default_rng(3)
    pos = rng.uniform(0.0, 6.6, (2000, 3)) + np.array([-3.3, -3.3, 0.6])
    for mode in (G.Engine.NSEARCH_PLAIN, G.Engine.NSEARCH_TILED):
        e = G.Engine(w["inp"])
        e.set_neighbor_search(mode)
        e.set_particles(pos)
        with pytest.raises(G.DemError) as err:
            e.step(1)
        assert err.value.code == -3
        e.close()


def test_radix_sort_edge_sizes():
    # tile boundaries of the sort (2048 keys per block) and a single particle
    for n in (1, 2, 2047, 2048, 2049, 5000):
        w = W.small_bed(n=max(n, 2))
        pos = w["pos"][:n]
        e = G.Engine(w["inp"])
        o = O.Oracle(w["inp"])
        e.set_particles(pos)
        o.set_particles(pos)
        e.rebuild()
        o.rebuild()
        assert np.array_equal(e.hash_index()[1], o.hash_index()[1]), n
        assert np.array_equal(e.nebor()[1], o.nebor()[1]), n
        e.close()


@pytest.mark.parametrize("n", [0, 1, 2, 63, 64, 65, 129, 700])
def test_ragged_sizes_step_bit_exact(n):
    # an empty system, a lone sphere and counts around the 64-particle wave
    # (a last wave with one particle, an exact number of waves, a few
    # waves per XCD): 300 steps of the default step kernel with fast random
    # velocities (several rebuilds requested on the device) equal the
    # oracle bit for bit
    w = W.small_bed(n=max(n, 2))
    pos = w["pos"][:n]
    vel = np.random.default_rng(n).normal(size=(n, 3)) * 50.0
    e = G.Engine(w["inp"])
    o = O.Oracle(w["inp"])
    e.set_particles(pos, vel)
    o.set_particles(pos, vel)
    e.step(300)
    o.step(300)
    assert e.iter == o.iter == 300 and e.total == n
    assert e.n_rebuild == o.n_rebuild
    for u, v in zip(e.particles(), o.particles()):
        assert np.array_equal(u, v)
    e.close()


def one_step_forces(w, vel=None, angv=None, precision=G.PREC_FP64):
    e, o = pair(w, precision, vel, angv)
    e.store_forces(True)
    e.step(1)
    o.step(1)
    return e, o


def normalized_force_error(e, o):
    F1, M1, Mr1 = e.forces()
    F2, M2, Mr2 = o.forces()
    sc = force_scale(o)
    errF = np.linalg.norm(F1 - F2, axis=1) / sc
    return errF, (F1, M1, Mr1), (F2, M2, Mr2), sc


def test_one_step_forces_fp64_dense_bed():
    w = W.small_bed(n=3000)
    rng = np.random.default_rng(5)
    vel = rng.normal(size=(3000, 3)) * 1e-3
    angv = rng.normal(size=(3000, 3)) * 1e-2
    e, o = one_step_forces(w, vel, angv)
    err, g, r, sc = normalized_force_error(e, o)
    assert err.max() <= 1e-12, err.max()
    # moments and rolling terms, relative to the same scale times the radius
    assert (np.linalg.norm(g[1] - r[1], axis=1) / sc).max() <= 1e-12
    assert (np.abs(g[2] - r[2]) / sc).max() <= 1e-12
    # the history written by that step
    h1, h2 = e.hist_pp(), o.hist_pp()
    assert np.array_equal(h1["stage"], h2["stage"])
    for f in ("nDsp", "nDsp_mx", "plstRad", "plstDfm"):
        assert np.allclose(h1[f], h2[f], rtol=1e-13, atol=1e-18), f
    assert np.allclose(h1["dispt"], h2["dispt"], rtol=1e-12, atol=1e-20)


def compare_state(e, o, pos_tol=1e-9, vel_rel=1e-7):
    p1, v1, w1 = e.particles()
    p2, v2, w2 = o.particles()
    vs = max(np.abs(v2).max(), 1e-12)
    ws = max(np.abs(w2).max(), 1e-12)
    assert np.abs(p1 - p2).max() <= pos_tol, np.abs(p1 - p2).max()
    assert np.abs(v1 - v2).max() <= vel_rel * vs, np.abs(v1 - v2).max() / vs
    assert np.abs(w1 - w2).max() <= 1e-6 * ws, np.abs(w1 - w2).max() / ws


def test_trajectory_fp64_dense_bed_300_steps():
    w = W.small_bed(n=3000)
    e, o = pair(w)
    e.step(300)
    o.step(300)
    compare_state(e, o)
    assert e.iter == o.iter == 300
    assert e.n_rebuild == o.n_rebuild
    n1, l1 = e.nebor()
    n2, l2 = o.nebor()
    assert np.array_equal(n1, n2) and np.array_equal(l1, l2)
    s = e.stats()
    assert s["n_contact"] > 3000  # a contact-heavy case


def test_fused_equals_separate_kernels():
    w = W.small_bed(n=2000)
    rng = np.random.default_rng(2)
    vel = rng.normal(size=(2000, 3)) * 1e-2
    a = G.Engine(w["inp"])
    b = G.Engine(w["inp"])
    for x in (a, b):
        x.set_particles(w["pos"], vel)
    a.step(100)
    b.step_unfused(100)
    for u, v in zip(a.particles(), b.particles()):
        assert np.array_equal(u, v)
    assert a.n_rebuild == b.n_rebuild


@pytest.mark.parametrize("kernel", [G.Engine.KERNEL_COMPACT, G.Engine.KERNEL_WAVE])
@pytest.mark.parametrize("precision", [G.PREC_FP64, G.PREC_MIXED])
def test_compacting_kernel_equals_loop_kernel(precision, kernel):
    # the compacting kernels sum each particle's contacts in slot order, so
    # they must reproduce the one-thread loop bit for bit, rebuilds included
    w = W.small_bed(n=3000)
    rng = np.random.default_rng(4)
    vel = rng.normal(size=(3000, 3)) * 50.0  # about 5 m/s: a rebuild every ~150 steps
    a = G.Engine(w["inp"], None, precision)
    b = G.Engine(w["inp"], None, precision)
    a.set_kernel(kernel)
    b.set_kernel(G.Engine.KERNEL_LOOP)
    for x in (a, b):
        x.set_particles(w["pos"], vel)
        x.store_forces(True)
    a.step(400)
    b.step(400)
    assert a.n_rebuild == b.n_rebuild and a.n_rebuild > 1
    for u, v in zip(a.particles(), b.particles()):
        assert np.array_equal(u, v)
    for u, v in zip(a.forces(), b.forces()):
        assert np.array_equal(u, v)
    ha, hb = a.hist_pp(), b.hist_pp()
    for f in ha:
        assert np.array_equal(ha[f], hb[f]), f


@pytest.mark.parametrize("kernel", [G.Engine.KERNEL_LOOP, G.Engine.KERNEL_COMPACT,
                                    G.Engine.KERNEL_WAVE])
def test_layout_refresh_changes_nothing(kernel):
    # a layout refresh only moves history records: with refreshes every 64
    # steps between rebuilds, the run is bit-identical to one without, and
    # both match the oracle
    w = W.small_bed(n=3000)
    rng = np.random.default_rng(5)
    vel = rng.normal(size=(3000, 3)) * 20.0
    a = G.Engine(w["inp"])
    b = G.Engine(w["inp"])
    o = O.Oracle(w["inp"])
    for x in (a, b):
        x.set_kernel(kernel)
        x.set_particles(w["pos"], vel)
        x.store_forces(True)
    o.set_particles(w["pos"], vel)
    a.set_relayout(64)
    b.set_relayout(0)
    a.step(640)
    b.step(640)
    o.step(640)
    assert a.n_relayout >= 3 and b.n_relayout == 0
    assert a.n_rebuild == b.n_rebuild == o.n_rebuild
    for u, v in zip(a.particles(), b.particles()):
        assert np.array_equal(u, v)
    for u, v in zip(a.forces(), b.forces()):
        assert np.array_equal(u, v)
    ha, hb = a.hist_pp(), b.hist_pp()
    for f in ha:
        assert np.array_equal(ha[f], hb[f]), f
    compare_state(a, o)


def test_layout_refresh_wet_polydisperse():
    w = W.wet_bed(n=3000, poly=True, reference_law=False, theta_deg=20.0, liquid=0.2)
    a = G.Engine(w["inp"], w["dia"])
    b = G.Engine(w["inp"], w["dia"])
    for x in (a, b):
        x.set_particles(w["pos"])
    a.set_relayout(64)
    b.set_relayout(0)
    a.step(400)
    b.step(400)
    assert a.n_relayout >= 3
    for u, v in zip(a.particles(), b.particles()):
        assert np.array_equal(u, v)
    ha, hb = a.hist_pp(), b.hist_pp()
    for f in ha:
        assert np.array_equal(ha[f], hb[f]), f


def test_morton_storage_equals_row_major_storage():
    # storage order changes memory locality only: lists (as original ids),
    # the reference's hash/cell tables and the trajectory are identical
    w = W.wet_bed(n=3000, poly=True, reference_law=False, theta_deg=20.0, liquid=0.2)
    a = G.Engine(w["inp"], w["dia"])
    b = G.Engine(w["inp"], w["dia"])
    a.set_order(True)
    b.set_order(False)
    for x in (a, b):
        x.set_particles(w["pos"])
    a.rebuild()
    b.rebuild()
    assert np.array_equal(a.hash_index()[0], b.hash_index()[0])
    assert np.array_equal(a.hash_index()[1], b.hash_index()[1])
    for u, v in zip(a.cells(), b.cells()):
        assert np.array_equal(u, v)
    a.step(200)
    b.step(200)
    for u, v in zip(a.particles(), b.particles()):
        assert np.array_equal(u, v)
    for u, v in zip(a.nebor(), b.nebor()):
        assert np.array_equal(u, v)


def test_dry_column_20000_steps():
    w = W.dry_column()
    e, o = pair(w)
    O.threads(16)
    e.step(w["steps"])
    o.step(w["steps"])
    compare_state(e, o, pos_tol=1e-8, vel_rel=1e-6)
    assert e.n_rebuild == o.n_rebuild


def test_rebuilds_triggered_on_device_match_oracle():
    # fast particles (about 1 m/s): the path-length trigger 0.3 * SearchGap
    # fires every ~100 steps, inside the 64-step launch chunks
    w = W.dry_column()
    rng = np.random.default_rng(11)
    vel = rng.normal(size=(2000, 3)) * 10.0
    e, o = pair(w, vel=vel)
    e.step(3000)
    o.step(3000)
    assert o.n_rebuild > 20
    assert e.n_rebuild == o.n_rebuild
    compare_state(e, o, pos_tol=1e-8, vel_rel=1e-6)


def test_wet_bed_capillary_parity():
    # monodisperse touching bed: bridges on every contact (reference law)
    w = W.wet_bed(n=3000)
    e, o = pair(w)
    e.step(200)
    o.step(200)
    compare_state(e, o, pos_tol=1e-9, vel_rel=1e-6)
    assert np.array_equal(e.lqdcnt(), o.lqdcnt())
    h1, h2 = e.hist_pp(o.K), o.hist_pp()
    assert np.array_equal(h1["lqdBrg"], h2["lqdBrg"])
    assert h2["lqdBrg"].sum() > 3000


def test_wet_polydisperse_levels_parity():
    # three size levels, cross-level search, theta = 20 deg law
    w = W.wet_bed(n=3000, poly=True, reference_law=False, theta_deg=20.0, liquid=0.2)
    e, o = pair(w)
    e.step(200)
    o.step(200)
    compare_state(e, o, pos_tol=1e-9, vel_rel=1e-6)
    n1, l1 = e.nebor(o.K)
    n2, l2 = o.nebor()
    assert np.array_equal(n1, n2) and np.array_equal(l1, l2)
    assert np.array_equal(e.lqdcnt(), o.lqdcnt())


def test_feed_then_rebuild_matches_oracle():
    w = W.dry_column()
    half = 1000
    e = G.Engine(w["inp"])
    o = O.Oracle(w["inp"])
    e.set_particles(w["pos"][:half])
    o.set_particles(w["pos"][:half])
    e.step(50)
    o.step(50)
    e.feed(w["pos"][half:])
    o.feed(w["pos"][half:])
    e.step(50)
    o.step(50)
    assert e.total == o.total == 2000
    compare_state(e, o)
    assert np.array_equal(e.nebor()[1], o.nebor()[1])


def test_overflow_is_reported_not_written():
    w = W.small_bed(n=200)
    pos = w["pos"].copy()
    pos[:60] = pos[0] + np.random.default_rng(0).normal(size=(60, 3)) * 0.01
    e = G.Engine(w["inp"])
    e.set_particles(pos)
    with pytest.raises(G.DemError) as err:
        e.step(1)
    assert err.value.code == -3
    # the truncated lists are never stepped on: later calls fail the same way
    with pytest.raises(G.DemError) as err:
        e.step(1)
    assert err.value.code == -3
    # until the state is replaced by one whose lists fit
    e.set_particles(w["pos"])
    it = e.iter
    e.step(2)
    assert e.iter == it + 2


@pytest.mark.parametrize("binary", [False, True])
def test_dump_load_round_trip(tmp_path, binary):
    w = W.small_bed(n=2000)
    a = G.Engine(w["inp"])
    a.set_particles(w["pos"])
    a.step(200)
    path = tmp_path / "heyi_dump.dat"
    a.dump(path, nBatch=4, binary=binary)
    assert G.load_para(path) == (2000, 4, 2000)
    b = G.Engine(w["inp"])
    assert b.load(path) == 4
    assert b.iter == a.iter
    # a restart rebuilds the lists (main.cpp:473) and remaps the restored
    # history; rebuild the original too so both sum in the same slot order
    a.rebuild()
    a.step(100)
    b.step(100)
    for u, v in zip(a.particles(), b.particles()):
        assert np.array_equal(u, v)
    # and the oracle restarted from the same dump's state agrees bit for bit
    if binary:
        o = O.Oracle(w["inp"])
        c = G.Engine(w["inp"])
        c.load(path)
        o.set_particles(*c.particles())
        o.set_contacts(*c.nebor(), c.hist_pp(), *c.walls(), c.hist_pw())
        o.set_lqdcnt(c.lqdcnt())
        c.step(100)
        o.step(100)
        for u, v in zip(c.particles(), o.particles()):
            assert np.array_equal(u, v)


def test_contact_stats_match_oracle_count():
    w = W.small_bed(n=3000)
    e, o = pair(w)
    e.step(50)
    o.step(50)
    s = e.stats()
    p = o.particles()[0]
    num, lst = o.nebor()
    nc = 0
    for i in range(len(p)):
        d = np.linalg.norm(p[i] - p[lst[i, :num[i]]], axis=1)
        nc += int((d - 1.0 < 0).sum())
    assert s["n_nebor"] == num.sum()
    assert s["n_contact"] == nc


# ---------------------------------------------------------------------------
# mixed precision: FP32 pair arithmetic, FP64 state and history.
# Bar: per-particle force error / (sum_j |f_ij| + m g) <= 1e-5 against the
# FP64 oracle after one step; positions within 1e-7 d after 300 steps.
# ---------------------------------------------------------------------------
def test_one_step_forces_mixed_dense_bed():
    w = W.small_bed(n=3000)
    rng = np.random.default_rng(5)
    vel = rng.normal(size=(3000, 3)) * 1e-3
    angv = rng.normal(size=(3000, 3)) * 1e-2
    e, o = one_step_forces(w, vel, angv, precision=G.PREC_MIXED)
    err, g, r, sc = normalized_force_error(e, o)
    print(f"mixed one-step force error: max {err.max():.3e} mean {err.mean():.3e}")
    assert err.max() <= 1e-5
    assert (np.linalg.norm(g[1] - r[1], axis=1) / sc).max() <= 1e-5
    h1, h2 = e.hist_pp(), o.hist_pp()
    assert np.array_equal(h1["stage"], h2["stage"])
    assert np.allclose(h1["nDsp"], h2["nDsp"], rtol=1e-13, atol=1e-18)


def test_trajectory_mixed_dense_bed_300_steps():
    w = W.small_bed(n=3000)
    e, o = pair(w, precision=G.PREC_MIXED)
    e.step(300)
    o.step(300)
    p1, v1, _ = e.particles()
    p2, v2, _ = o.particles()
    dp = np.abs(p1 - p2).max()
    dv = np.abs(v1 - v2).max() / np.abs(v2).max()
    print(f"mixed 300-step drift: max|dx| {dp:.3e} d, max|dv|/max|v| {dv:.3e}")
    assert dp <= 1e-7
    assert dv <= 1e-3
    assert e.n_rebuild == o.n_rebuild
