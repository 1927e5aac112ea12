"""The benchmark workloads C1, C3 and C4 on the GPU, at their full sizes,
against the oracle (the CPU restatement of the reference's loops), and the
product's periodic wrap of positions (basetrajectory.h:145-161) through
pyanalisi.Trajectory(..., pbc_wrap=True).

Each workload is generated exactly as bench.py generates it (same seeds and
cells), so these tests cover the inputs the benchmark lines are measured on:
  C1  1,000 atoms, cubic 32.2 A, 200 frames; Gofrt rmax 10, 200 bins, lags
      0..50, reset(150): the whole calculation equals the oracle's, in both
      arithmetic forms (gofrt.cpp:96-122 in calculatemultithread.h:114-139).
  C3  65,536 atoms, cubic 90 A, 1,000 frames; Gofrt rmax 15, 512 bins (the
      smallest bin, dr = 0.029 A, so the FP32 window has the least slack):
      the whole reset(936) calculation's lag-0 invariants and per-lag bound,
      one t0 x lag 0 and one (t0, lag 32) calc_single_th equal to the
      oracle, the FP64 form equal to the screened form on a late slice; the
      whole MSD (lags 0..498, centre-of-mass channel and drift removal,
      reset(499), msd.cpp:64-125) within 1e-10 of the oracle.
  C4  8,192 atoms, cubic 45 A, 60,000 frames, windows of 500..4000 frames:
      the block engine over windows of the workload (blockaverage.h:107-190)
      gives, per window, the counts of that window's own single calls (the
      FP64 form for the shortest, the screened form for the longest) and the
      MSD of its single call, bit for bit; the mean and variance over the
      windows equal the oracle's MediaVar (calcoliblocchi.h:21-65) bit for
      bit.
"""
import os
import sys

import numpy as np
import pytest

from analisi_amd import capi
from conftest import ROOT, is_same
from gpu_util import first_diff
from oracle import oracle as O

if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the benchmark's workload definitions)

NTHR = 16


def _self_channel(nt, t):
    half = nt * (nt + 1) // 2
    return half + half - (t + 1) * (t + 2) // 2 + t


def _lag0_invariants(c, type_id, nt, n):
    """lag 0 with rmin = 0: each self channel holds N_type per t0 in bin 0;
    each distinct channel counts every pair both ways (even counts)."""
    ntype = np.bincount(type_id, minlength=nt)
    half = nt * (nt + 1) // 2
    for t in range(nt):
        ch = _self_channel(nt, t)
        assert int(c[0, ch, 0]) == int(ntype[t]) * n
        assert int(c[0, ch, 1:].sum()) == 0
    assert np.all(c[0, :half] % 2 == 0)


def _oracle_traj(w, pbc_wrap=False):
    return O.Traj(w["pos"], w["types"], np.tile(w["box_l"], (w["F"], 1)), w["triclinic"], pbc_wrap=pbc_wrap)


@pytest.mark.gpu
@pytest.mark.parametrize("form", [capi.GOFRT_SCREENED, capi.GOFRT_FP64])
def test_c1_whole_workload_matches_oracle(form):
    w = bench.gofrt_workload("c1")
    ot = _oracle_traj(w)
    dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    print("C1: form %d, reset(%d), tmax %d" % (form, w["n"], w["tmax"]), flush=True)
    g = capi.DeviceGofrt(dt, w["rmin"], w["rmax"], w["nbin"], w["tmax"], 1, 1, form)
    g.reset(w["n"])
    got = g.calculate(0)
    ref, _ = O.gofrt(ot, w["rmin"], w["rmax"], w["nbin"], w["tmax"], NTHR, 1, 1, w["n"], 0, want_vdata=False)
    assert first_diff(got, ref) is None, first_diff(got, ref)
    _lag0_invariants(got, ot.type_id, ot.ntypes, w["n"])
    ps = g.pair_stats()
    assert ps["all"] == int(bench.n_updates(w["N"], w["n"], w["tmax"]))
    assert ps["evaluated"] <= ps["all"]
    if form == capi.GOFRT_FP64:
        assert ps["evaluated"] == ps["all"] == ps["fp64"]
    dt.close()


@pytest.fixture(scope="module")
def c3():
    w = bench.gofrt_workload("c3")
    ot = _oracle_traj(w)
    dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    yield w, ot, dt
    dt.close()


@pytest.mark.gpu
def test_c3_gofrt_whole_run_and_oracle(c3):
    w, ot, dt = c3
    N, nt = w["N"], ot.ntypes
    print("C3: whole reset(%d) screened calculate, 65,536 atoms" % w["n"], flush=True)
    g = capi.DeviceGofrt(dt, w["rmin"], w["rmax"], w["nbin"], w["tmax"], 1, 1, capi.GOFRT_SCREENED)
    leff = g.reset(w["n"])
    c = g.calculate(0)
    assert c.shape == (leff, nt * (nt + 1), w["nbin"])
    _lag0_invariants(c, ot.type_id, nt, w["n"])
    assert np.all(c.reshape(leff, -1).sum(axis=1) <= w["n"] * N * N)
    ps = g.pair_stats()
    assert ps["all"] == int(bench.n_updates(N, w["n"], w["tmax"]))
    assert 0 < ps["evaluated"] < ps["all"]
    assert g.stats()[0] == capi.GOFRT_SCREENED

    print("C3: one t0 x lag 0 against the oracle", flush=True)
    primo = 617
    ref, _ = O.gofrt(ot, w["rmin"], w["rmax"], w["nbin"], 1, NTHR, 1, 1, 1, primo, want_vdata=False)
    g1 = capi.DeviceGofrt(dt, w["rmin"], w["rmax"], w["nbin"], 1, 1, 1, capi.GOFRT_SCREENED)
    g1.reset(1)
    got = g1.calculate(primo)
    assert first_diff(got, ref) is None, first_diff(got, ref)

    print("C3: calc_single_th(lag 32, t0 5) over atoms [12345, 20537) against the oracle", flush=True)
    a0, a1, t, im = 12345, 20537, 32, 5
    ref = O.gofrt_single(ot, w["rmin"], w["rmax"], w["nbin"], leff, t, im, a0, a1, 300)
    got = g.calc_single_th(t, im, a0, a1, 300, None)
    assert first_diff(got, ref) is None, first_diff(got, ref)
    assert int(got[t].sum()) > 0

    print("C3: FP64 form equals the screened form on reset(2) at frame 900", flush=True)
    g64 = capi.DeviceGofrt(dt, w["rmin"], w["rmax"], w["nbin"], w["tmax"], 1, 1, capi.GOFRT_FP64)
    g64.reset(2)
    gsl = capi.DeviceGofrt(dt, w["rmin"], w["rmax"], w["nbin"], w["tmax"], 1, 1, capi.GOFRT_SCREENED)
    gsl.reset(2)
    a, b = g64.calculate(900), gsl.calculate(900)
    assert first_diff(a, b) is None, first_diff(a, b)


@pytest.mark.gpu
def test_c3_msd_whole_run_matches_oracle(c3):
    w, ot, dt = c3
    n, tmax = w["msd_n"], w["msd_tmax"]
    print("C3: MSD reset(%d), tmax %d, cm_msd and cm_self" % (n, tmax), flush=True)
    got = capi.msd(dt, 1, tmax, True, True, n, 0)
    ref = O.msd(ot, 1, tmax, NTHR, True, True, n, 0)
    assert got.shape == ref.shape == (min(n, tmax), 2, ot.ntypes)
    ok = is_same(got, ref, 1e-10)
    assert ok.all(), (np.argwhere(~ok)[:5], got[~ok][:5], ref[~ok][:5])
    assert np.all(got[0] == 0.0)
    # the drift-removed self part grows with the lag (sigma 0.08 A random walk)
    assert got[-1, 0, 0] > got[1, 0, 0] > 0


@pytest.mark.gpu
def test_c4_block_engine_windows_and_mediavar():
    w = bench.gofrt_workload("c4")
    traj, tid, nt = bench.device_traj(w, with_cm=False)
    starts, lens = bench.c4_windows(w)
    pick = sorted({int(np.argmin(lens)), int(np.argmax(lens)), 7, 40})
    st, ln = starts[pick], lens[pick]
    print("C4: windows %s, lengths %s, starts %s" % (pick, ln.tolist(), st.tolist()), flush=True)
    g = capi.DeviceGofrt(traj, w["rmin"], w["rmax"], w["nbin"], w["tmax"], 1, 1, capi.GOFRT_SCREENED)
    B = capi.DeviceBlocks(traj, st, ln, gofrt=g, msd=True, msd_skip=1, msd_tmax=w["msd_tmax"], use_graph=True)
    B.run()
    B.run()                       # a replay of the captured graph
    gmean, gvar, gcounts = B.gofrt_results()
    mmean, mvar, mper = B.msd_results()
    ishort, ilong = int(np.argmin(ln)), int(np.argmax(ln))
    print("C4: window %d against its FP64 single call" % pick[ishort], flush=True)
    g64 = capi.DeviceGofrt(traj, w["rmin"], w["rmax"], w["nbin"], w["tmax"], 1, 1, capi.GOFRT_FP64)
    g64.reset(int(ln[ishort]))
    one = g64.calculate(int(st[ishort]))
    assert first_diff(one, gcounts[ishort]) is None, first_diff(one, gcounts[ishort])
    print("C4: window %d against its screened single call" % pick[ilong], flush=True)
    gs = capi.DeviceGofrt(traj, w["rmin"], w["rmax"], w["nbin"], w["tmax"], 1, 1, capi.GOFRT_SCREENED)
    gs.reset(int(ln[ilong]))
    one = gs.calculate(int(st[ilong]))
    assert first_diff(one, gcounts[ilong]) is None, first_diff(one, gcounts[ilong])
    for k in range(len(pick)):
        _lag0_invariants(gcounts[k], tid, nt, int(ln[k]))
        m1 = capi.msd(traj, 1, w["msd_tmax"], False, False, int(ln[k]), int(st[k]))
        assert np.array_equal(m1, mper[k]), k
    mv = O.MediaVar(gmean.size)
    for k in range(len(pick)):
        mv.add(gcounts[k].ravel().astype(np.float64) * (1.0 / int(ln[k])))
    mv.end(len(pick))
    assert np.array_equal(gmean.ravel(), mv.mean) and np.array_equal(gvar.ravel(), mv.var)
    mv2 = O.MediaVar(mmean.size)
    for k in range(len(pick)):
        mv2.add(mper[k].ravel())
    mv2.end(len(pick))
    assert np.array_equal(mmean.ravel(), mv2.mean) and np.array_equal(mvar.ravel(), mv2.var)
    B.close()
    traj.close()


# --- the product's pbc_wrap --------------------------------------------------

def _wrap_case(triclinic):
    """Positions several cells away from the cell (both signs), so that the
    wrap's min-image loops take several steps per atom."""
    from analisi_amd import synth
    tilt = [2.1, -1.7, 1.3] if triclinic else None
    box_l = synth.lammps_box(13.0, tilt)
    pos = synth.random_walk(17, 240, 9, box_l, 0.3)
    shift = np.zeros_like(pos)
    shift[:, ::3] = 3.0 * 13.0 + 0.37
    shift[:, 1::3] = -5.0 * 13.0 - 0.11
    pos = pos + shift
    types = synth.types_by_fraction(240, [0.5, 0.3, 0.2])
    return pos, types, box_l


@pytest.mark.parametrize("triclinic", [False, True])
def test_pbc_wrap_positions_equal_oracle(triclinic):
    """pyanalisi.Trajectory(..., pbc_wrap=True) wraps every frame's positions
    around the cell centre exactly as the oracle's restatement of
    BaseTrajectory::pbc_wrap does (basetrajectory.h:145-161): bit for bit.
    Host-side: no device is needed."""
    from analisi_amd import pyanalisi as pa
    pos, types, box_l = _wrap_case(triclinic)
    F = pos.shape[0]
    fmt = pa.BoxFormat.LammpsTriclinic if triclinic else pa.BoxFormat.LammpsOrtho
    boxes = np.tile(box_l, (F, 1))
    t = pa.Trajectory(pos, np.zeros_like(pos), types.astype(np.int32), boxes, fmt, True, False)
    got = t.get_positions_copy()
    ot = O.Traj(pos, types, boxes, triclinic, pbc_wrap=True)
    assert np.array_equal(got, ot.pos)
    assert not np.array_equal(got, pos)          # the wrap moved atoms
    # the wrapped positions lie inside the cell (orthorhombic: each axis)
    if not triclinic:
        assert np.all(got >= -1e-12) and np.all(got <= 13.0 + 1e-12)
    assert t.getWrapPbc()


@pytest.mark.gpu
@pytest.mark.parametrize("triclinic", [False, True])
@pytest.mark.parametrize("form", [0, 1])
def test_pbc_wrap_gofrt_counts_equal_oracle(triclinic, form):
    """g(r,t) through a wrapped pyanalisi.Trajectory, as the Python and
    command-line callers run it (main.cpp:558, analysis.py:116), equals the
    oracle on the wrapped positions."""
    from analisi_amd import pyanalisi as pa
    pos, types, box_l = _wrap_case(triclinic)
    F = pos.shape[0]
    fmt = pa.BoxFormat.LammpsTriclinic if triclinic else pa.BoxFormat.LammpsOrtho
    boxes = np.tile(box_l, (F, 1))
    t = pa.Trajectory(pos, np.zeros_like(pos), types.astype(np.int32), boxes, fmt, True, False)
    g = pa.Gofrt(t, 0.0, 6.0, 60, 4, 1, 1, 1, False, form)
    g.reset(5)
    g.calculate(0)
    got = np.round(np.array(g) * 5).astype(np.uint64)
    ot = O.Traj(pos, types, boxes, triclinic, pbc_wrap=True)
    ref, _ = O.gofrt(ot, 0.0, 6.0, 60, 4, 4, 1, 1, 5, 0, want_vdata=False)
    assert first_diff(got, ref) is None, first_diff(got, ref)
    assert g.last_stats()[0] == form
