"""The pyanalisi surface (pybind11 over the C++ operator classes over the C
ABI) on the GPU, called the way the reference's own tests and the CLI call it:
tests/test_gk.py:4-12, tests/test_gofrt.py:4-13, tests/test_msd.py:4-11,
main.cpp:456-516 (block averages)."""
import os

import numpy as np
import pytest

from analisi_amd import synth
from conftest import GOLDEN, is_same
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    from analisi_amd import pyanalisi
    return pyanalisi


@pytest.fixture
def analisi_log(pa, gk_ref):
    return pa.ReadLog(np.array(gk_ref["log"], copy=True), list(gk_ref["headers"]))


def test_gk_reference_test_call(pa, analisi_log, gk_ref):
    gk = pa.GreenKubo(analisi_log, '', 1, ['c_flux[1]', 'c_vcm[1][1]'], False, 2000, 2, False, 0, 4, False, 1, 100)
    gk.reset(analisi_log.getNtimesteps() - 2000)
    gk.calculate(0)
    m = np.array(gk, copy=False)
    assert m.shape == (2000, 14)
    assert is_same(m, gk_ref["test_gk"]).all()


def test_gk_header_error(pa, analisi_log):
    with pytest.raises(RuntimeError, match="not found"):
        pa.GreenKubo(analisi_log, '', 1, ['nope'], False, 20, 1, False, 0, 1, False, 1, 100)


@pytest.mark.parametrize("name,heads", [("single", ["c_flux[1]"]), ("double", ["c_flux[1]", "c_vcm[1][1]"])])
def test_gk_block_average_matches_cli_text(pa, analisi_log, name, heads):
    """BlockAverageG + MediaVarCovar on the GPU path, printed as the CLI prints
    it (main.cpp:502-513, 6 significant digits): text-identical to the
    reference's tests/data/cli/GK_heat_* outputs."""
    mean, var = pa.gk_block_average(analisi_log, heads, 20, 1, 0, 1, False, 0)
    narr, s = mean.shape
    mine = [" ".join(sum([["%g" % mean[j, i], "%g" % var[j, i]] for j in range(narr)], [])) + " " for i in range(s)]
    with open(os.path.join(GOLDEN, "cli_GK_heat_%s.txt" % name)) as f:
        ref = [l.rstrip("\n") for l in f if l.strip() and not l.startswith("#")]
    norm = lambda x: x.replace("-nan", "nan")
    assert [norm(a) for a in mine] == [norm(b) for b in ref]


def _cell_traj(F, N, seed):
    """A trajectory in the reference's test layout (conftest.py:31-45): CellVectors
    boxes, types 0 except the last 16 atoms (types 1 and 2)."""
    cell = np.diag([4.14141307, 7.81544411, 6.44646321])   # tests/data/cells.npy[0]
    box_l = synth.lammps_box(np.diag(cell))
    pos = synth.random_walk(seed, N, F, box_l, 0.05)
    vel = np.zeros_like(pos)
    types = np.zeros(N, dtype="i")
    types[-16:-8] = 1
    types[-8:] = 2
    boxes = np.tile(cell, (F, 1, 1))
    return pos, vel, types, boxes, box_l


def test_gofrt_reference_test_call(pa, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)                 # debug=1 appends ./gofrt.dump
    pos, vel, types, boxes, box_l = _cell_traj(712, 56, 3)
    traj = pa.Trajectory(pos, vel, types, boxes, pa.BoxFormat.CellVectors, False, False)
    gofr = pa.Gofrt(traj, 0.0, 3.8, 200, 10, 4, 10, False, 1)
    gofr.reset(700)
    gofr.calculate(0)
    gr = np.array(gofr, copy=True)
    assert gr.shape == (10, 12, 200)
    assert os.path.exists("gofrt.dump")
    ot = O.Traj(pos, types, np.tile(box_l, (712, 1)), False)
    counts, vref = O.gofrt(ot, 0.0, 3.8, 200, 10, 4, 10, 1, 700, 0)
    np.testing.assert_array_equal(np.round(gr * 70).astype(np.int64), counts.astype(np.int64))
    assert is_same(gr, vref).all()              # count*incr vs the reference's repeated sums
    # the screen cannot be used (rmax > half the 4.14 A cell): AUTO picked FP64
    assert gofr.last_stats()[0] == 0


def test_gofrt_calc_single_th(pa):
    pos, vel, types, boxes, box_l = _cell_traj(30, 56, 4)
    traj = pa.Trajectory(pos, vel, types, boxes, pa.BoxFormat.CellVectors, False, False)
    g = pa.Gofrt(traj, 0.5, 2.0, 30, 5, 1, 1, 1, False)
    g.reset(20)
    g.calculate(0)                               # sets incr and zeroes
    full = np.array(g, copy=True)
    ot = O.Traj(pos, types, np.tile(box_l, (30, 1)), False)
    c = np.zeros((5, 12, 30), np.uint64)
    for t in range(5):
        for im in range(20):
            O.gofrt_single(ot, 0.5, 2.0, 30, 5, t, im, 0, 56, 0, c)
    np.testing.assert_array_equal(np.round(full * 20).astype(np.int64), c.astype(np.int64))


def test_msd_reference_test_call(pa):
    pos, vel, types, boxes, box_l = _cell_traj(700, 56, 5)
    traj = pa.Trajectory(pos, vel, types, boxes, pa.BoxFormat.CellVectors, False, False)
    msd = pa.MeanSquareDisplacement(traj, 10, 400, 4, True, False, False)
    msd.reset(300)
    msd.calculate(0)
    m = np.array(msd, copy=False)
    assert m.shape == (300, 2, 3)
    ot = O.Traj(pos, types, np.tile(box_l, (700, 1)), False)
    ref = O.msd(ot, 10, 400, 4, True, False, 300, 0)
    assert is_same(m, ref).all()
    with pytest.raises(RuntimeError, match="too short"):
        msd.reset(400)
        msd.calculate(0)


def test_block_averages_match_oracle(pa):
    F, N = 260, 120
    box_l = synth.lammps_box(11.0, [0.6, -0.4, 0.5])
    pos = synth.random_walk(8, N, F, box_l, 0.1)
    types = synth.types_by_fraction(N, [0.6, 0.4])
    traj = pa.Trajectory(pos, np.zeros_like(pos), types, np.tile(box_l, (F, 1)), pa.BoxFormat.LammpsTriclinic,
                         False, False)
    ot = O.Traj(pos, types, np.tile(box_l, (F, 1)), True)
    nb = 5
    # Gofrt: blocks of s = (F - nExtra)/nb, MediaVar over count*incr vectors
    mean, var = pa.gofrt_block_average(traj, nb, 0.0, 5.0, 25, 8, 2, 1)
    s = (F - O.n_extra_timesteps(F, nb, 8)) // nb
    mv = O.MediaVar(mean.size)
    for b in range(nb):
        _, v = O.gofrt(ot, 0.0, 5.0, 25, 8, 3, 2, 1, s, b * s)
        mv.add(v)
    mv.end(nb)
    assert is_same(mean.ravel(), mv.mean, 1e-10).all()
    assert is_same(var.ravel(), mv.var, 1e-8).all()
    # MSD
    mean, var = pa.msd_block_average(traj, nb, 3, 10, True, True)
    s = (F - O.n_extra_timesteps(F, nb, 10)) // nb
    mv = O.MediaVar(mean.size)
    for b in range(nb):
        mv.add(O.msd(ot, 3, 10, 2, True, True, s, b * s))
    mv.end(nb)
    assert is_same(mean.ravel(), mv.mean, 1e-9).all()
    np.testing.assert_allclose(var.ravel(), mv.var, rtol=1e-6, atol=1e-28)


def test_triclinic_cell_vectors_rotation(pa):
    """A triclinic cell given as rotated cell vectors: the trajectory recovers
    the LAMMPS cell (positive diagonal, triclinic.h:30-38) and rotates
    positions into it; g(r,t) equals the oracle on the rotated positions."""
    F, N = 12, 90
    Lc = np.array([[9.0, 0, 0], [1.1, 9.5, 0], [-0.8, 0.7, 10.0]])    # rows a, b, c (lower triangular)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    vecs = Lc @ R.T                                                     # rotated lattice vectors (rows)
    box_l = synth.lammps_box([9.0, 9.5, 10.0], [1.1, -0.8, 0.7])
    posl = synth.random_walk(9, N, F, box_l, 0.1)
    posr = posl @ R.T
    cellv = np.tile(vecs.T, (F, 1, 1))                                  # numpy layout: columns = vectors
    traj = pa.Trajectory(posr, np.zeros_like(posr), np.zeros(N, "i"), np.ascontiguousarray(cellv),
                         pa.BoxFormat.CellVectors, False, True)
    b = traj.get_box_copy()[0]
    np.testing.assert_allclose(b[3:], [4.5, 4.75, 5.0, 1.1, -0.8, 0.7], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(traj.get_positions_copy(), posl, atol=1e-11)
    g = pa.Gofrt(traj, 0.0, 4.0, 40, 3, 1, 1, 1, False)
    g.reset(8)
    g.calculate(0)
    rot = traj.get_positions_copy()
    ot = O.Traj(rot, np.zeros(N, np.int32), np.tile(box_l, (F, 1)), True)
    ot.box[:] = traj.get_box_copy()
    ot.pos[:] = rot
    counts, _ = O.gofrt(ot, 0.0, 4.0, 40, 3, 1, 1, 1, 8, 0)
    np.testing.assert_array_equal(np.round(np.array(g) * 8).astype(np.int64), counts.astype(np.int64))


def test_lammps_binary_operators_match_the_numpy_trajectory(pa, tmp_path):
    """Gofrt_lammps / MeanSquareDisplacement_lammps over a Traj window give
    what Gofrt / MeanSquareDisplacement give over the in-memory Trajectory of
    the same data (written with write_lammps_binary), for a calculation
    inside the loaded window (device offset primo - window start) and for one
    outside it (the window moves to primo first): counts bit for bit; MSD to
    the reference's is_same (1e-10), since the device sums atoms in a slot
    order taken from the first frame on the device, which differs between a
    window and the whole trajectory."""
    N, F = 300, 30
    tilt = [1.1, -0.7, 0.5]
    box_l = synth.lammps_box(13.0, tilt)
    pos = synth.random_walk(41, N, F, box_l, 0.3)
    vel = np.random.default_rng(2).normal(size=(F, N, 3))
    types = synth.types_by_fraction(N, [0.6, 0.3, 0.1]).astype(np.int32)
    tnp = pa.Trajectory(pos, vel, types, np.tile(box_l, (F, 1)), pa.BoxFormat.LammpsTriclinic, False, False)
    path = str(tmp_path / "traj.bin")
    tnp.write_lammps_binary(path, 0, -1)
    tl = pa.Traj(path)
    tl.setAccessWindowSize(16)
    tl.setAccessStart(4)
    n, tmax = 8, 6
    for form in (0, 1):   # FP64 and screened forms
        g0 = pa.Gofrt(tnp, 0.0, 5.5, 40, tmax, 1, 1, 1, False, form)
        g1 = pa.Gofrt_lammps(tl, 0.0, 5.5, 40, tmax, 1, 1, 1, False, form)
        g0.reset(n)
        g1.reset(n)
        for primo in (6, 0, 14):   # inside [4, 20), then two moves of the window
            g0.calculate(primo)
            g1.calculate(primo)
            assert np.array_equal(np.array(g0), np.array(g1)), (form, primo)
            assert tl.get_current_timestep() <= primo
    m0 = pa.MeanSquareDisplacement(tnp, 1, 5, 1, True, True)
    m1 = pa.MeanSquareDisplacement_lammps(tl, 1, 5, 1, True, True)
    m0.reset(9)
    m1.reset(9)
    for primo in (tl.get_current_timestep() + 1, 2):
        m0.calculate(primo)
        m1.calculate(primo)
        assert is_same(np.array(m0), np.array(m1)).all(), primo
    big = pa.MeanSquareDisplacement_lammps(tl, 1, 10, 1, False, False)
    big.reset(10)
    with pytest.raises(RuntimeError, match="not enough loaded timesteps"):
        big.calculate(0)


def test_greenkubo_file_matches_greenkubo_on_the_same_log(pa, gk_ref, tmp_path):
    """GreenKubo_file over ReadLog_file (the text log, as the CLI's -l reads
    it) equals GreenKubo over ReadLog (numpy) bit for bit, and both equal the
    reference's test_gk.csv golden to is_same; a "#traj:" charge-current
    column of a LAMMPS binary feeds GreenKubo_file like any other column."""
    log, headers = gk_ref["log"], list(gk_ref["headers"])
    p = str(tmp_path / "gk_integral.dat")
    with open(p, "w") as f:
        f.write("    ".join(headers) + "\n")
        for row in log:
            f.write(repr(float(row[0])) + "\t" + " ".join(repr(float(x)) for x in row[1:]) + "\n")
    rf = pa.ReadLog_file(p, None, 1, 1, 200, [])
    rn = pa.ReadLog(np.array(log, copy=True), headers)
    args = ('', 1, ['c_flux[1]', 'c_vcm[1][1]'], False, 2000, 2, False, 0, 4, False, 1, 100)
    g0, g1 = pa.GreenKubo(rn, *args), pa.GreenKubo_file(rf, *args)
    for g in (g0, g1):
        g.reset(rn.getNtimesteps() - 2000)
        g.calculate(0)
    m0, m1 = np.array(g0, copy=True), np.array(g1, copy=True)
    assert np.array_equal(np.nan_to_num(m0, nan=0.5), np.nan_to_num(m1, nan=0.5))
    assert is_same(m1, gk_ref["test_gk"]).all()

    N, F = 40, 300
    box_l = synth.lammps_box(9.0)
    pos = synth.random_walk(6, N, F, box_l, 0.1)
    vel = np.random.default_rng(1).normal(size=(F, N, 3))
    types = synth.types_by_fraction(N, [0.5, 0.5]).astype(np.int32)
    b = str(tmp_path / "t.bin")
    pa.Trajectory(pos, vel, types, np.tile(box_l, (F, 1)), pa.BoxFormat.LammpsOrtho, False, False) \
        .write_lammps_binary(b, 0, -1)
    lp = str(tmp_path / "log")
    with open(lp, "w") as f:
        f.write("Step c_a\n" + "".join("%d %r\n" % (k + 1, 0.5 * k) for k in range(F)))
    head = "#traj:JZ 2 1.0 -1.0"
    t = pa.Traj(b)
    r = pa.ReadLog_file(lp, t, 1, 1, 200, [head])
    r.calc_currents(t, 3)
    J = r.get_data_copy()[:, 1:4]
    gk = pa.GreenKubo_file(r, '', 1, [head], False, 50, 1, False, 0, 1)
    gk.reset(F - 50)
    gk.calculate(0)
    c = np.array(gk, copy=True)
    ref = np.array([np.mean(np.sum(J[:F - 50] * J[l:l + F - 50], axis=1)) / 3.0 for l in range(50)])
    np.testing.assert_allclose(c[:, 0], ref, rtol=1e-12, atol=1e-15)


def test_lammps_binary_block_averages_match_the_numpy_trajectory(pa, tmp_path):
    """gofrt_block_average / msd_block_average over a Traj (the window moves
    from block to block) equal the in-memory trajectory's: g(r,t) mean and
    variance bit for bit, MSD to is_same."""
    N, F = 200, 30
    box_l = synth.lammps_box(12.0)
    pos = synth.random_walk(43, N, F, box_l, 0.3)
    vel = np.zeros((F, N, 3))
    types = synth.types_by_fraction(N, [0.5, 0.5]).astype(np.int32)
    tnp = pa.Trajectory(pos, vel, types, np.tile(box_l, (F, 1)), pa.BoxFormat.LammpsOrtho, False, False)
    path = str(tmp_path / "traj.bin")
    tnp.write_lammps_binary(path, 0, -1)
    tl = pa.Traj(path)
    tl.setAccessWindowSize(20)
    tl.setAccessStart(0)
    m0, v0 = pa.gofrt_block_average(tnp, 2, 0.0, 5.0, 30, 6, 1, 1)
    m1, v1 = pa.gofrt_block_average(tl, 2, 0.0, 5.0, 30, 6, 1, 1)
    assert np.array_equal(m0, m1) and np.array_equal(np.nan_to_num(v0), np.nan_to_num(v1))
    m0, v0 = pa.msd_block_average(tnp, 2, 1, 5, False, False)
    m1, v1 = pa.msd_block_average(tl, 2, 1, 5, False, False)
    assert is_same(m0, m1).all() and is_same(v0, v1).all()


def _sh_close(got, want, tol=1e-12):
    scale = np.abs(want).max()
    assert scale > 0 and np.abs(got - want).max() <= tol * scale, np.abs(got - want).max() / scale


def test_spherical_correlations_reference_test_call(pa):
    """tests/test_sh.py:4-22 on a synthetic trajectory of the reference's test
    layout (types 0, then 8 atoms of type 1 and 8 of type 2): [(0.7, 3.5)]*9,
    4 bins, tmax 200, 4 threads, skip 90, buffer 10, debug, no neighbour list;
    reset(700), calculate(0).  The values equal the oracle's to 1e-12 of the
    largest (FP64, different summation order); calc(t), SphericalBase::calc,
    equals the oracle's density."""
    F, N = 905, 64
    box_l = synth.lammps_box(8.0)
    pos = synth.random_walk(61, N, F, box_l, 0.05)
    vel = np.zeros_like(pos)
    types = np.zeros(N, dtype=np.int32)
    types[-16:-8] = 1
    types[-8:] = 2
    t = pa.Trajectory(pos, vel, types, np.tile(box_l, (F, 1)), pa.BoxFormat.LammpsOrtho, False, False)
    rm = [(0.7, 3.5)] * 9
    sh = pa.SphericalCorrelations(t, rm, 4, 200, 4, 90, 10, True, [])
    sh.reset(700)
    sh.calculate(0)
    res = np.array(sh, copy=True)
    assert res.shape == (200, 3, 3, 4, 11)
    ot = O.Traj(pos, types, np.tile(box_l, (F, 1)), False)
    _sh_close(res, O.sph_correlations(ot, rm, 4, 200, 90, 700, 0))
    _sh_close(sh.calc(7).reshape(N, 3, 4, 121), O.sph_density(ot, 7, rm, 4))
    assert sh.getNumberOfExtraTimestepsNeeded(10) == min(F // 11 + 1, 200)
    with pytest.raises(RuntimeError, match="too short"):
        sh.calculate(6)                                   # 200 + 700 + 6 > 905
    with pytest.raises(RuntimeError, match="radial range for each pair"):
        pa.SphericalCorrelations(t, rm[:4], 4, 200, 4, 90, 10, True, [])
    with pytest.raises(RuntimeError, match="parameters for each atomic type"):
        pa.SphericalCorrelations(t, rm, 1, 200, 4, 90, 10, True, [(19, 16.0, 16.0)])
    few = pa.SphericalCorrelations(t, rm, 4, 3, 4, 90, 10, False, [])
    few.reset(10)
    with pytest.raises(RuntimeError, match="enough timesteps to parallelize"):
        few.calculate(0)                                  # leff 3 < 4 threads


def test_spherical_correlations_lammps_binary_window(pa, tmp_path):
    """SphericalCorrelations_lammps over a Traj window equals
    SphericalCorrelations over the in-memory Trajectory of the same data, for
    a start inside the window and one that moves it."""
    N, F = 90, 40
    box_l = synth.lammps_box(9.0, [0.8, -0.5, 0.3])
    pos = synth.random_walk(67, N, F, box_l, 0.2)
    vel = np.zeros_like(pos)
    types = synth.types_by_fraction(N, [0.7, 0.3]).astype(np.int32)
    tnp = pa.Trajectory(pos, vel, types, np.tile(box_l, (F, 1)), pa.BoxFormat.LammpsTriclinic, False, False)
    path = str(tmp_path / "sh.bin")
    tnp.write_lammps_binary(path, 0, -1)
    tl = pa.Traj(path)
    tl.setAccessWindowSize(30)
    tl.setAccessStart(0)
    rm = [(0.5, 3.2), (0.9, 3.6), (0.9, 3.6), (0.6, 3.0)]
    s0 = pa.SphericalCorrelations(tnp, rm, 3, 6, 1, 3, 10, False, [])
    s1 = pa.SphericalCorrelations_lammps(tl, rm, 3, 6, 1, 3, 10, False, [])
    s0.reset(14)
    s1.reset(14)
    for primo in (4, 17):
        s0.calculate(primo)
        s1.calculate(primo)
        _sh_close(np.array(s1), np.array(s0))


def test_steinhardt_reference_fixture_call(pa):
    """SteinhardtFixture of tests/src/test_sh.cpp:37-48 through the Python
    class: [(0.5, 3.0)]*4 for 2 types, 4 bins, histogram size 100, l = (4, 6),
    skip 17, histogram on; reset(all loaded timesteps), calculate(0).  Equal to
    the oracle; the q trajectory form (do_histogram False) too."""
    F, N = 40, 90
    box_l = synth.lammps_box(8.0)
    pos = synth.random_walk(71, N, F, box_l, 0.1)
    types = synth.types_by_fraction(N, [0.7, 0.3]).astype(np.int32)
    t = pa.Trajectory(pos, np.zeros_like(pos), types, np.tile(box_l, (F, 1)), pa.BoxFormat.LammpsOrtho, False, False)
    rm = [(0.5, 3.0)] * 4
    st = pa.SteinhardtOrderParameterHistogram_6(t, rm, 4, 100, [4, 6], 2, 17, True, [], False)
    st.reset(F)
    st.calculate(0)
    h = np.array(st, copy=True)
    assert h.shape == (4, 2, 2, 100, 100)
    ot = O.Traj(pos, types, np.tile(box_l, (F, 1)), False)
    np.testing.assert_allclose(h, O.steinhardt(ot, rm, 4, 6, 100, [4, 6], 17, F, 0, True), rtol=1e-13, atol=0)
    tq = pa.SteinhardtOrderParameterHistogram_6(t, rm, 4, 100, [4, 6], 2, 17, False, [], False)
    tq.reset(F)
    tq.calculate(0)
    q = np.array(tq, copy=True)
    assert q.shape == (4, F // 17, N, 2, 6)
    assert np.abs(q - O.steinhardt(ot, rm, 4, 6, 100, [4, 6], 17, F, 0, False)).max() <= 1e-12
    with pytest.raises(RuntimeError, match="averaged steinhardt"):
        pa.SteinhardtOrderParameterHistogram_6(t, rm, 4, 100, [4, 6], 2, 17, True, [], True)
    with pytest.raises(RuntimeError, match="set nbin to 1"):
        pa.SteinhardtOrderParameterHistogram_6(t, rm, 4, 100, [4, 6], 2, 17, True, [(19, 16.0, 16.0)], False)
    with pytest.raises(RuntimeError, match="too short"):
        st.calculate(1)


def test_lammps_window_read_ahead_changes_nothing(pa, tmp_path):
    """Block averages over a Traj with the next block's window read ahead on
    a host thread (the default: its reads, ordering and uploads overlap the
    current block's kernels) equal those with read-ahead off, bit for bit;
    prefetch(ts) then setAccessStart(ts) serves what a direct read serves."""
    N, F = 300, 60
    box_l = synth.lammps_box(12.0, [0.9, -0.6, 0.4])
    pos = synth.random_walk(47, N, F, box_l, 0.3)
    vel = np.random.default_rng(5).normal(size=(F, N, 3))
    types = synth.types_by_fraction(N, [0.6, 0.4]).astype(np.int32)
    tnp = pa.Trajectory(pos, vel, types, np.tile(box_l, (F, 1)), pa.BoxFormat.LammpsTriclinic, False, False)
    path = str(tmp_path / "ahead.bin")
    tnp.write_lammps_binary(path, 0, -1)
    tl = pa.Traj(path)
    tl.setAccessWindowSize(20)
    tl.setAccessStart(0)
    on = pa.gofrt_block_average(tl, 4, 0.0, 5.0, 30, 4, 1, 1)
    mon = pa.msd_block_average(tl, 4, 1, 4, True, True)
    tl.setPrefetch(False)
    off = pa.gofrt_block_average(tl, 4, 0.0, 5.0, 30, 4, 1, 1)
    moff = pa.msd_block_average(tl, 4, 1, 4, True, True)
    for a, b in zip(on + mon, off + moff):
        assert np.array_equal(np.nan_to_num(a), np.nan_to_num(b))
    ref = pa.gofrt_block_average(tnp, 4, 0.0, 5.0, 30, 4, 1, 1)
    assert np.array_equal(on[0], ref[0])
    tl.setPrefetch(True)
    tl.setAccessWindowSize(20)          # the block loops sized the window to their blocks
    tl.prefetch(33)
    tl.setAccessStart(33)
    direct = pa.Traj(path)
    direct.setAccessWindowSize(20)
    direct.setAccessStart(33)
    assert np.array_equal(np.array(tl), np.array(direct))
    g = pa.Gofrt_lammps(tl, 0.0, 5.0, 30, 3, 1, 1, 1, False)
    h = pa.Gofrt_lammps(direct, 0.0, 5.0, 30, 3, 1, 1, 1, False)
    g.reset(10)
    h.reset(10)
    g.calculate(35)
    h.calculate(35)
    assert np.array_equal(np.array(g), np.array(h))


def test_steinhardt_neighbour_list_fixture_call(pa):
    """SteinhardtFixtureNeigh of tests/src/test_sh.cpp:50-62 through the
    Python class: SANN lists of at most 19 and 17 neighbours within 4 A for the
    two types, one bin, l = (4, 6), skip 13, histogram off and on, plain and
    averaged order; equal to the oracle.  SphericalCorrelations with the same
    lists equals the oracle too."""
    F, N = 30, 90
    box_l = synth.lammps_box(16.0)      # ~6 atoms within 4 A: the lists of 19 and 17 hold them
    pos = synth.random_walk(73, N, F, box_l, 0.1)
    types = synth.types_by_fraction(N, [0.7, 0.3]).astype(np.int32)
    t = pa.Trajectory(pos, np.zeros_like(pos), types, np.tile(box_l, (F, 1)), pa.BoxFormat.LammpsOrtho, False, False)
    rm = [(0.5, 3.0)] * 4
    nls = [(19, 16.0, 16.0), (17, 16.0, 16.0)]
    ot = O.Traj(pos, types, np.tile(box_l, (F, 1)), False)
    st = pa.SteinhardtOrderParameterHistogram_6(t, rm, 1, 100, [4, 6], 2, 13, False, nls, False)
    st.reset(F)
    st.calculate(0)
    q = np.array(st, copy=True)
    assert q.shape == (1, F // 13, N, 2, 6)
    assert np.abs(q - O.steinhardt_sann(ot, nls, 6, 100, [4, 6], 13, F, 0, False)).max() <= 1e-12
    sh = pa.SteinhardtOrderParameterHistogram_6(t, rm, 1, 100, [4, 6], 2, 13, True, nls, False)
    sh.reset(F)
    sh.calculate(0)
    np.testing.assert_allclose(np.array(sh), O.steinhardt_sann(ot, nls, 6, 100, [4, 6], 13, F, 0, True),
                               rtol=1e-13, atol=0)
    sc = pa.SphericalCorrelations(t, rm, 2, 6, 1, 3, 10, False, nls)
    sc.reset(12)
    sc.calculate(2)
    _sh_close(np.array(sc), O.sph_correlations_sann(ot, nls, 2, 6, 3, 12, 2))
    for hist in (False, True):
        av = pa.SteinhardtOrderParameterHistogram_6(t, rm, 1, 100, [4, 6], 2, 13, hist, nls, True)
        av.reset(F)
        av.calculate(0)
        want = O.steinhardt_sann(ot, nls, 6, 100, [4, 6], 13, F, 0, hist, averaged=True)
        if hist:
            np.testing.assert_allclose(np.array(av), want, rtol=1e-13, atol=0)
        else:   # NaN where a neighbour has no neighbours of type_j (0/0, as the reference)
            assert np.isnan(want).any()
            np.testing.assert_allclose(np.array(av), want, rtol=0, atol=1e-12, equal_nan=True)


# ---------------------------------------------------------------------------
# The command line's block-averaged text outputs (main.cpp:388-581)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name,heads", [("single", ["c_flux[1]"]), ("double", ["c_flux[1]", "c_vcm[1][1]"])])
def test_cli_gk_text_equals_reference_files(pa, analisi_log, gk_ref, name, heads, tmp_path):
    """-l gk_integral.dat -H -a HEADS with the defaults (20 blocks, factors 1):
    the whole text the native printer writes over the GPU path (factor lines,
    the columns' legend, the rows) equals the reference's
    tests/data/cli/GK_heat_* line for line, a NaN's printed sign aside.  The
    text log read from a file (as the command line reads it) prints the same."""
    txt = pa.cli_gk_text(analisi_log, heads)
    with open(os.path.join(GOLDEN, "cli_GK_heat_%s.txt" % name)) as f:
        ref = f.read()
    norm = lambda x: x.replace("-nan", "nan").split("\n")
    assert norm(txt) == norm(ref)
    if name == "single":
        p = str(tmp_path / "gk_integral.dat")
        with open(p, "w") as f:
            f.write("    ".join(gk_ref["headers"]) + "\n")
            for row in gk_ref["log"]:
                f.write(repr(float(row[0])) + "\t" + " ".join(repr(float(x)) for x in row[1:]) + "\n")
        assert pa.cli_gk_text(pa.ReadLog_file(p, None, 1, 1, 200, []), heads) == txt


def _fmt_rows(mean, var, w):
    m, v = mean.reshape(-1, w), var.reshape(-1, w)
    return "".join("".join("%g %g " % (m[i, j], v[i, j]) for j in range(w)) + "\n" for i in range(m.shape[0]))


def test_cli_msd_text(pa):
    """-q, -Q and --mean-square-displacement-self with -s 10 -S 50 (the
    options of the reference's MSD_cm / MSD_normal files): each line is a
    lag's block-averaged mean and variance of every column, 6 significant
    digits, each value followed by a space; the -Q lines start with the -q
    lines' columns and lag 0 is all zeros, as in those files."""
    N, F = 64, 1200
    box_l = synth.lammps_box(10.0)
    pos = synth.random_walk(91, N, F, box_l, 0.1)
    types = synth.types_by_fraction(N, [0.5, 0.5]).astype(np.int32)
    t = pa.Trajectory(pos, np.zeros_like(pos), types, np.tile(box_l, (F, 1)), pa.BoxFormat.LammpsOrtho, False, False)
    nb, skip, stop = 20, 10, 50
    txt = {}
    for cm, me in ((False, False), (True, False), (False, True)):
        txt[cm, me] = pa.cli_msd_text(t, nb, skip, stop, cm, me)
        mean, var = pa.msd_block_average(t, nb, skip, stop, cm, me)
        assert txt[cm, me] == _fmt_rows(mean, var, 2 * (2 if cm else 1))
    q, Q = txt[False, False].split("\n")[:-1], txt[True, False].split("\n")[:-1]
    assert len(q) == len(Q) == stop
    assert q[0] == "0 " * 4 and Q[0] == "0 " * 8
    assert all(b.split()[:4] == a.split() for a, b in zip(q, Q))


def test_cli_gofrt_text(pa, tmp_path):
    """-g 12 -F 0 6 -S 6 -e 2 with 2 blocks: the columns' legend, then per
    lag (step 2) and bin the line "t r" with every channel's mean and
    variance, two empty lines after each lag.  The legend is pinned by
    physics: rmax exceeds the largest minimum-image distance of the 6 A cell,
    so every pair is counted, and the distinct column of g(a, b) sums over r
    to the ordered pair count 2 N_a N_b (a != b) or N_a (N_a - 1), the
    same-atom column of g(a, a) to N_a and of g(a, b != a) to 0; types of
    30/20/10 atoms make every count different.  Over a LAMMPS binary (which
    the command line wraps into the cell) the text is the same as over the
    wrapped in-memory trajectory."""
    N, F = 60, 40
    box_l = synth.lammps_box(6.0)
    pos = synth.random_walk(17, N, F, box_l, 0.2)
    types = np.repeat([0, 1, 2], [30, 20, 10]).astype(np.int32)
    sizes = [30, 20, 10]
    t = pa.Trajectory(pos, np.zeros_like(pos), types, np.tile(box_l, (F, 1)), pa.BoxFormat.LammpsOrtho, False, False)
    nbin, stop, every = 12, 6, 2
    txt = pa.cli_gofrt_text(t, 2, 0.0, 6.0, nbin, stop, 1, every)
    mean, var = pa.gofrt_block_average(t, 2, 0.0, 6.0, nbin, stop, 1, every)
    lines = txt.split("\n")
    legend = [l for l in lines if l.startswith("#")]
    assert len(legend) == 2 * 6 + 3 and lines[:len(legend)] == legend
    body = "\n".join(lines[len(legend):])
    want = ""
    for lag in range(0, mean.shape[0], every):
        for r in range(nbin):
            want += "%d %d" % (lag, r) + "".join(" %g %g" % (mean[lag, c, r], var[lag, c, r])
                                                  for c in range(mean.shape[1])) + "\n"
        want += "\n\n"
    assert body == want
    rows = [l.split() for l in lines[len(legend):] if l]
    for entry in legend[1:-2]:
        pair, col = entry[1:8], int(entry.rsplit(" ", 1)[1])   # "#g(a, b), ... index: col"
        a, b = int(entry[3]), int(entry[6])
        same = "same atom" in entry
        want_sum = (sizes[a] if a == b else 0) if same else (sizes[a] * (sizes[a] - 1) if a == b else 2 * sizes[a] * sizes[b])
        for lag in range(0, mean.shape[0], every):
            got = sum(float(r[col - 1]) for r in rows if int(r[0]) == lag)
            assert got == pytest.approx(want_sum, rel=1e-4, abs=1e-9), (pair, same, lag)
    path = str(tmp_path / "g.bin")
    t.write_lammps_binary(path, 0, -1)
    tl = pa.Traj(path)
    tl.setAccessWindowSize(F)          # nothing loaded yet: the printer turns the wrap on first, as the CLI does
    tw = pa.Trajectory(pos, np.zeros_like(pos), types, np.tile(box_l, (F, 1)), pa.BoxFormat.LammpsOrtho, True, False)
    assert pa.cli_gofrt_text(tl, 2, 0.0, 6.0, nbin, stop, 1, every) == pa.cli_gofrt_text(tw, 2, 0.0, 6.0, nbin, stop, 1, every)


def test_cli_gk_text_covariances_and_factors(pa, analisi_log, gk_ref):
    """-H with -F 2 3 and -z 0 2 1 3 (main.cpp:266-274, 430-513): every column
    times its factor (factor_conv * factor_intToCorr = 6 for the 4
    correlations, factor_conv = 2 for the rest), variances times its square,
    then the two covariances times both columns' factors, against the
    oracle's block statistics (MediaVarCovar, calcoliblocchi.h:72-202)
    printed with 6 significant digits; an odd -z list is an error."""
    heads = ["c_flux[1]", "c_vcm[1][1]"]
    txt = pa.cli_gk_text(analisi_log, heads, factor_conv=2.0, factor_intToCorr=3.0, covariance=[0, 2, 1, 3])
    lines = txt.split("\n")
    assert lines[0] == "#Factors: 2, 6" and lines[1] == "# factor integral, factor correlation" and lines[2] == "#2 3"
    rows = [l for l in lines if l and not l.startswith("#")]
    log = gk_ref["log"]
    hd = list(gk_ref["headers"])
    idx = [hd.index(h) for h in heads]
    T, nb = log.shape[0], 20
    s = (T - O.n_extra_timesteps(T, nb, 0)) // nb
    narr = 3 * 4 + 2
    mvc = O.MediaVarCovar(narr, s, [(0, 2), (1, 3)])
    for b in range(nb):
        mvc.add(O.gk(log, idx, 1, 0, False, 0, 1, s, b * s))
    mvc.end()
    f = np.array([6.0] * 4 + [2.0] * (narr - 4))
    assert len(rows) == s
    for i in (0, 1, 7, s // 2, s - 1):
        want = " ".join(sum([["%g" % (mvc.mean[j, i] * f[j]), "%g" % (mvc.var[j, i] * f[j] * f[j])] for j in range(narr)], []))
        want += " " + " ".join("%g" % (mvc.cov[k, i] * f[a] * f[b]) for k, (a, b) in enumerate([(0, 2), (1, 3)])) + " "
        assert rows[i].replace("-nan", "nan") == want.replace("-nan", "nan"), i
    with pytest.raises(RuntimeError, match="even number"):
        pa.cli_gk_text(analisi_log, heads, covariance=[0, 2, 1])
