"""g(r,t) on the GPU against the oracle: counts bit-exact for both arithmetic
forms, every cell shape, wrapped and unwrapped positions, skip/every,
calc_single_th atom ranges, error paths; and the full-size invariants."""
import numpy as np
import pytest

from analisi_amd import capi
from gpu_util import make_case, first_diff
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    # id, natoms, nframes, L, fractions, sigma, tilt, wrap, shift
    ("ortho", 300, 24, 14.0, [0.5, 0.5], 0.2, None, False, 0.0),
    ("ortho-unwrapped", 257, 20, 12.0, [0.6, 0.3, 0.1], 0.3, None, False, 37.0),
    ("tri", 301, 22, 13.0, [0.8, 0.1, 0.1], 0.2, [1.3, -1.1, 0.9], False, 0.0),
    ("tri-unwrapped", 190, 18, 11.0, [0.5, 0.5], 0.25, [-1.0, 0.8, -1.2], False, -29.0),
    ("tri-wrapped", 222, 18, 11.5, [1.0], 0.25, [0.7, 0.6, -0.5], True, 5.0),
]


@pytest.mark.parametrize("form", [capi.GOFRT_FP64, capi.GOFRT_SCREENED])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_counts_match_oracle(case, form):
    _, N, F, L, fr, sig, tilt, wrap, shift = case
    ot, dt = make_case(7, N, F, L, fr, sig, tilt, wrap, shift)
    rmin, rmax, nbin, tmax, skip, every, n = 0.3, 5.2, 57, 6, 2, 1, 11
    ref, _ = O.gofrt(ot, rmin, rmax, nbin, tmax, 4, skip, every, n, 1, want_vdata=False)
    g = capi.DeviceGofrt(dt, rmin, rmax, nbin, tmax, skip, every, form)
    g.reset(n)
    got = g.calculate(1)
    assert first_diff(got, ref) is None, first_diff(got, ref)
    used, cand, nre = g.stats()
    assert used == form
    if form == capi.GOFRT_SCREENED:
        nin = int(ref.sum())
        assert cand >= nin           # every counted pair was a candidate


def test_every_and_skip_and_lag_layout():
    ot, dt = make_case(3, 200, 40, 10.0, [0.5, 0.5], 0.2)
    for skip, every in [(1, 3), (5, 2), (7, 1)]:
        ref, _ = O.gofrt(ot, 0.0, 4.0, 40, 9, 2, skip, every, 23, 3, want_vdata=False)
        for form in (capi.GOFRT_FP64, capi.GOFRT_SCREENED):
            g = capi.DeviceGofrt(dt, 0.0, 4.0, 40, 9, skip, every, form)
            g.reset(23)
            got = g.calculate(3)
            assert first_diff(got, ref) is None, (skip, every, form, first_diff(got, ref))


def _scrambled_case(seed, N, F, L, fractions, tilt):
    """Every frame independent uniform random positions: each frame's slot
    order of the screen differs from every other frame's."""
    from analisi_amd import synth
    box_l = synth.lammps_box(L, tilt)
    pos = np.stack([synth.random_walk(seed + f, N, 1, box_l, 0.0)[0] for f in range(F)])
    types = synth.types_by_fraction(N, fractions)
    ot = O.Traj(pos, types, np.tile(box_l, (F, 1)), tilt is not None)
    return ot, capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)


@pytest.mark.parametrize("form", [capi.GOFRT_FP64, capi.GOFRT_SCREENED])
@pytest.mark.parametrize("scrambled", [False, True])
def test_calc_single_th_ranges(form, scrambled):
    """Atom ranges of calc_single_th select i atoms by their original index,
    also when the screen holds each frame in its own slot order."""
    if scrambled:
        ot, dt = _scrambled_case(5, 150, 12, 9.0, [0.3, 0.7], [0.5, 0.4, -0.3])
    else:
        ot, dt = make_case(5, 150, 12, 9.0, [0.3, 0.7], 0.2, [0.5, 0.4, -0.3])
    rmin, rmax, nbin, leff = 0.0, 4.0, 33, 5
    g = capi.DeviceGofrt(dt, rmin, rmax, nbin, leff, 1, 1, form)
    g.reset(6)
    acc = None
    ref = np.zeros((leff, 6, nbin), dtype=np.uint64)
    for (t, im, a0, a1) in [(0, 0, 0, 150), (3, 2, 10, 77), (4, 5, 149, 150), (2, 1, 0, 0)]:
        acc = g.calc_single_th(t, im, a0, a1, 2, acc)
        O.gofrt_single(ot, rmin, rmax, nbin, leff, t, im, a0, a1, 2, ref)
    assert first_diff(acc, ref) is None, first_diff(acc, ref)
    assert g.stats()[0] == form


def test_bin_edges_adversarial():
    """Pairs placed at distances within a few ulps of bin edges, at rmin and
    at rmax: the device bin must be the reference's float-rounded bin."""
    rmin, rmax, nbin = 0.5, 4.5, 40
    dr = (rmax - rmin) / nbin
    d = []
    for k in range(0, nbin + 1):
        e = rmin + k * dr
        d += [np.nextafter(e, 0), e, np.nextafter(e, 10), e * (1 - 2e-8), e * (1 + 2e-8)]
    d = np.array(d)
    M = len(d)
    # atom pairs along x: atom 2m at 0, atom 2m+1 at d[m]; far apart in y,z
    L = 400.0
    pos = np.zeros((2, 2 * M, 3))
    pos[:, 0::2, 1] = np.arange(M) * 6.0
    pos[:, 1::2, 1] = np.arange(M) * 6.0
    pos[:, 1::2, 0] = d
    types = np.zeros(2 * M, np.int32)
    box = np.tile([0, L, 0, L, 0, L], (2, 1)).astype(float)
    ot = O.Traj(pos, types, box, False)
    dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    ref, _ = O.gofrt(ot, rmin, rmax, nbin, 1, 1, 1, 1, 1, 0, want_vdata=False)
    for form in (capi.GOFRT_FP64, capi.GOFRT_SCREENED):
        g = capi.DeviceGofrt(dt, rmin, rmax, nbin, 1, 1, 1, form)
        g.reset(1)
        got = g.calculate(0)
        assert first_diff(got, ref) is None, (form, first_diff(got, ref))


@pytest.mark.parametrize("origin", [20.0, -6.0], ids=["inside", "corner"])
@pytest.mark.parametrize("tilt", [None, (2.5, -1.5, 1.0)], ids=["cubic", "tri"])
def test_bin_edges_adversarial_cluster(tilt, origin):
    """Pairs within a few ulps of bin edges, rmin and rmax, inside one compact
    cloud, so that wavefronts take the cluster screen: every mode (including
    the offset-difference evaluation of modes 2 and 4 and its radius-dependent
    window) must give the reference's float-rounded bins.  The corner cloud
    straddles the cell's corner, so pairs and offsets cross the periodic
    boundary in every direction."""
    from analisi_amd import synth
    rmin, rmax, nbin = 0.5, 4.5, 40
    dr = (rmax - rmin) / nbin
    d = []
    for k in range(0, nbin + 1):
        e = rmin + k * dr
        d += [np.nextafter(e, 0), e, np.nextafter(e, 10), e * (1 - 2e-8), e * (1 + 2e-8)]
    d = np.array(d)
    M = len(d)
    L = 60.0
    rng = np.random.default_rng(5)
    base = origin + 12.0 * rng.random((M, 3))
    pos = np.zeros((2, 2 * M, 3))
    pos[:, 0::2] = base
    pos[:, 1::2] = base
    pos[:, 1::2, 0] += d
    types = (np.arange(2 * M) % 2).astype(np.int32)
    box = np.tile(synth.lammps_box(L, None if tilt is None else np.array(tilt)), (2, 1))
    ot = O.Traj(pos, types, box, tilt is not None)
    dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    ref, _ = O.gofrt(ot, rmin, rmax, nbin, 1, 1, 1, 1, 1, 0, want_vdata=False)
    assert ref.sum() >= M  # every planted pair is inside [rmin, rmax] or at its edges
    g = capi.DeviceGofrt(dt, rmin, rmax, nbin, 1, 1, 1, capi.GOFRT_SCREENED)
    g.reset(1)
    for cl in (2, 1, 4, 3, 0):
        g.set_tuning(8, 0, 1 << 30, cl)
        got = g.calculate(0)
        assert first_diff(got, ref) is None, (cl, first_diff(got, ref))
        c, tot = g.cluster_stats()
        if cl:
            assert 0 < c <= tot, (cl, c, tot)
        _, _, rechecks = g.stats()
        assert rechecks > 0, cl


@pytest.mark.parametrize("tilt,origin", [(None, 20.0), ((2.5, -1.5, 1.0), -6.0)], ids=["cubic", "tri-corner"])
def test_recheck_two_stages(tilt, origin):
    """Pairs near every bin edge, rmin and rmax at two scales: 1e-6 and 1e-5
    of the distance (inside the FP32 window; mostly outside the error bound
    of the FP64 evaluation from the 32-bit fractions, which then decides
    them) and a few ulps to 2e-8 (inside that bound, so the reference's FP64
    sequence from the positions decides them).  Every cluster mode must give
    the reference's bins, and both stages must have run."""
    from analisi_amd import synth
    rmin, rmax, nbin = 0.5, 4.5, 40
    dr = (rmax - rmin) / nbin
    d = []
    for k in range(0, nbin + 1):
        e = rmin + k * dr
        d += [e * (1 - 1e-5), e * (1 - 1e-6), e * (1 + 1e-6), e * (1 + 1e-5)]
        d += [np.nextafter(e, 0), e, np.nextafter(e, 10), e * (1 - 2e-8), e * (1 + 2e-8)]
    d = np.array(d)
    M = len(d)
    L = 60.0
    rng = np.random.default_rng(11)
    base = origin + 12.0 * rng.random((M, 3))
    pos = np.zeros((2, 2 * M, 3))
    pos[:, 0::2] = base
    pos[:, 1::2] = base
    pos[:, 1::2, 0] += d
    types = (np.arange(2 * M) % 2).astype(np.int32)
    box = np.tile(synth.lammps_box(L, None if tilt is None else np.array(tilt)), (2, 1))
    ot = O.Traj(pos, types, box, tilt is not None)
    dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    ref, _ = O.gofrt(ot, rmin, rmax, nbin, 1, 1, 1, 1, 1, 0, want_vdata=False)
    g = capi.DeviceGofrt(dt, rmin, rmax, nbin, 1, 1, 1, capi.GOFRT_SCREENED)
    g.reset(1)
    for cl in (2, 1, 4, 3, 0):
        g.set_tuning(8, 0, 1 << 30, cl)
        got = g.calculate(0)
        assert first_diff(got, ref) is None, (cl, first_diff(got, ref))
        _, _, rechecks = g.stats()
        window, exact = g.recheck_stats()
        assert window == rechecks, (cl, window, rechecks)
        # the ulp-level pairs (three per edge, both orders) take the exact sequence
        assert 6 * (nbin + 1) <= exact < window, (cl, exact, window)


@pytest.mark.parametrize("tilt,origin", [(None, 20.0), ((2.5, -1.5, 1.0), -6.0)], ids=["cubic", "tri-corner"])
def test_fp32_window_edges(tilt, origin):
    """Pairs along random directions at 1e-6 to 1e-4 A on either side of
    every bin edge, rmin and rmax: around the FP32 window's width (1e-5 to
    1e-4 A here), so that some are decided by the FP32 evaluation alone,
    just outside the window, and some by the rechecks.  Every cluster mode
    and the 16-bit screen must give the reference's bins."""
    from analisi_amd import synth
    rmin, rmax, nbin = 0.5, 4.5, 40
    dr = (rmax - rmin) / nbin
    offs = np.array([1e-6, 3e-6, 1e-5, 3e-5, 1e-4])
    d = []
    for k in range(0, nbin + 1):
        e = rmin + k * dr
        d += list(e - offs) + list(e + offs)
    d = np.array(d)
    M = len(d)
    L = 60.0
    rng = np.random.default_rng(17)
    base = origin + 12.0 * rng.random((M, 3))
    n = rng.normal(size=(M, 3))
    n /= np.linalg.norm(n, axis=1)[:, None]
    pos = np.zeros((3, 2 * M, 3))
    pos[:, 0::2] = base
    pos[:, 1::2] = base + d[:, None] * n
    types = (np.arange(2 * M) % 2).astype(np.int32)
    box = np.tile(synth.lammps_box(L, None if tilt is None else np.array(tilt)), (3, 1))
    ot = O.Traj(pos, types, box, tilt is not None)
    dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    ref, _ = O.gofrt(ot, rmin, rmax, nbin, 2, 1, 1, 1, 2, 0, want_vdata=False)
    g = capi.DeviceGofrt(dt, rmin, rmax, nbin, 2, 1, 1, capi.GOFRT_SCREENED)
    g.reset(2)
    for cl in (2, 1, 4, 3, 0):
        g.set_tuning(8, 0, 1 << 30, cl)
        got = g.calculate(0)
        assert first_diff(got, ref) is None, (cl, first_diff(got, ref))
        _, cand, rechecks = g.stats()
        assert 0 < rechecks < cand, (cl, rechecks, cand)


@pytest.mark.parametrize("tilt,origin", [(None, 20.0), ((2.5, -1.5, 1.0), -6.0)], ids=["cubic", "tri-corner"])
def test_self_pairs_at_edges(tilt, origin):
    """Each atom moves, frame to frame, by a distance at a bin edge, rmin or
    rmax (ulps to 1e-4 A away, along a random direction), so that the pairs
    of an atom with itself at lag 1 sit where the FP32 window, the FP64
    stage and the exact sequence decide them.  The screened kernels bin
    every pair, self pairs included, into the distinct rows and then move
    each self pair to its self row with the reference's FP64 bin: the self
    and distinct channels must both be the reference's in every mode."""
    from analisi_amd import synth
    rmin, rmax, nbin = 0.5, 4.5, 40
    dr = (rmax - rmin) / nbin
    offs = np.array([0.0, 2e-8, 1e-6, 1e-5, 1e-4])
    d = []
    for k in range(0, nbin + 1):
        e = rmin + k * dr
        d += list(e - offs) + list(e + offs[1:])
    d = np.array(d)
    M = len(d)
    L = 60.0
    rng = np.random.default_rng(23)
    base = origin + 14.0 * rng.random((M, 3))
    n = rng.normal(size=(M, 3))
    n /= np.linalg.norm(n, axis=1)[:, None]
    pos = base[None, :, :] + np.arange(3)[:, None, None] * d[None, :, None] * n[None, :, :]
    types = (np.arange(M) % 2).astype(np.int32)
    box = np.tile(synth.lammps_box(L, None if tilt is None else np.array(tilt)), (3, 1))
    ot = O.Traj(pos, types, box, tilt is not None)
    dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    ref, _ = O.gofrt(ot, rmin, rmax, nbin, 2, 1, 1, 1, 2, 0, want_vdata=False)
    half = ref.shape[1] // 2
    assert ref[1, half:].sum() >= M // 2  # most planted self pairs are counted at lag 1
    g = capi.DeviceGofrt(dt, rmin, rmax, nbin, 2, 1, 1, capi.GOFRT_SCREENED)
    g.reset(2)
    for cl in (2, 1, 4, 3, 0):
        g.set_tuning(8, 0, 1 << 30, cl)
        got = g.calculate(0)
        assert first_diff(got, ref) is None, (cl, first_diff(got, ref))
        window, exact = g.recheck_stats()
        assert exact > 0 and window > exact, (cl, window, exact)


def test_too_short_raises():
    ot, dt = make_case(2, 50, 10, 8.0, [1.0], 0.1)
    g = capi.DeviceGofrt(dt, 0.0, 3.0, 10, 3, 1, 1)
    g.reset(8)
    g.calculate(0)                           # 3 + 8 + 0 <= 10 + 1
    with pytest.raises(capi.AnlError) as e:
        g.calculate(1)
    assert e.value.code == capi.ANL_ERR_TOO_SHORT
    assert "too short" in str(e.value)


def test_screened_refuses_large_rmax():
    ot, dt = make_case(2, 50, 4, 8.0, [1.0], 0.1)
    g = capi.DeviceGofrt(dt, 0.0, 4.5, 10, 1, 1, 1, capi.GOFRT_SCREENED)
    g.reset(2)
    with pytest.raises(capi.AnlError):
        g.calculate(0)
    g2 = capi.DeviceGofrt(dt, 0.0, 4.5, 10, 1, 1, 1, capi.GOFRT_AUTO)
    g2.reset(2)
    ref, _ = O.gofrt(ot, 0.0, 4.5, 10, 1, 1, 1, 1, 2, 0, want_vdata=False)
    assert first_diff(g2.calculate(0), ref) is None
    assert g2.stats()[0] == capi.GOFRT_FP64


@pytest.mark.parametrize("form", [capi.GOFRT_FP64, capi.GOFRT_SCREENED])
def test_launch_shapes_and_autotune_keep_counts(form):
    """Every launch shape (unroll, XCD order, t0 chunk down to one t0 per
    workgroup) gives the oracle's counts; the autotuner times the built-in
    candidates, keeps one of them and the counts stay the same."""
    ot, dt = make_case(12, 700, 30, 15.0, [0.7, 0.3], 0.15, [1.1, -0.9, 0.6])
    ref, _ = O.gofrt(ot, 0.0, 6.0, 48, 9, 4, 1, 1, 20, 0, want_vdata=False)
    g = capi.DeviceGofrt(dt, 0.0, 6.0, 48, 9, 1, 1, form)
    g.reset(20)
    for unroll, xcd, lp, cl in [(4, 0, 31, 1), (8, 1, 31, 0), (4, 1, 12, 2), (8, 0, 16, 0), (4, 0, 40, 1),
                                (8, 0, 30, 2), (8, 1, 29, 3), (8, 0, 18, 4)]:
        g.set_tuning(unroll, xcd, 1 << lp, cl)
        assert g.tuning() == (unroll, xcd, 1 << lp, cl)
        got = g.calculate(0)
        assert first_diff(got, ref) is None, (unroll, xcd, lp, cl, first_diff(got, ref))
    ms, best = g.autotune(0, 3)
    assert len(ms) == 9 and (ms > 0).all() and 0 <= best < 9
    assert first_diff(g.calculate(0), ref) is None
    ms, best = g.autotune(0, 2, [(8, 1, 20, 1), (4, 0, 22, 0)])
    assert g.tuning() == [(8, 1, 1 << 20, 1), (4, 0, 1 << 22, 0)][best]
    assert first_diff(g.calculate(0), ref) is None
    with pytest.raises(capi.AnlError):
        g.autotune(0, 2, [(5, 0, 20, 1)])
    with pytest.raises(capi.AnlError):
        g.autotune(0, 2, [(8, 0, 20, 5)])


# Boxes large against rmax, so that compact wavefronts take the cluster screen:
# id, natoms, nframes, L, fractions, sigma, tilt, wrap, unwrapped shift, rmax,
# first frame of sparse uniform random positions (None: none)
CLUSTER_CASES = [
    ("tri", 4000, 10, 40.0, [0.7, 0.3], 0.3, [3.1, -2.7, 2.2], False, 0.0, 7.0, None),
    ("ortho-unwrapped", 3000, 9, 36.0, [0.5, 0.5], 0.4, None, False, 61.0, 6.5, None),
    ("tri-wrapped", 2500, 8, 34.0, [1.0], 0.5, [-2.0, 1.5, -1.0], True, 0.0, 8.0, None),
    ("tri-mixed", 400, 12, 40.0, [0.6, 0.4], 0.3, [1.0, 1.0, -1.0], False, 0.0, 6.0, 6),
]


@pytest.mark.parametrize("case", CLUSTER_CASES, ids=[c[0] for c in CLUSTER_CASES])
def test_cluster_screen_matches_oracle(case):
    """The cluster (FP16 MFMA) screen gives the oracle's counts.  In the last
    case the first 6 frames hold all 400 atoms in an 8 A blob and the later
    frames spread them uniformly over the 40 A cell: 64 atoms of a sparse
    frame span about 20 A, their wavefronts fall back to the 16-bit screen
    in the same launch, while the blob frames' wavefronts use the cluster
    screen."""
    from analisi_amd import synth
    name, N, F, L, fr, sig, tilt, wrap, shift, rmax, scramble = case
    if scramble is None:
        ot, dt = make_case(21, N, F, L, fr, sig, tilt, wrap, shift)
    else:
        box_l = synth.lammps_box(L, tilt)
        pos = synth.random_walk(21, N, F, synth.lammps_box(8.0), sig * 0.1) + 16.0
        for f in range(scramble, F):
            pos[f] = synth.random_walk(1000 + f, N, 1, box_l, 0.0)[0]
        types = synth.types_by_fraction(N, fr)
        ot = O.Traj(pos, types, np.tile(box_l, (F, 1)), tilt is not None)
        dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    nbin, tmax, n = 64, 3, F - 3
    ref, _ = O.gofrt(ot, 0.0, rmax, nbin, tmax, 8, 1, 1, n, 1, want_vdata=False)
    g = capi.DeviceGofrt(dt, 0.0, rmax, nbin, tmax, 1, 1, capi.GOFRT_SCREENED)
    g.reset(n)
    for cl in (4, 3, 2, 1, 0):
        g.set_tuning(8, 0, 1 << 30, cl)
        got = g.calculate(1)
        assert first_diff(got, ref) is None, (cl, first_diff(got, ref))
        c, tot = g.cluster_stats()
        if cl:
            assert 0 < c <= tot, (c, tot)
            if scramble is not None:
                assert c < tot, (c, tot)
        else:
            assert tot == 0
