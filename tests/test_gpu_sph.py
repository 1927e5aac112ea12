"""Spherical-harmonic density correlations and Steinhardt order parameters on
the GPU (anl_sph_*, anl_steinhardt_calculate) against the oracle
(sphericalbase.cpp:19-66, sphericalcorrelations.cpp:62-270,
steinhardt.cpp:90-236).

FP64 throughout; the device sums in a different order than the reference
(densities in slot order, products as a Gram matrix over (atom, m), t1 sums
in chunks), so values agree to rounding: the tolerance is 1e-12 of each
array's largest magnitude.  Densities and correlations are checked against
the oracle at small sizes (orthorhombic and triclinic cells, 1-3 types,
skip/tmax/primo variations, several density chunks), and the full
2,048-atom workload through its own snapshots: every lag equals the numpy
contraction of the device densities of its frames.  Steinhardt q values agree
to 1e-12 and their histograms equal the oracle's."""
import numpy as np
import pytest

from analisi_amd import capi, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _close(got, want, tol=TOL):
    scale = max(np.abs(want).max(), 1e-300)
    err = np.abs(got - want).max() / scale
    assert err <= tol, (err, scale)


def _case(seed, N, F, L, frac, tri, sigma=0.25):
    tilt = (2.0 * synth.uniform(seed, 7, 3) - 1.0) * 0.15 * L if tri else None
    box_l = synth.lammps_box(L, tilt) if tri else synth.lammps_box(L)
    pos = synth.random_walk(seed, N, F, box_l, sigma)
    types = synth.types_by_fraction(N, frac)
    ot = O.Traj(pos, types, np.tile(box_l, (F, 1)), tri)
    dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    return ot, dt


def _ranges(nt, seed):
    u = synth.uniform(seed, 3, 2 * nt * nt).reshape(nt * nt, 2)
    lo = 0.4 + 0.6 * u[:, 0]
    return np.stack([lo, lo + 2.2 + 0.8 * u[:, 1]], axis=1)


@pytest.mark.parametrize("tri,frac", [(False, [1.0]), (False, [0.5, 0.5]), (True, [0.6, 0.3, 0.1])])
def test_snapshot_matches_oracle(tri, frac):
    ot, dt = _case(31, 70, 4, 8.5, frac, tri)
    rm = _ranges(ot.ntypes, 5)
    s = capi.DeviceSph(dt, rm, 3)
    for f in (0, 3):
        got = s.snapshot(f)
        want = O.sph_density(ot, f, rm, 3)
        assert np.count_nonzero(want) > 0
        _close(got, want)


@pytest.mark.parametrize("tri,frac,skip,tmax,n,primo", [
    (False, [0.5, 0.5], 1, 4, 6, 0),
    (False, [0.5, 0.5], 2, 3, 7, 2),
    (True, [0.6, 0.3, 0.1], 3, 0, 5, 1),
    (True, [1.0], 1, 1, 4, 3),
])
def test_correlations_match_oracle(tri, frac, skip, tmax, n, primo):
    ot, dt = _case(41, 60, 14, 8.0, frac, tri)
    rm = _ranges(ot.ntypes, 9)
    s = capi.DeviceSph(dt, rm, 3)
    got = s.calculate(n, tmax, skip, primo)
    want = O.sph_correlations(ot, rm, 3, tmax, skip, n, primo)
    assert got.shape == want.shape
    _close(got, want)


def test_chunked_densities_and_lower_lmax():
    """A budget of a few frames splits the t1 values into many chunks; lmax 6
    (the other instantiation of the reference) uses the same kernels."""
    ot, dt = _case(43, 50, 40, 7.5, [0.7, 0.3], False)
    rm = _ranges(ot.ntypes, 4)
    for lmax in (10, 6):
        s = capi.DeviceSph(dt, rm, 2, lmax)
        whole = s.calculate(30, 6, 1, 2)
        fbytes = 50 * 2 * 2 * (lmax + 1) ** 2 * 8
        s.set_budget(9 * fbytes)          # 6 lag frames + 4 t1 values per chunk
        chunked = s.calculate(30, 6, 1, 2)
        _close(chunked, whole, 1e-13)
        want = O.sph_correlations(ot, rm, 2, 6, 1, 30, 2, lmax)
        _close(chunked, want)


def test_errors():
    ot, dt = _case(47, 20, 8, 7.0, [1.0], False)
    s = capi.DeviceSph(dt, [(0.5, 3.0)], 2)
    with pytest.raises(capi.AnlError, match="too short") as e:
        s.calculate(5, 3, 1, 1)           # 3 + 5 + 1 > 8
    assert e.value.code == capi.ANL_ERR_TOO_SHORT
    s.calculate(5, 3, 1, 0)
    with pytest.raises(capi.AnlError, match="rmax > rmin"):
        capi.DeviceSph(dt, [(3.0, 0.5)], 2)
    with pytest.raises(capi.AnlError, match="lmax"):
        capi.DeviceSph(dt, [(0.5, 3.0)], 2, 11)


def test_fullsize_lags_equal_snapshot_contractions():
    """2,048 atoms (2 types), 0.1 atoms/A^3, 4 bins in [0.7, 3.5) A, lmax 10:
    C[dt][ti][tj][b][lmax-l] = mean over t1 and over atoms of type ti of
    sum_m S_t1 S_t1+dt, with S the device's own snapshots."""
    N, L = 2048, (2048 / 0.1) ** (1 / 3)
    ot, dt = _case(53, N, 40, L, [0.5, 0.5], False, 0.1)
    rm = [(0.7, 3.5)] * 4
    s = capi.DeviceSph(dt, rm, 4)
    n, tmax, skip, primo = 20, 8, 3, 5
    got = s.calculate(n, tmax, skip, primo)
    snaps = {f: s.snapshot(f) for f in range(primo, primo + n + tmax - 1)}
    # neighbours per atom near 0.1 * 4/3 pi (3.5^3 - 0.7^3): the bins are populated
    counts = snaps[primo][:, :, :, -1] * np.sqrt(4 * np.pi)
    assert 12 < counts.sum(axis=(1, 2)).mean() < 24
    t1s = list(range(0, n, skip))
    ti = ot.type_id
    lsz = [2 * l + 1 for l in range(10, -1, -1)]
    edges = np.concatenate([[0], np.cumsum(lsz)])
    want = np.zeros_like(got)
    for d in range(tmax):
        for t1 in t1s:
            p = snaps[primo + t1] * snaps[primo + t1 + d]
            msum = np.stack([p[..., edges[k]:edges[k + 1]].sum(axis=-1) for k in range(11)], axis=-1)
            for t in range(2):
                want[d, t] += msum[ti == t].mean(axis=0)
        want[d] /= len(t1s)
    _close(got, want, 1e-11)
    assert np.all(got[0, ..., 10] > 0)


# ---------------------------------------------------------------------------
# Steinhardt order parameters (steinhardt.cpp:90-236)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("tri,frac,skip,n,primo", [(False, [0.5, 0.5], 2, 7, 1), (True, [0.6, 0.3, 0.1], 3, 9, 0),
                                                  (False, [1.0], 1, 3, 4)])
def test_steinhardt_matches_oracle(tri, frac, skip, n, primo):
    """q_1..q_lmax of every (bin, frame, atom, type_j) within 1e-12 (q is at
    most 1); the (q4, q6) and (q2, q4, q6) histograms equal the oracle's."""
    ot, dt = _case(59, 80, 12, 8.5, frac, tri)
    rm = _ranges(ot.ntypes, 13)
    for lmax in (10, 6):
        s = capi.DeviceSph(dt, rm, 3, lmax)
        got = s.steinhardt(n, skip, primo, [4, 6], 16, do_histogram=False)
        want = O.steinhardt(ot, rm, 3, lmax, 16, [4, 6], skip, n, primo, False)
        assert got.shape == want.shape
        assert np.abs(got - want).max() <= 1e-12
        for ll, nb in (([4, 6], 16), ([2, 4, 6], 7), ([], 5)):
            h = s.steinhardt(n, skip, primo, ll, nb, do_histogram=True)
            hw = O.steinhardt(ot, rm, 3, lmax, nb, ll, skip, n, primo, True)
            assert h.shape == hw.shape and h.sum() > 0
            np.testing.assert_allclose(h, hw, rtol=1e-13, atol=0)


def test_steinhardt_fcc_and_errors():
    """A jittered fcc crystal (12 neighbours in (1.2, 1.6) A): q4 and q6 of
    every atom are the tabulated 0.191 and 0.575; errors as the reference."""
    g = np.array([[i, j, k] for i in range(3) for j in range(3) for k in range(3)], dtype=float)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    pos = (np.concatenate([(g + b) * 2.0 for b in basis]) + 0.1)[None]
    pos = pos + 1e-5 * synth.normal(3, 5, pos.size).reshape(pos.shape)
    pos = np.repeat(pos, 3, axis=0)
    N = pos.shape[1]
    ot = O.Traj(pos, np.zeros(N, dtype=np.int32), np.tile(synth.lammps_box(6.0), (3, 1)), False)
    dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    s = capi.DeviceSph(dt, [(1.2, 1.6)], 1, 6)
    q = s.steinhardt(3, 1, 0, [4, 6], 10, do_histogram=False)
    assert np.all(np.abs(q[..., 3] - 0.191) < 1e-3) and np.all(np.abs(q[..., 5] - 0.575) < 1e-3)
    h = s.steinhardt(3, 1, 0, [4, 6], 10)
    assert h[0, 0, 0, 1, 5] == pytest.approx(N * 1.0)   # floor(0.191*10) = 1, floor(0.575*10) = 5; 3 frames x incr 1/3
    with pytest.raises(capi.AnlError, match="too short"):
        s.steinhardt(3, 1, 1, [4, 6], 10)
    with pytest.raises(capi.AnlError, match="Wrong value of l"):
        s.steinhardt(3, 1, 0, [4, 8], 10)


# ---------------------------------------------------------------------------
# SANN neighbour lists (neighbour.cpp:8-101)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("tri", [False, True])
def test_sann_densities_correlations_steinhardt(tri):
    """Densities, correlations and Steinhardt q / histograms with SANN
    neighbour lists, plain and averaged (steinhardt.cpp:128-173), equal the
    oracle's (1e-12; histograms exactly)."""
    ot, dt = _case(79, 90, 12, 8.0, [0.6, 0.4], tri)
    spec = [(40, 3.3 ** 2, 0.0), (30, 3.6 ** 2, 0.0)]
    s = capi.DeviceSph(dt, _ranges(2, 3), 2, 10)
    s.set_neighbours(spec)
    for f in (0, 7):
        want, cnt = O.sph_density_sann(ot, f, spec, 2)
        assert cnt.max() >= 3
        _close(s.snapshot(f), want)
    _close(s.calculate(6, 4, 2, 1), O.sph_correlations_sann(ot, spec, 2, 4, 2, 6, 1))
    s1 = capi.DeviceSph(dt, _ranges(2, 3), 1, 6)
    s1.set_neighbours(spec)
    q = s1.steinhardt(9, 2, 2, [4, 6], 20, do_histogram=False)
    assert np.abs(q - O.steinhardt_sann(ot, spec, 6, 20, [4, 6], 2, 9, 2, False)).max() <= 1e-12
    h = s1.steinhardt(9, 2, 2, [4, 6], 20)
    np.testing.assert_allclose(h, O.steinhardt_sann(ot, spec, 6, 20, [4, 6], 2, 9, 2, True), rtol=1e-13, atol=0)
    for hist in (False, True):
        qa = s1.steinhardt(9, 2, 2, [4, 6], 20, do_histogram=hist, averaged=True)
        qw = O.steinhardt_sann(ot, spec, 6, 20, [4, 6], 2, 9, 2, hist, averaged=True)
        np.testing.assert_allclose(qa, qw, rtol=1e-13 if hist else 0, atol=0 if hist else 1e-12)
    with pytest.raises(capi.AnlError, match="neighbour list"):
        capi.DeviceSph(dt, _ranges(2, 3), 1, 6).steinhardt(9, 2, 2, [4, 6], 20, averaged=True)
    small = capi.DeviceSph(dt, _ranges(2, 3), 1, 6)
    small.set_neighbours([(3, 3.3 ** 2, 0.0), (30, 3.6 ** 2, 0.0)])
    with pytest.raises(capi.AnlError, match="Too many neighbours"):
        small.snapshot(0)


@pytest.mark.parametrize("late_first", [False, True])
def test_sann_capacity_follows_the_reference_pair_order(late_first):
    """A list that is exactly full when a later atom of its type comes within
    the other type's (larger) cutoff throws, as the reference's pair loop
    does (neighbour.cpp:15-47); the same atoms in another index order do not
    (the oracle test of the same case says why)."""
    from test_oracle_sph import _exactly_full_case
    pos, types, box_l = _exactly_full_case(late_first)
    ot = O.Traj(pos, types, np.tile(box_l, (2, 1)), False)
    dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    s = capi.DeviceSph(dt, _ranges(2, 3), 1, 6)
    s.set_neighbours([(10, 2.0 ** 2, 0.0), (2, 1.2 ** 2, 0.0)])
    if late_first:
        _close(s.snapshot(0), O.sph_density_sann(ot, 0, [(10, 4.0, 0.0), (2, 1.44, 0.0)], 1, 6)[0])
    else:
        with pytest.raises(capi.AnlError, match="Too many neighbours"):
            s.snapshot(0)


def test_steinhardt_histogram_equals_binned_full_order_q():
    """The histogram expands the densities only up to its largest l; binning
    the per-atom q (expanded to lmax) by the reference's rule
    (steinhardt.cpp:194-210: floorf of the float product, clamped, first l
    slowest, weight 1/int(n/skip), atoms with neighbours only) gives the same
    histogram, at 2,048 atoms and for l lists ending below lmax."""
    N, L = 2048, (2048 / 0.1) ** (1 / 3)
    ot, dt = _case(61, N, 12, L, [0.5, 0.5], False, 0.1)
    s = capi.DeviceSph(dt, [(0.7, 3.5)] * 4, 2, 10)
    n, skip, primo, nb = 10, 3, 1, 40
    q = s.steinhardt(n, skip, primo, [], nb, do_histogram=False)      # [bin][frame][atom][tj][l-1]
    has = np.any(q != 0, axis=-1)
    ti = ot.type_id
    for ll in ([4, 6], [2, 5, 8], [3]):
        h = s.steinhardt(n, skip, primo, ll, nb)
        want = np.zeros_like(h)
        idx = [np.minimum(np.floor((q[..., l - 1] * nb).astype(np.float32)).astype(np.int64), nb - 1) for l in ll]
        for b, f, a, tj in zip(*np.nonzero(has)):
            want[(b, ti[a], tj) + tuple(int(x[b, f, a, tj]) for x in idx)] += 1.0 / (n // skip)
        assert h.sum() > 0
        np.testing.assert_allclose(h, want, rtol=1e-12, atol=0)
