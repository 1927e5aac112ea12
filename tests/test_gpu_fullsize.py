"""The 16,384-atom triclinic g(r,t) workload at its full size (16,384 atoms,
types 80/10/10 %, 2,000 frames, 256 bins to 12 A; the bench's seed and cell)
on the GPU, checked through properties that hold at any size, and against the
oracle on the (t0, lag) pairs it can afford:
  * lag 0: every self channel holds N_type per t0, all in bin 0 (rmin = 0);
    every distinct channel is even (each pair counted both ways);
  * every lag: the counts of a lag add up to at most n0 N^2 ordered pairs;
  * the screened form (cluster and 16-bit screens) equals the plain FP64
    form bit for bit on a slice, and every wavefront-t0 unit clusters;
  * one t0 x lag 0 and one (t0, lag 16) single-thread call equal the oracle.
"""
import numpy as np
import pytest

from analisi_amd import capi, synth
from gpu_util import first_diff
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RMIN, RMAX, NBIN = 0.0, 12.0, 256


def _self_channel(nt, t):
    half = nt * (nt + 1) // 2
    return half + half - (t + 1) * (t + 2) // 2 + t


@pytest.fixture(scope="module")
def c2():
    seed, N, F, L = 2024, 16384, 2000, 56.0
    tilt = (2.0 * synth.uniform(seed, 7, 3) - 1.0) * 0.1 * L
    box_l = synth.lammps_box(L, tilt)
    pos = synth.random_walk(seed, N, F, box_l, 0.05)
    types = synth.types_by_fraction(N, [0.8, 0.1, 0.1])
    ot = O.Traj(pos, types, np.tile(box_l, (F, 1)), True)
    dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    return ot, dt


def test_fullsize_invariants_forms_and_oracle(c2):
    ot, dt = c2
    N, nt = ot.pos.shape[1], ot.ntypes
    tmax, n, primo = 17, 24, 1500   # late t0, after the atoms have diffused
    g = capi.DeviceGofrt(dt, RMIN, RMAX, NBIN, tmax, 1, 1, capi.GOFRT_SCREENED)
    g.reset(n)
    c = g.calculate(primo)
    assert c.shape == (tmax, nt * (nt + 1), NBIN)
    ntype = np.bincount(ot.type_id, minlength=nt)
    half = nt * (nt + 1) // 2
    for t in range(nt):
        ch = _self_channel(nt, t)
        assert int(c[0, ch, 0]) == int(ntype[t]) * n and int(c[0, ch, 1:].sum()) == 0
    assert np.all(c[0, :half] % 2 == 0)
    assert np.all(c.reshape(tmax, -1).sum(axis=1) <= n * N * N)
    clustered, units = g.cluster_stats()
    assert units > 0 and clustered == units

    g64 = capi.DeviceGofrt(dt, RMIN, RMAX, NBIN, tmax, 1, 1, capi.GOFRT_FP64)
    g64.reset(2)
    gsl = capi.DeviceGofrt(dt, RMIN, RMAX, NBIN, tmax, 1, 1, capi.GOFRT_SCREENED)
    gsl.reset(2)
    a, b = g64.calculate(primo), gsl.calculate(primo)
    assert first_diff(a, b) is None, first_diff(a, b)

    ref, _ = O.gofrt(ot, RMIN, RMAX, NBIN, 1, 8, 1, 1, 1, primo, want_vdata=False)
    g1 = capi.DeviceGofrt(dt, RMIN, RMAX, NBIN, 1, 1, 1, capi.GOFRT_SCREENED)
    g1.reset(1)
    got = g1.calculate(primo)
    assert first_diff(got, ref) is None, first_diff(got, ref)

    leff = tmax
    ref = np.zeros((leff, nt * (nt + 1), NBIN), dtype=np.uint64)
    O.gofrt_single(ot, RMIN, RMAX, NBIN, leff, 16, 3, 0, N, primo, ref)
    got = g.calc_single_th(16, 3, 0, N, primo, None)
    assert first_diff(got, ref) is None, first_diff(got, ref)


def test_fullsize_green_kubo_blocks():
    """The Green-Kubo block workload at its full size: 2 vector currents of
    2,000,000 AR(2) steps (the bench's seed), 32 uneven blocks x 20,000
    lags in one call.  The split-BF16 MFMA form stays within 1e-6 of each
    column's scale sqrt(C_aa(0) C_bb(0)) of the FP64 form for C(l) and within
    1e-6 (l + 1) of it for the running integral; each block of the batched
    FP64 call equals its own single call bit for bit (blocks 0, 13 and 31);
    the block mean is the plain mean of the blocks to rounding."""
    T, lmax, nb, seed = 2_000_000, 20_000, 32, 55
    J = synth.ar2_currents(seed, T, ncur=2, coupling=0.3)
    span = T - lmax
    u = synth.uniform(seed, 31, nb)
    lens = lmax + np.floor((span - nb * lmax) * u / u.sum()).astype(np.int64)
    lens[-1] += span - lens.sum()
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
    s = capi.DeviceSeries(J)
    f64, mean64, _ = s.gk_blocks([0, 3], 1, lmax, 1, False, 0, starts, lens, mfma=False)
    mf, _, _ = s.gk_blocks([0, 3], 1, lmax, 1, False, 0, starts, lens, mfma=True)
    assert f64.shape == (nb, lmax, 14)
    for b in range(nb):
        c0 = np.sqrt(np.abs(np.diag(f64[b, 0, :4].reshape(2, 2))))
        sig = np.outer(c0, c0).reshape(-1)
        dC = np.abs(mf[b, :, :4] - f64[b, :, :4]) / sig
        assert dC.max() <= 1e-6, (b, dC.max())
        dI = np.abs(mf[b, :, 4:12:2] - f64[b, :, 4:12:2]) / sig / (np.arange(lmax)[:, None] + 1.0)
        assert dI.max() <= 1e-6, (b, dI.max())
    for b in (0, 13, 31):
        one = s.gk([0, 3], 1, lmax, int(lens[b]), 1, False, 0, int(starts[b]))
        same = np.array_equal(np.nan_to_num(one, nan=0.25), np.nan_to_num(f64[b], nan=0.25))
        assert same, b
    fin = np.isfinite(mean64[:, :4])
    np.testing.assert_allclose(mean64[:, :4][fin], f64[:, :, :4].mean(axis=0)[fin], rtol=1e-9, atol=1e-12)
