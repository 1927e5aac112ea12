"""MSD and Green-Kubo on the GPU against the oracle and the reference's
goldens.  Tolerance: the reference's own is_same, 1e-10 relative
(tests/src/test_fixtures.h:110-122)."""
import numpy as np
import pytest

from analisi_amd import capi, synth
from conftest import is_same
from gpu_util import make_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cm_msd,cm_self", [(False, False), (True, False), (True, True), (False, True)])
def test_msd_matches_oracle(cm_msd, cm_self):
    ot, dt = make_case(9, 333, 60, 15.0, [0.5, 0.3, 0.2], 0.1, unwrap_shift=20.0)
    for skip, tmax, n, primo in [(1, 20, 30, 3), (4, 0, 25, 0), (3, 17, 40, 2)]:
        ref = O.msd(ot, skip, tmax, 3, cm_msd, cm_self, n, primo)
        got = capi.msd(dt, skip, tmax, cm_msd, cm_self, n, primo)
        assert got.shape == ref.shape
        assert is_same(got, ref).all(), np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300))
        if cm_msd:   # the centre-of-mass channel replays the running mean exactly
            assert np.array_equal(got[:, 1, :], ref[:, 1, :])


@pytest.mark.parametrize("cm_self", [False, True])
def test_msd_lag_tiles_and_autotune(cm_self):
    """Every launch shape (lag tile, wavefronts per workgroup) of the
    t0-stride-1 kernel gives the same bits (each (atom, lag) sum runs over t0
    in order, the workgroup sums its atoms in slot order); tile 0, the
    general kernel, agrees to rounding; all are within 1e-10 of the oracle.
    Lag counts that are not multiples of a workgroup's lag span (from 32 to
    192), an atom count that is not a multiple of 64 and n not a multiple of
    the tile exercise the partial tiles.  The autotuner keeps one of the
    candidates and the result stays the same."""
    ot, dt = make_case(21, 331, 90, 14.0, [0.6, 0.4], 0.12, unwrap_shift=-17.0)
    shapes = dt.MSD_SHAPES
    for tmax, n, primo in [(37, 41, 5), (0, 29, 0), (50, 40, 0)]:
        ref = O.msd(ot, 1, tmax, 3, True, cm_self, n, primo)
        outs = {}
        for tile, waves in shapes:
            dt.set_msd_tile(tile, waves)
            assert dt.msd_shape() == (tile, waves)
            outs[tile, waves] = capi.msd(dt, 1, tmax, True, cm_self, n, primo)
            assert is_same(outs[tile, waves], ref).all(), (tile, waves, tmax, n)
        for sh in shapes:
            if sh[0]:
                assert np.array_equal(outs[sh], outs[8, 4]), (sh, tmax, n)
        ms, best = dt.msd_autotune(1, tmax, True, cm_self, n, primo)
        assert len(ms) == len(shapes) and np.all(ms > 0) and dt.msd_shape() == shapes[best]
        assert np.array_equal(capi.msd(dt, 1, tmax, True, cm_self, n, primo), outs[dt.msd_shape()])
    with pytest.raises(capi.AnlError):
        dt.set_msd_tile(24, 12)     # 24 frames of registers do not fit 12 wavefronts


def test_msd_errors():
    ot, dt = make_case(9, 20, 20, 5.0, [1.0], 0.1)
    with pytest.raises(capi.AnlError) as e:
        capi.msd(dt, 1, 10, False, False, 10, 1)      # 10+10+1 > 20 (no +1 for MSD)
    assert e.value.code == capi.ANL_ERR_TOO_SHORT
    capi.msd(dt, 1, 10, False, False, 10, 0)


def _cols(ref, names):
    hd = list(ref["headers"])
    return [hd.index(h) for h in names]


def test_gk_golden_and_oracle(gk_ref):
    log = gk_ref["log"]
    idx = _cols(gk_ref, ["c_flux[1]", "c_vcm[1][1]"])
    s = capi.DeviceSeries(log)
    got = s.gk(idx, 1, 0, 5000, 1, False, 100, 0)
    assert is_same(got, gk_ref["gk_2comp"]).all()
    ref = O.gk(log, idx, 1, 0, False, 100, 1, 5000, 0)
    same = (got == ref) | (np.isnan(got) & np.isnan(ref))
    assert same.all(), "device GK differs from the oracle at %d entries" % (~same).sum()
    got4 = s.gk(idx, 1, 2000, log.shape[0] - 2000, 4, False, 0, 0)
    assert is_same(got4, gk_ref["test_gk"]).all()


@pytest.mark.parametrize("ncur,skip,n_seg,sub", [(1, 1, 1, False), (2, 3, 5, True), (3, 2, 2, False), (4, 1, 3, True)])
def test_gk_synthetic_matches_oracle(ncur, skip, n_seg, sub):
    J = synth.ar2_currents(11, 6000, ncur=ncur)
    data = np.concatenate([np.ones((6000, 1)), J], axis=1)     # a leading unused column
    idx = [1 + 3 * k for k in range(ncur)]
    s = capi.DeviceSeries(data)
    got = s.gk(idx, skip, 300, 4000, n_seg, sub, 50, 17)
    ref = O.gk(data, idx, skip, 300, sub, 50, n_seg, 4000, 17)
    same = (got == ref) | (np.isnan(got) & np.isnan(ref))
    if ncur <= 2:
        assert same.all()
    else:   # Schur complement of >= 2x2 blocks: same algorithm, order unpinned
        assert is_same(got[:, :3 * ncur * ncur], ref[:, :3 * ncur * ncur]).all()
        np.testing.assert_allclose(got[1:, -2:], ref[1:, -2:], rtol=1e-9)


def test_gk_too_short():
    s = capi.DeviceSeries(np.zeros((100, 3)))
    with pytest.raises(capi.AnlError) as e:
        s.gk([0], 1, 50, 60, 1, False, 0, 0)
    assert e.value.code == capi.ANL_ERR_TOO_SHORT
