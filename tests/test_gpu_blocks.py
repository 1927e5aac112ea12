"""Block engine (uneven blocks in one launch, device MediaVar, HIP graph) on
the GPU against the oracle's per-block Gofrt / MSD and MediaVar
(blockaverage.h:107-203, calcoliblocchi.h:21-65)."""
import numpy as np
import pytest

from analisi_amd import capi, synth
from conftest import is_same
from gpu_util import first_diff, make_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _blocks(seed, nb, F, lo, hi, reach):
    lens = (lo + np.floor(synth.uniform(seed, 20, nb) * (hi - lo + 1))).astype(np.int64)
    starts = np.floor(synth.uniform(seed, 21, nb) * (F - lens - reach)).astype(np.int64)
    return starts.astype(np.uint64), lens.astype(np.uint32)


@pytest.mark.parametrize("form", [capi.GOFRT_FP64, capi.GOFRT_SCREENED])
def test_uneven_blocks_match_oracle(form):
    F = 700
    ot, dt = make_case(21, 400, F, 16.0, [0.6, 0.3, 0.1], 0.08)
    primo, n = _blocks(5, 6, F, 60, 150, 40)
    gt, mt = 21, 31
    g = capi.DeviceGofrt(dt, 0.0, 6.0, 48, gt, 2, 1, form)
    eng = capi.DeviceBlocks(dt, primo, n, gofrt=g, msd=True, msd_skip=3, msd_tmax=mt, cm_msd=True, cm_self=True,
                            use_graph=True)
    eng.run()
    gmean, gvar, counts = eng.gofrt_results()
    mmean, mvar, mper = eng.msd_results()
    mvg = O.MediaVar(gmean.size)
    mvm = O.MediaVar(mmean.size)
    for b in range(len(primo)):
        c, _ = O.gofrt(ot, 0.0, 6.0, 48, gt, 3, 2, 1, int(n[b]), int(primo[b]), want_vdata=False)
        assert first_diff(counts[b], c) is None, (b, first_diff(counts[b], c))
        incr = 1.0 / (int(n[b]) // 2)
        mvg.add(c.astype(np.float64) * incr)
        m = O.msd(ot, 3, mt, 2, True, True, int(n[b]), int(primo[b]))
        assert is_same(mper[b], m).all()
        mvm.add(mper[b])
    mvg.end(len(primo))
    mvm.end(len(primo))
    # same inputs, same Welford order: bit-identical
    np.testing.assert_array_equal(gmean.ravel(), mvg.mean)
    np.testing.assert_array_equal(gvar.ravel(), mvg.var)
    np.testing.assert_array_equal(mmean.ravel(), mvm.mean)
    np.testing.assert_array_equal(mvar.ravel(), mvm.var)


def test_graph_replays_equal_eager():
    """Every replay of the captured graph equals an eager evaluation, with
    other eager device work and synchronisations between the replays."""
    F = 400
    ot, dt = make_case(22, 300, F, 15.0, [0.5, 0.5], 0.1, tilt=[0.7, -0.6, 0.5])
    primo, n = _blocks(6, 4, F, 50, 90, 30)
    g1 = capi.DeviceGofrt(dt, 0.0, 5.5, 40, 15, 1, 1, capi.GOFRT_AUTO)
    eager = capi.DeviceBlocks(dt, primo, n, gofrt=g1, msd=True, msd_skip=1, msd_tmax=20, use_graph=False)
    eager.run()
    ref_g = eager.gofrt_results()
    ref_m = eager.msd_results()
    g2 = capi.DeviceGofrt(dt, 0.0, 5.5, 40, 15, 1, 1, capi.GOFRT_AUTO)
    graph = capi.DeviceBlocks(dt, primo, n, gofrt=g2, msd=True, msd_skip=1, msd_tmax=20, use_graph=True)
    other = capi.DeviceGofrt(dt, 0.2, 3.0, 10, 5, 1, 1, capi.GOFRT_FP64)
    other.reset(20)
    for rep in range(3):
        graph.run()
        got_g = graph.gofrt_results()
        got_m = graph.msd_results()
        for a, b in zip(got_g, ref_g):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(got_m, ref_m):
            np.testing.assert_array_equal(a, b)
        other.calculate(1)                      # eager work between replays


def test_blocks_reject_unequal_lags_and_short():
    F = 120
    ot, dt = make_case(23, 50, F, 10.0, [1.0], 0.1)
    g = capi.DeviceGofrt(dt, 0.0, 4.0, 10, 30, 1, 1)
    with pytest.raises(capi.AnlError):       # n = 20 < tmax gives 20 lags, n = 40 gives 30
        capi.DeviceBlocks(dt, [0, 10], [20, 40], gofrt=g)
    with pytest.raises(capi.AnlError) as e:
        capi.DeviceBlocks(dt, [0, 90], [40, 40], gofrt=g)
    assert e.value.code == capi.ANL_ERR_TOO_SHORT


@pytest.mark.parametrize("graph", [True, False])
def test_several_engines_share_one_gofrt(graph):
    """Several block engines built from one Gofrt, replayed after all of them
    exist and in an interleaved order: each window equals its own single call
    (each engine keeps its own work list and block table on the device)."""
    F = 500
    ot, dt = make_case(23, 350, F, 15.0, [0.7, 0.3], 0.1, tilt=[0.5, -0.4, 0.3])
    g = capi.DeviceGofrt(dt, 0.0, 5.0, 32, 12, 1, 1, capi.GOFRT_AUTO)
    sets = [_blocks(7 + k, 3 + k, F, 40 + 10 * k, 90 + 20 * k, 20) for k in range(3)]
    engines = [capi.DeviceBlocks(dt, p, n, gofrt=g, msd=False, use_graph=graph) for p, n in sets]
    for order in ([2, 0, 1], [1, 2, 0]):
        for k in order:
            engines[k].run()
        g1 = capi.DeviceGofrt(dt, 0.0, 5.0, 32, 12, 1, 1, capi.GOFRT_FP64)  # single calls between replays
        for k, (p, n) in enumerate(sets):
            _, _, counts = engines[k].gofrt_results()
            for b in range(len(p)):
                g1.reset(int(n[b]))
                one = g1.calculate(int(p[b]))
                assert first_diff(counts[b], one) is None, (order, k, b, first_diff(counts[b], one))
