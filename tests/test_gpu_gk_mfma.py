"""Green-Kubo with the lag products on the matrix cores (split-BF16 MFMA)
against the FP64 oracle (greenkuboNcomponentionicfluid.cpp:156-356).

Tolerances, per correlation column ab with scale
sigma_ab = sqrt(C_aa(0) C_bb(0)) (the Cauchy-Schwarz bound of |C_ab(l)|):
  * AR(2) currents: |dC_ab(l)| <= 1e-6 sigma_ab, and the running integrals
    |dI_ab(l)| <= 1e-6 sigma_ab (l + 1), the bound their sum of C gives;
  * integer currents |J| <= 8: every value is exact in bf16 (lo = 0) and
    every FP32 chunk sum is an exact integer, so C agrees with the FP64 form
    to 1e-12 sigma_ab -- any misplaced tile row, column or lag shows up here.
"""
import numpy as np
import pytest

from analisi_amd import capi, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# ncur, skip, n_seg, tmax, n, primo
CASES = [
    (1, 1, 1, 0, 700, 0),        # leff = n, one chunk, one tile group
    (2, 1, 1, 300, 9000, 5),     # two FP32 chunks of 8192 t0
    (2, 3, 5, 2100, 4000, 17),   # skip and segments
    (3, 2, 2, 4500, 5000, 1),    # 142 lag tiles: two tile groups
    (4, 1, 3, 37, 1000, 3),      # four currents, short lags
]


def _scales(ref, ncur):
    c0 = ref[0, :ncur * ncur].reshape(ncur, ncur)
    d = np.sqrt(np.abs(np.diag(c0)))
    return np.outer(d, d).reshape(-1)


def _series(kind, ncur, rows, seed):
    if kind == "int":
        rng = np.random.default_rng(seed)
        J = rng.integers(-8, 9, size=(rows, 3 * ncur)).astype(np.float64)
    else:
        J = synth.ar2_currents(seed, rows, ncur=ncur)
    return np.ascontiguousarray(np.concatenate([np.full((rows, 1), 7.0), J], axis=1))


@pytest.mark.parametrize("kind", ["int", "ar2"])
@pytest.mark.parametrize("case", CASES, ids=[str(c) for c in CASES])
def test_gk_mfma_matches_fp64_oracle(kind, case):
    ncur, skip, n_seg, tmax, n, primo = case
    leff = n if (tmax == 0 or n < tmax) else tmax
    rows = leff + primo + n
    data = _series(kind, ncur, rows, 31 + ncur)
    idx = [1 + 3 * k for k in range(ncur)]
    sub = n_seg == 3
    s = capi.DeviceSeries(data)
    got = s.gk(idx, skip, tmax, n, n_seg, sub, 10, primo, mfma=True)
    ref = O.gk(data, idx, skip, tmax, sub, 10, n_seg, n, primo)
    assert got.shape == ref.shape
    nc = ncur * ncur
    sig = _scales(ref, ncur)
    dC = np.abs(got[:, :nc] - ref[:, :nc]) / sig
    tol = 1e-12 if kind == "int" else 1e-6
    assert dC.max() <= tol, (dC.max(), np.unravel_index(dC.argmax(), dC.shape))
    I_got, I_ref = got[:, nc:3 * nc:2], ref[:, nc:3 * nc:2]
    dI = np.abs(I_got - I_ref) / sig / (np.arange(leff)[:, None] + 1.0)
    assert dI.max() <= tol, dI.max()
    # the FP64 form of the same call is the oracle's, bit for bit (ncur <= 2)
    if ncur <= 2:
        f64 = s.gk(idx, skip, tmax, n, n_seg, sub, 10, primo)
        assert np.array_equal(np.nan_to_num(f64, nan=1.5), np.nan_to_num(ref, nan=1.5))


def test_gk_mfma_errors():
    s = capi.DeviceSeries(np.zeros((100, 3)))
    with pytest.raises(capi.AnlError) as e:
        s.gk([0], 1, 50, 60, 1, False, 0, 0, mfma=True)
    assert e.value.code == capi.ANL_ERR_TOO_SHORT
    with pytest.raises(capi.AnlError) as e:
        s.gk([1], 1, 5, 20, 1, False, 0, 0, mfma=True)
    assert e.value.code == capi.ANL_ERR_HEADER


def test_gk_mfma_pyanalisi_reference_call(gk_ref):
    """The reference's test call (tests/test_gk.py:4-12) with mfma=True,
    against the golden FP64 result (tests/test_gk.py snapshot)."""
    from analisi_amd import pyanalisi as pa
    log = pa.ReadLog(np.array(gk_ref["log"], copy=True), list(gk_ref["headers"]))
    gk = pa.GreenKubo(log, '', 1, ['c_flux[1]', 'c_vcm[1][1]'], False, 2000, 2, False, 0, 4, False, 1, 100,
                      mfma=True)
    gk.reset(log.getNtimesteps() - 2000)
    gk.calculate(0)
    m = np.array(gk, copy=True)
    ref = gk_ref["test_gk"]
    assert m.shape == ref.shape
    sig = _scales(ref, 2)
    dC = np.abs(m[:, :4] - ref[:, :4]) / sig
    assert dC.max() <= 1e-6, dC.max()


@pytest.mark.parametrize("mfma", [False, True])
def test_gk_blocks_match_single_calls_and_mediavar(mfma):
    """Uneven blocks in one call: every block equals its own single call (same
    kernels, bit for bit), and the block mean / variance equal the oracle's
    MediaVar over those blocks (calcoliblocchi.h:21-65, the arithmetic of
    MediaVarCovar without covariance pairs)."""
    J = synth.ar2_currents(77, 9000, ncur=2)
    s = capi.DeviceSeries(J)
    primo = [0, 1500, 2200, 4100, 7000]
    n = [1400, 700, 1900, 2800, 900]
    tmax = 350
    per, mean, var = s.gk_blocks([0, 3], 2, tmax, 2, True, 20, primo, n, mfma=mfma)
    assert per.shape == (5, tmax, 14)
    for b in range(5):
        one = s.gk([0, 3], 2, tmax, n[b], 2, True, 20, primo[b], mfma=mfma)
        assert np.array_equal(np.nan_to_num(per[b], nan=0.25), np.nan_to_num(one, nan=0.25)), b
        if not mfma:
            ref = O.gk(J, [0, 3], 2, tmax, True, 20, 2, n[b], primo[b])
            assert np.array_equal(np.nan_to_num(one, nan=0.25), np.nan_to_num(ref, nan=0.25)), b
    mv = O.MediaVar(per[0].size)
    for b in range(5):
        mv.add(per[b].ravel())
    mv.end(5)
    same = lambda a, b_: np.array_equal(np.nan_to_num(a, nan=0.25), np.nan_to_num(b_, nan=0.25))
    assert same(mean.ravel(), mv.mean) and same(var.ravel(), mv.var)


def test_gk_blocks_errors():
    s = capi.DeviceSeries(np.zeros((1000, 3)))
    with pytest.raises(capi.AnlError) as e:      # different leff
        s.gk_blocks([0], 1, 100, 1, False, 0, [0, 10], [200, 50])
    assert e.value.code == capi.ANL_ERR_ARG
    with pytest.raises(capi.AnlError) as e:      # block 1 runs past the end
        s.gk_blocks([0], 1, 100, 1, False, 0, [0, 800], [200, 200])
    assert e.value.code == capi.ANL_ERR_TOO_SHORT


@pytest.mark.parametrize("mfma", [False, True])
def test_gk_blocks_more_than_one_batch(mfma):
    """70 blocks in one call run as two batches of launches (at most 64
    blocks each); every block still equals its own single call."""
    J = synth.ar2_currents(78, 7400, ncur=1)
    s = capi.DeviceSeries(J)
    primo = [100 * b for b in range(70)]
    n = [100 + 20 * (b % 7) for b in range(70)]
    tmax = 50
    per, mean, var = s.gk_blocks([0], 1, tmax, 1, False, 0, primo, n, mfma=mfma)
    assert per.shape == (70, tmax, 5)
    for b in (0, 5, 63, 64, 69):
        one = s.gk([0], 1, tmax, n[b], 1, False, 0, primo[b], mfma=mfma)
        assert np.array_equal(np.nan_to_num(per[b], nan=0.25), np.nan_to_num(one, nan=0.25)), b


def test_gk_mfma_shapes_and_autotune():
    """Every launch shape of the MFMA lag products (tiles per wavefront 4|8,
    chunk steps 8|16, wavefronts per workgroup 2|4): shapes with the same
    chunk length give the same bits; every shape is exact on integer currents
    and within 1e-6 of the FP64 oracle on AR(2) currents; the autotuner keeps
    one of its candidates and the result is that shape's."""
    ncur, skip, n_seg, tmax, n, primo = 2, 1, 2, 1300, 20000, 3
    idx = [1, 4]
    shapes = [(q, cs, wg) for q in (4, 8) for cs in (8, 16) for wg in (2, 4)]
    for kind in ("int", "ar2"):
        data = _series(kind, ncur, tmax + primo + n, 41)
        ref = O.gk(data, idx, skip, tmax, False, 0, n_seg, n, primo)
        sig = _scales(ref, ncur)
        s = capi.DeviceSeries(data)
        outs = {}
        for sh in shapes:
            s.set_gk_shape(*sh)
            assert s.gk_shape() == sh
            outs[sh] = s.gk(idx, skip, tmax, n, n_seg, False, 0, primo, mfma=True)
            dC = np.abs(outs[sh][:, :4] - ref[:, :4]) / sig
            assert dC.max() <= (1e-12 if kind == "int" else 1e-6), (kind, sh, dC.max())
        same = lambda a, b: np.array_equal(np.nan_to_num(a, nan=0.5), np.nan_to_num(b, nan=0.5))
        for sh in shapes:
            base = (4, sh[1], 2)
            assert same(outs[sh], outs[base]), (kind, sh)
        ms, best = s.gk_autotune(idx, skip, tmax, n_seg, [primo, primo + 10], [n - 10, n - 10])
        assert len(ms) == 6 and np.all(ms > 0)
        chosen = s.gk_shape()
        assert chosen in shapes
        assert same(s.gk(idx, skip, tmax, n, n_seg, False, 0, primo, mfma=True), outs[chosen])
        s.close()
