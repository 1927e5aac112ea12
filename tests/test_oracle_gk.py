"""Pins the oracle's Green-Kubo path (greenkuboNcomponentionicfluid.cpp:156-356)
and block statistics (calcoliblocchi.h, blockaverage.h:107-190) against the
reference's own golden vectors, whose input (gk_integral.dat) is in the
snapshot."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, is_same
from oracle import oracle as O


def _cols(ref, names):
    hd = list(ref["headers"])
    return [hd.index(h) for h in names]


def test_gk_2comp_golden(gk_ref):
    # test_gk.cpp:8-35: GK(c_flux[1], c_vcm[1][1]), lmax 0, reset(5000), calculate(0)
    out = O.gk(gk_ref["log"], _cols(gk_ref, ["c_flux[1]", "c_vcm[1][1]"]), 1, 0, False, 100, 1, 5000, 0)
    assert out.shape == gk_ref["gk_2comp"].shape
    assert is_same(out, gk_ref["gk_2comp"]).all()
    # k_einst is NaN at lag 1 (0/0 Schur complement), as in the golden
    assert np.isnan(out[1, 13]) and np.isnan(gk_ref["gk_2comp"][1, 13])


def test_gk_nseg4_golden(gk_ref):
    # test_gk.py:4-12: lmax 2000, n_seg 4, reset(T-2000), 2 threads
    log = gk_ref["log"]
    out = O.gk(log, _cols(gk_ref, ["c_flux[1]", "c_vcm[1][1]"]), 1, 2000, False, 0, 4, log.shape[0] - 2000, 0)
    assert is_same(out, gk_ref["test_gk"]).all()


@pytest.mark.parametrize("name,heads", [("single", ["c_flux[1]"]), ("double", ["c_flux[1]", "c_vcm[1][1]"])])
def test_cli_block_average_text(gk_ref, name, heads):
    """main.cpp:456-516: 20 blocks of (T - nExtra)/20 steps, MediaVarCovar,
    printed with iostream's default 6 significant digits.  Text-identical."""
    log = gk_ref["log"]
    idx = _cols(gk_ref, heads)
    T, nb = log.shape[0], 20
    s = (T - O.n_extra_timesteps(T, nb, 0)) // nb
    narr = 3 * len(idx) ** 2 + 2
    mvc = O.MediaVarCovar(narr, s)
    for b in range(nb):
        mvc.add(O.gk(log, idx, 1, 0, False, 0, 1, s, b * s))
    mvc.end()
    mine = []
    for i in range(s):
        parts = []
        for j in range(narr):
            parts += ["%g" % mvc.mean[j, i], "%g" % mvc.var[j, i]]
        mine.append(" ".join(parts) + " ")
    with open(os.path.join(GOLDEN, "cli_GK_heat_%s.txt" % name)) as f:
        ref = [l.rstrip("\n") for l in f if l.strip() and not l.startswith("#")]
    assert len(ref) == s
    norm = lambda x: x.replace("-nan", "nan")
    assert [norm(a) for a in mine] == [norm(b) for b in ref]


def test_gk_integrals_closed_form():
    """Constant current J=(c,c,c): C(l)=c^2 for every lag, trapezoid I(l)=c^2 l,
    Einstein E(l) = sum_k c^2 k = c^2 l(l+1)/2 - c^2 l/2 ... divided by l."""
    c = 2.0
    data = np.full((64, 3), c)
    out = O.gk(data, [0], 1, 16, False, 0, 1, 32, 0)
    C = out[:, 0]
    np.testing.assert_array_equal(C, c * c)
    l = np.arange(16)
    np.testing.assert_array_equal(out[:, 1], c * c * l)          # I(l)
    # E(l) = sum_{k=1..l} (C(k-1)k + C(k)k)/2 = c^2 l(l+1)/2, then /l
    np.testing.assert_allclose(out[1:, 2], c * c * (l[1:] + 1) / 2.0, rtol=1e-15)
    np.testing.assert_array_equal(out[1:, 3], out[1:, 1])        # k_gk = I for one current


def test_gk_schur_three_currents_vs_numpy():
    """n=3 currents: k = M00 - M01 inv(M11) M10 of the symmetrised integral
    matrix, against numpy.linalg (the bit-level pin against Eigen's own
    evaluation is test_schur_matches_eigen below)."""
    from analisi_amd import synth
    J = synth.ar2_currents(3, 4000, ncur=3)
    out = O.gk(J, [0, 3, 6], 1, 200, False, 0, 1, 3000, 0)
    n = 3
    for l in (5, 50, 199):
        I = out[l, n * n:3 * n * n:2].reshape(n, n)
        M = (I + I.T) / 2
        k = M[0, 0] - M[0, 1:] @ np.linalg.inv(M[1:, 1:]) @ M[1:, 0]
        assert abs(out[l, 3 * n * n] - k) <= 1e-9 * max(1.0, abs(k))


def test_mediavar_closed_form():
    rng = np.random.default_rng(0)
    blocks = rng.normal(size=(7, 50))
    mv = O.MediaVar(50)
    for b in blocks:
        mv.add(b)
    mv.end(7)
    np.testing.assert_allclose(mv.mean, blocks.mean(0), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(mv.var, blocks.var(0, ddof=1) / 7, rtol=1e-12)


def _eigen_or_skip(mats):
    ref = O.ref_schur00(mats)
    if ref is None:
        pytest.skip("oracle/_ref/ref_schur not built (it needs the reference's vendored Eigen headers)")
    return ref


def test_schur_matches_eigen():
    """The oracle's (0,0) Schur complement against Eigen 3.3.9's evaluation of
    the reference's expression (greenkuboNcomponentionicfluid.cpp:309-316,
    oracle/ref_schur.cpp): bit-identical up to 5 currents (a 4 x 4 inverse,
    where Eigen's LU and triangular solves run element by element); from 6
    currents Eigen's triangular solve updates through its packed block kernel,
    whose sums are ordered differently, and the two agree to a few ulps."""
    rng = np.random.default_rng(2024)
    for n in range(1, 9):
        mats = []
        for t in range(300):
            A = rng.normal(size=(n, n)) * 10.0 ** rng.uniform(-3, 3)
            M = (A + A.T) / 2
            if t % 3:
                M += n * np.abs(M).max() * np.eye(n)   # well conditioned
            mats.append(M)
        ref = _eigen_or_skip(mats)
        got = np.array([O.schur00(M) for M in mats])
        if n <= 5:
            assert (got == ref).all(), (n, int((got != ref).sum()))
        else:
            np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0)


@pytest.mark.parametrize("ncur", [3, 4, 5])
def test_gk_kappa_matches_eigen(ncur):
    """The whole GK epilogue for 3 to 5 currents: the symmetrised integral
    matrix built as the reference builds it (each product's integral added
    as half to both mirrored entries, greenkuboNcomponentionicfluid.cpp:
    293-306) and its complement by Eigen must equal the oracle's k_gk bit for
    bit at every lag."""
    from analisi_amd import synth
    J = synth.ar2_currents(7 + ncur, 3000, ncur=ncur)
    idx = [3 * k for k in range(ncur)]
    out = O.gk(J, idx, 1, 120, False, 0, 1, 2500, 0)
    N = ncur * ncur
    mats = []
    for l in range(1, out.shape[0]):
        M = np.zeros((ncur, ncur))
        i = 0
        for j1 in range(ncur):
            for j2 in range(ncur):
                v = out[l, N + 2 * i]
                M[j1, j2] += v / 2.0   # matr[j2 * n + j1], column-major (row j1, column j2)
                M[j2, j1] += v / 2.0
                i += 1
        mats.append(M)
    ref = _eigen_or_skip(mats)
    got = out[1:, 3 * N]
    assert (got == ref).all(), int((got != ref).sum())
