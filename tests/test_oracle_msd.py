"""Pins the oracle's MSD path (msd.cpp:40-146, trajectory_numpy.cpp:201-225)
with closed forms; the reference's MSD goldens have no input in the snapshot."""
import numpy as np
import pytest

from oracle import oracle as O
from analisi_amd import synth


def _traj(pos, types, L=20.0):
    F = pos.shape[0]
    box = np.tile(synth.lammps_box(L), (F, 1))
    return O.Traj(pos, types, box, False)


def test_linear_motion_exact():
    """x_i(t) = x_i0 + v_i t with dyadic values and |v|^2 = 9/16 for every atom:
    every term equals 9/16 lag^2 exactly, so the running mean is exact."""
    N, F = 12, 40
    rng = np.random.default_rng(0)
    perms = np.array([[1, 2, 2], [2, 1, 2], [2, 2, 1], [-1, 2, -2], [2, -2, 1], [-2, -1, 2]]) / 4.0
    v = perms[rng.integers(0, 6, N)]
    x0 = np.round(rng.uniform(0, 20, (N, 3)) * 64) / 64
    pos = x0[None] + v[None] * np.arange(F)[:, None, None]
    types = synth.types_by_fraction(N, [0.5, 0.5])
    tr = _traj(pos, types)
    out = O.msd(tr, 1, 15, 3, False, False, 20, 0)
    lag = np.arange(15)[:, None]
    np.testing.assert_array_equal(out[:, 0, :], 9.0 / 16.0 * lag ** 2 * np.ones((1, 2)))


def test_cm_self_removes_uniform_translation_and_cm_channel():
    """Every atom of a type translated by the same u(t): cm_self MSD = 0 and the
    centre-of-mass channel (f_cm=2) is |u(t0+l)-u(t0)|^2."""
    N, F = 10, 30
    rng = np.random.default_rng(1)
    types = synth.types_by_fraction(N, [0.6, 0.4])
    x0 = rng.uniform(0, 10, (N, 3))
    u = np.stack([np.array([0.5, -0.25, 0.125]) * t for t in range(F)])
    pos = x0[None] + u[:, None, :]
    tr = _traj(pos, types)
    out = O.msd(tr, 1, 10, 2, True, True, 12, 3)
    assert np.abs(out[:, 0, :]).max() < 1e-24
    lag = np.arange(10)
    expect = (0.25 + 0.0625 + 0.015625) * lag ** 2
    np.testing.assert_allclose(out[:, 1, 0], expect, rtol=1e-13, atol=1e-25)
    np.testing.assert_allclose(out[:, 1, 1], expect, rtol=1e-13, atol=1e-25)


def test_thread_invariance_bitwise_and_skip():
    """Each lag is computed by one thread (PARALLEL_SPLIT_TIME), so results do
    not depend on nthreads (cf. msd_1a == msd_3a == msd_4a in the reference)."""
    box = synth.lammps_box(15.0)
    pos = synth.random_walk(4, 40, 60, box, 0.1)
    types = synth.types_by_fraction(40, [0.5, 0.25, 0.25])
    tr = _traj(pos, types, 15.0)
    a = O.msd(tr, 3, 20, 1, True, False, 35, 4)
    b = O.msd(tr, 3, 20, 3, True, False, 35, 4)
    c = O.msd(tr, 3, 20, 4, True, False, 35, 4)
    assert np.array_equal(a, b) and np.array_equal(a, c)
    # statistical closed form of a random walk: MSD(l) ~ 3 sigma^2 l
    d = O.msd(tr, 1, 20, 2, False, False, 35, 0)
    np.testing.assert_allclose(d[1:, 0, 0] / np.arange(1, 20), 3 * 0.01, rtol=0.35)


def test_msd_length_checks():
    box = synth.lammps_box(5.0)
    pos = synth.random_walk(3, 4, 20, box, 0.1)
    tr = _traj(pos, np.zeros(4, np.int32), 5.0)
    O.msd(tr, 1, 10, 1, False, False, 10, 0)                 # 10+10+0 <= 20
    with pytest.raises(RuntimeError):
        O.msd(tr, 1, 10, 1, False, False, 10, 1)             # 21 > 20 (no +1, msd.cpp:54)


def test_type_ids_and_cm():
    types = np.array([7, 3, 7, 9, 3], dtype=np.int32)
    tid = np.zeros(5, np.int32)
    nt = O.lib().orc_type_ids(O._p(types, O._ip), 5, O._p(tid, O._ip))
    assert nt == 3 and tid.tolist() == [1, 0, 1, 2, 0]
    pos = np.arange(2 * 5 * 3, dtype=np.float64).reshape(2, 5, 3)
    cm = np.zeros((2, 3, 3))
    O.lib().orc_cm(O._p(pos), O._p(tid, O._ip), 2, 5, 3, O._p(cm))
    for t in range(3):
        np.testing.assert_allclose(cm[:, t], pos[:, tid == t].mean(axis=1), rtol=1e-15)
