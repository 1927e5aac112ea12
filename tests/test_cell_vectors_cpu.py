"""CPU checks of the trajectory's cell-vector input (pyanalisi Trajectory with
BoxFormat.CellVectors): the LAMMPS box, the rotation and the rotated
positions and velocities must be those of the reference's Eigen computation
(triclinic.h:10-64), bit for bit.  The expected values come from
oracle/_ref/ref_cell, which runs those Eigen calls with the reference's
vendored Eigen headers; the tests skip where it is not built."""
import numpy as np
import pytest

from oracle import oracle as O


def _cells(seed, F):
    rng = np.random.default_rng(seed)
    cells = []
    for f in range(F):
        c = np.diag(rng.uniform(8, 15, 3)) + rng.normal(scale=2.0, size=(3, 3))
        if f % 5 == 1:
            c[0, 1] = 0.0          # a zero tilt
        if f % 5 == 2:
            c[:, 0] *= -1          # negative diagonal entries of R: reflections
        if f % 5 == 3:
            c[1] = -c[1]
        if f % 7 == 4:
            c[0, 0] = 0.0          # a zero pivot (the reflector's sign rule at c0 = 0)
        cells.append(c)
    cells[0] = np.diag([9.0, 10.0, 11.0])  # the diagonal shortcut
    return np.array(cells), rng


@pytest.mark.parametrize("seed", [5, 6])
def test_cell_vectors_match_eigen(seed):
    from analisi_amd import pyanalisi as pa
    F, N = 60, 9
    cells, rng = _cells(seed, F)
    pos = rng.normal(scale=6.0, size=(F, N, 3))
    vel = rng.normal(size=(F, N, 3))
    ref = O.ref_cell(list(cells), [np.concatenate([pos[f], vel[f]]) for f in range(F)])
    if ref is None:
        pytest.skip("oracle/_ref/ref_cell not built (it needs the reference's vendored Eigen headers)")
    t = pa.Trajectory(pos, vel, np.zeros(N, np.int32), cells, pa.BoxFormat.CellVectors, False, True)
    box, rot = t.get_box_copy(), t.get_rotation_matrix()
    P, V = t.get_positions_copy(), t.get_velocities_copy()
    assert box.shape == (F, 9) and rot.shape == (F, 3, 3)
    for f in range(F):
        b, q, v = ref[f]
        np.testing.assert_array_equal(box[f], b, err_msg="box, frame %d" % f)
        np.testing.assert_array_equal(rot[f].ravel(), q, err_msg="rotation, frame %d" % f)
        np.testing.assert_array_equal(P[f], v[:N], err_msg="positions, frame %d" % f)
        np.testing.assert_array_equal(V[f], v[N:], err_msg="velocities, frame %d" % f)


def test_cell_vectors_rotation_is_a_rotation():
    """Independent of Eigen: Q is orthogonal, the box's diagonal is positive,
    and the cell vectors (the columns of the numpy cell, which is Q R) rotated
    as row vectors are the lower-triangular LAMMPS cell rows (a along x, b in
    the xy plane)."""
    from analisi_amd import pyanalisi as pa
    F, N = 20, 3
    cells, rng = _cells(7, F)
    pos = np.ascontiguousarray(cells.transpose(0, 2, 1))  # rows: the cell vectors
    t = pa.Trajectory(pos, np.zeros_like(pos), np.zeros(N, np.int32), cells, pa.BoxFormat.CellVectors, False, True)
    box, rot, P = t.get_box_copy(), t.get_rotation_matrix(), t.get_positions_copy()
    for f in range(1, F):
        Q = rot[f].reshape(3, 3).T  # stored column-major
        np.testing.assert_allclose(Q.T @ Q, np.eye(3), atol=1e-14)
        assert box[f][3] > 0 and box[f][4] > 0 and box[f][5] > 0
        L = np.array([[2 * box[f][3], 0, 0], [box[f][6], 2 * box[f][4], 0], [box[f][7], box[f][8], 2 * box[f][5]]])
        np.testing.assert_allclose(P[f], L, atol=1e-12 * np.abs(cells[f]).max())


def test_degenerate_cells_match_eigen():
    """Eigen's reflector skips a column whose tail is at most the smallest
    normal double (tau = 0): a zero column, a denormal tilt, tiny pivots;
    and a cell whose R needs all three reflections."""
    from analisi_amd import pyanalisi as pa
    cells = [np.array([[0, 1., 2.], [0, 3., 1.], [0, 0.5, 4.]]),
             np.array([[5., 1e-320, 0], [0, 6., 2.], [0, 0, 7.]]),
             np.array([[-5., 0, 0], [1., -6., 0], [2., 3., -7.]]),
             np.array([[0, 0, 3.], [0, 4., 0], [5., 0, 0]]),
             np.array([[1e-300, 2., 0], [3., 1e-300, 0], [0, 0, 1.]])]
    N = 2
    rng = np.random.default_rng(0)
    pos = rng.normal(size=(len(cells), N, 3))
    vel = rng.normal(size=(len(cells), N, 3))
    ref = O.ref_cell(cells, [np.concatenate([pos[f], vel[f]]) for f in range(len(cells))])
    if ref is None:
        pytest.skip("oracle/_ref/ref_cell not built (it needs the reference's vendored Eigen headers)")
    t = pa.Trajectory(pos, vel, np.zeros(N, np.int32), np.array(cells), pa.BoxFormat.CellVectors, False, True)
    box, rot, P = t.get_box_copy(), t.get_rotation_matrix(), t.get_positions_copy()
    for f in range(len(cells)):
        b, q, v = ref[f]
        np.testing.assert_array_equal(box[f], b, err_msg="box, cell %d" % f)
        np.testing.assert_array_equal(rot[f].ravel(), q, err_msg="rotation, cell %d" % f)
        np.testing.assert_array_equal(P[f], v[:N], err_msg="positions, cell %d" % f)
