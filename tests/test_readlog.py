"""The text log reader ReadLog_file (readlog.cpp:26-343): header and row
rules, the Step column, skip, and the "#traj:JZ N q1 ... qN" charge current
computed from a LAMMPS-binary trajectory (readlog.cpp:253-293).  Host code
only (no GPU)."""
import os

import numpy as np
import pytest

from analisi_amd import pyanalisi as pa
from analisi_amd import synth
from conftest import GOLDEN
from oracle import lammps_bin as L


def test_gk_integral_text_reads_as_the_reference_data(tmp_path):
    """tests/data/gk_integral.dat rebuilt from its float64 copy (shortest
    round-trip decimals, a tab after the first column as in the file): the
    reader returns the same doubles."""
    ref = np.load(os.path.join(GOLDEN, "gk_reference.npz"))
    log, headers = ref["log"], [str(h) for h in ref["headers"]]
    p = str(tmp_path / "gk_integral.dat")
    with open(p, "w") as f:
        f.write("    ".join(headers) + "\n")
        for row in log:
            f.write(repr(float(row[0])) + "\t" + " ".join(repr(float(x)) for x in row[1:]) + "\n")
    r = pa.ReadLog_file(p, None, 1, 4, 200, [])
    assert r.getNtimesteps() == log.shape[0]
    np.testing.assert_array_equal(r.get_data_copy(), log)
    assert [r.timestep(i) for i in (0, 1, 19999)] == [1, 2, 20000]   # no Step column: the row count


def test_rules_by_hand(tmp_path):
    text = ("LAMMPS (29 Oct 2020)\n"
            "  Time spent = 0.1 s\n"
            "Step Temp  c_J[1]\tPress\n"
            "10 1.5 -2e-3 7\n"
            "11 1.25 4.5E+1 8\r\n"          # a Windows line end
            "12 3 1\n"                      # too few fields: dropped
            "WARNING: something\n"          # letters: not a row
            "13 0.5 6 -9\n"
            "0 2.5 7 10\n"                  # Step 0: the row count is its timestep
            "14 1 1 1\n"
            "Loop time of 3.0\n")
    p = str(tmp_path / "log.lammps")
    open(p, "w").write(text)
    r = pa.ReadLog_file(p)
    np.testing.assert_array_equal(r.get_data_copy(),
                                  [[1.5, -2e-3, 7], [1.25, 45.0, 8], [0.5, 6, -9], [2.5, 7, 10], [1, 1, 1]])
    assert [r.timestep(i) for i in range(5)] == [10, 11, 13, 4, 14]
    r2 = pa.ReadLog_file(p, None, 2)                # every second accepted row
    np.testing.assert_array_equal(r2.get_data_copy(), [[1.5, -2e-3, 7], [0.5, 6, -9], [1, 1, 1]])
    with pytest.raises(RuntimeError, match="Cannot find the column headers"):
        open(str(tmp_path / "n.txt"), "w").write("1 2 3\n4 5 6\n")
        pa.ReadLog_file(str(tmp_path / "n.txt"))
    with pytest.raises(RuntimeError, match="reserved for step number"):
        pa.ReadLog_file(p, None, 1, 1, 200, ["Step"])
    with pytest.raises(RuntimeError, match="It must be in the form"):
        pa.ReadLog_file(p, None, 1, 1, 200, ["#traj:JZ 3 1 2"])


def test_traj_charge_current(tmp_path):
    """Columns "#traj:JZ 2 q0 q1" = 0 + v_cm(type 0) q0 + v_cm(type 1) q1 per
    row, v_cm the per-type running mean of the velocities read from the
    binary (trajectory.cpp:604-617), filled block by block."""
    N, F = 50, 12
    box_l = synth.lammps_box(10.0)
    pos = synth.random_walk(4, N, F, box_l, 0.2)
    vel = np.random.default_rng(9).normal(size=(F, N, 3))
    types = synth.types_by_fraction(N, [0.7, 0.3]).astype(np.int32)
    b = str(tmp_path / "t.bin")
    pa.Trajectory(pos, vel, types, np.tile(box_l, (F, 1)), pa.BoxFormat.LammpsOrtho, False, False) \
        .write_lammps_binary(b, 0, -1)
    lp = str(tmp_path / "log")
    with open(lp, "w") as f:
        f.write("Step c_a\n" + "".join("%d %r\n" % (k + 1, 0.5 * k) for k in range(F)))
    head = "#traj:JZ 2 1.5 -0.5"
    t = pa.Traj(b)
    r = pa.ReadLog_file(lp, t, 1, 1, 200, [head])
    np.testing.assert_array_equal(r.get_data_copy()[:, 1:], 0.0)
    r.calc_currents(t, 5)
    w = L.window(b, 0, F)
    # the oracle's running mean is the plain mean up to rounding
    for ty in (0, 1):
        np.testing.assert_allclose(w["vcm"][:, ty], vel[:, types == ty].mean(axis=1), rtol=1e-12, atol=1e-14)
    J = (0.0 + w["vcm"][:, 0] * 1.5) + w["vcm"][:, 1] * -0.5
    np.testing.assert_array_equal(r.get_data_copy()[:, 1:], J)
    np.testing.assert_array_equal(r.get_data_copy()[:, 0], 0.5 * np.arange(F))
    bad = pa.ReadLog_file(lp, t, 1, 1, 200, ["#traj:JZ 3 1 2 3"])
    with pytest.raises(RuntimeError, match="number of coefficients must be equal"):
        bad.calc_currents(t, 2)


def test_threads_and_blocks_do_not_change_the_table(tmp_path):
    """Rows parsed in blocks of nthreads * nbatch lines over several threads
    are the rows of a one-thread read, in file order, with the same skip
    counting across block boundaries (bad and text lines interleaved)."""
    rng = np.random.default_rng(5)
    lines = ["some preamble", "Step a b\tc"]
    for k in range(1000):
        if k % 37 == 5:
            lines.append("%d 1 2" % k)                  # too few fields
        elif k % 53 == 7:
            lines.append("text in between")
        else:
            lines.append("%d %r %r %r" % (k, *(float(x) for x in rng.normal(size=3))))
    p = str(tmp_path / "log")
    open(p, "w").write("\n".join(lines) + "\n")
    one = pa.ReadLog_file(p, None, 3, 1, 200, [])
    many = pa.ReadLog_file(p, None, 3, 3, 7, [])
    assert one.getNtimesteps() == many.getNtimesteps() > 300
    np.testing.assert_array_equal(one.get_data_copy(), many.get_data_copy())
    assert [one.timestep(i) for i in range(one.getNtimesteps())] == \
           [many.timestep(i) for i in range(many.getNtimesteps())]
