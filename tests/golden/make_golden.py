"""Copy the reference's own test data for the hot path into small fixtures.

Reads (as data, never as code) from the reference checkout and writes
tests/golden/*.npz and *.txt.  Run once in the build container:

    python tests/golden/make_golden.py /root/reference

Contents
  gk_reference.npz
    log          gk_integral.dat as float64 [20000, 7] (tests/data/gk_integral.dat)
    headers      its column names
    gk_2comp     tests/cpp_regression_data/gk/gk_2comp, float64 [5000, 14]
                 (GK c_flux[1] c_vcm[1][1], reset(5000), test_gk.cpp:8-35)
    test_gk      tests/test_gk/test_gk.csv as float64 [2000, 14]
                 (test_gk.py:4-12: lmax 2000, n_seg 4, reset(T-2000))
  cli_GK_heat_single.txt, cli_GK_heat_double.txt
                 tests/data/cli/GK_heat_* (stdout of the CLI, test_cli.sh:31-32)
  cli_MSD_normal.txt, cli_MSD_cm.txt
                 tests/data/cli/MSD_normal (-Q -s 10 -S 50) and MSD_cm
                 (-q -s 10 -S 50), test_cli.sh:25-26; their input
                 lammps2020.bin is not in the snapshot, so they fix the
                 layout only (the -Q columns start with the -q ones)
  gofr_reference.npz
    test_gofr    tests/test_gofrt/test_gofr.csv as float64 [10, 12, 200]
                 (test_gofrt.py:4-13; its input positions.npy is not in the
                 snapshot, so only structural facts of it are usable)
  lammps/        the LAMMPS-binary trajectory data, copied byte for byte:
    10steps.bin, .pos, .vel, .cel, .species
                 tests/data/10steps.*: a binary trajectory (432 atoms, 10
                 triclinic frames) and the cp.x text files it was made from
    dump4.lammpstrj
                 the first 4 timesteps of tools/lammps/dump.lammpstrj
    water125.pos, .vel, .cel, .species
                 tools/cp/water125_testanalisi.* (cp.x text trajectory)
    test_lammps2analisi.csv, test_cp2analisi.csv
                 tests/test_write_binary/*.csv: positions, boxes and
                 velocities read back from the binaries the two converters
                 write (test_write_binary.py:59-83)
"""
import os
import shutil
import sys

import numpy as np


def read_csv_column(path):
    vals = []
    with open(path) as f:
        next(f)
        for line in f:
            cell = line.split(",")[1].strip()
            # pandas writes NaN as an empty cell (k_einst at lag 1 is NaN)
            vals.append(float(cell) if cell else float("nan"))
    return np.array(vals, dtype=np.float64)


def main(ref):
    out = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(ref, "tests/data/gk_integral.dat")) as f:
        headers = f.readline().split()
        rows = [[float(x) for x in line.split()] for line in f if line.strip()]
    log = np.array(rows, dtype=np.float64)
    raw = np.fromfile(os.path.join(ref, "tests/cpp_regression_data/gk/gk_2comp"), dtype=np.float64)
    gk_2comp = raw.reshape(5000, 14)
    test_gk = read_csv_column(os.path.join(ref, "tests/test_gk/test_gk.csv")).reshape(2000, 14)
    np.savez_compressed(os.path.join(out, "gk_reference.npz"), log=log, headers=np.array(headers),
                        gk_2comp=gk_2comp, test_gk=test_gk)
    for name in ("GK_heat_single", "GK_heat_double", "MSD_normal", "MSD_cm"):
        shutil.copyfile(os.path.join(ref, "tests/data/cli", name), os.path.join(out, "cli_" + name + ".txt"))
    gofr = read_csv_column(os.path.join(ref, "tests/test_gofrt/test_gofr.csv")).reshape(10, 12, 200)
    np.savez_compressed(os.path.join(out, "gofr_reference.npz"), test_gofr=gofr)
    lmp = os.path.join(out, "lammps")
    os.makedirs(lmp, exist_ok=True)
    for ext in ("bin", "pos", "vel", "cel", "species"):
        shutil.copyfile(os.path.join(ref, "tests/data/10steps." + ext), os.path.join(lmp, "10steps." + ext))
    for ext in ("pos", "vel", "cel", "species"):
        shutil.copyfile(os.path.join(ref, "tools/cp/water125_testanalisi." + ext), os.path.join(lmp, "water125." + ext))
    with open(os.path.join(ref, "tools/lammps/dump.lammpstrj")) as f:
        lines = f.readlines()
    natoms = int(lines[3])
    with open(os.path.join(lmp, "dump4.lammpstrj"), "w") as f:
        f.writelines(lines[:4 * (natoms + 9)])
    for name in ("test_lammps2analisi.csv", "test_cp2analisi.csv"):
        shutil.copyfile(os.path.join(ref, "tests/test_write_binary", name), os.path.join(lmp, name))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
