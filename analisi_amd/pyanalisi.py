"""The pyanalisi surface of the pair-correlation path (drop-in for
``import pyanalisi`` in code that uses Trajectory, Traj, ReadLog,
ReadLog_file, BoxFormat, Gofrt, MeanSquareDisplacement,
SphericalCorrelations, SteinhardtOrderParameterHistogram_<l>, their
``_lammps`` forms over Traj, GreenKubo and GreenKubo_file)::

    from analisi_amd import pyanalisi
    traj = pyanalisi.Trajectory(pos, vel, types, box, pyanalisi.BoxFormat.CellVectors, False, False)
    g = pyanalisi.Gofrt(traj, 0.0, 3.8, 200, 10, 4, 10, False, 1)
    g.reset(700); g.calculate(0); gr = np.array(g, copy=True)
"""
from ._pyanalisi import (BoxFormat, Gofrt, Gofrt_lammps, GreenKubo, GreenKubo_file,  # noqa: F401
                         MeanSquareDisplacement, MeanSquareDisplacement_lammps, ReadLog, ReadLog_file,
                         SphericalCorrelations, SphericalCorrelations_lammps, Traj, Trajectory, gk_block_average,
                         gofrt_block_average, has_mmap, info, msd_block_average, cli_gk_text, cli_msd_text,
                         cli_gofrt_text)
from . import _pyanalisi as _native

# SteinhardtOrderParameterHistogram_<l> and _<l>_lammps, l = 2..10
for _name in dir(_native):
    if _name.startswith("SteinhardtOrderParameterHistogram_"):
        globals()[_name] = getattr(_native, _name)
del _name
