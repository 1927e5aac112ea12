"""Shared helpers of the GPU tests: seeded trajectories in both the oracle's
and the device library's form."""
import numpy as np

from analisi_amd import capi, synth
from oracle import oracle as O


def make_case(seed, natoms, nframes, L, fractions, sigma, tilt=None, pbc_wrap=False, unwrap_shift=0.0):
    box_l = synth.lammps_box(L, tilt)
    pos = synth.random_walk(seed, natoms, nframes, box_l, sigma)
    if unwrap_shift:
        pos[:, ::5] += unwrap_shift
    types = synth.types_by_fraction(natoms, fractions)
    tri = tilt is not None
    ot = O.Traj(pos, types, np.tile(box_l, (nframes, 1)), tri, pbc_wrap=pbc_wrap)
    dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    return ot, dt


def first_diff(a, b):
    bad = np.argwhere(a != b)
    if len(bad) == 0:
        return None
    k = tuple(bad[0])
    return "first mismatch at %s: %s vs %s (%d mismatches)" % (k, a[k], b[k], len(bad))
