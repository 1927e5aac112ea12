"""g(r,t) screened form on seeded random configurations against the oracle:
cell shape (cubic, orthorhombic, triclinic with tilts up to a quarter of the
cell), rmin, rmax, bin count, types, unwrapped shifts and diffusion drawn per
case; every cluster-screen mode and the 16-bit screen must give the
reference's counts bit for bit (gofrt.cpp:96-122).  Depending on the
case, from none to most of the wavefronts are compact enough for the
cluster screen (the rest take the 16-bit screen in the same launch), so
the FP32 windows of both candidate evaluations, the FP64 stage from the
fractions and the exact rechecks all run on pairs that no case places by
hand."""
import numpy as np
import pytest

from analisi_amd import capi, synth
from gpu_util import first_diff
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def draw_case(seed):
    rng = np.random.default_rng(1000 + seed)
    L = rng.uniform(30.0, 44.0, size=3) if seed % 3 == 2 else np.full(3, rng.uniform(30.0, 44.0))
    tilt = None
    if seed % 2 == 1:
        tilt = rng.uniform(-0.25, 0.25, size=3) * L.min()
    N = int(rng.integers(1200, 2600))
    nt = int(rng.integers(1, 4))
    fr = rng.dirichlet(np.ones(nt)) * 0.9 + 0.1 / nt
    fr = fr / fr.sum()
    # smallest cell width, so that rmax stays below half of it (the screen's domain)
    A = np.array([[L[0], 0, 0], [0, L[1], 0], [0, 0, L[2]]], dtype=float)
    if tilt is not None:
        A[0, 1], A[0, 2], A[1, 2] = tilt
    Ai = np.linalg.inv(A)
    wmin = 1.0 / np.linalg.norm(Ai, axis=1).max()
    rmax = rng.uniform(3.0, 0.3 * wmin)
    rmin = rng.uniform(0.0, 1.5) if seed % 4 else 0.0
    nbin = int(rng.integers(16, 400))
    sigma = rng.uniform(0.05, 0.6)
    shift = rng.choice([0.0, 3.7 * L.max(), -2.2 * L.max()])
    return dict(seed=seed, L=L, tilt=tilt, N=N, fr=fr, rmin=rmin, rmax=rmax, nbin=nbin, sigma=sigma, shift=shift)


@pytest.mark.parametrize("seed", range(8))
def test_screened_matches_oracle_on_random_cases(seed):
    c = draw_case(seed)
    F, n, tmax = 7, 3, 3
    box_l = synth.lammps_box(c["L"], c["tilt"])
    pos = synth.random_walk(77 + seed, c["N"], F, box_l, c["sigma"])
    if c["shift"]:
        pos[:, ::7] += c["shift"]
    types = synth.types_by_fraction(c["N"], c["fr"])
    ot = O.Traj(pos, types, np.tile(box_l, (F, 1)), c["tilt"] is not None)
    dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    print("case", {k: (np.round(v, 3).tolist() if isinstance(v, np.ndarray) else v) for k, v in c.items()}, flush=True)
    ref, _ = O.gofrt(ot, c["rmin"], c["rmax"], c["nbin"], tmax, 4, 1, 1, n, 1, want_vdata=False)
    g = capi.DeviceGofrt(dt, c["rmin"], c["rmax"], c["nbin"], tmax, 1, 1, capi.GOFRT_SCREENED)
    g.reset(n)
    for cl in (2, 1, 4, 3, 0):
        g.set_tuning(8, 0, 1 << 30, cl)
        got = g.calculate(1)
        assert first_diff(got, ref) is None, (seed, cl, first_diff(got, ref))
        used, cand, rechecks = g.stats()
        assert used == capi.GOFRT_SCREENED
        window, exact = g.recheck_stats()
        assert window == rechecks and exact <= window
        c_units, units = g.cluster_stats()
        print("  mode %d: %d candidates, %d rechecks (%d exact), %d of %d units clustered"
              % (cl, cand, rechecks, exact, c_units, units), flush=True)


@pytest.mark.parametrize("tilt", [None, (3.0, -2.0, 1.5)], ids=["cubic", "tri"])
def test_screened_far_from_origin(tilt):
    """Unwrapped positions thousands of cell lengths from the origin: the
    reference's FP64 minimum image then takes thousands of steps per pair and
    its rounding (on coordinates near 1e5 A) is no longer negligible next to
    the 32-bit quantisation; the FP64 evaluation from the fractions bounds it
    from the largest position (ref_vector_rounding), so the counts must stay
    the reference's in every mode."""
    L, N, F, n, tmax = 36.0, 1500, 5, 2, 2
    rmin, rmax, nbin = 0.4, 7.5, 300
    box_l = synth.lammps_box(L, tilt)
    pos = synth.random_walk(91, N, F, box_l, 0.3)
    rng = np.random.default_rng(92)
    shift = rng.integers(-3000, 3000, size=(N, 3)) * L  # whole cell lengths, as unwrapped runs accumulate
    pos += shift[None, :, :]
    types = synth.types_by_fraction(N, [0.5, 0.5])
    ot = O.Traj(pos, types, np.tile(box_l, (F, 1)), tilt is not None)
    dt = capi.DeviceTrajectory(ot.pos, ot.box, ot.type_id, ot.ntypes, ot.cm)
    ref, _ = O.gofrt(ot, rmin, rmax, nbin, tmax, 16, 1, 1, n, 0, want_vdata=False)
    assert ref.sum() > 0
    g = capi.DeviceGofrt(dt, rmin, rmax, nbin, tmax, 1, 1, capi.GOFRT_SCREENED)
    g.reset(n)
    for cl in (2, 1, 4, 3, 0):
        g.set_tuning(8, 0, 1 << 30, cl)
        got = g.calculate(0)
        assert first_diff(got, ref) is None, (cl, first_diff(got, ref))
        window, exact = g.recheck_stats()
        print("  mode %d: %d rechecks, %d exact" % (cl, window, exact), flush=True)
