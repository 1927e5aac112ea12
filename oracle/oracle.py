"""ctypes front of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by ``bench.py``'s cpu_baseline leg.  The product package
``analisi_amd`` never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u64 = ctypes.c_uint64
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class OrcTraj(ctypes.Structure):
    _fields_ = [
        ("pos", _dp), ("box", _dp), ("type_id", _ip), ("cm", _dp),
        ("nframes", _u64), ("natoms", _u64), ("ntypes", _u64), ("box_stride", _u64),
        ("triclinic", ctypes.c_int32), ("pad_", ctypes.c_int32), ("loaded", _u64),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_type_ids.restype = _u64
        L.orc_type_ids.argtypes = [_ip, _u64, _ip]
        L.orc_lammps_to_internal.argtypes = [_dp]
        L.orc_boxes_from_lammps.argtypes = [_dp, _u64, ctypes.c_int, _dp]
        L.orc_cm.argtypes = [_dp, _ip, _u64, _u64, _u64, _dp]
        L.orc_min_image.argtypes = [_dp, _dp, _dp, ctypes.c_int]
        L.orc_d2_min_image.restype = ctypes.c_double
        L.orc_d2_min_image.argtypes = [_dp, _dp, _dp, ctypes.c_int, _dp]
        L.orc_pbc_wrap.argtypes = [_dp, _dp, _u64, _u64, _u64, ctypes.c_int]
        L.orc_gofrt_channel.restype = ctypes.c_uint
        L.orc_gofrt_channel.argtypes = [ctypes.c_uint] * 3 + [ctypes.c_int]
        L.orc_gofrt_bin.restype = ctypes.c_int
        L.orc_gofrt_bin.argtypes = [ctypes.c_double] * 3
        L.orc_gofrt_calculate.restype = ctypes.c_int
        L.orc_gofrt_calculate.argtypes = [ctypes.POINTER(OrcTraj), ctypes.c_double, ctypes.c_double,
                                          ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                          ctypes.c_uint, ctypes.c_uint, _u64, _u64p, _dp]
        L.orc_gofrt_single.argtypes = [ctypes.POINTER(OrcTraj), ctypes.c_double, ctypes.c_double,
                                       ctypes.c_uint, ctypes.c_uint, ctypes.c_int, ctypes.c_int,
                                       _u64, _u64, _u64, _u64p]
        L.orc_n_extra_timesteps.restype = ctypes.c_uint
        L.orc_n_extra_timesteps.argtypes = [_u64, ctypes.c_uint, ctypes.c_uint]
        L.orc_msd_calculate.restype = ctypes.c_int
        L.orc_msd_calculate.argtypes = [ctypes.POINTER(OrcTraj), ctypes.c_uint, ctypes.c_uint,
                                        ctypes.c_uint, ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                        _u64, _dp]
        L.orc_schur00.restype = ctypes.c_double
        L.orc_schur00.argtypes = [_dp, ctypes.c_int]
        L.orc_gk_calculate.restype = ctypes.c_int
        L.orc_gk_calculate.argtypes = [_dp, _u64, _u64, _up, ctypes.c_uint, ctypes.c_uint,
                                       ctypes.c_uint, ctypes.c_int, ctypes.c_uint, ctypes.c_uint,
                                       ctypes.c_uint, ctypes.c_uint, _dp]
        L.orc_mv_update.argtypes = [_dp, _dp, _dp, _u64, ctypes.c_uint]
        L.orc_mv_end.argtypes = [_dp, _u64, ctypes.c_uint]
        L.orc_mvc_update.argtypes = [_dp, _dp, _dp, _dp, _u64, ctypes.c_uint, _up, ctypes.c_uint,
                                     ctypes.c_uint, _dp, _dp]
        L.orc_mvc_end.argtypes = [_dp, _dp, _u64, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint]
        L.orc_sph_ylm.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int, _dp]
        L.orc_sph_density.argtypes = [ctypes.POINTER(OrcTraj), _u64, _dp, ctypes.c_uint, ctypes.c_int, _dp]
        L.orc_sph_correlations.restype = ctypes.c_int
        L.orc_sph_correlations.argtypes = [ctypes.POINTER(OrcTraj), _dp, ctypes.c_uint, ctypes.c_int,
                                           ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, _u64, _dp]
        L.orc_sph_corr_from_densities.argtypes = [ctypes.POINTER(OrcTraj), _dp, ctypes.c_uint, ctypes.c_int,
                                                  ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, _dp]
        L.orc_steinhardt.restype = ctypes.c_int
        L.orc_steinhardt.argtypes = [ctypes.POINTER(OrcTraj), _dp, ctypes.c_uint, ctypes.c_int, ctypes.c_uint, _up,
                                     ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, _u64, ctypes.c_int, _dp]
        L.orc_sph_density_sann.restype = ctypes.c_int
        L.orc_sph_density_sann.argtypes = [ctypes.POINTER(OrcTraj), _u64, _up, _dp, ctypes.c_uint, ctypes.c_int, _dp,
                                           _up]
        L.orc_sph_correlations_sann.restype = ctypes.c_int
        L.orc_sph_correlations_sann.argtypes = [ctypes.POINTER(OrcTraj), _up, _dp, ctypes.c_uint, ctypes.c_int,
                                                ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, _u64, _dp]
        L.orc_steinhardt_sann.restype = ctypes.c_int
        L.orc_steinhardt_sann.argtypes = [ctypes.POINTER(OrcTraj), _up, _dp, ctypes.c_int, ctypes.c_uint, _up,
                                          ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, _u64, ctypes.c_int,
                                          ctypes.c_int, _dp]
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class Traj:
    """What BaseTrajectory serves: positions/box in internal format, type ids
    and the per-type centre of mass of the raw positions.

    ``box_lammps`` is the LammpsOrtho [F,6] or LammpsTriclinic [F,9] array;
    ``pbc_wrap`` wraps positions as trajectory_numpy.cpp:186-194 does."""

    def __init__(self, pos, types, box_lammps, triclinic, pbc_wrap=False):
        L = lib()
        raw = np.ascontiguousarray(pos, dtype=np.float64)
        F, N, _ = raw.shape
        self.types = np.ascontiguousarray(types, dtype=np.int32)
        self.type_id = np.zeros(N, dtype=np.int32)
        self.ntypes = int(L.orc_type_ids(_p(self.types, _ip), N, _p(self.type_id, _ip)))
        stride = 9 if triclinic else 6
        bl = np.ascontiguousarray(box_lammps, dtype=np.float64).reshape(F, stride)
        self.box = np.empty_like(bl)
        L.orc_boxes_from_lammps(_p(bl), F, int(bool(triclinic)), _p(self.box))
        self.cm = np.empty((F, self.ntypes, 3), dtype=np.float64)
        L.orc_cm(_p(raw), _p(self.type_id, _ip), F, N, self.ntypes, _p(self.cm))
        self.pos = raw.copy()
        if pbc_wrap:
            L.orc_pbc_wrap(_p(self.pos), _p(self.box), stride, F, N, int(bool(triclinic)))
        self.nframes, self.natoms, self.triclinic, self.stride = F, N, bool(triclinic), stride
        self._s = OrcTraj(_p(self.pos), _p(self.box), _p(self.type_id, _ip), _p(self.cm),
                          F, N, self.ntypes, stride, int(bool(triclinic)), 0, F)

    @property
    def ref(self):
        return ctypes.byref(self._s)


def gofrt(tr, rmin, rmax, nbin, tmax, nthreads, skip, every, n, primo, want_vdata=True):
    """Returns (counts[leff, nch, nbin] uint64, vdata float64 or None)."""
    L = lib()
    leff = n if (n < tmax or tmax == 0) else tmax
    nch = tr.ntypes * (tr.ntypes + 1)
    counts = np.zeros((leff, nch, nbin), dtype=np.uint64)
    vdata = np.zeros((leff, nch, nbin), dtype=np.float64) if want_vdata else None
    rc = L.orc_gofrt_calculate(tr.ref, rmin, rmax, nbin, tmax, nthreads, skip, every, n, primo,
                               _p(counts, _u64p), _p(vdata) if want_vdata else None)
    if rc != 0:
        raise RuntimeError("trajectory is too short for this kind of calculation")
    return counts, vdata


def gofrt_single(tr, rmin, rmax, nbin, leff, t, imedia, a0, a1, primo, counts=None):
    L = lib()
    nch = tr.ntypes * (tr.ntypes + 1)
    if counts is None:
        counts = np.zeros((leff, nch, nbin), dtype=np.uint64)
    L.orc_gofrt_single(tr.ref, rmin, rmax, nbin, leff, t, imedia, a0, a1, primo, _p(counts, _u64p))
    return counts


def msd(tr, skip, tmax, nthreads, cm_msd, cm_self, n, primo):
    L = lib()
    leff = n if (n < tmax or tmax == 0) else tmax
    fcm = 2 if cm_msd else 1
    out = np.zeros((leff, fcm, tr.ntypes), dtype=np.float64)
    rc = L.orc_msd_calculate(tr.ref, skip, tmax, nthreads, int(cm_msd), int(cm_self), n, primo, _p(out))
    if rc == -1:
        raise RuntimeError("trajectory is too short for this kind of calculation")
    if rc == -2:
        raise RuntimeError("there are not enough loaded timesteps inside the trajectory object")
    return out


def gk(data, idx_j, skip, lmax, subtract_mean, start_mean, n_seg, n, primo):
    L = lib()
    data = np.ascontiguousarray(data, dtype=np.float64)
    T, ncols = data.shape
    idx = np.ascontiguousarray(idx_j, dtype=np.uint32)
    ncur = len(idx)
    leff = n if (n < lmax or lmax == 0) else lmax
    out = np.zeros((leff, 3 * ncur * ncur + 2), dtype=np.float64)
    rc = L.orc_gk_calculate(_p(data), T, ncols, _p(idx, _up), ncur, skip, lmax, int(bool(subtract_mean)),
                            start_mean, n_seg, n, primo, _p(out))
    if rc != 0:
        raise RuntimeError("trajectory is too short for this kind of calculation")
    return out


def sph_ylm(v, lmax=10):
    """Real harmonics of the direction of v in the reference memory order."""
    out = np.empty((lmax + 1) ** 2, dtype=np.float64)
    lib().orc_sph_ylm(float(v[0]), float(v[1]), float(v[2]), lmax, _p(out))
    return out


def _rminmax(rminmax, nt):
    rm = np.ascontiguousarray(np.asarray(rminmax, dtype=np.float64).reshape(-1, 2))
    assert rm.shape[0] == nt * nt
    return rm


def sph_density(tr, frame, rminmax, nbin, lmax=10):
    """SphericalBase::calc of one frame: [N][nt][nbin][(lmax+1)^2]."""
    rm = _rminmax(rminmax, tr.ntypes)
    out = np.empty((tr.natoms, tr.ntypes, nbin, (lmax + 1) ** 2), dtype=np.float64)
    lib().orc_sph_density(tr.ref, frame, _p(rm), nbin, lmax, _p(out))
    return out


def sph_correlations(tr, rminmax, nbin, tmax, skip, n, primo, lmax=10):
    """SphericalCorrelations reset(n) + calculate(primo): [leff][nt][nt][nbin][lmax+1]."""
    rm = _rminmax(rminmax, tr.ntypes)
    leff = n if (n < tmax or tmax == 0) else tmax
    nt = tr.ntypes
    out = np.zeros((leff, nt, nt, nbin, lmax + 1), dtype=np.float64)
    rc = lib().orc_sph_correlations(tr.ref, _p(rm), nbin, lmax, tmax, skip, n, primo, _p(out))
    if rc != 0:
        raise RuntimeError("trajectory is too short for this kind of calculation")
    return out


def sph_corr_from_densities(tr, dens, nbin, leff, skip, n, lmax=10):
    """The t1/lag loops of sph_correlations over densities dens[f] of frames
    primo + f (f < n + leff - 1), as sph_density returns them."""
    dens = np.ascontiguousarray(dens, dtype=np.float64)
    nt = tr.ntypes
    out = np.zeros((leff, nt, nt, nbin, lmax + 1), dtype=np.float64)
    lib().orc_sph_corr_from_densities(tr.ref, _p(dens), nbin, lmax, leff, skip, n, _p(out))
    return out


def steinhardt(tr, rminmax, nbin, lmax, nbin_st, l_list, skip, n, primo, do_hist):
    """Steinhardt reset(n) + calculate(primo): histogram [nbin][nt][nt][nbin_st]*len(l_list)
    (values count * incr) or q_l [nbin][n/skip][N][nt][lmax] (l = 1..lmax)."""
    rm = _rminmax(rminmax, tr.ntypes)
    ll = np.ascontiguousarray(l_list, dtype=np.uint32)
    nt, skip = tr.ntypes, max(skip, 1)
    shape = (nbin, nt, nt) + (nbin_st,) * len(ll) if do_hist else (nbin, n // skip, tr.natoms, nt, lmax)
    out = np.zeros(shape, dtype=np.float64)
    rc = lib().orc_steinhardt(tr.ref, _p(rm), nbin, lmax, nbin_st, _p(ll, _up), len(ll), skip, n, primo,
                              int(bool(do_hist)), _p(out))
    if rc != 0:
        raise RuntimeError("trajectory is too short for this kind of calculation")
    return out


def _neigh_spec(spec, nt):
    """[(max neighbours, cutoff^2, skin^2)] per type -> (uint32 counts, float64 cutoff^2)."""
    assert len(spec) == nt
    return (np.ascontiguousarray([s[0] for s in spec], dtype=np.uint32),
            np.ascontiguousarray([s[1] for s in spec], dtype=np.float64))


_TOO_MANY = "Too many neighbours in shell!"


def sph_density_sann(tr, frame, spec, nbin, lmax=10):
    """SphericalBase::calc with a SANN neighbour list: (density, counts[N][nt])."""
    nn, c2 = _neigh_spec(spec, tr.ntypes)
    out = np.empty((tr.natoms, tr.ntypes, nbin, (lmax + 1) ** 2), dtype=np.float64)
    cnt = np.zeros((tr.natoms, tr.ntypes), dtype=np.uint32)
    if lib().orc_sph_density_sann(tr.ref, frame, _p(nn, _up), _p(c2), nbin, lmax, _p(out), _p(cnt, _up)):
        raise RuntimeError(_TOO_MANY)
    return out, cnt


def sph_correlations_sann(tr, spec, nbin, tmax, skip, n, primo, lmax=10):
    nn, c2 = _neigh_spec(spec, tr.ntypes)
    leff = n if (n < tmax or tmax == 0) else tmax
    nt = tr.ntypes
    out = np.zeros((leff, nt, nt, nbin, lmax + 1), dtype=np.float64)
    rc = lib().orc_sph_correlations_sann(tr.ref, _p(nn, _up), _p(c2), nbin, lmax, tmax, skip, n, primo, _p(out))
    if rc == -1:
        raise RuntimeError("trajectory is too short for this kind of calculation")
    if rc == -2:
        raise RuntimeError(_TOO_MANY)
    return out


def steinhardt_sann(tr, spec, lmax, nbin_st, l_list, skip, n, primo, do_hist, averaged=False):
    """Steinhardt with a SANN neighbour list (one radial bin), optionally averaged."""
    nn, c2 = _neigh_spec(spec, tr.ntypes)
    ll = np.ascontiguousarray(l_list, dtype=np.uint32)
    nt, skip = tr.ntypes, max(skip, 1)
    shape = (1, nt, nt) + (nbin_st,) * len(ll) if do_hist else (1, n // skip, tr.natoms, nt, lmax)
    out = np.zeros(shape, dtype=np.float64)
    rc = lib().orc_steinhardt_sann(tr.ref, _p(nn, _up), _p(c2), lmax, nbin_st, _p(ll, _up), len(ll), skip, n,
                                   primo, int(bool(do_hist)), int(bool(averaged)), _p(out))
    if rc == -1:
        raise RuntimeError("trajectory is too short for this kind of calculation")
    if rc == -2:
        raise RuntimeError(_TOO_MANY)
    return out


def n_extra_timesteps(nframes, n_b, lmax):
    return int(lib().orc_n_extra_timesteps(nframes, n_b, lmax))


class MediaVar:
    """calcoliblocchi.h:21-65 on flat vectors."""

    def __init__(self, length):
        self.mean = np.zeros(length)
        self.var = np.zeros(length)
        self.iblock = 0

    def add(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        lib().orc_mv_update(_p(self.mean), _p(self.var), _p(x), x.size, self.iblock)
        self.iblock += 1

    def end(self, n_b):
        lib().orc_mv_end(_p(self.var), self.var.size, n_b)


class MediaVarCovar:
    """calcoliblocchi.h:72-202."""

    def __init__(self, n_columns, size_block, cvar=()):
        self.nc = n_columns
        self.sb = size_block
        self.cv = np.ascontiguousarray(np.array(cvar, dtype=np.uint32).reshape(-1))
        self.ncv = len(cvar)
        self.mean = np.zeros((n_columns, size_block))
        self.var = np.zeros((n_columns, size_block))
        self.cov = np.zeros((max(self.ncv, 1), size_block))
        self.dx = np.zeros(n_columns)
        self.dx2 = np.zeros(n_columns)
        self.iblock = 0

    def add(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        self.iblock += 1
        lib().orc_mvc_update(_p(self.mean), _p(self.var), _p(self.cov), _p(x), self.sb, self.nc,
                             _p(self.cv, _up) if self.ncv else None, self.ncv, self.iblock,
                             _p(self.dx), _p(self.dx2))

    def end(self):
        lib().orc_mvc_end(_p(self.var), _p(self.cov), self.sb, self.nc, self.ncv, self.iblock)


def schur00(M):
    """The (0,0) Schur complement of the square matrix M, as the oracle's GK
    epilogue computes it (greenkuboNcomponentionicfluid.cpp:309-316)."""
    M = np.asfortranarray(np.asarray(M, dtype=np.float64))
    col = np.ascontiguousarray(M.T).ravel()  # column-major order
    return lib().orc_schur00(_p(col), M.shape[0])


REF_SCHUR = os.path.join(_HERE, "_ref", "ref_schur")


def ref_schur00(mats):
    """The same complement of each matrix by Eigen's evaluation of the
    reference's expression (oracle/_ref/ref_schur, built from ref_schur.cpp
    against the reference's vendored Eigen); None when it is not built."""
    if not os.path.exists(REF_SCHUR):
        return None
    import struct
    parts = [struct.pack("<i", len(mats))]
    for M in mats:
        M = np.asarray(M, dtype=np.float64)
        parts.append(struct.pack("<i", M.shape[0]))
        parts.append(np.ascontiguousarray(M.T).tobytes())  # column-major
    r = subprocess.run([REF_SCHUR], input=b"".join(parts), capture_output=True, check=True)
    return np.frombuffer(r.stdout, dtype=np.float64)


REF_CELL = os.path.join(_HERE, "_ref", "ref_cell")


def ref_cell(cells, vecs):
    """The reference's lammps cell from cell vectors by Eigen
    (triclinic.h:10-64; oracle/_ref/ref_cell, built from ref_cell.cpp against
    the reference's vendored Eigen): for each frame's numpy cell [3][3] and
    row vectors [n][3], the LAMMPS box (9 values), Q (column-major) and the
    rotated vectors.  None when it is not built."""
    if not os.path.exists(REF_CELL):
        return None
    import struct
    parts = [struct.pack("<i", len(cells))]
    for c, v in zip(cells, vecs):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        parts += [np.ascontiguousarray(c, dtype=np.float64).tobytes(), struct.pack("<i", len(v)), v.tobytes()]
    r = subprocess.run([REF_CELL], input=b"".join(parts), capture_output=True, check=True)
    flat = np.frombuffer(r.stdout, dtype=np.float64)
    out, o = [], 0
    for v in vecs:
        n = len(np.asarray(v).reshape(-1, 3))
        out.append((flat[o:o + 9], flat[o + 9:o + 18], flat[o + 18:o + 18 + 3 * n].reshape(n, 3)))
        o += 18 + 3 * n
    return out
