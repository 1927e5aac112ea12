"""ctypes binding of the C ABI in include/analisi_hip.h (libanalisi_hip.so).

This is the Python-side binding a maintainer would add next to pyanalisi's
pybind11 module; the pybind11 classes in ``analisi_amd._pyanalisi`` go
through the same C ABI from C++.  Loading fails loudly when the library is
missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libanalisi_hip.so")

ANL_OK = 0
ANL_ERR_ARG = 1
ANL_ERR_TOO_SHORT = 2
ANL_ERR_NOT_LOADED = 3
ANL_ERR_HIP = 4
ANL_ERR_NO_DEVICE = 5
ANL_ERR_HEADER = 6

GOFRT_FP64 = 0
GOFRT_SCREENED = 1
GOFRT_AUTO = 2
GK_FP64 = 0
GK_MFMA = 1

# (name, restype, argtypes) of every exported entry point
_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_i = ctypes.c_int
_d = ctypes.c_double
_pp = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES = [
    ("anl_last_error", ctypes.c_char_p, []),
    ("anl_init", _i, [_i]),
    ("anl_info", ctypes.c_char_p, []),
    ("anl_traj_create", _i, [_vp, _vp, _u32, _vp, _u32, _vp, _u64, _u64, _pp]),
    ("anl_traj_destroy", _i, [_vp]),
    ("anl_gofrt_create", _i, [_vp, _d, _d, _u32, _u32, _u32, _u32, _i, _pp]),
    ("anl_gofrt_destroy", _i, [_vp]),
    ("anl_gofrt_reset", _i, [_vp, _u32, ctypes.POINTER(_u32)]),
    ("anl_gofrt_calculate", _i, [_vp, _u64, _vp]),
    ("anl_gofrt_calculate_async", _i, [_vp, _u64, _vp, _vp]),
    ("anl_gofrt_calc_single_th", _i, [_vp, _i, _i, _i, _i, _u64, _vp]),
    ("anl_gofrt_calculate_blocks_async", _i, [_vp, _vp, _vp, _u32, _vp, _vp]),
    ("anl_gofrt_cluster_stats", _i, [_vp, ctypes.POINTER(_u64), ctypes.POINTER(_u64)]),
    ("anl_gofrt_recheck_stats", _i, [_vp, ctypes.POINTER(_u64), ctypes.POINTER(_u64)]),
    ("anl_gofrt_set_tuning", _i, [_vp, _i, _i, _u64, _i]),
    ("anl_gofrt_get_tuning", _i, [_vp, _vp, _vp, _vp, _vp]),
    ("anl_gofrt_autotune", _i, [_vp, _u64, _u32, _vp, _u32, _vp, _vp]),
    ("anl_gofrt_last_stats", _i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_u64), ctypes.POINTER(_u64)]),
    ("anl_gofrt_pair_stats", _i, [_vp] + [ctypes.POINTER(_u64)] * 5),
    ("anl_gofrt_screen_info", _i, [_vp, _d, ctypes.POINTER(_i), ctypes.POINTER(_u32)] + [ctypes.POINTER(_d)] * 4),
    ("anl_blocks_create", _i, [_vp, _vp, _vp, _u32, _vp, _i, _u32, _u32, _i, _i, _i, _pp]),
    ("anl_blocks_run", _i, [_vp, _vp]),
    ("anl_blocks_gofrt_results", _i, [_vp, _vp, _vp, _vp, ctypes.POINTER(_u32)]),
    ("anl_blocks_msd_results", _i, [_vp, _vp, _vp, _vp, ctypes.POINTER(_u32)]),
    ("anl_blocks_destroy", _i, [_vp]),
    ("anl_msd_calculate", _i, [_vp, _u32, _u32, _i, _i, _u32, _u64, _vp]),
    ("anl_msd_calculate_async", _i, [_vp, _u32, _u32, _i, _i, _u32, _u64, _vp, _vp]),
    ("anl_msd_calculate_noalloc", _i, [_vp, _u32, _u32, _i, _i, _u32, _u64, _vp, _vp]),
    ("anl_msd_set_tuning", _i, [_vp, _u32, _u32]),
    ("anl_msd_get_tuning", _i, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    ("anl_msd_autotune", _i, [_vp, _u32, _u32, _i, _i, _u32, _u64, _vp, _u32, _vp, ctypes.POINTER(_u32)]),
    ("anl_series_create", _i, [_vp, _u64, _u32, _pp]),
    ("anl_series_destroy", _i, [_vp]),
    ("anl_gk_calculate", _i, [_vp, _vp, _u32, _u32, _u32, _u32, _u32, _i, _u32, _u64, _vp]),
    ("anl_gk_calculate_mfma", _i, [_vp, _vp, _u32, _u32, _u32, _u32, _u32, _i, _u32, _u64, _vp]),
    ("anl_gk_calculate_blocks_async", _i, [_vp, _vp, _u32, _u32, _u32, _u32, _i, _u32, _vp, _vp, _u32, _i, _vp, _vp,
                                          _vp, _vp]),
    ("anl_gk_calculate_async", _i, [_vp, _vp, _u32, _u32, _u32, _u32, _u32, _i, _u32, _u64, _vp, _vp]),
    ("anl_gk_calculate_mfma_async", _i, [_vp, _vp, _u32, _u32, _u32, _u32, _u32, _i, _u32, _u64, _vp, _vp]),
    ("anl_gk_set_tuning", _i, [_vp, _u32, _u32, _u32]),
    ("anl_gk_get_tuning", _i, [_vp] + [ctypes.POINTER(_u32)] * 3),
    ("anl_gk_autotune", _i, [_vp, _vp, _u32, _u32, _u32, _u32, _vp, _vp, _u32, _vp, _u32, _vp, ctypes.POINTER(_u32)]),
    ("anl_mediavar_update_async", _i, [_vp, _vp, _vp, _u64, _u32, _vp]),
    ("anl_mediavar_end_async", _i, [_vp, _u64, _u32, _vp]),
    ("anl_counts_to_double_async", _i, [_vp, _d, _u64, _vp, _vp]),
    ("anl_sph_create", _i, [_vp, _u32, _vp, _u32, _pp]),
    ("anl_sph_destroy", _i, [_vp]),
    ("anl_sph_calculate", _i, [_vp, _u64, _u32, _u32, _u32, _vp]),
    ("anl_sph_snapshot", _i, [_vp, _u64, _vp]),
    ("anl_sph_set_budget", _i, [_vp, _u64]),
    ("anl_sph_set_neighbours", _i, [_vp, _vp, _vp]),
    ("anl_steinhardt_calculate", _i, [_vp, _u64, _u32, _u32, _vp, _u32, _u32, _i, _i, _vp]),
]

_lib = None


class AnlError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libanalisi_hip.so is not built (python -m analisi_amd.build); "
                              "analisi_amd has no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != ANL_OK:
        raise AnlError(rc, lib().anl_last_error().decode())


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class DeviceTrajectory:
    """anl_traj: device mirror of positions (internal layout), boxes (internal
    format), type ids and per-type centres of mass."""

    def __init__(self, pos, box_internal, type_id, ntypes, pos_cm=None):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        F, N, _ = self.pos.shape
        self.box = np.ascontiguousarray(box_internal, dtype=np.float64).reshape(F, -1)
        self.type_id = np.ascontiguousarray(type_id, dtype=np.int32)
        self.cm = None if pos_cm is None else np.ascontiguousarray(pos_cm, dtype=np.float64)
        self.ntypes = int(ntypes)
        self.nframes, self.natoms = F, N
        h = ctypes.c_void_p()
        check(lib().anl_traj_create(_ptr(self.pos), _ptr(self.box), self.box.shape[1], _ptr(self.type_id),
                                    self.ntypes, _ptr(self.cm), F, N, ctypes.byref(h)))
        self.handle = h

    # (lag tile, wavefronts per workgroup) of the MSD window kernel, in the
    # library's default candidate order (tile 0: eight lags per thread)
    MSD_SHAPES = [(8, 4), (12, 4), (16, 4), (24, 4), (16, 8), (24, 8), (12, 12), (16, 12), (8, 16), (0, 4)]

    def set_msd_tile(self, lag_tile, waves=4):
        """Launch shape of the MSD kernel: lag tile (8, 12, 16, 24; 0 = eight
        lags per thread) and wavefronts per workgroup (see MSD_SHAPES)."""
        check(lib().anl_msd_set_tuning(self.handle, lag_tile, waves))

    def msd_tile(self):
        return self.msd_shape()[0]

    def msd_shape(self):
        v, w = _u32(), _u32()
        check(lib().anl_msd_get_tuning(self.handle, ctypes.byref(v), ctypes.byref(w)))
        return v.value, w.value

    def msd_autotune(self, skip, tmax, cm_msd, cm_self, n, primo, candidates=None):
        """Times the MSD calculation with each launch shape (pairs (lag tile,
        wavefronts); None = MSD_SHAPES) and keeps the fastest.  Returns
        (times in ms per candidate, index of the best)."""
        cand = None if candidates is None else np.ascontiguousarray(candidates, dtype=np.uint32).reshape(-1)
        nc = len(self.MSD_SHAPES) if cand is None else cand.size // 2
        ms = np.zeros(nc, dtype=np.float32)
        best = _u32()
        check(lib().anl_msd_autotune(self.handle, skip, tmax, int(bool(cm_msd)), int(bool(cm_self)), n, primo,
                                     _ptr(cand), 0 if cand is None else nc, _ptr(ms), ctypes.byref(best)))
        return ms, best.value

    def close(self):
        if self.handle:
            lib().anl_traj_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# number of launch shapes anl_gofrt_autotune times when given no candidates
GOFRT_DEFAULT_SHAPES = 9


def screen_info(box9, rmax):
    """Screened-form parameters of a constant cell (internal-format box of 9
    values) computed on the host: dict with valid, threshold, rho_cap, dq,
    q2_max, threshold_max (see anl_gofrt_screen_info)."""
    box = np.ascontiguousarray(box9, dtype=np.float64).reshape(9)
    valid, thr = _i(), _u32()
    d = [_d() for _ in range(4)]
    check(lib().anl_gofrt_screen_info(_ptr(box), rmax, ctypes.byref(valid), ctypes.byref(thr),
                                      *[ctypes.byref(x) for x in d]))
    return dict(valid=bool(valid.value), threshold=thr.value, rho_cap=d[0].value, dq=d[1].value,
                q2_max=d[2].value, threshold_max=d[3].value)


class DeviceGofrt:
    def __init__(self, traj, rmin, rmax, nbin, tmax=0, skip=1, every=1, form=GOFRT_AUTO):
        self.traj = traj
        self.nbin = int(nbin)
        self.nch = traj.ntypes * (traj.ntypes + 1)
        h = ctypes.c_void_p()
        check(lib().anl_gofrt_create(traj.handle, rmin, rmax, nbin, tmax, skip, every, form, ctypes.byref(h)))
        self.handle = h
        self.leff = 0

    def reset(self, n):
        leff = _u32()
        check(lib().anl_gofrt_reset(self.handle, n, ctypes.byref(leff)))
        self.leff = leff.value
        return self.leff

    def calculate(self, primo):
        out = np.zeros((self.leff, self.nch, self.nbin), dtype=np.uint64)
        check(lib().anl_gofrt_calculate(self.handle, primo, _ptr(out)))
        return out

    def calc_single_th(self, t, imedia, a0, a1, primo, counts=None):
        if counts is None:
            counts = np.zeros((self.leff, self.nch, self.nbin), dtype=np.uint64)
        check(lib().anl_gofrt_calc_single_th(self.handle, t, imedia, a0, a1, primo, _ptr(counts)))
        return counts

    def stats(self):
        """(form used, candidates, FP64 rechecks) of the last calculate."""
        f = _i()
        c = _u64()
        r = _u64()
        check(lib().anl_gofrt_last_stats(self.handle, ctypes.byref(f), ctypes.byref(c), ctypes.byref(r)))
        return f.value, c.value, r.value

    def pair_stats(self):
        """Ordered pairs of the last calculate: dict with evaluated (some
        distance test looked at the pair), all (the reference's count),
        tile, screen16 and fp64 (evaluated by each test)."""
        v = [_u64() for _ in range(5)]
        check(lib().anl_gofrt_pair_stats(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(zip(("evaluated", "all", "tile", "screen16", "fp64"), (x.value for x in v)))

    def recheck_stats(self):
        """(window, exact): candidates re-evaluated in FP64 from their fractions (FP32 distance near an edge),
        and those of them that took the reference's FP64 sequence from the positions (still within the FP64
        bound of an edge), in the last screened calculate."""
        w, x = _u64(), _u64()
        check(lib().anl_gofrt_recheck_stats(self.handle, ctypes.byref(w), ctypes.byref(x)))
        return w.value, x.value

    def cluster_stats(self):
        """(clustered, all) (wavefront, t0) units of the last screened calculate."""
        c, t = _u64(), _u64()
        check(lib().anl_gofrt_cluster_stats(self.handle, ctypes.byref(c), ctypes.byref(t)))
        return c.value, t.value

    def set_tuning(self, unroll, xcd, pairs_per_wg, cluster=2):
        """cluster: 0 off; 1 cluster screen over the groups the bounding-sphere table leaves; 2 the same with
        the step's columns cached in LDS as FP32 offsets; 3 cluster screen over every j in order; 4 as 3 with
        the column cache."""
        check(lib().anl_gofrt_set_tuning(self.handle, unroll, int(bool(xcd)), pairs_per_wg, int(cluster)))

    def tuning(self):
        """(unroll, xcd, pairs per workgroup, cluster) of the pair kernel's launch shape."""
        u, x, p, c = _i(), _i(), _u64(), _i()
        check(lib().anl_gofrt_get_tuning(self.handle, ctypes.byref(u), ctypes.byref(x), ctypes.byref(p),
                                         ctypes.byref(c)))
        return u.value, x.value, p.value, c.value

    def autotune(self, primo, nlag_probe, candidates=None):
        """Times candidate launch shapes [(unroll, xcd, log2 pairs, cluster), ...]
        on the first nlag_probe lags of reset()'s calculation at primo; keeps
        the fastest.  Returns (times in ms per candidate, index of the best)."""
        if candidates is None:
            cand, nc = None, GOFRT_DEFAULT_SHAPES
        else:
            cand = np.ascontiguousarray(candidates, dtype=np.int32).reshape(-1, 4)
            nc = cand.shape[0]
        ms = np.zeros(nc, dtype=np.float32)
        best = _u32()
        check(lib().anl_gofrt_autotune(self.handle, primo, nlag_probe, None if cand is None else _ptr(cand),
                                       0 if cand is None else nc, _ptr(ms), ctypes.byref(best)))
        return ms, best.value

    def close(self):
        if self.handle:
            lib().anl_gofrt_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBlocks:
    """anl_blocks: Gofrt (through a DeviceGofrt's parameters) and/or MSD of
    uneven blocks in one launch each, with the block mean and variance."""

    def __init__(self, traj, primo, n, gofrt=None, msd=False, msd_skip=1, msd_tmax=0, cm_msd=False,
                 cm_self=False, use_graph=True):
        self.traj = traj
        self.primo = np.ascontiguousarray(primo, dtype=np.uint64)
        self.n = np.ascontiguousarray(n, dtype=np.uint32)
        self.nb = len(self.primo)
        self.gofrt = gofrt
        self.cm_msd = cm_msd
        h = ctypes.c_void_p()
        check(lib().anl_blocks_create(traj.handle, _ptr(self.primo), _ptr(self.n), self.nb,
                                      gofrt.handle if gofrt is not None else None, int(bool(msd)), msd_skip,
                                      msd_tmax, int(bool(cm_msd)), int(bool(cm_self)), int(bool(use_graph)),
                                      ctypes.byref(h)))
        self.handle = h

    def run(self, stream=None):
        check(lib().anl_blocks_run(self.handle, stream))

    def gofrt_results(self):
        leff = _u32()
        check(lib().anl_blocks_gofrt_results(self.handle, None, None, None, ctypes.byref(leff)))
        g = self.gofrt
        shape = (leff.value, g.nch, g.nbin)
        mean, var = np.zeros(shape), np.zeros(shape)
        counts = np.zeros((self.nb,) + shape, dtype=np.uint64)
        check(lib().anl_blocks_gofrt_results(self.handle, _ptr(mean), _ptr(var), _ptr(counts), ctypes.byref(leff)))
        return mean, var, counts

    def msd_results(self):
        leff = _u32()
        check(lib().anl_blocks_msd_results(self.handle, None, None, None, ctypes.byref(leff)))
        shape = (leff.value, 2 if self.cm_msd else 1, self.traj.ntypes)
        mean, var = np.zeros(shape), np.zeros(shape)
        per = np.zeros((self.nb,) + shape)
        check(lib().anl_blocks_msd_results(self.handle, _ptr(mean), _ptr(var), _ptr(per), ctypes.byref(leff)))
        return mean, var, per

    def close(self):
        if self.handle:
            lib().anl_blocks_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def msd(traj, skip, tmax, cm_msd, cm_self, n, primo):
    leff = n if (n < tmax or tmax == 0) else tmax
    out = np.zeros((leff, 2 if cm_msd else 1, traj.ntypes), dtype=np.float64)
    check(lib().anl_msd_calculate(traj.handle, skip, tmax, int(bool(cm_msd)), int(bool(cm_self)), n, primo,
                                  _ptr(out)))
    return out


class DeviceSph:
    """anl_sph: spherical-harmonic density correlations over a DeviceTrajectory;
    rminmax is a list of ntypes^2 (rmin, rmax) pairs."""

    def __init__(self, traj, rminmax, nbin, lmax=10):
        self.traj = traj
        self.nbin, self.lmax = int(nbin), int(lmax)
        rm = np.ascontiguousarray(np.asarray(rminmax, dtype=np.float64).reshape(-1, 2))
        if rm.shape[0] != traj.ntypes ** 2:
            raise ValueError("rminmax must hold ntypes^2 (rmin, rmax) pairs")
        h = ctypes.c_void_p()
        check(lib().anl_sph_create(traj.handle, self.lmax, _ptr(rm), self.nbin, ctypes.byref(h)))
        self.handle = h

    def set_neighbours(self, spec):
        """SANN neighbour lists: spec = [(max neighbours, cutoff^2, skin^2)] per type."""
        nn = np.ascontiguousarray([x[0] for x in spec], dtype=np.uint32)
        c2 = np.ascontiguousarray([x[1] for x in spec], dtype=np.float64)
        check(lib().anl_sph_set_neighbours(self.handle, _ptr(nn), _ptr(c2)))

    def set_budget(self, nbytes):
        check(lib().anl_sph_set_budget(self.handle, int(nbytes)))

    def calculate(self, n, tmax, skip, primo):
        leff = n if (n < tmax or tmax == 0) else tmax
        nt = self.traj.ntypes
        out = np.zeros((leff, nt, nt, self.nbin, self.lmax + 1), dtype=np.float64)
        check(lib().anl_sph_calculate(self.handle, primo, n, tmax, skip, _ptr(out)))
        return out

    def steinhardt(self, n, skip, primo, l_list, nbin_st, do_histogram=True, averaged=False):
        """Steinhardt order parameters of frames primo + i, i = 0, skip, .. < n - n % skip."""
        nt, N = self.traj.ntypes, self.traj.natoms
        ll = np.ascontiguousarray(l_list, dtype=np.uint32)
        skip = max(int(skip), 1)
        shape = (self.nbin, nt, nt) + (nbin_st,) * len(ll) if do_histogram else \
            (self.nbin, n // skip, N, nt, self.lmax)
        out = np.zeros(shape, dtype=np.float64)
        check(lib().anl_steinhardt_calculate(self.handle, primo, n, skip, _ptr(ll), len(ll), nbin_st,
                                             int(bool(do_histogram)), int(bool(averaged)), _ptr(out)))
        return out

    def snapshot(self, frame):
        nt = self.traj.ntypes
        out = np.zeros((self.traj.natoms, nt, self.nbin, (self.lmax + 1) ** 2), dtype=np.float64)
        check(lib().anl_sph_snapshot(self.handle, frame, _ptr(out)))
        return out

    def close(self):
        if self.handle:
            lib().anl_sph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Plain device memory (hipMalloc) for the stream-ordered entry points."""
    _hip = None

    def __init__(self, nbytes):
        if DeviceBuffer._hip is None:
            DeviceBuffer._hip = ctypes.CDLL("libamdhip64.so")
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        if DeviceBuffer._hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(self.nbytes, 8))) != 0:
            raise AnlError(ANL_ERR_HIP, "hipMalloc of %d bytes failed" % self.nbytes)

    def to_host(self, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        if DeviceBuffer._hip.hipDeviceSynchronize() != 0 or DeviceBuffer._hip.hipMemcpy(
                ctypes.c_void_p(out.ctypes.data), self.ptr, ctypes.c_size_t(out.nbytes), 2) != 0:
            raise AnlError(ANL_ERR_HIP, "device to host copy failed")
        return out

    def free(self):
        if self.ptr:
            DeviceBuffer._hip.hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceSeries:
    def __init__(self, data):
        self.data = np.ascontiguousarray(data, dtype=np.float64)
        h = ctypes.c_void_p()
        check(lib().anl_series_create(_ptr(self.data), self.data.shape[0], self.data.shape[1], ctypes.byref(h)))
        self.handle = h

    def gk(self, idx_j, skip, tmax, n, n_seg, subtract_mean, start_mean, primo, mfma=False):
        """GreenKubo reset(n) + calculate(primo); mfma selects the split-BF16
        matrix-core form of the lag products."""
        idx = np.ascontiguousarray(idx_j, dtype=np.uint32)
        ncur = len(idx)
        leff = n if (n < tmax or tmax == 0) else tmax
        out = np.zeros((leff, 3 * ncur * ncur + 2), dtype=np.float64)
        fn = lib().anl_gk_calculate_mfma if mfma else lib().anl_gk_calculate
        check(fn(self.handle, _ptr(idx), ncur, skip, tmax, n, n_seg, int(bool(subtract_mean)), start_mean, primo,
                 _ptr(out)))
        return out

    def gk_blocks(self, idx_j, skip, tmax, n_seg, subtract_mean, start_mean, primo, n, mfma=False, stats=True):
        """GreenKubo of every block b = reset(n[b]) + calculate(primo[b]) in one
        call; returns (per-block [nb][leff][narr], mean, variance of the mean)."""
        idx = np.ascontiguousarray(idx_j, dtype=np.uint32)
        ncur = len(idx)
        pr = np.ascontiguousarray(primo, dtype=np.uint64)
        nn = np.ascontiguousarray(n, dtype=np.uint32)
        nb = len(pr)
        leff = int(nn[0]) if (int(nn[0]) < tmax or tmax == 0) else tmax
        narr = 3 * ncur * ncur + 2
        out = DeviceBuffer(nb * leff * narr * 8)
        mean = DeviceBuffer(leff * narr * 8) if stats else None
        var = DeviceBuffer(leff * narr * 8) if stats else None
        check(lib().anl_gk_calculate_blocks_async(self.handle, _ptr(idx), ncur, skip, tmax, n_seg,
                                                  int(bool(subtract_mean)), start_mean, _ptr(pr), _ptr(nn), nb,
                                                  GK_MFMA if mfma else GK_FP64, out.ptr,
                                                  mean.ptr if stats else None, var.ptr if stats else None, None))
        res = out.to_host((nb, leff, narr))
        if not stats:
            return res, None, None
        return res, mean.to_host((leff, narr)), var.to_host((leff, narr))

    def set_gk_shape(self, tiles_per_wave, chunk_steps, waves_per_group):
        check(lib().anl_gk_set_tuning(self.handle, tiles_per_wave, chunk_steps, waves_per_group))

    def gk_shape(self):
        v = [_u32() for _ in range(3)]
        check(lib().anl_gk_get_tuning(self.handle, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def gk_autotune(self, idx_j, skip, tmax, n_seg, primo, n, candidates=None):
        """Times the MFMA lag products of the blocks (primo[b], n[b]) with each
        shape (tiles per wavefront, chunk steps, wavefronts per workgroup) and
        keeps the fastest.  Returns (times in ms, index of the best)."""
        idx = np.ascontiguousarray(idx_j, dtype=np.uint32)
        pr = np.ascontiguousarray(primo, dtype=np.uint64).reshape(-1)
        nn = np.ascontiguousarray(n, dtype=np.uint32).reshape(-1)
        cand = None if candidates is None else np.ascontiguousarray(candidates, dtype=np.uint32).reshape(-1, 3)
        nc = 6 if cand is None else cand.shape[0]
        ms = np.zeros(nc, dtype=np.float32)
        best = _u32()
        check(lib().anl_gk_autotune(self.handle, _ptr(idx), len(idx), skip, tmax, n_seg, _ptr(pr), _ptr(nn), len(pr),
                                    _ptr(cand), 0 if cand is None else nc, _ptr(ms), ctypes.byref(best)))
        return ms, best.value

    def close(self):
        if self.handle:
            lib().anl_series_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
