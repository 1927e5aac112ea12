"""The allocation-free MSD entry point (anl_msd_calculate_noalloc) through
the C ABI against anl_msd_calculate, the entry point it is a variant of:
bit for bit (same kernels), on two calls in a row, the second with a smaller
input, with both MSD kernels (t0 stride 1 and 2) and the centre-of-mass
options; and a calculation too large for its fixed scratch fails with
ANL_ERR_ARG before anything is launched."""
import ctypes

import numpy as np
import pytest

from analisi_amd import capi
from gpu_util import make_case

pytestmark = pytest.mark.gpu


class _Dev:
    """A device buffer from the HIP runtime (the test allocates; the entry
    point under test does not)."""

    def __init__(self, nbytes):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(max(nbytes, 8))) == 0
        self.nbytes = nbytes

    def to_host(self, shape):
        out = np.zeros(shape, dtype=np.float64)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(ctypes.c_void_p(out.ctypes.data), self.p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def _noalloc(dt, skip, tmax, cm_msd, cm_self, n, primo):
    leff = n if (n < tmax or tmax == 0) else tmax
    shape = (leff, 2 if cm_msd else 1, dt.ntypes)
    buf = _Dev(int(np.prod(shape)) * 8)
    try:
        capi.check(capi.lib().anl_msd_calculate_noalloc(dt.handle, skip, tmax, int(cm_msd), int(cm_self), n, primo,
                                                         buf.p, None))
        return buf.to_host(shape)
    finally:
        buf.free()


@pytest.mark.parametrize("skip", [1, 2])
@pytest.mark.parametrize("cm_msd,cm_self", [(False, False), (True, True)])
def test_msd_noalloc_equals_msd_calculate(skip, cm_msd, cm_self):
    big = make_case(13, 2000, 260, 30.0, [0.5, 0.3, 0.2], 0.1, unwrap_shift=11.0)[1]
    small = make_case(14, 130, 40, 9.0, [0.7, 0.3], 0.1)[1]
    for dt, tmax, n, primo in [(big, 120, 130, 4), (small, 9, 25, 1)]:
        want = capi.msd(dt, skip, tmax, cm_msd, cm_self, n, primo)
        got = _noalloc(dt, skip, tmax, cm_msd, cm_self, n, primo)
        assert np.array_equal(got, want), (skip, cm_msd, cm_self, dt.natoms)


def test_msd_noalloc_rejects_too_large():
    """12,800 atoms of 200 types, 60 lags: 60 x 200 chunks x 200 types x 8 B
    = 19.2 MB of partial sums, more than the 16 MiB scratch."""
    ot, dt = make_case(15, 12800, 130, 60.0, [0.005] * 200, 0.05)
    with pytest.raises(capi.AnlError) as e:
        _noalloc(dt, 1, 60, False, False, 60, 0)
    assert e.value.code == capi.ANL_ERR_ARG and "scratch" in str(e.value)
    # the same trajectory with 40 lags fits (12.8 MB) and equals anl_msd_calculate
    assert np.array_equal(_noalloc(dt, 1, 40, False, False, 40, 0), capi.msd(dt, 1, 40, False, False, 40, 0))
