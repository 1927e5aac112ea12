"""CPU-side checks of the C ABI: the library loads and exports every entry
point include/analisi_hip.h declares; without a GPU the calls report
ANL_ERR_NO_DEVICE instead of computing anything."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def _declared():
    with open(os.path.join(ROOT, "include", "analisi_hip.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(anl_\w+)\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from analisi_amd import capi
    L = capi.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == {s[0] for s in capi.SIGNATURES}


def test_no_device_is_an_error_not_a_fallback():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    from analisi_amd import capi
    rc = capi.lib().anl_init(0)
    assert rc == capi.ANL_ERR_NO_DEVICE
    assert "gfx950" in capi.lib().anl_last_error().decode()


def _box9(L, tilt=(0.0, 0.0, 0.0)):
    """Internal-format box of a cell with edge lengths L (origin 0)."""
    return [0.0, 0.0, 0.0, L / 2, L / 2, L / 2, tilt[0], tilt[1], tilt[2]]


def test_cluster_screen_stays_in_fp16_range():
    """The cluster screen's tiles hold |p|^2 and |q|^2 in FP16 and mark
    inactive rows and columns past N with 60000: a cell wide enough for a
    large rmax must turn the cluster screen off instead of letting the
    threshold reach the sentinel (rmax ~ 250 A in a 600 A cell gives a
    threshold ~ 62,500 without the range cap) or |q|^2 overflow FP16."""
    from analisi_amd import capi
    # the bench's 90 A cubic cell, rmax 15: cluster screen on, far inside range
    s = capi.screen_info(_box9(90.0), 15.0)
    assert s["valid"] and s["rho_cap"] > 20.0
    assert s["q2_max"] < 16400 and s["threshold_max"] < 16500
    # 600 A cell, rmax 250: the 16-bit screen is valid, the cluster screen is off
    s = capi.screen_info(_box9(600.0), 250.0)
    assert s["valid"] and s["rho_cap"] == 0.0 and s["threshold_max"] == 0.0
    # 520 A cell, rmax 100: only the FP16 range limits rho (128 - 100 - dq)
    s = capi.screen_info(_box9(520.0), 100.0)
    assert s["valid"] and 27.0 < s["rho_cap"] < 28.0
    # rmax at half the cell width: no screen at all
    s = capi.screen_info(_box9(20.0), 10.0)
    assert not s["valid"] and s["rho_cap"] == 0.0
    # sweep of cubic and triclinic cells: whenever the cluster screen is on,
    # the largest tile value and threshold stay far below 65504 and 60000
    for L in (12.0, 30.0, 56.0, 90.0, 200.0, 400.0, 700.0, 1500.0):
        for frac in (0.05, 0.2, 0.35, 0.45, 0.49):
            for tilt in ((0.0, 0.0, 0.0), (0.1 * L, -0.05 * L, 0.08 * L)):
                s = capi.screen_info(_box9(L, tilt), frac * L)
                if s["rho_cap"] > 0:
                    assert s["valid"]
                    assert s["rho_cap"] + frac * L + s["dq"] <= 128.0
                    assert s["q2_max"] <= 16400.0, (L, frac, tilt, s)
                    assert s["threshold_max"] <= 16500.0, (L, frac, tilt, s)


def test_gofrt_kernels_have_no_static_lds():
    """The screened g(r,t) kernel's offset evaluation carries LDS byte
    addresses in 16-bit halves of its queue entries (rows and cached columns
    lie in the first 2 x 256 x 16 bytes of the dynamic shared memory), which
    holds while the dynamic shared memory starts at address 0: no kernel of
    the file may declare static shared arrays, only the one dynamic one."""
    with open(os.path.join(ROOT, "analisi_amd", "csrc", "gofrt.hip")) as f:
        src = f.read()
    decls = re.findall(r"^[^/\n]*__shared__[^;]*;", src, flags=re.M)
    assert decls, "no shared memory declaration found"
    for d in decls:
        assert re.fullmatch(r"\s*extern __shared__ unsigned char smem\[\];", d), d
