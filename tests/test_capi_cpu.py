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
