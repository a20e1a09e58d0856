"""C-ABI checks that need no GPU: the library loads and exports every entry
point include/diskfft.h declares; calls fail loudly (no CPU fallback) when no
device is visible."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dfk_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import diskfft_amd
    L = ctypes.CDLL(diskfft_amd.LIBPATH)
    names = _declared("diskfft.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(diskfft_amd.SYMBOLS) == names


def test_model_library_exports():
    import diskfft_amd
    L = ctypes.CDLL(diskfft_amd.LIBPATH)
    names = _declared("diskfft_model.h")
    assert sorted(diskfft_amd.MODEL_SYMBOLS) == names
    assert all(hasattr(L, n) for n in names)


def test_no_device_means_loud_failure():
    import diskfft_amd
    if diskfft_amd.device_count() > 0:
        pytest.skip("a device is visible; this checks the no-device path")
    with pytest.raises(diskfft_amd.DiskfftError) as ei:
        diskfft_amd.Ensemble(diskfft_amd.disk_params(16, 16))
    assert ei.value.code == diskfft_amd.DFK_EHIP
    assert "device" in str(ei.value)


def test_missing_library_raises(monkeypatch, tmp_path):
    import diskfft_amd
    monkeypatch.setattr(diskfft_amd, "_lib", None)
    monkeypatch.setattr(diskfft_amd, "LIBPATH", str(tmp_path / "nope.so"))
    with pytest.raises(ImportError):
        diskfft_amd.lib()
