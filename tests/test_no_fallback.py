"""The device path has no CPU fallback: without the built library, or without a
gfx950 device, creating a device simulation raises instead of computing."""
import os

import pytest

import cases
from demystification_dfn_amd import _abi, device


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(device, "LIB_PATH", str(tmp_path / "libdfn.so"))
    monkeypatch.setattr(device, "_fns", None)
    monkeypatch.setattr(device, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        device.DeviceSimulation(cases.single_fault_small())


@pytest.mark.skipif(not os.path.exists(device.LIB_PATH), reason="device library not built")
@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_no_device_is_a_hip_error():
    with pytest.raises(_abi.DfnError) as e:
        device.DeviceSimulation(cases.single_fault_small())
    assert e.value.code == _abi.DFN_ERR_HIP
