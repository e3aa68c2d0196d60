"""The C ABI libraries load on any machine and export every declared symbol;
the Python ctypes mirror matches the C struct layout."""

import ctypes as C
import os
import re

import claysampler_amd as ca
from claysampler_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^(?:int|const char\*|size_t|const int16_t\*|const clay_plan_view\*)\s+(clay_\w+)\(",
                                 text, re.M)))


def test_hip_library_exports_every_declared_entry_point():
    lib = C.CDLL(ca.HIP_LIB)
    names = declared("clay_hip.h")
    assert len(names) >= 13
    for name in names:
        assert hasattr(lib, name), name


def test_host_library_exports_every_declared_entry_point():
    lib = C.CDLL(ca.HOST_LIB)
    names = declared("clay_host.h")
    assert "clay_session_add_job" in names
    for name in names:
        assert hasattr(lib, name), name


def test_ctypes_layout_matches_c():
    for name, struct in abi.ABI_STRUCTS.items():
        assert ca.host().clay_abi_sizeof(name.encode()) == C.sizeof(struct), name


def test_create_without_device_fails_cleanly():
    """No HIP device here: creation must fail with a message, never fall back."""
    if os.path.exists("/dev/kfd"):
        return
    lib = ca.hip()
    handle = C.c_void_p()
    cfg = abi.HipConfig(0, abi.POLICY_NATURAL, 48, 0, abi.default_audio_config(), 0, 0, 64, 0, 0)
    assert lib.clay_hip_create(C.byref(cfg), C.byref(handle)) != abi.OK
    assert lib.clay_hip_last_error(None)


def test_sinc_table_of_the_render_core_matches_the_oracle():
    import numpy as np

    import oracle

    for taps in (16, 32, 48, 64):
        assert np.array_equal(ca.sinc_table(taps), oracle.sinc_table(taps))
