"""The C ABI library loads and exports every declared symbol (CPU only)."""

import ctypes
import os
import re

import celeritas_amd

HEADER = os.path.join(celeritas_amd.INCLUDE_DIR, "celeritas_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(celer_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(celeritas_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_no_device_is_reported_cleanly():
    # Loading and querying never initializes a transport path on the CPU
    assert celeritas_amd.lib().celer_amd_version() == b"0.1.0"
    if not celeritas_amd.device_available():
        p = celeritas_amd.lib().celer_params_create(b"{}")
        assert not p
        assert b"no HIP device" in celeritas_amd.lib().celer_amd_last_error()
