"""CPU: the C-ABI library loads without a GPU and exports every symbol that
include/ugu_hip.h declares; the Python signatures cover the same set."""
import os
import re

from ugu_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "ugu_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ugu_hip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load_library()
    names = header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_header():
    assert sorted(_abi.SIGNATURES) == header_functions()


def test_abi_version():
    assert _abi.load_library().ugu_hip_abi_version() == 4


def test_create_without_device_returns_null():
    # no GPU in this container (or an invalid ordinal): creation fails cleanly
    lib = _abi.load_library()
    assert not lib.ugu_hip_create(1 << 20)


def test_struct_sizes_match_header_layout():
    import ctypes
    assert ctypes.sizeof(_abi.UguCamera) == 4 * (2 + 4 + 9 + 3 + 9 + 3)
    assert ctypes.sizeof(_abi.UguOptions) == 28
    assert ctypes.sizeof(_abi.UguOutputs) == 40
    assert ctypes.sizeof(_abi.UguTiling) == 16
