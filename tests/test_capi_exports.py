"""The C-ABI library loads without a GPU and exports every entry point that
include/kumo_hip.h declares (no compute call is made here)."""
import ctypes
import os
import re

from kumo_hip import binding

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "kumo_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void)\s+(kumo_hip_\w+)\(", text, re.M)))


def test_header_declares_the_binding_exports():
    assert sorted(binding.EXPORTS) == declared()


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(binding.LIBRARY)
    missing = [name for name in declared() if not hasattr(lib, name)]
    assert not missing, missing


def test_default_options_without_device():
    o = binding.options()
    assert o.block_size == 128 and o.backend == 0 and o.precision == 0
    assert o.screening_threshold == 1e-12 and o.potential_tolerance == 1e-8 and o.fukui_delta == 0.1
