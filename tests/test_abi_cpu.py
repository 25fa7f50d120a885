"""The C ABI library loads without a GPU and exports every entry point that
include/percolate_hip.h declares; host-only entry points and argument
checks work without a device."""
import ctypes
import os
import re

import pytest

from pypercolate_amd import _native as nat

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      'include', 'percolate_hip.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(pc_[a-z0-9_]+)\s*\(', text)))


def test_header_declares_functions():
    names = declared_functions()
    assert 'pc_sweep' in names and 'pc_microcanonical_sums' in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    lib = nat.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_binding_table_matches_header():
    assert sorted(nat.SIGNATURES) == declared_functions()


def test_host_only_entry_points():
    lib = nat.load_library()
    assert lib.pc_abi_version() >= 1
    assert lib.pc_record_bytes(65536) == 4
    assert lib.pc_record_bytes(65537) == 8
    a, b = ctypes.c_int64(), ctypes.c_int64()
    assert lib.pc_binomial_window(32512, 0.5, ctypes.byref(a), ctypes.byref(b)) == 0
    sigma = (32512 * 0.25) ** 0.5
    assert a.value <= 16256 - 12 * sigma and a.value + b.value - 1 >= 16256 + 12 * sigma
    assert lib.pc_binomial_window(10, 0.0, ctypes.byref(a), ctypes.byref(b)) == 0
    assert a.value == 0 and b.value == 11


def test_argument_errors_without_gpu():
    lib = nat.load_library()
    rc = lib.pc_sweep(0, 1, None, None, 0, None, 1, 1, None, 1, None, None, 0, 0, 0,
                      None, None)
    assert rc == nat.PC_EINVAL
    assert b'num_nodes' in lib.pc_last_error()
    rc = lib.pc_sweep(16, 1, None, None, 0, None, 1, 1, None, 1, None, None, 0, 99,
                      0, None, None)
    assert rc == nat.PC_EINVAL and b'merge_rounds' in lib.pc_last_error()
    rc = lib.pc_sweep(20000, 1, None, None, 0, None, 1, 1, None, 1, None, None, 0,
                      0, nat.PC_STATE_LDS, None, None)
    assert rc == nat.PC_EINVAL
    with pytest.raises(ValueError):
        nat.check(rc, 'pc_sweep')
    # both wave flags at once is not a mode
    rc = lib.pc_sweep(16, 1, None, None, 0, None, 1, 1, None, 1, None, None, 0,
                      0, nat.PC_SWEEP_WAVES_MASK, None, None)
    assert rc == nat.PC_EINVAL and b'state_mode' in lib.pc_last_error()
    rc = lib.pc_binomial_window(10, 1.5, ctypes.byref(ctypes.c_int64()),
                                ctypes.byref(ctypes.c_int64()))
    assert rc == nat.PC_EINVAL
    rc = lib.pc_microcanonical_sums(1 << 22, 10, 1, None, 1, None, None, None,
                                    None, 0, None)
    assert rc == nat.PC_EOVERFLOW
