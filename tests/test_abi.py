"""The C ABI library loads and exports every entry point include/ragdoll_hip.h declares."""
import ctypes
import os
import re

import pytest

import ragdoll_amd
from ragdoll_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "ragdoll_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rgh_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = abi.load_library()
    names = declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(abi.EXPORTS)


def test_status_strings():
    L = abi.load_library()
    assert L.rgh_status_str(0) == b"ok"
    assert L.rgh_status_str(3) == b"invalid dimensions"
    assert L.rgh_status_str(999) == b"unknown status"


def test_argument_errors_return_codes_before_touching_a_device():
    L = abi.load_library()
    h = ctypes.c_void_p()
    assert L.rgh_create(None, ctypes.byref(h)) == 1
    for w, hh, prof, qp in [(0, 16, 0, 26), (17, 16, 0, 26), (16, 16, 7, 26), (16, 16, 0, 0), (16, 16, 0, 52)]:
        cfg = abi.RgConfig(w, hh, prof, qp, 0, 0, 0, 0)
        assert L.rgh_create(ctypes.byref(cfg), ctypes.byref(h)) in (2, 3)
    assert L.rgh_destroy(None) == 1
    assert L.rgh_analyse_p_frame(None, None, None, None, None) == 1
    assert L.rgh_analyse_batch(None, 0, None) == 1


def test_record_layouts_match_between_product_and_oracle():
    from oracle_lib import DECISION_DTYPE, LEVELS_PER_MB
    assert DECISION_DTYPE == abi.DECISION_DTYPE
    assert LEVELS_PER_MB == abi.LEVELS_PER_MB
    assert ctypes.sizeof(abi.RgConfig) == 36  # nine int32 fields of rgh_config
    assert ctypes.sizeof(abi.RgSliceConfig) == 24  # six int32 fields of rgh_slice_config
    assert ctypes.sizeof(abi.RgStreamConfig) == 28  # seven int32 fields of rgh_stream_config


def test_compact_preset_is_baseline_only():
    # With CABAC the reference's Compact preset codes B pictures: refused before any device use.
    L = abi.load_library()
    h = ctypes.c_void_p()
    for prof in (1, 2):
        cfg = abi.RgConfig(64, 48, prof, 26, 0, 0, 0, abi.FLAG_COMPACT, 0)
        assert L.rgh_create(ctypes.byref(cfg), ctypes.byref(h)) == 2


def test_partitions_do_not_combine_with_other_presets():
    # RGH_FLAG_PARTITIONS extends the Balanced search only: refused with Compact or Fast before
    # any device use.
    L = abi.load_library()
    h = ctypes.c_void_p()
    for extra in (abi.FLAG_COMPACT, abi.FLAG_FAST):
        cfg = abi.RgConfig(64, 48, 0, 26, 0, 0, 0, abi.FLAG_PARTITIONS | extra, 0)
        assert L.rgh_create(ctypes.byref(cfg), ctypes.byref(h)) == 1  # RGH_ERR_INVALID_ARGUMENT


def test_writers_refuse_unknown_presets():
    import numpy as np
    L = abi.load_library()
    dec = np.zeros(4, abi.DECISION_DTYPE)
    lv = np.zeros((4, abi.LEVELS_PER_MB), np.int16)
    out = np.zeros(4096, np.uint8)
    n = ctypes.c_int64()
    cfg = abi.RgSliceConfig(2, 2, 26, 1, 2, 7)
    assert L.rgh_write_p_slice_cavlc(ctypes.byref(cfg), abi._ptr(dec), abi._ptr(lv), abi._ptr(out), 4096,
                                     ctypes.byref(n)) == 1
    scfg = abi.RgStreamConfig(32, 32, 0, 26, 30, 1, 9)
    assert L.rgh_write_headers(ctypes.byref(scfg), abi._ptr(out), 4096, ctypes.byref(n)) == 1
