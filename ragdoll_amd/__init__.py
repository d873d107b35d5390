"""ragdoll_amd: MI355X-native H.264 P-frame analysis for ragdoll's Balanced encoder.

The compute lives in `libragdoll_hip.so` (HIP kernels for gfx950 behind the C ABI of
include/ragdoll_hip.h). This module is a thin ctypes binding of that ABI for Python callers
(tests, benchmark, training data generation); it never falls back to a CPU path: when the
library is missing or no GPU is visible, its calls raise.
"""
import ctypes
import os

import numpy as np

from .abi import (DECISION_DTYPE, LEVELS_PER_MB, PROFILES, RghError, StreamAnalyzer, load_library,
                  analyse_batch, analyse_frames, write_p_slice_cavlc, write_idr_slice_cavlc, write_p_slice_cabac,
                  write_idr_slice_cabac, write_headers)

__all__ = ["DECISION_DTYPE", "LEVELS_PER_MB", "PROFILES", "RghError", "StreamAnalyzer",
           "load_library", "analyse_batch", "analyse_frames", "write_p_slice_cavlc", "write_idr_slice_cavlc",
           "write_p_slice_cabac", "write_idr_slice_cabac", "write_headers"]
