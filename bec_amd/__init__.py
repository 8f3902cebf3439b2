"""MI355X-native split-step Wigner GPE propagator (gfx950 HIP kernels behind a C ABI).

The package is a thin ctypes layer over libbec_amd.so (include/bec_api.h).
"""
from .api import (BecError, CalcParams, FFTPlan, GsOptions, ModelParams, Solver, default_model, derive,
                  device_count, ground_states_batched, ground_states_two_component, gs_options, lib,
                  params1d, params2d, params2d_from_model, run_1d, run_2d, time_copy, cells_1d)
from .api import (FFT_FORWARD, FFT_INVERSE, FP32, FP64, GS_GRADIENT, GS_RULE_REFERENCE, GS_RULE_RESIDUAL,
                  GS_SPLIT_STEP, PULSE_HALF_PI, PULSE_PI, SPACE_HYBRID, SPACE_K, SPACE_X)

__all__ = ["BecError", "CalcParams", "FFTPlan", "GsOptions", "ModelParams", "Solver", "default_model", "derive",
           "device_count", "ground_states_batched", "ground_states_two_component", "gs_options", "lib", "params1d", "params2d", "params2d_from_model", "run_1d", "run_2d", "cells_1d", "time_copy", "FFT_FORWARD", "FFT_INVERSE", "FP32",
           "FP64", "GS_GRADIENT", "GS_RULE_REFERENCE", "GS_RULE_RESIDUAL", "GS_SPLIT_STEP", "PULSE_HALF_PI",
           "PULSE_PI", "SPACE_HYBRID", "SPACE_K", "SPACE_X"]
