"""MI355X-native per-timestep hot path of the Unstaggered_PIC teaching codes.

The compute path is the HIP library ``libupic.so`` (C ABI in include/upic.h);
:mod:`unstaggered_pic_amd.lib` binds it.  Nothing in this package falls back
to a CPU implementation.
"""
from .abi import (CODE_P1, CODE_P2, CODE_P3, P3_LANDAU, P3_TWO_STREAM, P3_WEIBEL, PREC_FP64, PREC_MIXED,
                  Diagnostics, HostState, P3Case, Params, PathStats, StepDiagnostics, StepInfo, make_params)

__all__ = ["CODE_P1", "CODE_P2", "CODE_P3", "P3_LANDAU", "P3_TWO_STREAM", "P3_WEIBEL", "PREC_FP64", "PREC_MIXED",
           "Diagnostics", "HostState", "P3Case", "Params", "PathStats", "StepDiagnostics", "StepInfo", "make_params"]
