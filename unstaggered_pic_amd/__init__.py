"""MI355X-native per-timestep hot path of the Unstaggered_PIC teaching codes.

The compute path is the HIP library ``libupic.so`` (C ABI in include/upic.h);
:mod:`unstaggered_pic_amd.lib` binds it.  Nothing in this package falls back
to a CPU implementation.
"""
from .abi import (CODE_P1, CODE_P2, CODE_P3, PREC_FP64, PREC_MIXED, Diagnostics, HostState, Params, PathStats,
                  StepInfo, make_params)

__all__ = ["CODE_P1", "CODE_P2", "CODE_P3", "PREC_FP64", "PREC_MIXED", "Diagnostics", "HostState", "Params",
           "PathStats", "StepInfo", "make_params"]
