"""MI355X-native derivative-carrying MLP engine (DNNF90's hot path on gfx950).

The product is libdnnf90_amd.so (include/dnnf90.h, include/dnnf90_amd.h);
this package holds its sources (csrc/), the build script, a ctypes mirror
of the C ABI (api) and a torch.autograd.Function over it (autograd).
"""
from .api import (ACT_ERF, ACT_GELU, ACT_J0, ACT_J1, ACT_SIN, ACT_TANH, PREC_BF16X3, PREC_BF16X6, PREC_FP32,
                  PREC_FP64, BatchWork, Committee, DNNF90Error, Grad, KalmanFilter, Net, Residual, TableSet,
                  TWork, Work, adam_step, current_alpha)

__all__ = ["ACT_ERF", "ACT_GELU", "ACT_J0", "ACT_J1", "ACT_SIN", "ACT_TANH", "PREC_BF16X3", "PREC_BF16X6",
           "PREC_FP32", "PREC_FP64", "BatchWork", "Committee", "DNNF90Error", "Grad", "KalmanFilter", "Net",
           "Residual",
           "TableSet", "TWork", "Work", "adam_step", "current_alpha"]
