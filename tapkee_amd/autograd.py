"""KL(P||Q) as a torch.autograd.Function on ROCm tensors.

forward  : KL(P||Q) = sum_ij p_ij log(p_ij / q_ij) over the CSR pattern of P,
           q_ij = (1 + |y_i - y_j|^2)^-1 / Z, and in the same call its
           gradient (tka_kl_grad_dev): at theta == 0 one all-pairs pass on
           the device gives the repulsion and Z, which the KL reuses; at
           theta > 0 the Barnes-Hut tree gives both.
backward : dKL/dY = 4 * (F_attr - F_rep / Z), the gradient the forward left
           (the one tapkee returns, tsne.hpp:325-407, times 4), scaled by the
           incoming gradient.
Nothing of size N^2 is stored: the forward keeps Y's gradient (N x d).
float32 and float64 tensors are both accepted (P's values in Y's dtype).

`TSNELoss` wraps it for a fixed P built from data by the engine, so any
torch optimiser or a parametric map X -> Y can be driven by it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from . import build_p as _build_p

if _lib._lib is not None and not _lib.loaded_after_torch:
    raise ImportError("the engine was loaded before torch: import torch (or tapkee_amd.autograd) first so "
                      "that both use one HIP runtime")


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def _check_inputs(Y, row_ptr, col, val):
    if not Y.is_cuda:
        raise _lib.DeviceError("KLDivergence runs on the GPU: Y must be a device tensor")
    if Y.dtype not in (torch.float32, torch.float64) or val.dtype != Y.dtype:
        raise _lib.WrongParameterError("Y must be float32 or float64 and P's values of the same dtype")
    if row_ptr.dtype != torch.int32 or col.dtype != torch.int32:
        raise _lib.WrongParameterError("row_ptr and col must be int32")
    if Y.dim() != 2 or row_ptr.numel() != Y.shape[0] + 1:
        raise _lib.WrongParameterError("row_ptr must have N+1 entries for Y of shape (N, d)")


class KLDivergence(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Y, row_ptr, col, val, theta: float = 0.0):
        _check_inputs(Y, row_ptr, col, val)
        Y = Y.contiguous()
        val = val.contiguous()
        kl = C.c_double(0.0)
        z = C.c_double(0.0)
        N, d = Y.shape
        dY = torch.empty_like(Y)
        prec = _lib.FP64 if Y.dtype == torch.float64 else _lib.FP32
        _lib.check(_lib.lib().tka_kl_grad_dev(Y.data_ptr(), N, d, row_ptr.data_ptr(), col.data_ptr(),
                                              val.data_ptr(), float(theta), prec, C.byref(kl), dY.data_ptr(),
                                              C.byref(z), _stream(Y)))
        ctx.save_for_backward(dY)
        return torch.tensor(kl.value, dtype=Y.dtype, device=Y.device)

    @staticmethod
    def backward(ctx, grad_out):
        (dY,) = ctx.saved_tensors
        return 4.0 * dY * grad_out, None, None, None, None


def kl_divergence(Y, row_ptr, col, val, theta: float = 0.0):
    return KLDivergence.apply(Y, row_ptr, col, val, theta)


class TSNELoss(torch.nn.Module):
    """KL(P||Q) for the P of data X (N x D), built on the GPU by the engine."""

    def __init__(self, X, perplexity: float = 30.0, theta: float = 0.0, dist_dtype: str = "bf16",
                 device="cuda"):
        super().__init__()
        row, col, val = _build_p(np.asarray(X, dtype=np.float64), perplexity, dist_dtype=dist_dtype)
        self.theta = float(theta)
        self.register_buffer("row_ptr", torch.from_numpy(row).to(device))
        self.register_buffer("col", torch.from_numpy(col).to(device))
        self.register_buffer("val", torch.from_numpy(val.astype(np.float32)).to(device))
        self.register_buffer("val64", torch.from_numpy(val).to(device))

    def forward(self, Y):
        val = self.val64 if Y.dtype == torch.float64 else self.val
        return KLDivergence.apply(Y, self.row_ptr, self.col, val, self.theta)

