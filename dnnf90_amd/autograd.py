"""torch.autograd.Function over the C ABI.

forward(x, w) -> T[N, nout, na]: every carried derivative of every output,
computed by the gfx950 forward kernels; backward(dL/dT) -> dL/dw by the
hand-written adjoint kernels (the exact analytical reverse of
lib/train_module.f90:360-450).  Like the reference, the adjoint stops at the
first hidden layer, so no gradient flows to x (lib/train_module.f90:430).

This is what the Python comparison scripts (bench/post/compare_frameworks.py,
bench/compare/grad_cost.py) consume: any loss written in torch on T seeds the
adjoint through autograd.

One work space serves every call of a DerivMLP.  The reference's contract is
that a backward follows the matching forward on the same work space
(lib/train_module.f90:341-355); a loss built from several forwards (interior
and boundary points summed into one loss) breaks it, so every forward is
stamped, and a backward whose forward is no longer the one held in the work
space first runs that forward again on its saved inputs.  The kernels are
deterministic, so the re-run reproduces the planes bit for bit.
"""
from __future__ import annotations

import torch

from .api import PREC_FP64, BatchWork, Grad, Net


class _DerivForward(torch.autograd.Function):
    # the work space launches on its own torch stream, which waits for the
    # caller's stream before (x, the seed) and is waited for after (T): the
    # ABI's NULL stream means "the work space's own stream", so torch's
    # default stream cannot be handed over directly
    @staticmethod
    def forward(ctx, x, w, layer):
        x = x.contiguous()
        T = layer._run_forward(x, w, want=True)
        ctx.layer = layer
        ctx.stamp = layer._held
        ctx.save_for_backward(x, w)
        return T

    @staticmethod
    def backward(ctx, gT):
        layer = ctx.layer
        x, w = ctx.saved_tensors
        if layer._held != ctx.stamp:
            # another forward ran on the work space since this one: restore it
            layer._run_forward(x, w, want=False)
            ctx.stamp = layer._held
        if gT.shape != (x.shape[0], layer.net.dims[-1], layer.net.na):
            raise ValueError(f"seed of shape {tuple(gT.shape)} for a forward of {x.shape[0]} points")
        layer.grad.zero()
        cur = torch.cuda.current_stream(gT.device)
        seed = gT.contiguous().to(torch.float64)
        layer.stream.wait_stream(cur)
        layer.bwork.backward(seed, layer.grad)
        cur.wait_stream(layer.stream)
        g = torch.from_numpy(layer.grad.get()).to(gT.device)
        return None, g, None


class DerivMLP(torch.nn.Module):
    """MLP whose forward returns all carried input derivatives.

    w is a flat float64 Parameter in the ABI layout: for l = 2..nlayer,
    ndim(l) rows of ndim(l-1)+1 values, bias first.  budget (device bytes)
    switches the adjoint to per-layer recompute of h once keeping every
    layer's state would pass it.
    """

    def __init__(self, net: Net, maxpt: int, prec: int = PREC_FP64, device="cuda", budget=0):
        super().__init__()
        self.net = net
        self.w = torch.nn.Parameter(torch.from_numpy(net.weights).to(device))
        self.bwork = BatchWork(net, maxpt, prec, budget)
        self.stream = torch.cuda.Stream(device=device)
        self.bwork.set_stream(self.stream.cuda_stream)
        self.grad = Grad(net)
        self._wver = None
        self._stamps = 0     # forwards run so far
        self._held = 0       # stamp of the forward whose planes the work space holds

    def _sync_weights(self, w):
        ver = (w.data_ptr(), w._version)
        if ver != self._wver:
            self.net.weights = w.detach().cpu().numpy()
            self._wver = ver

    def _run_forward(self, x, w, want):
        self._sync_weights(w)
        cur = torch.cuda.current_stream(x.device)
        T = None
        if want:
            T = torch.empty((x.shape[0], self.net.dims[-1], self.net.na), dtype=torch.float64,
                            device=x.device)
        self.stream.wait_stream(cur)
        self.bwork.forward(x, T, want=want)
        cur.wait_stream(self.stream)
        self._stamps += 1
        self._held = self._stamps
        return T

    def forward(self, x):
        if x.shape[0] > self.bwork.maxpt:
            raise ValueError(f"batch of {x.shape[0]} exceeds maxpt={self.bwork.maxpt}")
        return _DerivForward.apply(x.to(torch.float64), self.w, self)
