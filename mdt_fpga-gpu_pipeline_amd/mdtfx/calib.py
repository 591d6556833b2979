"""Gain/delay calibration through the F-engine and X-engine.

No counterpart in the reference (an extension of the FX chain).  The
F-engine applies per-input gains g_a and delays tau_a to the channelised
signal (mdt_fx_set_gains); the visibilities become V'_ab = G_a conj(G_b) V_ab
with G_a(c) = g_a exp(-2 pi i k_c tau_a / N).  CalibLoss evaluates
L = sum |V' - M|^2 on the GPU (mdt_calib_loss_grad) and its backward is the
hand-written adjoint kernel; PyTorch is only the autograd harness here.
"""
import ctypes

import torch

from . import _ptr, _stream, check, lib


class CalibPlan:
    def __init__(self, NN, N):
        self._h = ctypes.c_void_p()
        check(lib().mdt_calib_plan_create(ctypes.byref(self._h), NN, N))
        self.NN, self.N = NN, N
        self.nb = NN * (NN + 1) // 2

    def _sync_stream(self):
        check(lib().mdt_calib_set_stream(self._h, _stream()))

    def _check(self, V, M, g, tau):
        if V.numel() < self.N * self.nb or (M is not None and M.numel() < self.N * self.nb):
            raise ValueError("visibilities need N x NN(NN+1)/2 entries")
        if g.numel() < self.NN or (tau is not None and tau.numel() < self.NN):
            raise ValueError("gains and delays need NN entries")

    def apply(self, V, g, tau):
        """V' = G_a conj(G_b) V (the gained F-engine's visibilities)."""
        self._check(V, None, g, tau)
        self._sync_stream()
        Vp = torch.empty_like(V)
        check(lib().mdt_calib_apply(self._h, _ptr(V, torch.complex64), _ptr(g, torch.complex64),
                                    _ptr(tau, torch.float32) if tau is not None else None,
                                    _ptr(Vp, torch.complex64)))
        return Vp

    def loss_grad(self, V, M, g, tau, want_grad=True):
        self._check(V, M, g, tau)
        self._sync_stream()
        loss = torch.empty((), dtype=torch.float32, device="cuda")
        gg = torch.empty(self.NN, dtype=torch.complex64, device="cuda") if want_grad else None
        gt = torch.empty(self.NN, dtype=torch.float32, device="cuda") if want_grad else None
        check(lib().mdt_calib_loss_grad(self._h, _ptr(V, torch.complex64), _ptr(M, torch.complex64),
                                        _ptr(g, torch.complex64),
                                        _ptr(tau, torch.float32) if tau is not None else None,
                                        _ptr(loss, torch.float32),
                                        _ptr(gg, torch.complex64) if want_grad else None,
                                        _ptr(gt, torch.float32) if want_grad else None))
        return loss, gg, gt

    def destroy(self):
        if self._h:
            check(lib().mdt_calib_plan_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class CalibLoss(torch.autograd.Function):
    """L(g, tau) = sum_{c, a>=b} |G_a conj(G_b) V_ab - M_ab|^2.

    forward: the GPU loss kernel; backward: the GPU adjoint kernel, returning
    grad_output * (dL/dRe g + i dL/dIm g) and grad_output * dL/dtau."""

    @staticmethod
    def forward(ctx, g, tau, V, M, plan):
        g = g.detach().contiguous()
        tau = tau.detach().contiguous()
        loss, _, _ = plan.loss_grad(V, M, g, tau, want_grad=False)
        ctx.save_for_backward(g, tau, V, M)
        ctx.plan = plan
        return loss

    @staticmethod
    def backward(ctx, gout):
        g, tau, V, M = ctx.saved_tensors
        _, gg, gt = ctx.plan.loss_grad(V, M, g, tau, want_grad=True)
        return gout * gg, gout * gt, None, None, None


def calib_loss(g, tau, V, M, plan):
    return CalibLoss.apply(g, tau, V, M, plan)
