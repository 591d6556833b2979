"""The allocation-free calibration entry point (mdt_calib_loss_grad_noalloc,
for callers that cannot allocate device memory), called through the C ABI:
the same loss and adjoint as mdt_calib_loss_grad with a plan, bit for bit,
on two calls in a row -- the second with a smaller shape, on the same static
workspace -- and within the FP64 oracle's bounds; shapes outside the limits
are refused before any work is queued."""
import ctypes

import numpy as np
import pytest
import torch

import mdtfx
import oracle as orc
from mdtfx.calib import CalibPlan

pytestmark = pytest.mark.gpu


def _case(NN, N, seed):
    rng = np.random.default_rng(seed)
    nb = NN * (NN + 1) // 2
    V = (rng.standard_normal((N, nb)) + 1j * rng.standard_normal((N, nb))).astype(np.complex64)
    M = (rng.standard_normal((N, nb)) + 1j * rng.standard_normal((N, nb))).astype(np.complex64)
    g = (rng.uniform(0.8, 1.2, NN) * np.exp(1j * rng.uniform(-np.pi, np.pi, NN))).astype(np.complex64)
    tau = rng.uniform(-4, 4, NN).astype(np.float32)
    return V, M, g, tau


def _noalloc(NN, N, V, M, g, tau):
    loss = torch.full((), np.nan, dtype=torch.float32, device="cuda")
    gg = torch.full((NN,), complex(np.nan, np.nan), dtype=torch.complex64, device="cuda")
    gt = torch.full((NN,), np.nan, dtype=torch.float32, device="cuda")
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    st = mdtfx.lib().mdt_calib_loss_grad_noalloc(NN, N, p(V), p(M), p(g), p(tau), p(loss), p(gg), p(gt),
                                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 0, mdtfx.lib().mdt_last_error_string()
    torch.cuda.synchronize()
    return float(loss.cpu()), gg.cpu().numpy(), gt.cpu().numpy()


def test_noalloc_equals_plan_entry_twice():
    for NN, N, seed in ((150, 512, 1), (40, 96, 2)):  # the second call smaller
        print(f"no-alloc calibration NN={NN} N={N}", flush=True)
        V, M, g, tau = _case(NN, N, seed)
        args = [torch.from_numpy(a).cuda() for a in (V, M, g, tau)]
        plan = CalibPlan(NN, N)
        l0, g0, t0 = plan.loss_grad(*args)
        torch.cuda.synchronize()
        l0, g0, t0 = float(l0.cpu()), g0.cpu().numpy(), t0.cpu().numpy()
        l1, g1, t1 = _noalloc(NN, N, *args)
        assert l1 == l0 and np.array_equal(g1, g0) and np.array_equal(t1, t0)
        L64, gg64, gt64 = orc.calib_loss_grad(V.astype(np.complex128), M.astype(np.complex128),
                                              g.astype(np.complex128), tau.astype(np.float64), NN, N)
        assert abs(l1 - L64) / L64 < 1e-6
        assert np.max(np.abs(g1 - gg64)) / np.max(np.abs(gg64)) < 1e-5
        assert np.max(np.abs(t1 - gt64)) / np.max(np.abs(gt64)) < 1e-5
        plan.destroy()


def test_noalloc_refuses_shapes_beyond_the_limits():
    L = mdtfx.lib()
    d = ctypes.c_void_p(256)  # never dereferenced: validation fails first
    assert L.mdt_calib_loss_grad_noalloc(4097, 1024, d, d, d, d, d, d, d, None) == mdtfx.ERR_INVALID_VALUE
    assert L.mdt_calib_loss_grad_noalloc(64, 16385, d, d, d, d, d, d, d, None) == mdtfx.ERR_INVALID_VALUE
    assert L.mdt_calib_loss_grad_noalloc(64, 1024, d, d, d, d, None, None, None, None) == mdtfx.ERR_NULL_POINTER
