"""Gain/delay calibration (the calibration case: 64 inputs, 4-tap PFB, 1024
channels, 512 spectra): the F-engine gain stage, the GPU loss and the
hand-written adjoint, against the FP64 oracle (analytic gradient and central
finite differences)."""
import numpy as np
import pytest
import torch

import mdtfx
import oracle as orc
from mdtfx.calib import CalibPlan, calib_loss

pytestmark = pytest.mark.gpu

NN, N, T, NS = 64, 1024, 4, 512


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def c5():
    rng = np.random.default_rng(55)
    plan = mdtfx.FxPlan(NN, N, T, NS, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP32)
    x8 = mdtfx.synth_i8(NN, N, NS + T - 1, seed=55, sigma=16.0, rho=0.5, max_delay=4)
    V = plan.dump(x8, plan.alloc_xinput(), plan.alloc_vis())
    g_true = (rng.uniform(0.8, 1.2, NN) * np.exp(1j * rng.uniform(-np.pi, np.pi, NN))).astype(np.complex64)
    tau_true = rng.uniform(-4, 4, NN).astype(np.float32)
    g = (g_true * (1 + 0.05 * (rng.standard_normal(NN) + 1j * rng.standard_normal(NN)))).astype(np.complex64)
    tau = (tau_true + 0.1 * rng.standard_normal(NN)).astype(np.float32)
    cal = CalibPlan(NN, N)
    M = cal.apply(V, torch.from_numpy(g_true).cuda(), torch.from_numpy(tau_true).cuda())
    return dict(plan=plan, x8=x8, V=V, M=M, g=g, tau=tau, g_true=g_true, tau_true=tau_true, cal=cal)


def _autonorm(v, v64):
    a, b = np.tril_indices(NN)
    diag = np.array([orc.tri_index(i, i) for i in range(NN)])
    auto = np.abs(np.real(v64[:, diag]))
    return float(np.max(np.abs(v - v64) / np.sqrt(auto[:, a] * auto[:, b])))


def test_fengine_gain_stage_matches_apply(c5):
    """V' from the gained F-engine == G_a conj(G_b) V (FP32 on both sides)"""
    plan = c5["plan"]
    g = torch.from_numpy(c5["g"]).cuda()
    tau = torch.from_numpy(c5["tau"]).cuda()
    plan.set_gains(g, tau)
    Vfe = _host(plan.dump(c5["x8"], plan.alloc_xinput(), plan.alloc_vis()))
    plan.set_gains(None, None)
    Vap = _host(c5["cal"].apply(c5["V"], g, tau))
    V64 = orc.calib_apply(_host(c5["V"]).astype(np.complex128), c5["g"].astype(np.complex128),
                          c5["tau"].astype(np.float64), NN, N)
    assert _autonorm(Vap, V64) < 1e-6
    assert _autonorm(Vfe, V64) < 1e-5


def test_loss_and_adjoint_vs_fp64(c5):
    V = c5["V"]
    V64 = _host(V).astype(np.complex128)
    M64 = _host(c5["M"]).astype(np.complex128)
    g64 = c5["g"].astype(np.complex128)
    t64 = c5["tau"].astype(np.float64)
    L64, gg64, gt64 = orc.calib_loss_grad(V64, M64, g64, t64, NN, N)
    loss, gg, gt = c5["cal"].loss_grad(V, c5["M"], torch.from_numpy(c5["g"]).cuda(),
                                       torch.from_numpy(c5["tau"]).cuda())
    L, gg, gt = float(_host(loss)), _host(gg), _host(gt)
    assert abs(L - L64) / L64 < 1e-6, (L, L64)
    eg = np.max(np.abs(gg - gg64)) / np.max(np.abs(gg64))
    et = np.max(np.abs(gt - gt64)) / np.max(np.abs(gt64))
    print(f"adjoint vs FP64 analytic: gain {eg:.2e}, delay {et:.2e}")
    assert eg < 1e-6
    assert et < 1e-5


def test_adjoint_vs_finite_differences(c5):
    """GPU adjoint vs FP64 central finite differences of the FP64 loss"""
    V64 = _host(c5["V"]).astype(np.complex128)
    M64 = _host(c5["M"]).astype(np.complex128)
    g64 = c5["g"].astype(np.complex128)
    t64 = c5["tau"].astype(np.float64)
    _, gg, gt = c5["cal"].loss_grad(c5["V"], c5["M"], torch.from_numpy(c5["g"]).cuda(),
                                    torch.from_numpy(c5["tau"]).cuda())
    gg, gt = _host(gg), _host(gt)
    h = 1e-5
    sel = np.arange(0, NN, 5)  # 13 of the 64 inputs: 78 FP64 loss evaluations
    fd_g = np.zeros(NN, np.complex128)
    fd_t = np.zeros(NN)
    L = lambda g, t: orc.calib_loss_grad(V64, M64, g, t, NN, N)[0]
    for a in sel:
        for unit in (1.0, 1j):
            gp, gm = g64.copy(), g64.copy()
            gp[a] += h * unit
            gm[a] -= h * unit
            d = (L(gp, t64) - L(gm, t64)) / (2 * h)
            fd_g[a] += d * (1.0 if unit == 1.0 else 1j)
        tp, tm = t64.copy(), t64.copy()
        tp[a] += h
        tm[a] -= h
        fd_t[a] = (L(g64, tp) - L(g64, tm)) / (2 * h)
    eg = np.max(np.abs(gg[sel] - fd_g[sel])) / np.max(np.abs(fd_g[sel]))
    et = np.max(np.abs(gt[sel] - fd_t[sel])) / np.max(np.abs(fd_t[sel]))
    print(f"adjoint vs FP64 finite differences: gain {eg:.2e}, delay {et:.2e}")
    assert eg < 1e-6
    assert et < 1e-5


def test_autograd_function(c5):
    g = torch.from_numpy(c5["g"]).cuda().requires_grad_(True)
    tau = torch.from_numpy(c5["tau"]).cuda().requires_grad_(True)
    loss = calib_loss(g, tau, c5["V"], c5["M"], c5["cal"])
    (3.0 * loss).backward()
    _, gg, gt = c5["cal"].loss_grad(c5["V"], c5["M"], g.detach(), tau.detach())
    assert torch.allclose(g.grad, 3.0 * gg) and torch.allclose(tau.grad, 3.0 * gt)
    # a few gradient-descent steps lower the loss
    with torch.no_grad():
        l0 = float(loss)
        g2 = g - 1e-3 * g.grad / g.grad.abs().max()
        t2 = tau - 1e-3 * tau.grad / tau.grad.abs().max()
        l1 = float(calib_loss(g2, t2, c5["V"], c5["M"], c5["cal"]))
    assert l1 < l0


def test_loss_zero_at_truth(c5):
    loss, _, _ = c5["cal"].loss_grad(c5["V"], c5["M"], torch.from_numpy(c5["g_true"]).cuda(),
                                     torch.from_numpy(c5["tau_true"]).cuda(), want_grad=False)
    assert float(_host(loss)) == 0.0


@pytest.mark.parametrize("xprec,N_,T_", [(x, n, t) for n, t in ((1024, 4), (4096, 8))
                                         for x in (mdtfx.XPREC_FP32, mdtfx.XPREC_FP16_SPLIT, mdtfx.XPREC_FP8)]
                         + [(mdtfx.XPREC_FP32, 256, 4), (mdtfx.XPREC_FP16_SPLIT, 256, 8)])
def test_gain_stage_in_every_xinput_layout(xprec, N_, T_):
    """the F-engine's gain/delay stage writes the same samples in every X
    input layout (its own compiled variant per layout), and the X-engine
    gives the same visibilities"""
    NN_, NS_ = 24, 40
    print(f"gain stage layouts xprec={xprec} N={N_}", flush=True)
    rng = np.random.default_rng(7)
    g = torch.from_numpy((rng.uniform(0.5, 1.5, NN_) * np.exp(1j * rng.uniform(-3, 3, NN_))).astype(np.complex64)).cuda()
    tau = torch.from_numpy(rng.uniform(-6, 6, NN_).astype(np.float32)).cuda()
    x8 = mdtfx.synth_i8(NN_, N_, NS_ + T_ - 1, seed=8, sigma=14.0, rho=0.3, max_delay=4)
    ref = vref = None
    for layout in (mdtfx.XIN_ROWS, mdtfx.XIN_PAIRS, mdtfx.XIN_BLOCKS):
        p = mdtfx.FxPlan(NN_, N_, T_, NS_, mdtfx.WINDOW_BH4, xprec)
        p.set_xinput_layout(layout)
        p.set_gains(g, tau)
        buf = p.fengine(x8, p.alloc_xinput())
        got = _host(p.xinput_view(buf)).view(np.uint16)
        v = _host(p.xengine(buf, p.alloc_vis())).view(np.uint32)
        # FP8 per-channel format (4096 channels): exponents from power sums
        # added in a fixed order; the block-scaled format (1024) has one layout
        e = _host(p.channel_exponents()) if xprec == mdtfx.XPREC_FP8 and not p.f8block else np.zeros(N_, np.int32)
        if ref is None:
            ref, vref, eref = got, v, e
            # the gains are applied: the first input's samples differ from a plan without them
            p0 = mdtfx.FxPlan(NN_, N_, T_, NS_, mdtfx.WINDOW_BH4, xprec)
            p0.set_xinput_layout(layout)
            plain = _host(p0.xinput_view(p0.fengine(x8, p0.alloc_xinput()))).view(np.uint16)
            assert not np.array_equal(plain, got)
            continue
        assert np.array_equal(got, ref), layout
        assert np.array_equal(e, eref), layout
        assert np.array_equal(v, vref), layout


def test_gain_scale_follows_the_fengine_launch(c5):
    """FP16 split with gains large enough to change the FP16 pre-scale
    (|g| = 3): the X-engine divides by the scale the F-engine used, also when
    the gains are switched off between the F- and X-engine calls, and gains
    updated in place are picked up at the next launch (no host read-back)"""
    rng = np.random.default_rng(56)
    plan = mdtfx.FxPlan(NN, N, T, NS, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP16_SPLIT)
    s0 = plan.split_scale()
    g_big = (3.0 * np.exp(1j * rng.uniform(-np.pi, np.pi, NN))).astype(np.complex64)
    tau = torch.from_numpy(c5["tau"]).cuda()
    g = torch.from_numpy(c5["g"]).cuda()
    plan.set_gains(g, tau)
    g.copy_(torch.from_numpy(g_big).cuda())  # in place, after set_gains
    xin = plan.fengine(c5["x8"], plan.alloc_xinput())
    s1 = plan.split_scale()
    plan.set_gains(None, None)               # between the F- and the X-engine
    v = _host(plan.xengine(xin, plan.alloc_vis()))
    assert s1 * 2 <= s0, (s0, s1)            # the gains did change the scale
    V64 = orc.calib_apply(_host(c5["V"]).astype(np.complex128), g_big.astype(np.complex128),
                          c5["tau"].astype(np.float64), NN, N)
    assert _autonorm(v, V64) < 1e-5
    # gains off again: back to the plan's own scale and the ungained result
    v0 = _host(plan.dump(c5["x8"], plan.alloc_xinput(), plan.alloc_vis()))
    assert plan.split_scale() == s0
    assert _autonorm(v0, _host(c5["V"]).astype(np.complex128)) < 1e-5


@pytest.mark.parametrize("NN_,N_", [(150, 64), (3, 40)])
def test_adjoint_tiles_and_determinism(NN_, N_):
    """inputs over several 64-input tiles (and a tiny array): loss and adjoint
    on random visibilities against the FP64 oracle, and bit-identical when
    run twice (fixed-order reductions)"""
    rng = np.random.default_rng(NN_)
    nb = NN_ * (NN_ + 1) // 2
    V = (rng.standard_normal((N_, nb)) + 1j * rng.standard_normal((N_, nb))).astype(np.complex64)
    M = (rng.standard_normal((N_, nb)) + 1j * rng.standard_normal((N_, nb))).astype(np.complex64)
    g = (rng.uniform(0.8, 1.2, NN_) * np.exp(1j * rng.uniform(-np.pi, np.pi, NN_))).astype(np.complex64)
    tau = rng.uniform(-4, 4, NN_).astype(np.float32)
    cal = CalibPlan(NN_, N_)
    args = [torch.from_numpy(a).cuda() for a in (V, M, g, tau)]
    l1, g1, t1 = (_host(x) for x in cal.loss_grad(*args))
    l2, g2, t2 = (_host(x) for x in cal.loss_grad(*args))
    assert l1 == l2 and np.array_equal(g1, g2) and np.array_equal(t1, t2)
    L64, gg64, gt64 = orc.calib_loss_grad(V.astype(np.complex128), M.astype(np.complex128),
                                          g.astype(np.complex128), tau.astype(np.float64), NN_, N_)
    assert abs(float(l1) - L64) / L64 < 1e-6
    assert np.max(np.abs(g1 - gg64)) / np.max(np.abs(gg64)) < 1e-5
    assert np.max(np.abs(t1 - gt64)) / np.max(np.abs(gt64)) < 1e-5
    Vp = _host(cal.apply(args[0], args[2], args[3]))
    Vp64 = orc.calib_apply(V.astype(np.complex128), g.astype(np.complex128), tau.astype(np.float64), NN_, N_)
    assert np.max(np.abs(Vp - Vp64)) / np.max(np.abs(Vp64)) < 1e-6
