"""GPU parity at the full sizes of the two measured workloads, called through
the C ABI, with the same synthetic data the measurement uses:

* scaled array: 256 inputs, 8-tap PFB, 4096 channels, 2048 spectra per dump
  (FP16 hi+lo split and FP32 X-engines, FP64 cross-dump accumulator);
* wide-band: 512 inputs, 4-tap PFB, 1024 channels, 8192 spectra per dump
  (FP8 e4m3 X-engine; 4.3e9 samples, more than 32-bit element indices hold).

At these sizes the FP64 oracle chain runs on a few channels spread over the
band (the tone bin included), with the same bars as the reduced-size tests;
the rest of the output is checked through properties that hold at any size:
every visibility written (the output is NaN-filled first), autocorrelations
real and positive, |V_ab| <= sqrt(V_aa V_bb) on every cross baseline of
every channel (Cauchy-Schwarz holds exactly for the correlation of whatever
samples the X-engine multiplied, so only its FP32 summation may break it),
the FP16 split and FP32 paths within the error bar of each other everywhere,
a repeated dump bit-identical, and the FP64 accumulator after two dumps of
the same samples equal to twice one dump, exactly."""
import numpy as np
import pytest
import torch

import mdtfx
import oracle as orc

pytestmark = pytest.mark.gpu

SEED, RHO, MAXDELAY = 2026, 0.3, 8


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _nan_vis(plan):
    return torch.full((plan.N, plan.nb), complex(np.nan, np.nan), dtype=torch.complex64, device="cuda")


def _baselines(NN):
    """triangle order (a >= b, row by row): a, b of every column, the
    autocorrelation columns and the cross-correlation columns"""
    a, b = np.tril_indices(NN)
    diag = np.array([orc.tri_index(i, i) for i in range(NN)])
    assert np.array_equal(diag, np.nonzero(a == b)[0])
    cross = np.nonzero(a != b)[0]
    return tuple(torch.from_numpy(v).cuda() for v in (a, b, diag, cross))


def _autonorm_err(vis, v64, NN):
    """per channel: max over a >= b of |V - V64| / sqrt(V64_aa V64_bb)"""
    a, b = np.tril_indices(NN)
    diag = np.array([orc.tri_index(i, i) for i in range(NN)])
    auto = np.real(v64[:, diag])
    return np.max(np.abs(vis - v64) / np.sqrt(auto[:, a] * auto[:, b]), axis=1)


def _invariants(vis, NN, tol):
    """every element finite; autocorrelations real (|Im| <= tol Re) and
    positive; the largest coherence |V_ab| / sqrt(V_aa V_bb) over every
    channel and cross baseline returned (Cauchy-Schwarz: <= 1 up to the
    summation error)"""
    assert bool(torch.isfinite(vis.real).all() and torch.isfinite(vis.imag).all()), "unwritten or non-finite output"
    a, b, diag, cross = _baselines(NN)
    auto = vis[:, diag]
    assert bool((auto.real > 0).all())
    im = float((auto.imag.abs() / auto.real).max())
    assert im <= tol, im
    ar = auto.real.double()
    ac, bc = a[cross], b[cross]
    worst = 0.0
    for c0 in range(0, vis.shape[0], 256):  # channel blocks: bounded temporaries
        v = vis[c0:c0 + 256][:, cross]
        norm = torch.sqrt(ar[c0:c0 + 256][:, ac] * ar[c0:c0 + 256][:, bc])
        worst = max(worst, float((v.abs().double() / norm).max()))
    return worst


def _cross_err(v1, v2, NN):
    """max over channels and a >= b of |V1 - V2| / sqrt(V2_aa V2_bb), on the GPU"""
    a, b, diag, _ = _baselines(NN)
    ar = v2[:, diag].real.double()
    worst = 0.0
    for c0 in range(0, v1.shape[0], 256):
        norm = torch.sqrt(ar[c0:c0 + 256][:, a] * ar[c0:c0 + 256][:, b])
        d = (v1[c0:c0 + 256].to(torch.complex128) - v2[c0:c0 + 256].to(torch.complex128)).abs()
        worst = max(worst, float((d / norm).max()))
    return worst


def test_scaled_array_full_size():
    NN, N, T, Ns = 256, 4096, 8, 2048
    tone = 205  # bin of the 0.05 cycles/sample tone
    print(f"full size: scaled array NN={NN} N={N} T={T} Ns={Ns}", flush=True)
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=SEED, sigma=20.0, rho=RHO, max_delay=MAXDELAY,
                        tones=((tone / N, 30.0),))
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP16_SPLIT)
    xin = plan.alloc_xinput()
    acc = torch.zeros((N, plan.nb), dtype=torch.complex128, device="cuda")
    v16 = plan.dump(x8, xin, _nan_vis(plan), acc)
    v16b = plan.dump(x8, xin, _nan_vis(plan), acc)
    torch.cuda.synchronize()
    assert torch.equal(v16.view(torch.float32), v16b.view(torch.float32)), "repeated dump differs"
    assert torch.equal(acc, 2 * v16.to(torch.complex128)), "FP64 accumulator != twice one dump"
    del v16b, acc
    cs16 = _invariants(v16, NN, 1e-5)

    plan.set_precision(mdtfx.XPREC_FP32)
    del xin
    torch.cuda.empty_cache()
    xin = plan.alloc_xinput()
    v32 = plan.dump(x8, xin, _nan_vis(plan))
    del xin
    torch.cuda.empty_cache()
    cs32 = _invariants(v32, NN, 1e-5)
    e16_32 = _cross_err(v16, v32, NN)

    W = _host(plan.window()).astype(np.float64)
    chans = np.unique(np.concatenate([[tone - 1, tone, tone + 1], np.linspace(0, N - 1, 6).astype(int)]))
    v64 = orc.fx_chain_f64(W, _host(x8), NN, N, T, Ns, chans)
    ci = torch.from_numpy(chans).cuda()
    e16 = _autonorm_err(_host(v16[ci]), v64, NN)
    e32 = _autonorm_err(_host(v32[ci]), v64, NN)
    print(f"vs FP64 on channels {chans.tolist()}: fp16-split {e16.max():.3e}, fp32 {e32.max():.3e}; "
          f"fp16-split vs fp32 over all channels {e16_32:.3e}; largest coherence "
          f"{cs16:.8f} / {cs32:.8f}")
    assert e16.max() < 1e-5 and e32.max() < 1e-5
    assert e16_32 < 2e-5
    assert cs16 <= 1 + 1e-5 and cs32 <= 1 + 1e-5


def test_wideband_fp8_full_size():
    NN, N, T, Ns = 512, 1024, 4, 8192
    tone, amp, sigma = 300, 12.0, 4.0
    print(f"full size: wide-band FP8 NN={NN} N={N} T={T} Ns={Ns}", flush=True)
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=SEED, sigma=sigma, rho=RHO, max_delay=MAXDELAY,
                        tones=((tone / N, amp),))
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP8)
    xin = plan.alloc_xinput()
    v8 = plan.dump(x8, xin, _nan_vis(plan))
    torch.cuda.synchronize()
    del xin
    torch.cuda.empty_cache()
    # FP8 products are exact and summed in FP32: Cauchy-Schwarz to ~2^-16
    cs8 = _invariants(v8, NN, 1e-3)

    W = _host(plan.window()).astype(np.float64)
    chans = np.unique(np.concatenate([[tone - 1, tone, tone + 1], np.linspace(0, N - 1, 6).astype(int)]))
    v64 = orc.fx_chain_f64(W, _host(x8), NN, N, T, Ns, chans)
    e8 = _autonorm_err(_host(v8[torch.from_numpy(chans).cuda()]), v64, NN)
    i_t = int(np.nonzero(chans == tone)[0][0])
    noise = np.array([i for i in range(len(chans)) if abs(int(chans[i]) - tone) > 1])
    diag = np.array([orc.tri_index(i, i) for i in range(NN)])
    auto = np.real(v64[:, diag]).mean(axis=1)
    dr_db = 10 * np.log10(auto[i_t] / np.median(auto[noise]))
    print(f"tone {dr_db:.1f} dB above noise; FP8 vs FP64: tone {e8[i_t]:.3e}, noise max {e8[noise].max():.3e}; "
          f"largest coherence {cs8:.6f}")
    assert dr_db >= 40.0
    # e4m3 rounding averages over the spectra in noise channels (~6 sigma bound);
    # a bin-centred tone repeats its rounding error in every spectrum (<= 2 x 2^-4)
    bound = 0.35 / np.sqrt(Ns)
    assert e8[noise].max() < bound, (e8, bound)
    assert e8[i_t] < 2 * 2.0 ** -4 + bound
    assert cs8 <= 1 + 1e-3
