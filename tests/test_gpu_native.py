"""GPU parity of the native fused path (int8 -> fused F-engine -> triangle
X-engine) against the oracle and against the composition of the drop-in
kernels, called through the C ABI."""
import numpy as np
import pytest
import torch

import mdtfx
import oracle as orc

pytestmark = pytest.mark.gpu


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _tri_autonorm_err(vis, v64, NN):
    """max over (c, a >= b) of |V - V64| / sqrt(V64_aa V64_bb)"""
    a, b = np.tril_indices(NN)
    diag = np.array([orc.tri_index(i, i) for i in range(NN)])
    auto = np.real(v64[:, diag])
    norm = np.sqrt(auto[:, a] * auto[:, b])
    return float(np.max(np.abs(vis - v64) / norm))


@pytest.mark.parametrize("NN,N,T,Ns", [(40, 1024, 4, 256), (3, 32, 4, 11), (5, 64, 8, 17),
                                       (4, 2048, 4, 9), (3, 4096, 8, 16), (37, 4096, 8, 69), (3, 8192, 4, 5),
                                       (3, 16384, 4, 3), (6, 256, 5, 12), (40, 256, 4, 100), (70, 256, 8, 200),
                                       (6, 1024, 16, 40), (5, 2048, 5, 33), (3, 4096, 12, 20), (300, 512, 7, 70),
                                       (4, 1024, 128, 9)])
@pytest.mark.parametrize("layout", [mdtfx.XIN_ROWS, mdtfx.XIN_PAIRS, mdtfx.XIN_BLOCKS])
def test_fused_fengine_equals_dropin_composition(NN, N, T, Ns, layout):
    """fused PFB+FFT+corner turn == pps_batch_i8 -> fft -> reorder, bit for
    bit, in every X input layout (persistent and generic F-engine kernels)"""
    print(f"fengine NN={NN} N={N} T={T} Ns={Ns} layout={layout}", flush=True)
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_HAMMING)
    plan.set_xinput_layout(layout)
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=N + T, sigma=16.0, tones=((0.1234, 20.0),))
    xin = plan.xinput_view(plan.fengine(x8, plan.alloc_xinput()))
    W = plan.window()
    F = mdtfx.pps_batch_i8(W, x8, NN, Ns)
    fp = mdtfx.FftPlan(N, Ns * NN)
    fp.set_stream(torch.cuda.current_stream().cuda_stream)
    fp.exec_c2c(F, F)
    ref = _host(F).transpose(2, 1, 0)  # [Ns][NN][N] -> [N][NN][Ns]
    got = _host(xin)
    assert np.array_equal(got.view(np.uint32), np.ascontiguousarray(ref).view(np.uint32))


def test_native_chain_40in_hamming_vs_fp64():
    """The 40-input Hamming case: 40 inputs, 4-tap Hamming PFB, 1024 channels, 256
    spectra, int8 N(0, 16^2) + 2 tones -> 820 baselines x 1024 channels."""
    NN, N, T, Ns = 40, 1024, 4, 256
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_HAMMING)
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=1, sigma=16.0, tones=((0.1, 12.0), (-0.3, 8.0)))
    xin = plan.alloc_xinput()
    vis = plan.dump(x8, xin, plan.alloc_vis())
    V = _host(vis)
    W = _host(plan.window()).astype(np.float64)
    v64 = orc.fx_chain_f64(W, _host(x8), NN, N, T, Ns, np.arange(N))
    err = _tri_autonorm_err(V, v64, NN)
    print(f"40-input case: max autocorrelation-normalised error vs FP64: {err:.3e}")
    assert err < 1e-5
    # FP32 port on the CPU agrees to the same bar
    v32 = orc.fx_chain_f32(W.astype(np.float32), _host(x8), NN, N, T, Ns)
    assert _tri_autonorm_err(v32, v64, NN) < 1e-5


@pytest.mark.parametrize("xprec", [mdtfx.XPREC_FP32, mdtfx.XPREC_FP16_SPLIT, mdtfx.XPREC_FP8])
def test_xinput_layouts_agree(xprec):
    """every X input layout (rows with any padding, channel pairs, blocks, the
    per-precision default) holds the same samples and gives the same
    visibilities, bit for bit (FP8: the per-channel format at 2048 channels,
    the block-scaled one has a single layout)"""
    NN, N, T, Ns = 6, (2048 if xprec == mdtfx.XPREC_FP8 else 512), 8, 40
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=4, sigma=14.0, rho=0.2)
    ref = vref = eref = None
    q = 4 if xprec == mdtfx.XPREC_FP8 else 1  # FP8 path: rows pitches rounded up to multiples of 4
    for layout, pads in [(mdtfx.XIN_ROWS, (8, 64)), (mdtfx.XIN_ROWS, (0, 0)), (mdtfx.XIN_ROWS, (2, 2)),
                         (mdtfx.XIN_PAIRS, None), (mdtfx.XIN_BLOCKS, None), (mdtfx.XIN_AUTO, None)]:
        print(f"xinput layout={layout} pads={pads} xprec={xprec}", flush=True)
        p = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, xprec)
        p.set_xinput_layout(layout)
        if pads is not None:
            p.set_xinput_padding(*pads)
            rp, cp = p.xinput_pitch()
            assert rp == -(-(Ns + (Ns & 1) + pads[0]) // q) * q and cp == -(-(NN * rp + pads[1]) // q) * q
            assert p.xinput_strides() == (1 << 30, 0, 1, cp, rp)
        elif layout != mdtfx.XIN_AUTO:
            with pytest.raises(mdtfx.MdtError):
                p.xinput_pitch()  # rows layout only
        buf = p.fengine(x8, p.alloc_xinput())
        got = _host(p.xinput_view(buf)).view(np.uint16)
        v = _host(p.xengine(buf, p.alloc_vis())).view(np.uint32)
        e = _host(p.channel_exponents()) if xprec == mdtfx.XPREC_FP8 else None
        if ref is None:
            ref, vref, eref = got, v, e
            continue
        assert np.array_equal(got, ref)
        # FP8: the per-channel exponents come from power sums added in a fixed order
        if e is not None:
            assert np.array_equal(e, eref)
        assert np.array_equal(v, vref)


@pytest.mark.parametrize("xprec,NN,G,B", [(mdtfx.XPREC_FP32, 256, 1, 1 << 30), (mdtfx.XPREC_FP16_SPLIT, 256, 2, 8),
                                          (mdtfx.XPREC_FP16_SPLIT, 64, 2, 8), (mdtfx.XPREC_FP16_SPLIT, 300, 1, 16),
                                          (mdtfx.XPREC_FP8, 512, 1, 32), (mdtfx.XPREC_FP8, 64, 1, 32)])
def test_xinput_auto_layout(xprec, NN, G, B):
    """the default layout: rows for FP32, channel pairs for the FP16 split up
    to 256 inputs, 128-byte spectrum blocks above that and for FP8"""
    plan = mdtfx.FxPlan(NN, 1024, 4, 64, mdtfx.WINDOW_BH4, xprec)
    Bq, bp, Gq, cp, rp = plan.xinput_strides()
    assert (Gq, Bq) == (G, B)
    plan.set_precision(mdtfx.XPREC_FP32)  # the default follows the precision
    assert plan.xinput_strides()[2] == 1 and plan.xinput_strides()[1] == 0


@pytest.mark.parametrize("N,T", [(1024, 4), (4096, 8)])
@pytest.mark.parametrize("xprec", [mdtfx.XPREC_FP32, mdtfx.XPREC_FP16_SPLIT, mdtfx.XPREC_FP8])
@pytest.mark.parametrize("layout", [mdtfx.XIN_PAIRS, mdtfx.XIN_BLOCKS])
def test_xinput_block_geometry(layout, xprec, N, T):
    """channel pairs: sample (k, e, s) at (s // B) * block_pitch + (k // 2) *
    chan_pitch + e * 2B + (k % 2) * B + s % B, B = 64 bytes of samples;
    blocks: at (s // B) * block_pitch + k * chan_pitch + e * B + s % B, B =
    128 bytes of samples.  The F-engine writes every sample where the formula
    puts it and nothing past Ns (checked on the raw buffer)."""
    NN, Ns = 5, 37
    if xprec == mdtfx.XPREC_FP8 and N == 1024:
        N = 2048  # the per-channel FP8 format; the block-scaled one (512, 1024) is in tests/test_gpu_fp8.py
    print(f"xinput geometry layout={layout} xprec={xprec} N={N}", flush=True)
    esize = 4 if xprec == mdtfx.XPREC_FP8 else 8
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, xprec)
    plan.set_xinput_layout(layout)
    B, bp, G, cp, rp = plan.xinput_strides()
    if layout == mdtfx.XIN_PAIRS:
        assert (B, G, rp, cp, bp) == (64 // esize, 2, 128 // esize, NN * 128 // esize, (N // 2) * NN * 128 // esize)
    else:
        assert (B, G, rp, cp, bp) == (128 // esize, 1, 128 // esize, NN * 128 // esize, N * NN * 128 // esize)
    assert plan.xinput_bytes() == -(-Ns // B) * bp * esize
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=9, sigma=12.0)
    buf = plan.alloc_xinput()
    buf.view(torch.uint8).fill_(0xA5)
    plan.fengine(x8, buf)
    raw = _host(buf).view(np.uint32 if esize == 4 else np.uint64)
    rows = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, xprec)
    rows.set_xinput_layout(mdtfx.XIN_ROWS)
    want = np.ascontiguousarray(_host(rows.xinput_view(rows.fengine(x8, rows.alloc_xinput()))))
    want = want.view(np.uint32 if esize == 4 else np.uint64).reshape(N, NN, Ns)
    k, e, s = np.meshgrid(np.arange(N), np.arange(NN), np.arange(Ns), indexing="ij")
    off = (s // B) * bp + (k // G) * cp + e * rp + (k % G) * B + s % B
    assert np.array_equal(raw[off], want)
    filler = np.uint32(0xA5A5A5A5) if esize == 4 else np.uint64(0xA5A5A5A5A5A5A5A5)
    tail = np.ones(raw.shape, bool)
    tail[off.ravel()] = False
    assert np.all(raw[tail] == filler)


@pytest.mark.parametrize("NN,N,T,Ns", [(3, 32, 4, 2), (65, 256, 4, 70), (130, 128, 8, 40),
                                       (57, 256, 4, 70), (64, 128, 4, 33), (33, 64, 4, 20)])
def test_native_xengine_edges(NN, N, T, Ns):
    """tile edges (NN not a multiple of 64, Ns not a multiple of 32)"""
    print(f"native NN={NN} N={N} T={T} Ns={Ns}", flush=True)
    plan = mdtfx.FxPlan(NN, N, T, Ns)
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=NN, sigma=10.0, rho=0.3)
    vis = plan.dump(x8, plan.alloc_xinput(), torch.full((N, plan.nb), complex(np.nan, np.nan), dtype=torch.complex64, device="cuda"))
    V = _host(vis)
    assert np.all(np.isfinite(V))
    W = _host(plan.window()).astype(np.float64)
    chans = np.arange(N)
    v64 = orc.fx_chain_f64(W, _host(x8), NN, N, T, Ns, chans)
    assert _tri_autonorm_err(V, v64, NN) < 1e-5


def test_fp64_accumulator_two_dumps():
    NN, N, T, Ns = 8, 64, 4, 16
    plan = mdtfx.FxPlan(NN, N, T, Ns)
    acc = torch.zeros((N, plan.nb), dtype=torch.complex128, device="cuda")
    vs = []
    for d in range(2):
        x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=10 + d)
        vs.append(_host(plan.dump(x8, plan.alloc_xinput(), plan.alloc_vis(), acc)).astype(np.complex128))
    assert np.array_equal(_host(acc), vs[0] + vs[1])


def test_streaming_history_equals_long_run():
    """D dumps with the (T-1)-spectrum carry (client.cu:1170) == one long run"""
    NN, N, T, Ns, D = 4, 256, 8, 16, 3
    long_x8 = mdtfx.synth_i8(NN, N, D * Ns + T - 1, seed=77, sigma=12.0)
    long_plan = mdtfx.FxPlan(NN, N, T, D * Ns)
    long_xin = _host(long_plan.xinput_view(long_plan.fengine(long_x8, long_plan.alloc_xinput())))
    plan = mdtfx.FxPlan(NN, N, T, Ns)
    buf = torch.empty((Ns + T - 1, NN, N, 2), dtype=torch.int8, device="cuda")
    buf[:T - 1] = long_x8[:T - 1]
    for d in range(D):
        buf[T - 1:] = long_x8[T - 1 + d * Ns: T - 1 + (d + 1) * Ns]
        xin = np.ascontiguousarray(_host(plan.xinput_view(plan.fengine(buf, plan.alloc_xinput()))))
        assert np.array_equal(xin.view(np.uint32), np.ascontiguousarray(long_xin[:, :, d * Ns:(d + 1) * Ns]).view(np.uint32))
        mdtfx.history_shift(buf, NN, N, T, Ns)


def test_synth_deterministic_and_statistics():
    a = _host(mdtfx.synth_i8(16, 1024, 32, seed=5, sigma=16.0))
    b = _host(mdtfx.synth_i8(16, 1024, 32, seed=5, sigma=16.0))
    c = _host(mdtfx.synth_i8(16, 1024, 32, seed=6, sigma=16.0))
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    s = a.astype(np.float64)
    assert abs(s.mean()) < 0.2
    assert abs(s.std() - 16.0 / np.sqrt(2)) < 0.3  # sigma^2 split over I and Q
    # common component: correlation coefficient rho between inputs
    r = _host(mdtfx.synth_i8(2, 1024, 64, seed=9, sigma=20.0, rho=0.3)).astype(np.float64)
    z0 = r[:, 0, :, 0] + 1j * r[:, 0, :, 1]
    z1 = r[:, 1, :, 0] + 1j * r[:, 1, :, 1]
    rho = np.abs(np.mean(z0 * np.conj(z1))) / np.sqrt(np.mean(np.abs(z0) ** 2) * np.mean(np.abs(z1) ** 2))
    assert abs(rho - 0.3) < 0.02


def test_synth_beyond_32bit_grid():
    """the wide-band input (8195 frames x 512 inputs x 1024 channels = 4.3e9 samples,
    more than a 32-bit dispatch grid holds): every frame is generated, and
    frame f equals a one-frame synthesis started at t0 = f N"""
    NN, N, F = 512, 1024, 8195
    assert F * NN * N > 2 ** 32
    x = mdtfx.synth_i8(NN, N, F, seed=3, sigma=4.0, rho=0.3, max_delay=8, tones=((0.3, 12.0),))
    for f in (0, 4096, F - 1):
        one = mdtfx.synth_i8(NN, N, 1, seed=3, sigma=4.0, rho=0.3, max_delay=8, tones=((0.3, 12.0),), t0=f * N)
        assert torch.equal(x[f], one[0]), f
    rms = x[-64:].float().pow(2).mean().sqrt().item()
    assert abs(rms - x[:64].float().pow(2).mean().sqrt().item()) < 0.05 * rms
    del x
    torch.cuda.empty_cache()


@pytest.mark.parametrize("NN,N,T,Ns", [(40, 1024, 4, 256), (256, 4096, 8, 64), (130, 512, 8, 100), (3, 64, 4, 9),
                                       (300, 256, 4, 40),  # > 256 inputs: 128-input tiles, ragged
                                       (40, 1024, 16, 64), (64, 2048, 6, 48)])  # tap groups of four
def test_fp16_split_vs_fp64(NN, N, T, Ns):
    """FP16 hi+lo split X-engine within the 1e-5 autocorrelation-normalised bar
    (the 256-input scaled-array shape at reduced integration length)."""
    print(f"fp16split NN={NN} N={N} T={T} Ns={Ns}", flush=True)
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP16_SPLIT)
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=NN + N, sigma=20.0, rho=0.3, max_delay=8,
                        tones=((0.05, 30.0),))
    xin = plan.fengine(x8, plan.alloc_xinput())
    v16 = _host(plan.xengine(xin, torch.full((N, plan.nb), complex(np.nan, np.nan), dtype=torch.complex64, device="cuda")))
    assert np.all(np.isfinite(v16))
    plan.set_precision(mdtfx.XPREC_FP32)
    xin32 = plan.fengine(x8, plan.alloc_xinput())
    v32 = _host(plan.xengine(xin32, plan.alloc_vis()))
    W = _host(plan.window()).astype(np.float64)
    chans = np.unique(np.linspace(0, N - 1, min(N, 24)).astype(int))
    v64 = orc.fx_chain_f64(W, _host(x8), NN, N, T, Ns, chans)
    e16 = _tri_autonorm_err(v16[chans], v64, NN)
    e32 = _tri_autonorm_err(v32[chans], v64, NN)
    print(f"max normalised error vs FP64: fp16-split {e16:.3e}  fp32 {e32:.3e}")
    assert e16 < 1e-5 and e32 < 1e-5


def test_fp16_split_epilogue_format():
    """FP16-split X input: each sample is (hi_re, hi_im, lo_re, lo_im) of x*s,
    s an exact power of two, hi = fp16(x*s), lo = fp16(x*s - hi)."""
    NN, N, T, Ns = 5, 1024, 8, 16
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=12, sigma=20.0, tones=((0.2, 60.0),))
    p32 = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP32)
    p16 = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP16_SPLIT)
    x32 = np.ascontiguousarray(_host(p32.xinput_view(p32.fengine(x8, p32.alloc_xinput()))))
    raw = np.ascontiguousarray(_host(p16.xinput_view(p16.fengine(x8, p16.alloc_xinput()))))
    h = raw.view(np.float16).reshape(N, NN, Ns, 4).astype(np.float64)
    hi = h[..., 0] + 1j * h[..., 1]
    lo = h[..., 2] + 1j * h[..., 3]
    ref = x32.astype(np.complex128)
    big = np.abs(ref) > 1e-2 * np.abs(ref).max()
    s = np.median(np.abs(hi[big]) / np.abs(ref[big]))
    sc = 2.0 ** np.round(np.log2(s))
    assert abs(s / sc - 1) < 1e-3
    # hi is the FP16 rounding of x*s; lo the FP16 rounding of the remainder
    xs = ref * sc
    assert np.array_equal(hi.real, xs.real.astype(np.float16).astype(np.float64))
    assert np.array_equal(hi.imag, xs.imag.astype(np.float16).astype(np.float64))
    err = np.abs(hi + lo - xs)
    assert np.max(err[big] / np.abs(xs[big])) < 2.0 ** -20
    assert np.max(np.abs(xs)) < 32768.0  # headroom below the FP16 maximum


def test_front_end_refuses_undersized_buffers():
    """the kernels take raw pointers; the front end checks extents first"""
    NN, N, T, Ns = 4, 256, 4, 8
    plan = mdtfx.FxPlan(NN, N, T, Ns)
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=1)
    small = torch.empty(plan.xinput_bytes() // 8 - 1, dtype=torch.complex64, device="cuda")  # one sample short
    with pytest.raises(ValueError):
        plan.fengine(x8, small)
    plan.set_xinput_layout(mdtfx.XIN_ROWS)  # rows: padded, larger than N x NN x Ns
    with pytest.raises(ValueError):
        plan.fengine(x8, torch.empty((N, NN * Ns), dtype=torch.complex64, device="cuda"))
    with pytest.raises(ValueError):
        plan.fengine(x8[:-1], plan.alloc_xinput())
    with pytest.raises(ValueError):
        plan.xengine(plan.alloc_xinput(), torch.empty((N, plan.nb - 1), dtype=torch.complex64, device="cuda"))
