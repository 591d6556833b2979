"""FP8 (OCP e4m3) X-engine on v_mfma_scale_f32_32x32x64_f8f6f4, both sample
formats, with a tone 40 dB above the noise:

* block-scaled (512 and 1024 channels): the F-engine writes e4m3 (re, im) of
  x / 2^eb, eb one exponent per (input, 16 spectra) over all channels, and the
  exponents feed the MFMA's scale operands.  Pinned bit for bit: the
  exponents against the rule applied in numpy to the FP32 path's samples (the
  same FIR and FFT arithmetic), the e4m3 codes against the numpy quantiser
  (tests/e4m3.py, itself pinned to torch's float8_e4m3fn) of those samples;
* per-channel (2048 and 4096 channels): FP16 samples and channel powers from
  the F-engine, a power-of-two RMS scale per channel in the X-engine; the
  exponents against numpy's from the FP16 X input.

Then, for both: the visibilities against an FP64 correlation of the same
e4m3 samples (what remains is the FP8 MFMA's own summation: one
v_mfma_scale_f32_32x32x64_f8f6f4 sums its 64 exact products with ~2^-16 of
sum|p| error, measured on gfx950, chained MFMAs accumulate in FP32; hence a
2e-4 bound where an operand mix-up gives O(1)); against the FP64 oracle of the
whole chain and the FP32 path (noise channels inside e4m3's statistical
bound, the bin-centred tone inside the coherent bound 2 x 2^-4 plus it);
repeated full dumps bit-identical.
"""
import numpy as np
import pytest
import torch

import mdtfx
import oracle as orc
from e4m3 import E4M3_MAX, e4m3_rne  # pinned against torch's float8_e4m3fn (test_oracle.py)

pytestmark = pytest.mark.gpu


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _autonorm(v, v64, NN):
    a, b = np.tril_indices(NN)
    diag = np.array([orc.tri_index(i, i) for i in range(NN)])
    auto = np.real(v64[:, diag])
    return np.max(np.abs(v - v64) / np.sqrt(auto[:, a] * auto[:, b]), axis=1)


def block_exponent(maxabs):
    """numpy restatement of the block rule: the smallest integer eb with
    maxabs <= 448 2^eb, -127 for an all-zero block, within [-127, 127]"""
    m, E = np.frexp(np.asarray(maxabs, np.float32))
    eb = np.where(m <= 0.875, E - 9, E - 8)
    return np.clip(np.where(maxabs > 0, eb, -127), -127, 127).astype(np.int64)


def _decode(codes):
    """e4m3fn byte codes -> float64 (torch's decoding)"""
    c = torch.from_numpy(np.ascontiguousarray(codes))
    return c.view(torch.float8_e4m3fn).to(torch.float32).numpy().astype(np.float64)


def _x8(NN, N, T, Ns, tone_bin, tone_amp, sigma, seed, burst=None):
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=seed, sigma=sigma, rho=0.3, max_delay=4,
                        tones=((tone_bin / N, tone_amp),))
    if burst is not None:  # frames f0..f1 replaced by a strong burst (RFI-like)
        f0, f1, s2 = burst
        x8[f0:f1] = mdtfx.synth_i8(NN, N, f1 - f0, seed=seed + 1, sigma=s2, rho=0.3, max_delay=4)
    return x8


BLOCK_CASES = [
    # NN, N, T, Ns, tone bin, tone amplitude, sigma, burst
    (512, 1024, 4, 64, 300, 40.0, 4.0, None),    # wide-band shape, 64 of its 8192 spectra
    (96, 512, 8, 100, 77, 40.0, 4.0, None),      # partial 128-row tile, 100 = 6 x 16 + 4 spectra
    (64, 1024, 4, 96, 100, 1.0, 1.0, (40, 58, 60.0)),   # a burst ~60x the noise in 18 frames
    # a burst in frames 20-29: spectra 17-29, so the two 16-spectrum halves of
    # the first 32-spectrum MFMA step get different exponents
    (64, 1024, 4, 96, 100, 1.0, 1.0, (20, 30, 60.0)),
    # whole-channel kernel: 300 inputs (a ragged 32-row block, all four
    # workgroups of a channel), 70 spectra (the last 64-spectrum chunk half
    # past the dump); 200 inputs (one triangle workgroup), 48 spectra
    (300, 512, 4, 70, 100, 40.0, 4.0, None),
    (200, 1024, 4, 48, 500, 40.0, 4.0, None),
    (64, 1024, 12, 64, 321, 40.0, 4.0, None),   # 12 taps: the FIR in groups of four
]


@pytest.fixture(scope="module", params=BLOCK_CASES,
                ids=lambda c: f"NN{c[0]}_N{c[1]}_T{c[2]}_Ns{c[3]}" + (f"_burst{c[7][0]}" if c[7] else ""))
def blockcase(request):
    NN, N, T, Ns, kt, amp, sigma, burst = request.param
    print(f"fp8 block NN={NN} N={N} T={T} Ns={Ns} tone bin {kt} burst {burst}", flush=True)
    x8 = _x8(NN, N, T, Ns, kt, amp, sigma, seed=NN + Ns, burst=burst)
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP8)
    assert plan.f8block
    xin = plan.fengine(x8, plan.alloc_xinput())
    nanvis = torch.full((N, plan.nb), complex(np.nan, np.nan), dtype=torch.complex64, device="cuda")
    v8 = _host(plan.xengine(xin, nanvis))
    codes = _host(plan.xinput_view(xin))            # uint8 [N][NN][Ns][2]
    eb = _host(plan.xinput_scales(xin)).astype(np.int64) - 127  # [nblk][NN]
    p32 = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP32)
    xin32 = p32.fengine(x8, p32.alloc_xinput())
    x32 = _host(p32.xinput_view(xin32))             # complex64 [N][NN][Ns], the same FIR + FFT
    v32 = _host(p32.xengine(xin32, p32.alloc_vis()))
    chans = np.unique(np.concatenate([[kt - 1, kt, kt + 1], np.linspace(0, N - 1, 6).astype(int)]))
    return dict(NN=NN, N=N, T=T, Ns=Ns, kt=kt, plan=plan, x8=x8, xin=xin, v8=v8, codes=codes, eb=eb,
                x32=x32, v32=v32, chans=chans, burst=burst)


def test_block_exponents_and_codes_bit_exact(blockcase):
    c = blockcase
    NN, Ns, x32 = c["NN"], c["Ns"], c["x32"]
    nblk = -(-Ns // 16)
    m = np.maximum(np.abs(x32.real), np.abs(x32.imag))    # [N][NN][Ns] float32
    mx = np.zeros((nblk, NN), np.float32)
    for b in range(nblk):
        mx[b] = m[:, :, 16 * b: 16 * b + 16].max(axis=(0, 2))
    want_eb = block_exponent(mx)
    assert np.array_equal(c["eb"], want_eb), np.argwhere(c["eb"] != want_eb)[:5]
    sc = 2.0 ** want_eb[np.arange(Ns) // 16].T[None]       # [1][NN][Ns]
    q = np.stack([e4m3_rne(x32.real.astype(np.float64) / sc), e4m3_rne(x32.imag.astype(np.float64) / sc)], -1)
    assert np.max(np.abs(q)) <= E4M3_MAX
    got = _decode(c["codes"])
    bad = np.argwhere(got != q)
    assert bad.size == 0, (bad[:5], got[tuple(bad[0])], q[tuple(bad[0])])
    if c["burst"] is not None:  # the burst's blocks get their own, larger exponents
        assert want_eb.max() - np.median(want_eb) >= 2


def test_block_vs_fp64_of_quantised_samples(blockcase):
    c = blockcase
    NN, Ns, chans = c["NN"], c["Ns"], c["chans"]
    q = _decode(c["codes"][chans])                         # [ch][NN][Ns][2]
    sc = 2.0 ** c["eb"][np.arange(Ns) // 16].T[None]       # [1][NN][Ns]
    z = (q[..., 0] + 1j * q[..., 1]) * sc
    a, b = np.tril_indices(NN)
    for i, ch in enumerate(chans):
        V = (z[i] @ z[i].conj().T) / Ns
        err = _autonorm(c["v8"][ch][None, :], V[a, b][None, :], NN)[0]
        assert err < 2e-4, (ch, err)


def test_block_vs_fp64_oracle_and_fp32(blockcase):
    c = blockcase
    NN, N, T, Ns, chans, kt = c["NN"], c["N"], c["T"], c["Ns"], c["chans"], c["kt"]
    W = _host(c["plan"].window()).astype(np.float64)
    v64 = orc.fx_chain_f64(W, _host(c["x8"]), NN, N, T, Ns, chans)
    diag = np.array([orc.tri_index(i, i) for i in range(NN)])
    auto = np.real(v64[:, diag]).mean(axis=1)
    i_t = int(np.nonzero(chans == kt)[0][0])
    noise = np.array([i for i in range(len(chans)) if abs(int(chans[i]) - kt) > 1])
    dr_db = 10 * np.log10(auto[i_t] / np.median(auto[noise]))
    e8 = _autonorm(c["v8"][chans], v64, NN)
    e32 = _autonorm(c["v32"][chans], v64, NN)
    e8_32 = _autonorm(c["v8"][chans], c["v32"][chans].astype(np.complex128), NN)
    print(f"tone {dr_db:.1f} dB above noise; FP8 vs FP64: tone {e8[i_t]:.3e}, noise max "
          f"{e8[noise].max():.3e}; FP8 vs FP32 {e8_32.max():.3e}; FP32 vs FP64 {e32.max():.3e}")
    assert e32.max() < 1e-5
    # e4m3: relative rounding error <= 2^-4, RMS ~2^-4/sqrt(3), averaging
    # over the spectra: ~0.04 sqrt(2/Ns_eff) per baseline, bound at ~6 sigma;
    # channels below their block's largest sample keep e4m3's relative
    # precision as long as they stay normal (2^-6 of 448).  Ns_eff counts the
    # spectra that carry the power: a burst in 18 frames leaves about 18.
    p = (np.abs(c["x32"]) ** 2).astype(np.float64).sum(axis=(0, 1))
    ns_eff = p.sum() ** 2 / (p ** 2).sum()
    assert ns_eff <= Ns * (1 + 1e-9)
    if c["burst"] is None:
        assert ns_eff > 0.9 * Ns
    else:
        assert ns_eff < 0.4 * Ns
    bound = 0.35 / np.sqrt(ns_eff)
    assert e8[noise].max() < bound, (e8, bound)
    # the bin-centred tone's rounding error repeats in every spectrum
    assert e8[i_t] < 2.0 * 2.0 ** -4 + bound, (e8[i_t], bound)
    if c["burst"] is None:
        assert dr_db >= 30.0


def test_block_determinism_and_acc64(blockcase):
    """F- and X-engine again on the same samples: the same bits (no sums in
    any order), and the FP64 accumulator holds twice one dump"""
    c = blockcase
    plan = c["plan"]
    acc = torch.zeros((c["N"], plan.nb), dtype=torch.complex128, device="cuda")
    xin2 = plan.alloc_xinput()
    v1 = _host(plan.dump(c["x8"], xin2, plan.alloc_vis(), acc))
    v2 = _host(plan.dump(c["x8"], xin2, plan.alloc_vis(), acc))
    assert np.array_equal(_host(xin2), _host(c["xin"]))
    assert np.array_equal(v1, c["v8"]) and np.array_equal(v2, c["v8"])
    assert np.array_equal(_host(acc), 2.0 * v1.astype(np.complex128))


def test_block_format_ignores_layout_settings():
    """the block-scaled format has one layout (blocks of 32 spectra); a layout
    set on the plan applies when the precision changes"""
    plan = mdtfx.FxPlan(8, 512, 4, 40, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP8)
    B, bp, G, cp, rp = plan.xinput_strides()
    assert (B, G, rp, cp, bp) == (32, 1, 32, 8 * 32, 512 * 8 * 32)
    nbytes = plan.xinput_bytes()
    plan.set_xinput_layout(mdtfx.XIN_ROWS)
    assert plan.xinput_strides() == (B, bp, G, cp, rp) and plan.xinput_bytes() == nbytes
    assert nbytes == ((2 * 2 * bp + 255) // 256) * 256 + 3 * 8  # two blocks of 32 spectra, then 3 x 8 exponents
    plan.set_precision(mdtfx.XPREC_FP32)
    assert plan.xinput_strides()[1] == 0  # rows
    with pytest.raises(mdtfx.MdtError):
        mdtfx.FxPlan(8, 512, 4, 40, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP8).channel_exponents()


CHAN_CASES = [
    # NN, N, T, Ns, tone bin, tone amplitude, sigma
    (64, 2048, 4, 64, 600, 40.0, 4.0),
    (40, 4096, 8, 40, 1234, 40.0, 4.0),
    (40, 2048, 6, 40, 700, 40.0, 4.0),   # 6 taps, zero-padded to 8
]


@pytest.fixture(scope="module", params=CHAN_CASES, ids=lambda c: f"NN{c[0]}_N{c[1]}_T{c[2]}_Ns{c[3]}")
def chancase(request):
    NN, N, T, Ns, kt, amp, sigma = request.param
    print(f"fp8 per-channel NN={NN} N={N} T={T} Ns={Ns} tone bin {kt}", flush=True)
    x8 = _x8(NN, N, T, Ns, kt, amp, sigma, seed=NN + Ns)
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP8)
    assert not plan.f8block
    xin = plan.fengine(x8, plan.alloc_xinput())
    e = _host(plan.channel_exponents())
    nanvis = torch.full((N, plan.nb), complex(np.nan, np.nan), dtype=torch.complex64, device="cuda")
    v8 = _host(plan.xengine(xin, nanvis))
    xh = _host(plan.xinput_view(xin)).astype(np.float64)  # [N][NN][Ns][2] FP16 values of x*s
    chans = np.unique(np.concatenate([[kt - 1, kt, kt + 1], np.linspace(0, N - 1, 6).astype(int)]))
    return dict(NN=NN, N=N, T=T, Ns=Ns, kt=kt, plan=plan, x8=x8, xin=xin, e=e, v8=v8, xh=xh, chans=chans)


def test_channel_exponents(chancase):
    """the exponents from the F-engine's channel powers, now summed in a
    fixed order: numpy's from the FP16 X input, except where log2 sits at a
    rounding tie (the power is summed in FP32 on the device)"""
    c = chancase
    xh, e = c["xh"], c["e"]
    power = np.sum(xh[..., 0] ** 2 + xh[..., 1] ** 2, axis=(1, 2))
    t = np.log2(32.0 / np.sqrt(power / (c["NN"] * c["Ns"])))
    ref = np.clip(np.rint(t), -30, 30)
    tie = np.abs(np.abs(t - np.floor(t)) - 0.5) < 1e-3
    assert np.all((e == ref) | tie), np.nonzero((e != ref) & ~tie)
    assert e[c["kt"]] < np.median(e) - 2  # the tone channel gets a smaller exponent


def test_channel_vs_fp64_of_quantised_samples(chancase):
    c = chancase
    NN, Ns, chans = c["NN"], c["Ns"], c["chans"]
    s = c["plan"].split_scale()
    a, b = np.tril_indices(NN)
    for ch in chans:
        ec = float(c["e"][ch])
        lim = np.float64(np.float16(min(E4M3_MAX * 2.0 ** -ec, 65504.0)))
        y = np.clip(c["xh"][ch], -lim, lim) * 2.0 ** ec
        q = e4m3_rne(y)
        z = q[..., 0] + 1j * q[..., 1]
        V = (z @ z.conj().T) / Ns / (s * 2.0 ** ec) ** 2
        err = _autonorm(c["v8"][ch][None, :], V[a, b][None, :], NN)[0]
        assert err < 2e-4, (ch, err)


def test_channel_vs_fp64_oracle(chancase):
    c = chancase
    NN, N, T, Ns, chans, kt = c["NN"], c["N"], c["T"], c["Ns"], c["chans"], c["kt"]
    W = _host(c["plan"].window()).astype(np.float64)
    v64 = orc.fx_chain_f64(W, _host(c["x8"]), NN, N, T, Ns, chans)
    e8 = _autonorm(c["v8"][chans], v64, NN)
    i_t = int(np.nonzero(chans == kt)[0][0])
    noise = np.array([i for i in range(len(chans)) if abs(int(chans[i]) - kt) > 1])
    bound = 0.35 / np.sqrt(Ns)
    assert e8[noise].max() < bound, (e8, bound)
    assert e8[i_t] < 2.0 * 2.0 ** -4 + bound


def test_channel_determinism(chancase):
    """the whole dump again: the channel powers are summed in a fixed order,
    so the exponents and visibilities are the same bits"""
    c = chancase
    plan = c["plan"]
    v = _host(plan.dump(c["x8"], plan.alloc_xinput(), plan.alloc_vis()))
    assert np.array_equal(_host(plan.channel_exponents()), c["e"])
    assert np.array_equal(v, c["v8"])


def test_fp8_shape_limits():
    for N, T in ((256, 4), (8192, 4), (1024, 3)):
        with pytest.raises(mdtfx.MdtError):
            mdtfx.FxPlan(8, N, T, 16, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP8)
    mdtfx.FxPlan(8, 1024, 5, 16, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP8)  # any tap count in [4, 128]
    plan = mdtfx.FxPlan(8, 256, 4, 16, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP32)
    with pytest.raises(mdtfx.MdtError):
        plan.set_precision(mdtfx.XPREC_FP8)
    with pytest.raises(mdtfx.MdtError):
        plan.channel_exponents()
