"""FP8 (OCP e4m3) X-engine with per-channel power-of-two RMS scaling: the wide-band
shape (512 inputs, 4-tap PFB, 1024 channels) at a reduced integration length,
and a ragged shape (partial tiles, partial K chunk), with a tone 40 dB above
the noise.

Three checks, each able to fail on its own:
* the channel exponents against numpy's from the FP16 X input;
* the GPU visibilities against an FP64 correlation of the same e4m3 samples
  (a numpy RNE e4m3 quantiser applied with the GPU's exponents): this pins
  the MFMA operand arrangement, the conjugation, the scaling and the
  quantiser.  What remains is the FP8 MFMA's own summation: measured on
  gfx950, one v_mfma_scale_f32_32x32x64_f8f6f4 sums its 64 exact products
  with ~2^-16 of sum|p| error (products far below the largest in their group
  of 8 are truncated), while chained MFMAs accumulate in FP32; hence a 2e-4
  bound where an operand mix-up gives O(1);
* the error against the FP64 oracle of the whole chain and against the FP32
  path: in noise channels inside e4m3's statistical bound (rounding errors
  average over the spectra); in the tone channel inside the coherent bound
  2 x 2^-4 + the statistical one, since a bin-centred tone gives every
  spectrum the same value and so the same rounding error.  Per-channel
  scaling puts both the tone channel and the channels 40 dB below it at
  full e4m3 precision.
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


def _run(NN, N, T, Ns, tone_bin, tone_amp, sigma, seed):
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP8)
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=seed, sigma=sigma, rho=0.3, max_delay=4,
                        tones=((tone_bin / N, tone_amp),))
    xin = plan.fengine(x8, plan.alloc_xinput())
    e = _host(plan.channel_exponents())
    nanvis = torch.full((N, plan.nb), complex(np.nan, np.nan), dtype=torch.complex64, device="cuda")
    v8 = _host(plan.xengine(xin, nanvis))
    xh = _host(plan.xinput_view(xin)).astype(np.float64)  # [N][NN][Ns][2] FP16 values of x*s
    return plan, x8, xin, e, v8, xh


CASES = [
    # NN, N, T, Ns, tone bin, tone amplitude, sigma
    (512, 1024, 4, 64, 300, 40.0, 4.0),   # wide-band shape, 64 of its 8192 spectra
    (96, 512, 8, 100, 77, 40.0, 4.0),     # partial 128-row tile, 100 = 3 x 32 + 4 spectra
]


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"NN{c[0]}_N{c[1]}_T{c[2]}_Ns{c[3]}")
def fp8case(request):
    NN, N, T, Ns, kt, amp, sigma = request.param
    print(f"fp8 NN={NN} N={N} T={T} Ns={Ns} tone bin {kt}", flush=True)
    plan, x8, xin, e, v8, xh = _run(NN, N, T, Ns, kt, amp, sigma, seed=NN + Ns)
    # channels for the FP64 checks: the tone bin and its neighbours, spread noise channels
    chans = np.unique(np.concatenate([[kt - 1, kt, kt + 1], np.linspace(0, N - 1, 6).astype(int)]))
    return dict(NN=NN, N=N, T=T, Ns=Ns, kt=kt, plan=plan, x8=x8, xin=xin, e=e, v8=v8, xh=xh, chans=chans)


def test_fp8_channel_exponents(fp8case):
    c = fp8case
    xh, e = c["xh"], c["e"]
    power = np.sum(xh[..., 0] ** 2 + xh[..., 1] ** 2, axis=(1, 2))
    t = np.log2(32.0 / np.sqrt(power / (c["NN"] * c["Ns"])))
    ref = np.clip(np.rint(t), -30, 30)
    tie = np.abs(np.abs(t - np.floor(t)) - 0.5) < 1e-3  # power summed in FP32 on the GPU
    assert np.all((e == ref) | tie), np.nonzero((e != ref) & ~tie)
    # the tone channel gets a smaller exponent than the noise channels
    assert e[c["kt"]] < np.median(e) - 2


def test_fp8_vs_fp64_of_quantised_samples(fp8case):
    c = fp8case
    NN, Ns, chans = c["NN"], c["Ns"], c["chans"]
    s = c["plan"].split_scale()
    a, b = np.tril_indices(NN)
    for ch in chans:
        ec = float(c["e"][ch])
        lim = np.float64(np.float16(min(E4M3_MAX * 2.0 ** -ec, 65504.0)))
        y = np.clip(c["xh"][ch], -lim, lim) * 2.0 ** ec
        q = e4m3_rne(y)
        assert np.max(np.abs(q)) <= E4M3_MAX
        z = q[..., 0] + 1j * q[..., 1]  # [NN][Ns]
        V = (z @ z.conj().T) / Ns / (s * 2.0 ** ec) ** 2
        vref = V[a, b][None, :]
        err = _autonorm(c["v8"][ch][None, :], vref, NN)[0]
        assert err < 2e-4, (ch, err)


def test_fp8_vs_fp64_oracle_and_fp32(fp8case):
    c = fp8case
    NN, N, T, Ns, chans, kt = c["NN"], c["N"], c["T"], c["Ns"], c["chans"], c["kt"]
    plan = c["plan"]
    W = _host(plan.window()).astype(np.float64)
    v64 = orc.fx_chain_f64(W, _host(c["x8"]), NN, N, T, Ns, chans)
    diag = np.array([orc.tri_index(i, i) for i in range(NN)])
    auto = np.real(v64[:, diag]).mean(axis=1)
    i_t = int(np.nonzero(chans == kt)[0][0])
    noise = np.array([i for i in range(len(chans)) if abs(int(chans[i]) - kt) > 1])
    dr_db = 10 * np.log10(auto[i_t] / np.median(auto[noise]))
    e8 = _autonorm(c["v8"][chans], v64, NN)
    p32 = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP32)
    v32 = _host(p32.dump(c["x8"], p32.alloc_xinput(), p32.alloc_vis()))
    e32 = _autonorm(v32[chans], v64, NN)
    e8_32 = _autonorm(c["v8"][chans], v32[chans].astype(np.complex128), NN)
    print(f"tone {dr_db:.1f} dB above noise; FP8 vs FP64: tone {e8[i_t]:.3e}, noise max "
          f"{e8[noise].max():.3e}; FP8 vs FP32 {e8_32.max():.3e}; FP32 vs FP64 {e32.max():.3e}")
    assert dr_db >= 40.0
    assert e32.max() < 1e-5
    # e4m3: relative rounding error <= 2^-4, RMS ~2^-4/sqrt(3); over Ns samples
    # the normalised error is ~0.04 sqrt(2/Ns) per baseline; bound at ~6 sigma
    bound = 0.35 / np.sqrt(Ns)
    assert e8[noise].max() < bound, (e8, bound)
    # the tone's own rounding error repeats in every spectrum: |dV|/|V| <= 2 x 2^-4
    assert e8[i_t] < 2.0 * 2.0 ** -4 + bound, (e8[i_t], bound)
    assert e8_32[noise].max() < bound


def test_fp8_acc64_and_determinism(fp8case):
    c = fp8case
    plan = c["plan"]
    acc = torch.zeros((c["N"], plan.nb), dtype=torch.complex128, device="cuda")
    v1 = _host(plan.xengine(c["xin"], plan.alloc_vis(), acc))
    v2 = _host(plan.xengine(c["xin"], plan.alloc_vis(), acc))
    assert np.array_equal(v1, c["v8"]) and np.array_equal(v2, c["v8"])
    a = _host(acc)
    assert np.array_equal(a, 2.0 * v1.astype(np.complex128))


def test_fp8_shape_limits():
    for N, T in ((256, 4), (8192, 4), (1024, 5)):
        with pytest.raises(mdtfx.MdtError):
            mdtfx.FxPlan(8, N, T, 16, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP8)
    plan = mdtfx.FxPlan(8, 256, 4, 16, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP32)
    with pytest.raises(mdtfx.MdtError):
        plan.set_precision(mdtfx.XPREC_FP8)
    with pytest.raises(mdtfx.MdtError):
        plan.channel_exponents()
