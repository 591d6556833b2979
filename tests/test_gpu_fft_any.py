"""The drop-in FFT at lengths that are not powers of two: cufftPlan1d takes
any Nchannels the client accepts, 32..16384 (MDT_OneStation_client.h:56-57,
MDT_OneStation_client.cu:286, 622).  7-smooth lengths run mixed-radix
Stockham passes (radix 8/4/2, 3, 5, 7); other lengths Bluestein's chirp-z
convolution (power-of-two length M >= 2N-1; N > 8192 uses the split M = 32768
form).  Checked against the FP64 oracle (a brute-force DFT at these lengths,
itself pinned to numpy.fft golden vectors) and the numpy golden vectors
directly, forward and inverse, in place and out of place, with partial
workgroups of transforms; then the reference's four-call chain at 1000 and
1009 channels against the FP64 chain oracle."""
import numpy as np
import pytest
import torch

import mdtfx
import oracle as orc

pytestmark = pytest.mark.gpu


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _fft_err(X, ref):
    """max over rows of the relative L2 error"""
    return float(np.max(np.linalg.norm(X - ref, axis=-1) / np.linalg.norm(ref, axis=-1)))


# mixed radix: 35 = 5*7, 48, 63, 96, 100, 112, 125 = 5^3, 135 (35-135: the
# compiled short kernels, except 100 and 112), 243 = 3^5, 343 = 7^3, 1000,
# 1536, 3000, 6561 = 3^8, 9000, 11025 = 3^2 5^2 7^2, 16000;  Bluestein: 33 =
# 3*11 (M = 128), 4099 (prime, M = 16384), 8191 (prime, M = 16384), 12289
# (prime, split), 16383 = 3*43*127 (split)
LENGTHS = [33, 35, 48, 63, 96, 100, 112, 125, 135, 243, 343, 1000, 1536, 3000, 4099, 6561, 8191, 9000, 11025,
           12289, 16000, 16383]


@pytest.mark.parametrize("N", LENGTHS)
def test_fft_any_length_vs_fp64(N):
    print(f"fft any N={N}", flush=True)
    rng = np.random.default_rng(N)
    batch = 7 if N <= 4096 else 2
    x = (rng.standard_normal((batch, N)) + 1j * rng.standard_normal((batch, N))).astype(np.complex64)
    plan = mdtfx.FftPlan(N, batch)
    xi = torch.from_numpy(x).cuda()
    X = _host(plan.exec_c2c(xi, torch.empty_like(xi), mdtfx.FFT_FORWARD))
    ref = orc.fft_f64(x.astype(np.complex128), -1)
    err = _fft_err(X, ref)
    bound = 2.5e-7 * np.log2(N)
    print(f"  forward relative L2 error {err:.3e} (bound {bound:.3e})", flush=True)
    assert err < bound, err
    Xi = _host(plan.exec_c2c(xi, xi, mdtfx.FFT_INVERSE))  # in place
    erri = _fft_err(Xi, orc.fft_f64(x.astype(np.complex128), +1))
    assert erri < bound, erri
    plan.destroy()


@pytest.mark.parametrize("N,fpb", [(144, 14), (189, 10), (243, 8), (1000, 2), (1536, 1), (3000, 1), (6561, 1),
                                   (7000, 1), (8000, 1), (257, 4), (509, 4), (1009, 2), (2039, 1)])
def test_fft_any_streaming_many_rows(N, fpb):
    """Mixed-radix lengths of 144..8192 points and Bluestein lengths of
    257..2048 points (convolutions of 512..4096) run in persistent workgroups
    that load the next group of fpb rows while transforming the current one:
    batches of 3 * 1024 groups + 1 give every workgroup several groups and a
    partial last group.  Sampled rows against the FP64 oracle, forward out of
    place and inverse in place, and the same input row transformed
    bit-identically in the first group and in the partial last one (mixed
    radix: radix 8 first with the first pass from registers; 243 is not,
    6561 = 3^8 has radix 3 only; Bluestein: primes 257 (M = 512), 509
    (1024), 1009 (2048), 2039 (4096))."""
    batch = 3 * 1024 * fpb + 1
    print(f"fft any stream N={N} batch={batch}", flush=True)
    rng = np.random.default_rng(N + 11)
    x = (rng.standard_normal((batch, N)) + 1j * rng.standard_normal((batch, N))).astype(np.complex64)
    x[-1] = x[0]
    plan = mdtfx.FftPlan(N, batch)
    plan.set_stream(torch.cuda.current_stream().cuda_stream)
    xi = torch.from_numpy(x).cuda()
    X = _host(plan.exec_c2c(xi, torch.empty_like(xi), mdtfx.FFT_FORWARD))
    rows = np.unique(np.concatenate([[0, 1, batch - 2, batch - 1], rng.integers(0, batch, 20)]))
    bound = 2.5e-7 * np.log2(N)
    err = _fft_err(X[rows], orc.fft_f64(x[rows].astype(np.complex128), -1))
    print(f"  forward relative L2 error {err:.3e} (bound {bound:.3e})", flush=True)
    assert err < bound, err
    assert np.array_equal(X[0].view(np.uint32), X[-1].view(np.uint32))
    Xi = _host(plan.exec_c2c(xi, xi, mdtfx.FFT_INVERSE))
    erri = _fft_err(Xi[rows], orc.fft_f64(x[rows].astype(np.complex128), +1))
    assert erri < bound, erri
    assert np.array_equal(Xi[0].view(np.uint32), Xi[-1].view(np.uint32))
    plan.destroy()


def test_fft_any_plans_keep_their_lds():
    """plans with large LDS needs (8000 points streamed: 132 KiB; 16000 one-shot:
    136 KiB) still run after smaller plans of the same kernels are created"""
    rng = np.random.default_rng(5)
    big = {N: mdtfx.FftPlan(N, 3) for N in (8000, 16000)}
    small = [mdtfx.FftPlan(N, 3) for N in (1000, 10125)]
    for N, plan in big.items():
        x = (rng.standard_normal((3, N)) + 1j * rng.standard_normal((3, N))).astype(np.complex64)
        xi = torch.from_numpy(x).cuda()
        X = _host(plan.exec_c2c(xi, torch.empty_like(xi)))
        assert _fft_err(X, orc.fft_f64(x.astype(np.complex128), -1)) < 2.5e-7 * np.log2(N), N
        plan.destroy()
    for plan in small:
        plan.destroy()


@pytest.mark.parametrize("N", [1000, 1536, 3000, 4099])
def test_fft_any_numpy_golden(golden, N):
    x = golden[f"fft_in_N{N}"].astype(np.complex64)
    plan = mdtfx.FftPlan(N, 1)
    xi = torch.from_numpy(x).cuda()
    X = _host(plan.exec_c2c(xi, torch.empty_like(xi)))
    assert _fft_err(X, golden[f"fft_out_N{N}"]) < 2.5e-7 * np.log2(N)
    Xi = _host(plan.exec_c2c(xi, torch.empty_like(xi), mdtfx.FFT_INVERSE))
    assert _fft_err(Xi, golden[f"ifft_out_N{N}"]) < 2.5e-7 * np.log2(N)


@pytest.mark.parametrize("N,k0", [(1000, 123), (4099, 2000), (12289, 7)])
def test_fft_any_tone_and_impulse(N, k0):
    """an integer-bin tone lands in one bin, an impulse gives a flat spectrum"""
    n = np.arange(N)
    x = np.stack([np.exp(2j * np.pi * k0 * n / N), (n == 0).astype(complex)]).astype(np.complex64)
    plan = mdtfx.FftPlan(N, 2)
    X = _host(plan.exec_c2c(torch.from_numpy(x).cuda(), torch.empty((2, N), dtype=torch.complex64, device="cuda")))
    assert abs(X[0, k0] - N) < 1e-4 * N
    X[0, k0] = 0
    assert np.max(np.abs(X[0])) < 1e-4 * np.sqrt(N) * np.log2(N)
    assert np.max(np.abs(X[1] - 1)) < 1e-5


def test_fft_length_limits():
    for N in (12, 31, 16385, 20000):
        with pytest.raises(mdtfx.MdtError):
            mdtfx.FftPlan(N, 1)
    # powers of two below the client's minimum stay available
    mdtfx.FftPlan(8, 1).destroy()
    # the fused plan keeps its power-of-two transforms
    with pytest.raises(mdtfx.MdtError):
        mdtfx.FxPlan(16, 1000, 4, 16)


@pytest.mark.parametrize("N", [1000, 1009])
def test_dropin_chain_non_power_of_two(N):
    """the client's rank-2 sequence (PpS_Batch -> cufftExecC2C ->
    ReorderFOutput -> cublasCgemmBatched, MDT_OneStation_client.cu:1083-1128)
    at 38 inputs (the shipped configuration's 19 elements x 2 pols) and N =
    1000 (mixed radix) or 1009 (prime: Bluestein), every channel against the
    FP64 chain oracle"""
    NN, T, Ns = 38, 4, 48
    print(f"drop-in chain NN={NN} N={N} T={T} Ns={Ns}", flush=True)
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=N, sigma=16.0, rho=0.3, max_delay=4, tones=((0.2, 20.0),))
    W = mdtfx.pfb_window(N, T, mdtfx.WINDOW_BH4)
    f_in = torch.view_as_complex(x8.float().contiguous())
    F = mdtfx.pps_batch(W, f_in, NN, Ns)
    fft = mdtfx.FftPlan(N, Ns * NN)
    fft.exec_c2c(F, F, mdtfx.FFT_FORWARD)
    XIn = mdtfx.reorder_foutput(F, NN, N, Ns)
    XOut = torch.empty((N, NN, NN), dtype=torch.complex64, device="cuda")
    Aptr = torch.tensor([XIn.data_ptr() + c * NN * Ns * 8 for c in range(N)], dtype=torch.int64, device="cuda")
    Cptr = torch.tensor([XOut.data_ptr() + c * NN * NN * 8 for c in range(N)], dtype=torch.int64, device="cuda")
    xe = mdtfx.XEngine()
    xe.cgemm_batched(mdtfx.OP_N, mdtfx.OP_C, NN, NN, Ns, 1.0 / Ns, Aptr, NN, Aptr, NN, 0.0, Cptr, NN, N)
    a, b = np.tril_indices(NN)
    V = _host(XOut)[:, b, a]  # column-major [c][col b][row a] -> triangle (a >= b)
    v64 = orc.fx_chain_f64(_host(W).astype(np.float64), _host(x8), NN, N, T, Ns, np.arange(N))
    diag = np.array([orc.tri_index(i, i) for i in range(NN)])
    auto = np.real(v64[:, diag])
    err = float(np.max(np.abs(V - v64) / np.sqrt(auto[:, a] * auto[:, b])))
    print(f"  drop-in chain vs FP64: {err:.3e}", flush=True)
    assert err < 1e-5, err
