"""GPU parity of the drop-in kernels against the CPU oracle, called through the
C ABI (libmdtfx.so).  Integer/permutation/FIR work is compared bit for bit;
FFT and X-engine results against FP64 with the tolerances written below."""
import numpy as np
import pytest
import torch

import mdtfx
import oracle as orc

pytestmark = pytest.mark.gpu


def _rng(seed):
    return np.random.default_rng(seed)


def _to_gpu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


# ----------------------------------------------------------------- a1 window
@pytest.mark.parametrize("N,T", [(32, 4), (1024, 8), (4096, 8), (2048, 4), (100, 7)])
@pytest.mark.parametrize("kind", [mdtfx.WINDOW_BH4, mdtfx.WINDOW_HAMMING])
def test_window_vs_oracle(N, T, kind):
    W = _host(mdtfx.pfb_window(N, T, kind))
    ref = orc.window_f32(N, T, kind)
    # FP64 evaluation rounded once: GPU and host libm may differ in the last
    # FP64 bit, which can move the FP32 rounding by at most one ulp.
    ulp = np.abs(W.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1
    assert np.mean(ulp == 0) > 0.999
    TN = N * T
    assert W.ravel()[TN // 2] == 1.0


def test_window_exblackman_name():
    """the client's PFB_Window_ExBlackman call (client.cu:589) is the BH4 window"""
    for N, T in ((1024, 8), (2048, 4)):
        a = _host(mdtfx.pfb_window_exblackman(N, T))
        b = _host(mdtfx.pfb_window(N, T, mdtfx.WINDOW_BH4))
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ----------------------------------------------------------------- a2 FIR
@pytest.mark.parametrize("N,T,NN,Ns", [(64, 4, 3, 5), (1024, 8, 5, 40), (96, 5, 4, 33), (256, 16, 3, 70),
                                       (2048, 4, 6, 64)])
def test_pps_batch_bit_exact(N, T, NN, Ns):
    print(f"pps N={N} T={T} NN={NN} Ns={Ns}", flush=True)
    rng = _rng(N + T)
    W = mdtfx.pfb_window(N, T)
    x = (rng.standard_normal((Ns + T - 1, NN, N)) * 30 + 1j * rng.standard_normal((Ns + T - 1, NN, N)) * 30).astype(np.complex64)
    y = _host(mdtfx.pps_batch(W, _to_gpu(x), NN, Ns))
    ref = orc.pps_f32(_host(W), x, NN, Ns)
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))


def test_pps_batch_i8_all_pairs():
    """every int8 (I, Q) pair: the int8 FIR (the F-engine's input stage) ==
    the float FIR (PpS_Batch, MDT_device_functions.cu:113-161) on the same
    values, bit for bit; 5 frames, each a different permutation of the 2^16
    pairs over 4 inputs x 16384 channels"""
    N, T, NN, Ns = 16384, 4, 4, 2
    rng = _rng(16)
    pairs = np.stack(np.meshgrid(np.arange(-128, 128), np.arange(-128, 128), indexing="ij"), -1).reshape(-1, 2)
    x8 = np.stack([pairs[rng.permutation(65536)] for _ in range(Ns + T - 1)]).astype(np.int8).reshape(Ns + T - 1, NN, N, 2)
    W = mdtfx.pfb_window(N, T, mdtfx.WINDOW_BH4)
    y8 = _host(mdtfx.pps_batch_i8(W, _to_gpu(x8), NN, Ns))
    xf = (x8[..., 0].astype(np.float32) + 1j * x8[..., 1].astype(np.float32)).astype(np.complex64)
    yf = _host(mdtfx.pps_batch(W, _to_gpu(xf), NN, Ns))
    assert np.array_equal(y8.view(np.uint32), yf.view(np.uint32))
    ref = orc.pps_i8_f32(_host(W), x8, NN, Ns)
    assert np.array_equal(y8.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("N,T,NN,Ns", [(64, 4, 3, 5), (1024, 8, 5, 40), (4096, 8, 3, 9), (128, 6, 3, 4)])
def test_pps_batch_i8_bit_exact(N, T, NN, Ns):
    print(f"pps_i8 N={N} T={T} NN={NN} Ns={Ns}", flush=True)
    rng = _rng(3 * N + T)
    W = mdtfx.pfb_window(N, T, mdtfx.WINDOW_HAMMING)
    x8 = rng.integers(-128, 128, size=(Ns + T - 1, NN, N, 2), dtype=np.int8)
    y = _host(mdtfx.pps_batch_i8(W, _to_gpu(x8), NN, Ns))
    ref = orc.pps_i8_f32(_host(W), x8, NN, Ns)
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))


# ----------------------------------------------------------------- a3 FFT
def _fft_err(X, ref):
    """relative L2 error per transform row, max over rows"""
    num = np.linalg.norm(X - ref, axis=-1)
    den = np.linalg.norm(ref, axis=-1)
    return float(np.max(num / den))


@pytest.mark.parametrize("N", [8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384])
def test_fft_vs_fp64(N):
    print(f"fft N={N}", flush=True)
    rng = _rng(N)
    batch = max(3, min(300, (1 << 20) // N))
    x = (rng.standard_normal((batch, N)) + 1j * rng.standard_normal((batch, N))).astype(np.complex64)
    plan = mdtfx.FftPlan(N, batch)
    plan.set_stream(torch.cuda.current_stream().cuda_stream)
    xi = _to_gpu(x)
    out = torch.empty_like(xi)
    X = _host(plan.exec_c2c(xi, out, mdtfx.FFT_FORWARD))
    ref = orc.fft_f64(x.astype(np.complex128), -1)
    err = _fft_err(X, ref)
    # FP32 Stockham with FP64-rounded twiddles: error grows ~ eps * log2(N)
    assert err < 2.5e-7 * np.log2(N), err
    # inverse, in place
    Xi = _host(plan.exec_c2c(xi, xi, mdtfx.FFT_INVERSE))
    refi = orc.fft_f64(x.astype(np.complex128), +1)
    assert _fft_err(Xi, refi) < 2.5e-7 * np.log2(N)
    plan.destroy()


@pytest.mark.parametrize("N,fpb", [(256, 8), (512, 4), (1024, 4), (2048, 2), (4096, 1), (8192, 1)])
def test_fft_streaming_many_rows(N, fpb):
    """Batches large enough that each persistent workgroup runs several groups
    of rows, with a partial last group (batch = 3 * 1024 groups * fpb + 1):
    sampled rows against the FP64 oracle, forward out of place and inverse in
    place, and the same input row transformed bit-identically in the first
    group and in the partial last one."""
    batch = 3 * 1024 * fpb + 1
    print(f"fft stream N={N} batch={batch}", flush=True)
    rng = _rng(N + 7)
    x = (rng.standard_normal((batch, N)) + 1j * rng.standard_normal((batch, N))).astype(np.complex64)
    x[-1] = x[0]
    plan = mdtfx.FftPlan(N, batch)
    plan.set_stream(torch.cuda.current_stream().cuda_stream)
    xi = _to_gpu(x)
    X = _host(plan.exec_c2c(xi, torch.empty_like(xi), mdtfx.FFT_FORWARD))
    rows = np.unique(np.concatenate([[0, 1, batch - 2, batch - 1], rng.integers(0, batch, 60)]))
    ref = orc.fft_f64(x[rows].astype(np.complex128), -1)
    assert _fft_err(X[rows], ref) < 2.5e-7 * np.log2(N)
    assert np.array_equal(X[0].view(np.uint32), X[-1].view(np.uint32))
    Xi = _host(plan.exec_c2c(xi, xi, mdtfx.FFT_INVERSE))
    refi = orc.fft_f64(x[rows].astype(np.complex128), +1)
    assert _fft_err(Xi[rows], refi) < 2.5e-7 * np.log2(N)
    assert np.array_equal(Xi[0].view(np.uint32), Xi[-1].view(np.uint32))
    plan.destroy()


def test_fft_numpy_golden(golden):
    for N in [8, 32, 1024, 2048, 4096]:
        x = golden[f"fft_in_N{N}"].astype(np.complex64)
        plan = mdtfx.FftPlan(N, 1)
        xi = _to_gpu(x)
        X = _host(plan.exec_c2c(xi, torch.empty_like(xi)))
        assert _fft_err(X, golden[f"fft_out_N{N}"]) < 2.5e-7 * np.log2(N)


def test_fft_tone_single_bin():
    N, k0 = 4096, 1234
    n = np.arange(N)
    x = np.exp(2j * np.pi * k0 * n / N).astype(np.complex64)[None]
    plan = mdtfx.FftPlan(N, 1)
    X = _host(plan.exec_c2c(_to_gpu(x), torch.empty((1, N), dtype=torch.complex64, device="cuda")))[0]
    assert abs(X[k0] - N) < 1e-3 * N
    X[k0] = 0
    assert np.max(np.abs(X)) < 1e-3 * np.sqrt(N)


# ----------------------------------------------------------------- a4 corner turn
@pytest.mark.parametrize("NN,N,Ns", [(3, 32, 2), (38, 1024, 7), (40, 2048, 3), (70, 100, 5), (40, 100, 4), (64, 96, 3),
                                      (128, 64, 5), (40, 101, 3), (130, 64, 3), (4, 16384, 300)])
def test_reorder_bit_exact(NN, N, Ns):
    print(f"reorder NN={NN} N={N} Ns={Ns}", flush=True)
    rng = _rng(NN * N)
    bits = rng.integers(0, 2**63, size=(Ns, NN, N), dtype=np.int64)
    bits[0, 0, 0] = np.int64(0x7FC000017FC00001)  # NaN payloads
    bits[-1, -1, -1] = np.int64(-0x7FFFFFFF80000000)  # -0 halves
    F = _to_gpu(bits.view(np.complex64))
    X = _host(mdtfx.reorder_foutput(F, NN, N, Ns))
    ref = orc.reorder(bits, NN, N, Ns)
    assert np.array_equal(X.view(np.int64), ref)


# ----------------------------------------------------------------- a5 X-engine
def _autonorm_err(V, V64):
    """max |V - V64| / sqrt(V64_aa V64_bb) over a full [c][b][a] matrix set"""
    d = np.real(np.einsum("caa->ca", V64))
    norm = np.sqrt(d[:, :, None] * d[:, None, :])  # [c][b][a] symmetric
    return float(np.max(np.abs(V - V64) / norm))


@pytest.mark.parametrize("NN,Ns,N", [(3, 2, 32), (38, 64, 32), (40, 256, 32), (64, 33, 32), (100, 129, 32), (130, 40, 32),
                                     (40, 100, 32), (48, 33, 32), (4, 70, 32),  # partial chunks, NN 48 / 4
                                     (50, 70, 32), (58, 64, 32), (18, 40, 32)])  # three waves (NN 49-64), two blocks
def test_herk_vs_fp64(NN, Ns, N):
    print(f"herk NN={NN} Ns={Ns} N={N}", flush=True)
    rng = _rng(NN + Ns)
    x = (rng.standard_normal((N, Ns, NN)) * 100 + 1j * rng.standard_normal((N, Ns, NN)) * 100).astype(np.complex64)
    xe = mdtfx.XEngine()
    out = torch.full((N, NN, NN), complex(np.nan, np.nan), dtype=torch.complex64, device="cuda")
    V = _host(xe.herk(NN, Ns, 1.0 / Ns, _to_gpu(x), out, N))
    V64 = orc.xengine_full(x, NN, Ns, N)
    assert np.all(np.isfinite(V)), "unwritten output elements"
    assert _autonorm_err(V, V64) < 1e-5


@pytest.mark.parametrize("NN,Ns,N", [(38, 1024, 4), (40, 64, 6), (5, 3, 2), (64, 96, 3), (52, 40, 2)])
def test_cgemm_batched_reference_call(NN, Ns, N):
    """The exact reference call: cublasCgemmBatched(h, OP_N, OP_C, NN, NN, Ns,
    1/Ns, XInPtr, NN, XInPtr, NN, 0, XOutPtr, NN, N) (client.cu:656-660, 1128)."""
    print(f"cgemm_ref NN={NN} Ns={Ns} N={N}", flush=True)
    rng = _rng(7 * NN + N)
    x = (rng.standard_normal((N, Ns, NN)) + 1j * rng.standard_normal((N, Ns, NN))).astype(np.complex64)
    xin = _to_gpu(x)
    xout = torch.full((N, NN, NN), complex(np.nan, np.nan), dtype=torch.complex64, device="cuda")
    Aptr = torch.tensor([xin.data_ptr() + c * NN * Ns * 8 for c in range(N)], dtype=torch.int64, device="cuda")
    Cptr = torch.tensor([xout.data_ptr() + c * NN * NN * 8 for c in range(N)], dtype=torch.int64, device="cuda")
    xe = mdtfx.XEngine()
    xe.set_stream(torch.cuda.current_stream().cuda_stream)
    xe.cgemm_batched(mdtfx.OP_N, mdtfx.OP_C, NN, NN, Ns, 1.0 / Ns, Aptr, NN, Aptr, NN, 0.0, Cptr, NN, N)
    V = _host(xout)
    V64 = orc.xengine_full(x, NN, Ns, N)
    assert np.all(np.isfinite(V))
    assert _autonorm_err(V, V64) < 1e-5


@pytest.mark.parametrize("offset", [0, 1])
def test_cgemm_batched_hermitian_small(offset):
    """The Hermitian call (one pointer table for A and B) at a small even NN
    with complex alpha and beta, a padded ldc, a partial last chunk of
    samples, and (offset 1) table entries one element past 16-byte alignment."""
    rng = _rng(13 + offset)
    m, k, ldc, batch = 40, 100, 44, 3
    A = (rng.standard_normal((batch, k * m + 1)) + 1j * rng.standard_normal((batch, k * m + 1))).astype(np.complex64)
    C0 = (rng.standard_normal((batch, m * ldc)) + 1j * rng.standard_normal((batch, m * ldc))).astype(np.complex64)
    al, be = 0.5 - 0.25j, -1.0 + 2.0j
    Ad, Cd = _to_gpu(A), _to_gpu(C0)
    Aptr = torch.tensor([Ad.data_ptr() + (i * (k * m + 1) + offset) * 8 for i in range(batch)], dtype=torch.int64,
                        device="cuda")
    Cptr = torch.tensor([Cd.data_ptr() + i * m * ldc * 8 for i in range(batch)], dtype=torch.int64, device="cuda")
    xe = mdtfx.XEngine()
    xe.cgemm_batched(mdtfx.OP_N, mdtfx.OP_C, m, m, k, al, Aptr, m, Aptr, m, be, Cptr, ldc, batch)
    C = _host(Cd)
    for i in range(batch):
        Ai = np.ascontiguousarray(A[i][offset:offset + k * m])
        ref = orc.cgemm_nc(m, m, k, al, Ai, m, Ai, m, be, C0[i], ldc)
        got = C[i].reshape(m, ldc).T[:m]
        exp = ref.reshape(m, ldc).T[:m]
        scale = np.sqrt(k) * 2 + np.abs(C0[i]).max() * 3
        assert np.max(np.abs(got - exp)) < 1e-5 * scale
        assert np.array_equal(C[i].reshape(m, ldc).T[m:], C0[i].reshape(m, ldc).T[m:])  # padding untouched


def test_cgemm_batched_general():
    """A != B pointer tables, complex alpha/beta, m != n, padded leading dims."""
    rng = _rng(11)
    m, n, k, lda, ldb, ldc, batch = 70, 45, 37, 72, 50, 75, 3
    A = (rng.standard_normal((batch, k * lda)) + 1j * rng.standard_normal((batch, k * lda))).astype(np.complex64)
    B = (rng.standard_normal((batch, k * ldb)) + 1j * rng.standard_normal((batch, k * ldb))).astype(np.complex64)
    C0 = (rng.standard_normal((batch, n * ldc)) + 1j * rng.standard_normal((batch, n * ldc))).astype(np.complex64)
    al, be = 0.5 - 0.25j, -1.0 + 2.0j
    Ad, Bd, Cd = _to_gpu(A), _to_gpu(B), _to_gpu(C0)
    tab = lambda t, w: torch.tensor([t.data_ptr() + i * w * 8 for i in range(batch)], dtype=torch.int64, device="cuda")
    xe = mdtfx.XEngine()
    xe.cgemm_batched(mdtfx.OP_N, mdtfx.OP_C, m, n, k, al, tab(Ad, k * lda), lda, tab(Bd, k * ldb), ldb, be,
                     tab(Cd, n * ldc), ldc, batch)
    C = _host(Cd)
    for i in range(batch):
        ref = orc.cgemm_nc(m, n, k, al, A[i], lda, B[i], ldb, be, C0[i], ldc)
        got = C[i].reshape(n, ldc).T[:m]
        exp = ref.reshape(n, ldc).T[:m]
        scale = np.sqrt(k) * 2 + np.abs(C0[i]).max() * 3
        assert np.max(np.abs(got - exp)) < 1e-5 * scale
        # padding rows untouched
        assert np.array_equal(C[i].reshape(n, ldc).T[m:], C0[i].reshape(n, ldc).T[m:])


# ----------------------------------------------------------------- a6 history
def test_history_shift():
    NN, N, T, Ns = 5, 64, 8, 10
    rng = _rng(5)
    x = (rng.standard_normal((Ns + T - 1, NN, N)) + 1j * rng.standard_normal((Ns + T - 1, NN, N))).astype(np.complex64)
    d = _to_gpu(x)
    mdtfx.history_shift(d, NN, N, T, Ns)
    y = _host(d)
    assert np.array_equal(y[:T - 1], x[Ns:Ns + T - 1])
    assert np.array_equal(y[T - 1:], x[T - 1:])
    # overlapping case Ns < T-1, int8 elements
    Ns2 = 3
    x8 = rng.integers(-128, 128, size=(Ns2 + T - 1, NN, N, 2), dtype=np.int8)
    d8 = _to_gpu(x8)
    mdtfx.history_shift(d8, NN, N, T, Ns2)
    assert np.array_equal(_host(d8)[:T - 1], x8[Ns2:Ns2 + T - 1])


# ----------------------------------------------------------------- a7 unpack
def test_unpack_u32_exhaustive():
    """All 2^32 words, GPU vs oracle (client.cu:898-907), bit for bit."""
    chunk = 1 << 27
    for start in range(0, 1 << 32, chunk):
        print(f"unpack chunk {start:#x}", flush=True)
        w = torch.arange(start, start + chunk, dtype=torch.int64, device="cuda").to(torch.int32)
        f = _host(mdtfx.unpack_u32(w))
        ref = orc.unpack_u32(np.arange(start, start + chunk, dtype=np.uint64).astype(np.uint32))
        assert np.array_equal(f.view(np.uint32), ref.view(np.uint32)), hex(start)
    w = torch.tensor([-2**31], dtype=torch.int32, device="cuda")
    assert _host(mdtfx.unpack_u32(w))[0] == 4096.0


# ----------------------------------------------------------------- HBM probe (measurement aid)
def test_hbm_probe_copy_write_read():
    """the probe kernels behind bench.py's measured HBM ceilings move the bytes
    they are said to move: copy is exact, write stores its pattern, read and
    bad arguments behave"""
    print("hbm probe", flush=True)
    n = (1 << 20) + 48
    src = torch.randint(0, 256, (n,), dtype=torch.uint8, device="cuda")
    dst = torch.zeros_like(src)
    mdtfx.hbm_probe(mdtfx.PROBE_COPY, src, dst, 7)
    torch.cuda.synchronize()
    assert torch.equal(src, dst)
    mdtfx.hbm_probe(mdtfx.PROBE_WRITE, None, dst, 5)
    torch.cuda.synchronize()
    w = dst.view(torch.int32).view(-1, 4).cpu().numpy()
    assert np.array_equal(w[:, 0], np.arange(w.shape[0], dtype=np.int32)) and not w[:, 1:].any()
    mdtfx.hbm_probe(mdtfx.PROBE_READ, src, None, 3)
    torch.cuda.synchronize()
    with pytest.raises(mdtfx.MdtError):
        mdtfx.hbm_probe(mdtfx.PROBE_COPY, src[1:17], dst[:16], 1)  # misaligned source
    with pytest.raises(mdtfx.MdtError):
        mdtfx.hbm_probe(7, src, dst, 1)

