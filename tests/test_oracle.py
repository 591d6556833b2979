"""Pins the CPU oracle before it is trusted as the checker of the HIP path.

Each check is chosen so that a plausible slip in the oracle (dropped term,
wrong sign, swapped index, transposed operand) fails at least one of them:
library routines (numpy.fft, scipy windows, numpy.einsum), closed forms,
invariants, and brute force on tiny inputs.  The reference ships no tests or
golden vectors of its own, so these are the pins.
"""
import numpy as np
import pytest

import oracle as orc


# ----------------------------------------------------------------- window
@pytest.mark.parametrize("N,T", [(64, 4), (1024, 8), (2048, 4)])
@pytest.mark.parametrize("kind,name", [(orc.WINDOW_BH4, "bh4"), (orc.WINDOW_HAMMING, "hamming")])
def test_window_matches_scipy(golden, N, T, kind, name):
    W = orc.window_f64(N, T, kind)
    ref = golden[f"window_{name}_N{N}_T{T}"]
    assert np.max(np.abs(W - ref)) < 1e-15


@pytest.mark.parametrize("N,T", [(32, 4), (1024, 8), (4096, 8), (100, 6)])
def test_window_closed_form_values(N, T):
    W = orc.window_f64(N, T, orc.WINDOW_BH4).ravel()
    TN = N * T
    # centre tap: sinc(0) = 1 and w(TN/2) = a0 + a1 + a2 + a3 = 1
    assert abs(W[TN // 2] - 1.0) < 1e-15
    # sinc zeros at integer arguments (devfn.cu:70): exactly zero
    for i in range(T):
        if i * N != TN // 2:
            assert W[i * N] == 0.0
    # symmetry of the periodic window about TN/2
    i = np.arange(1, TN)
    assert np.max(np.abs(W[i] - W[TN - i])) < 1e-15
    W32 = orc.window_f32(N, T, orc.WINDOW_BH4).ravel()
    assert np.array_equal(W32, W.astype(np.float32))


def test_sinpi_cospi_exact_points():
    for n in range(-6, 7):
        assert orc.sinpi(n) == 0.0
        assert orc.cospi(n + 0.5) == 0.0
        assert orc.cospi(n) == (1.0 if n % 2 == 0 else -1.0)
    xs = np.linspace(-3.3, 3.3, 101)
    assert max(abs(orc.sinpi(x) - np.sin(np.pi * x)) for x in xs) < 1e-14
    assert max(abs(orc.cospi(x) - np.cos(np.pi * x)) for x in xs) < 1e-14


# ----------------------------------------------------------------- PpS_Batch
def _rand_c64(rng, shape, scale=10.0):
    return (rng.standard_normal(shape) * scale + 1j * rng.standard_normal(shape) * scale).astype(np.complex64)


def test_pps_impulse_and_constant():
    N, T, NN, Ns = 32, 4, 3, 5
    W = orc.window_f32(N, T)
    x = np.zeros((Ns + T - 1, NN, N), np.complex64)
    # impulse in frame f=2 of input 1: out[s][1] = W[2 - s] for s <= 2 (tap i = f - s)
    x[2, 1, :] = 1.0
    y = orc.pps_f32(W, x, NN, Ns)
    for s in range(Ns):
        i = 2 - s
        exp = W[i] if 0 <= i < T else np.zeros(N, np.float32)
        assert np.array_equal(y[s, 1].real, exp)
        assert np.all(y[s, 1].imag == 0)
        assert np.all(y[s, 0] == 0) and np.all(y[s, 2] == 0)
    # constant input j: out = j * sum_i W[i] (order of fmaf sums)
    x[:] = 1j
    y = orc.pps_f32(W, x, NN, Ns)
    ref = np.zeros(N, np.float32)
    for i in range(T):
        ref = (W[i].astype(np.float64) + ref).astype(np.float32)
    assert np.allclose(y.imag, ref[None, None, :], rtol=0, atol=1e-6)
    assert np.all(y.real == 0)


def test_pps_f32_vs_f64_and_i8(golden):
    rng = np.random.default_rng(1)
    N, T, NN, Ns = 64, 4, 3, 6
    W32 = orc.window_f32(N, T)
    x8 = rng.integers(-128, 128, size=(Ns + T - 1, NN, N, 2), dtype=np.int8)
    xc = (x8[..., 0].astype(np.float32) + 1j * x8[..., 1].astype(np.float32)).astype(np.complex64)
    a = orc.pps_f32(W32, xc, NN, Ns)
    b = orc.pps_i8_f32(W32, x8, NN, Ns)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    c = orc.pps_f64(W32.astype(np.float64), xc.astype(np.complex128), NN, Ns)
    # an explicit numpy FIR with the tap loop: sum_i W[i] x[s+i]
    ref = np.zeros((Ns, NN, N), np.complex128)
    for i in range(T):
        ref += W32[i].astype(np.float64)[None, None, :] * xc[i:i + Ns].astype(np.complex128)
    assert np.max(np.abs(c - ref)) < 1e-10
    assert np.max(np.abs(a - ref)) < 1e-3  # FP32 rounding of |x| <= 128*T


def test_pps_linearity():
    rng = np.random.default_rng(2)
    N, T, NN, Ns = 32, 5, 2, 4
    W = orc.window_f64(N, T)
    x = _rand_c64(rng, (Ns + T - 1, NN, N)).astype(np.complex128)
    y = _rand_c64(rng, (Ns + T - 1, NN, N)).astype(np.complex128)
    a, b = 0.7 - 0.2j, -1.3 + 0.5j
    lhs = orc.pps_f64(W, a * x + b * y, NN, Ns)
    rhs = a * orc.pps_f64(W, x, NN, Ns) + b * orc.pps_f64(W, y, NN, Ns)
    assert np.max(np.abs(lhs - rhs)) < 1e-10


# ----------------------------------------------------------------- FFT
@pytest.mark.parametrize("N", [8, 32, 1024, 2048, 4096, 1000, 1536, 3000, 4099])
def test_fft_f64_matches_numpy_golden(golden, N):
    x = golden[f"fft_in_N{N}"]
    X = orc.fft_f64(x, -1)
    ref = golden[f"fft_out_N{N}"]
    assert np.max(np.abs(X - ref)) / np.max(np.abs(ref)) < 1e-13
    Xi = orc.fft_f64(x, +1)
    refi = golden[f"ifft_out_N{N}"]
    assert np.max(np.abs(Xi - refi)) / np.max(np.abs(refi)) < 1e-13


@pytest.mark.parametrize("N", [4, 8, 16, 64, 12])
def test_dft_bruteforce_small(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((3, N)) + 1j * rng.standard_normal((3, N))
    n = np.arange(N)
    M = np.exp(-2j * np.pi * np.outer(n, n) / N)
    ref = x @ M.T
    assert np.max(np.abs(orc.dft_f64(x) - ref)) < 1e-12
    assert np.max(np.abs(orc.fft_f64(x) - ref)) < 1e-12


def test_fft_invariants():
    rng = np.random.default_rng(5)
    N = 256
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    X = orc.fft_f64(x)
    # Parseval: sum |X|^2 = N sum |x|^2
    assert abs(np.sum(np.abs(X) ** 2) - N * np.sum(np.abs(x) ** 2)) / (N * np.sum(np.abs(x) ** 2)) < 1e-13
    # integer-bin tone lands in exactly one bin (forward sign -1)
    k0 = 37
    tone = np.exp(2j * np.pi * k0 * np.arange(N) / N)
    T = orc.fft_f64(tone)
    assert abs(T[k0] - N) < 1e-9
    T[k0] = 0
    assert np.max(np.abs(T)) < 1e-9
    # impulse -> flat; shift theorem
    imp = np.zeros(N, complex)
    imp[0] = 1
    assert np.max(np.abs(orc.fft_f64(imp) - 1)) < 1e-15
    d = 5
    Xs = orc.fft_f64(np.roll(x, d))
    assert np.max(np.abs(Xs - X * np.exp(-2j * np.pi * np.arange(N) * d / N))) < 1e-10
    # FP32 oracle FFT close to FP64
    X32 = orc.fft_f32(x.astype(np.complex64))
    assert np.linalg.norm(X32 - X) / np.linalg.norm(X) < 1e-6


# ----------------------------------------------------------------- corner turn
def test_reorder_index_formula_and_bits():
    rng = np.random.default_rng(6)
    NN, N, Ns = 5, 7, 3
    bits = rng.integers(0, 2**32, size=(Ns, NN, N, 2), dtype=np.uint64).astype(np.uint32)
    bits[0, 0, 0, 0] = 0x7FC00001  # NaN payload
    bits[0, 0, 0, 1] = 0x80000000  # -0
    F = bits.view(np.uint64).reshape(Ns, NN, N)  # 8-byte elements (complex64)
    X = orc.reorder(F, NN, N, Ns)
    for s in range(Ns):
        for e in range(NN):
            for c in range(N):
                assert X[c, s, e] == F[s, e, c]
    # inverse permutation restores the input
    assert np.array_equal(np.transpose(X, (1, 2, 0)), F)


# ----------------------------------------------------------------- X-engine
def test_xengine_matches_einsum_golden(golden):
    x = golden["xeng_in"]
    Nc, Ns, NN = x.shape
    V = orc.xengine_full(x, NN, Ns, Nc)
    assert np.max(np.abs(V - golden["xeng_out"])) < 1e-12


def test_xengine_invariants():
    rng = np.random.default_rng(7)
    Nc, Ns, NN = 3, 50, 6
    x = _rand_c64(rng, (Nc, Ns, NN))
    x[:, :, 4] = x[:, :, 1]  # identical inputs
    V = orc.xengine_full(x, NN, Ns, Nc)  # [c][b][a]
    for c in range(Nc):
        M = V[c].T  # M[a][b]
        assert np.max(np.abs(M - M.conj().T)) < 1e-12  # Hermitian
        d = np.real(np.diag(M))
        assert np.all(d >= 0) and np.max(np.abs(np.imag(np.diag(M)))) < 1e-12
        assert np.all(np.abs(M) ** 2 <= np.outer(d, d) * (1 + 1e-12) + 1e-12)
        assert abs(M[4, 1] - M[1, 1]) < 1e-9
    # the lower triangle extraction matches bl(a, b) = a(a+1)/2 + b
    tri = orc.full_to_tri(V)
    assert tri[0, orc.tri_index(3, 2)] == V[0, 2, 3]


def test_cgemm_nc_general():
    rng = np.random.default_rng(8)
    m, n, k, lda, ldb, ldc = 4, 3, 5, 6, 4, 5
    A = _rand_c64(rng, (k * lda,))
    B = _rand_c64(rng, (k * ldb,))
    C0 = _rand_c64(rng, (n * ldc,))
    al, be = 0.5 - 0.25j, -1.0 + 2.0j
    out = orc.cgemm_nc(m, n, k, al, A, lda, B, ldb, be, C0, ldc)
    Am = A.reshape(k, lda).T[:m].astype(np.complex128)  # column-major m x k
    Bm = B.reshape(k, ldb).T[:n].astype(np.complex128)
    Cm = C0.reshape(n, ldc).T[:m].astype(np.complex128)
    ref = al * Am @ Bm.conj().T + be * Cm
    got = out.reshape(n, ldc).T[:m]
    assert np.max(np.abs(got - ref)) < 1e-9


# ----------------------------------------------------------------- unpack
def test_unpack_special_words():
    w = np.array([0, 1, 0x7FFFFFFF, 0x80000000, 0x80000001, 0xFFFFFFFF, 0x00080000, 0xFFF80000], np.uint32)
    got = orc.unpack_u32(w)
    exp = np.array([0.0, 2.0**-19, np.float32((2**31 - 1) * 2.0**-19), 4096.0,
                    -np.float32((2**31 - 1) * 2.0**-19), -(2.0**-19), 1.0, -1.0], np.float32)
    assert np.array_equal(got, exp)
    assert got[3] == 4096.0  # the reference quirk: 0x80000000 decodes positive


@pytest.mark.slow
def test_unpack_exhaustive_2_32():
    n, first = orc.unpack_exhaustive_mismatches()
    assert n == 0, f"first mismatching word {first:#x}"


# ----------------------------------------------------------------- depacketiser
def test_depacketise_positions():
    NN, N, Ns, T, S = 76, 32, 4, 4, 2  # two stations of 38 inputs
    words = 4 + 2 * 10 * S
    pk = np.zeros((3, words), np.uint32)
    # packet: station 1, snap 2 (offset 10 + 10), counter 0 -> samples 0, 1 (history, stack 1)
    pk[0, :3] = (1, 2, 0)
    # packet: station 0, snap 3 (offset 30 and 8 inputs), counter = begins/S -> first steady samples
    begins = (T - 1) * N
    pk[1, :3] = (0, 3, begins // S)
    pk[2, :3] = (0, 0, begins // S + N // S)  # sample N -> spectrum 1
    for p in range(3):
        pk[p, 4:] = np.arange(words - 4, dtype=np.uint32) * (1 << 19)  # word value k -> float k
    st, ns = orc.depacketise(pk, S, NN, N, Ns, T, first_block=True)
    # packet 0: sample i, input j -> stack 1, spectrum 0, element 38 + 20 + j, channel i
    for i in range(S):
        for j in range(10):
            v = st[1, 0, 38 + 20 + j, i]
            k = 2 * (i * 10 + j)
            assert v == complex(k, k + 1)
    # packet 1 (snap 3 has 8 inputs): stack 0, spectrum 0, element 30 + j
    for i in range(S):
        for j in range(8):
            k = 2 * (i * 8 + j)
            assert st[0, 0, 30 + j, i] == complex(k, k + 1)
    for i in range(S):
        assert st[0, 1, 0, i] == complex(2 * i * 10, 2 * i * 10 + 1)
    assert list(ns) == [2 * S, S, 0]


# ----------------------------------------------------------------- whole chain
def test_chain_f64_matches_numpy_golden(golden):
    NN, N, T, Ns = (int(v) for v in golden["chain_shape"])
    x8, W, ref = golden["chain_x8"], golden["chain_W"], golden["chain_vis"]
    chans = np.arange(N)
    vis = orc.fx_chain_f64(W, x8, NN, N, T, Ns, chans)
    scale = np.sqrt(np.max(np.abs(ref)))
    assert np.max(np.abs(vis - ref)) < 1e-11 * scale ** 2
    # subset of channels gives the same rows
    sub = orc.fx_chain_f64(W, x8, NN, N, T, Ns, np.array([5, 0, 63]))
    assert np.array_equal(sub, vis[[5, 0, 63]])


def test_chain_f32_close_to_f64(golden):
    NN, N, T, Ns = (int(v) for v in golden["chain_shape"])
    x8, W = golden["chain_x8"], golden["chain_W"]
    W32 = W.astype(np.float32)
    v32 = orc.fx_chain_f32(W32, x8, NN, N, T, Ns)
    v64 = orc.fx_chain_f64(W32.astype(np.float64), x8, NN, N, T, Ns, np.arange(N))
    a, b = np.tril_indices(NN)
    auto = np.real(v64[:, [orc.tri_index(i, i) for i in range(NN)]])
    norm = np.sqrt(auto[:, a] * auto[:, b])
    err = np.max(np.abs(v32 - v64) / norm)
    assert err < 1e-5, err


# ----------------------------------------------------------------- calibration
def _calib_case(rng, NN=4, N=8):
    nb = NN * (NN + 1) // 2
    V = rng.standard_normal((N, nb)) + 1j * rng.standard_normal((N, nb))
    M = rng.standard_normal((N, nb)) + 1j * rng.standard_normal((N, nb))
    g = rng.uniform(0.8, 1.2, NN) * np.exp(1j * rng.uniform(-np.pi, np.pi, NN))
    tau = rng.uniform(-4, 4, NN)
    return V, M, g, tau


def test_calib_apply_closed_form():
    rng = np.random.default_rng(9)
    NN, N = 3, 8
    V, M, g, tau = _calib_case(rng, NN, N)
    Vp = orc.calib_apply(V, g, tau, NN, N)
    for c in range(N):
        k = c if c < N // 2 else c - N
        for a in range(NN):
            for b in range(a + 1):
                ph = np.exp(-2j * np.pi * k * (tau[a] - tau[b]) / N)
                exp = g[a] * np.conj(g[b]) * ph * V[c, orc.tri_index(a, b)]
                assert abs(Vp[c, orc.tri_index(a, b)] - exp) < 1e-12
    L, _, _ = orc.calib_loss_grad(V, M, g, tau, NN, N)
    assert abs(L - np.sum(np.abs(Vp - M) ** 2)) < 1e-10 * L


def test_calib_grad_vs_finite_differences():
    rng = np.random.default_rng(10)
    NN, N = 4, 8
    V, M, g, tau = _calib_case(rng, NN, N)
    L, gg, gt = orc.calib_loss_grad(V, M, g, tau, NN, N)
    h = 1e-6
    for a in range(NN):
        for d, unit in ((0, 1.0), (1, 1j)):
            gp, gm = g.copy(), g.copy()
            gp[a] += h * unit
            gm[a] -= h * unit
            fd = (orc.calib_loss_grad(V, M, gp, tau, NN, N)[0] - orc.calib_loss_grad(V, M, gm, tau, NN, N)[0]) / (2 * h)
            an = gg[a].real if d == 0 else gg[a].imag
            assert abs(fd - an) < 1e-6 * max(1.0, abs(an)), (a, d, fd, an)
        tp, tm = tau.copy(), tau.copy()
        tp[a] += h
        tm[a] -= h
        fd = (orc.calib_loss_grad(V, M, g, tp, NN, N)[0] - orc.calib_loss_grad(V, M, g, tm, NN, N)[0]) / (2 * h)
        assert abs(fd - gt[a]) < 1e-6 * max(1.0, abs(gt[a])), (a, fd, gt[a])


def test_xengine_f64_branch_pinned(golden):
    """orc_xengine_full_f64 (the complex128 branch of oracle.xengine_full):
    on the golden input widened to complex128 it must give the einsum golden
    output, and on values that FP32 represents it must give bit for bit what
    the complex64 branch (orc_xengine_full_from_f32) gives; on complex128
    values that FP32 cannot hold, it must follow a numpy einsum of those
    values (a branch that rounded its input to FP32 would miss by ~1e-8)."""
    x = golden["xeng_in"]
    Nc, Ns, NN = x.shape
    V = orc.xengine_full(x.astype(np.complex128), NN, Ns, Nc)
    assert np.max(np.abs(V - golden["xeng_out"])) < 1e-12
    rng = np.random.default_rng(11)
    y = _rand_c64(rng, (5, 33, 7))
    a = orc.xengine_full(y, 7, 33, 5)
    b = orc.xengine_full(y.astype(np.complex128), 7, 33, 5)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    z = rng.standard_normal((3, 9, 4)) + 1j * rng.standard_normal((3, 9, 4))  # full FP64 mantissas
    ref = np.einsum("csa,csb->cba", z, np.conj(z)) / 9
    got = orc.xengine_full(z, 4, 9, 3)
    assert np.max(np.abs(got - ref)) < 1e-14
    z32 = orc.xengine_full(z.astype(np.complex64), 4, 9, 3)
    assert np.max(np.abs(z32 - ref)) > 1e-9  # the pin can tell the two branches apart
    # alpha is applied: alpha = 2 doubles every entry of the default 1/Ns
    assert np.allclose(orc.xengine_full(z, 4, 9, 3, alpha=2.0), 18.0 * ref, rtol=0, atol=1e-13)


def test_e4m3_quantiser_matches_torch_float8():
    """the test-side e4m3 quantiser (tests/e4m3.py) against torch's
    float8_e4m3fn conversion: random values over the whole range, every
    exactly representable value, every midpoint between neighbours (ties to
    even), and the subnormals, both signs.  All inputs are float32 values, so
    torch's float32 -> float8 conversion rounds once, as the quantiser does."""
    torch = pytest.importorskip("torch")
    from e4m3 import E4M3_MAX, e4m3_rne
    # every finite e4m3fn value (254 of the 256 codes are finite, 0x7f/0xff are NaN)
    codes = torch.arange(256, dtype=torch.int32).to(torch.uint8).view(torch.float8_e4m3fn)
    vals = codes.to(torch.float32).numpy()
    vals = np.unique(vals[np.isfinite(vals)]).astype(np.float64)
    assert vals.max() == E4M3_MAX and vals.min() == -E4M3_MAX and len(vals) == 253  # +-0 merge
    pos = vals[vals >= 0]
    mids = (pos[:-1] + pos[1:]) / 2  # exact ties
    rng = np.random.default_rng(12)
    rand = np.concatenate([
        rng.uniform(-E4M3_MAX, E4M3_MAX, 200000),
        np.exp2(rng.uniform(-12, np.log2(E4M3_MAX), 100000)) * rng.choice([-1, 1], 100000),
        rng.uniform(-2.0 ** -6, 2.0 ** -6, 10000),  # subnormal range
    ])
    x = np.concatenate([vals, mids, -mids, rand, np.nextafter(mids, 0), np.nextafter(mids, 1e9)])
    x = x[np.abs(x) <= E4M3_MAX].astype(np.float32)
    want = torch.from_numpy(x).to(torch.float8_e4m3fn).to(torch.float32).numpy().astype(np.float64)
    got = e4m3_rne(x.astype(np.float64))
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, (x[bad[:5]], got[bad[:5]], want[bad[:5]])
    assert x.size > 310000
