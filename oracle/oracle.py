"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never the thing
measured or shipped.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.

Each wrapper names the C function it calls; the C source cites the reference
file:line it restates.  Arrays are numpy; complex data is passed as
complex64/complex128 (interleaved re, im) so the byte layouts match the
reference's cuComplex buffers.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

WINDOW_BH4 = 0
WINDOW_HAMMING = 1


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_native = None


def lib_native():
    """The same C source built for this host's own CPU (gcc -O3 -march=native
    -fopenmp -ffp-contract=off), into a temporary directory: bench.py's CPU
    baseline times this build.  Falls back to the portable prebuilt library
    (returned with build = "prebuilt -march=x86-64-v3") when no compiler runs
    here.  Returns (library, build description)."""
    global _native
    if _native is None:
        import tempfile
        flags = ["-O3", "-march=native", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c11"]
        out = os.path.join(tempfile.mkdtemp(prefix="mdt_oracle_native_"), "liboracle_native.so")
        try:
            subprocess.run(["gcc", *flags, "-shared", "-o", out, os.path.join(_HERE, "mdt_oracle.c"), "-lm"],
                           check=True, capture_output=True, timeout=120)
            L = ctypes.CDLL(out)
            _declare(L)
            _native = (L, "gcc " + " ".join(flags))
        except (OSError, subprocess.SubprocessError):
            _native = (lib(), "prebuilt -march=x86-64-v3")
    return _native


def fx_chain_f32_with(L, W, x8, NN, N, T, Ns, nthreads=0):
    """orc_fx_chain_f32 through a given build of the oracle (lib_native)"""
    W = _c(W, np.float32)
    x8 = _c(x8, np.int8)
    nb = NN * (NN + 1) // 2
    vis = np.empty((N, nb), np.complex64)
    L.orc_fx_chain_f32(NN, N, T, Ns, _p(W), _p(x8), _p(vis), int(nthreads))
    return vis


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _declare(L):
    vp, ci, cl, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double
    sig = {
        "orc_sinpi": (cd, [cd]),
        "orc_cospi": (cd, [cd]),
        "orc_window_f64": (None, [ci, ci, ci, vp]),
        "orc_window_f32": (None, [ci, ci, ci, vp]),
        "orc_pps_f32": (None, [ci, ci, vp, vp, vp, ci, ci]),
        "orc_pps_i8_f32": (None, [ci, ci, vp, vp, vp, ci, ci]),
        "orc_pps_f64": (None, [ci, ci, vp, vp, vp, ci, ci]),
        "orc_dft_f64": (None, [ci, cl, vp, vp, ci]),
        "orc_fft_f64": (None, [ci, cl, vp, vp, ci]),
        "orc_fft_f32": (None, [ci, cl, vp, vp, ci]),
        "orc_reorder": (None, [ci, ci, ci, ci, vp, vp]),
        "orc_xengine_full_f64": (None, [ci, ci, ci, cd, vp, vp]),
        "orc_xengine_full_from_f32": (None, [ci, ci, ci, cd, vp, vp]),
        "orc_cgemm_nc_f64": (None, [ci, ci, ci, vp, vp, ci, vp, ci, vp, vp, ci, vp]),
        "orc_unpack_u32": (ctypes.c_float, [ctypes.c_uint32]),
        "orc_unpack_u32_array": (None, [vp, cl, vp]),
        "orc_count_unpack_mismatch_all": (ci, [vp, ctypes.POINTER(cl)]),
        "orc_depacketise": (None, [vp, ci, ci, ci, ci, ci, ci, ci, ci, vp, vp]),
        "orc_fx_chain_f64": (None, [ci, ci, ci, ci, vp, vp, vp, ci, vp]),
        "orc_fx_chain_f32": (None, [ci, ci, ci, ci, vp, vp, vp, ci]),
        "orc_calib_loss_grad_f64": (cd, [ci, ci, vp, vp, vp, vp, vp, vp]),
        "orc_calib_apply_f64": (None, [ci, ci, vp, vp, vp, vp]),
        "orc_num_threads": (ci, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def sinpi(x):
    return lib().orc_sinpi(float(x))


def cospi(x):
    return lib().orc_cospi(float(x))


def window_f64(N, T, kind=WINDOW_BH4):
    """orc_window_f64: devfn.cu:52-88 in FP64, layout [T][N]."""
    W = np.empty(T * N, np.float64)
    lib().orc_window_f64(N, T, kind, _p(W))
    return W.reshape(T, N)


def window_f32(N, T, kind=WINDOW_BH4):
    W = np.empty(T * N, np.float32)
    lib().orc_window_f32(N, T, kind, _p(W))
    return W.reshape(T, N)


def pps_f32(W, x, NN, Ns):
    """orc_pps_f32 (devfn.cu:113-161). x: complex64 [(Ns+T-1)][NN][N]."""
    T, N = W.shape
    W = _c(W, np.float32)
    x = _c(x, np.complex64)
    assert x.size == (Ns + T - 1) * NN * N
    out = np.empty((Ns, NN, N), np.complex64)
    lib().orc_pps_f32(N, T, _p(W), _p(x), _p(out), NN, Ns)
    return out


def pps_i8_f32(W, x8, NN, Ns):
    """orc_pps_i8_f32. x8: int8 [(Ns+T-1)][NN][N][2]."""
    T, N = W.shape
    W = _c(W, np.float32)
    x8 = _c(x8, np.int8)
    assert x8.size == (Ns + T - 1) * NN * N * 2
    out = np.empty((Ns, NN, N), np.complex64)
    lib().orc_pps_i8_f32(N, T, _p(W), _p(x8), _p(out), NN, Ns)
    return out


def pps_f64(W, x, NN, Ns):
    T, N = W.shape
    W = _c(W, np.float64)
    x = _c(x, np.complex128)
    out = np.empty((Ns, NN, N), np.complex128)
    lib().orc_pps_f64(N, T, _p(W), _p(x), _p(out), NN, Ns)
    return out


def dft_f64(x, sign=-1):
    x = _c(x, np.complex128)
    N = x.shape[-1]
    out = np.empty_like(x)
    lib().orc_dft_f64(N, x.size // N, _p(x), _p(out), sign)
    return out


def fft_f64(x, sign=-1):
    x = _c(x, np.complex128)
    N = x.shape[-1]
    out = np.empty_like(x)
    lib().orc_fft_f64(N, x.size // N, _p(x), _p(out), sign)
    return out


def fft_f32(x, sign=-1):
    x = _c(x, np.complex64)
    N = x.shape[-1]
    out = np.empty_like(x)
    lib().orc_fft_f32(N, x.size // N, _p(x), _p(out), sign)
    return out


def reorder(F, NN, N, Ns):
    """orc_reorder (devfn.cu:183-208): [Ns][NN][N] -> [N][Ns][NN], byte-exact."""
    F = np.ascontiguousarray(F)
    out = np.empty(F.size, F.dtype)
    lib().orc_reorder(NN, N, Ns, F.dtype.itemsize, _p(F), _p(out))
    return out.reshape(N, Ns, NN)


def xengine_full(XIn, NN, Ns, N, alpha=None):
    """orc_xengine_full_*: client.cu:1128 semantics.  Returns complex128
    [N][NN(col b)][NN(row a)] i.e. flat index c*NN*NN + a + b*NN."""
    if alpha is None:
        alpha = 1.0 / Ns
    out = np.empty(N * NN * NN, np.complex128)
    if XIn.dtype == np.complex64:
        XIn = _c(XIn, np.complex64)
        lib().orc_xengine_full_from_f32(NN, Ns, N, alpha, _p(XIn), _p(out))
    else:
        XIn = _c(XIn, np.complex128)
        lib().orc_xengine_full_f64(NN, Ns, N, alpha, _p(XIn), _p(out))
    return out.reshape(N, NN, NN)


def cgemm_nc(m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc):
    """orc_cgemm_nc_f64: one entry of cublasCgemm(OP_N, OP_C) in FP64."""
    al = np.array([alpha.real, alpha.imag], np.float64)
    be = np.array([beta.real, beta.imag], np.float64)
    A = _c(A, np.complex64)
    B = _c(B, np.complex64)
    out = np.zeros(ldc * n, np.complex128)
    if C0 is not None:
        C0 = _c(C0, np.complex64)
        c0p = _p(C0)
    else:
        c0p = None
    lib().orc_cgemm_nc_f64(m, n, k, _p(al), _p(A), lda, _p(B), ldb, _p(be), c0p, ldc, _p(out))
    return out


def unpack_u32(w):
    w = _c(w, np.uint32)
    out = np.empty(w.shape, np.float32)
    lib().orc_unpack_u32_array(_p(w), w.size, _p(out))
    return out


def unpack_exhaustive_mismatches():
    first = ctypes.c_long(-1)
    n = lib().orc_count_unpack_mismatch_all(None, ctypes.byref(first))
    return n, first.value


def depacketise(pkts, SperElement, NN, N, Ns, T, first_block):
    """orc_depacketise (client.cu:846-1005). pkts: uint32 [npkts][words]."""
    pkts = _c(pkts, np.uint32)
    npk, words = pkts.shape
    stacks = np.zeros(2 * Ns * NN * N, np.complex64)
    ns = np.zeros(3, np.int64)  # samples to stack 0, to stack 1, skipped
    lib().orc_depacketise(_p(pkts), npk, words, SperElement, NN, N, Ns, T,
                          1 if first_block else 0, _p(stacks), _p(ns))
    return stacks.reshape(2, Ns, NN, N), ns


def fx_chain_f64(W, x8, NN, N, T, Ns, chans):
    """orc_fx_chain_f64: int8 -> FP64 triangle visibilities for chans."""
    W = _c(W, np.float64)
    x8 = _c(x8, np.int8)
    ch = _c(chans, np.int32)
    nb = NN * (NN + 1) // 2
    vis = np.empty((len(ch), nb), np.complex128)
    lib().orc_fx_chain_f64(NN, N, T, Ns, _p(W), _p(x8), _p(ch), len(ch), _p(vis))
    return vis


def fx_chain_f32(W, x8, NN, N, T, Ns, nthreads=0):
    """orc_fx_chain_f32: the FP32 restatement (CPU baseline port)."""
    W = _c(W, np.float32)
    x8 = _c(x8, np.int8)
    nb = NN * (NN + 1) // 2
    vis = np.empty((N, nb), np.complex64)
    lib().orc_fx_chain_f32(NN, N, T, Ns, _p(W), _p(x8), _p(vis), int(nthreads))
    return vis


def calib_loss_grad(V, M, g, tau, NN, N):
    V = _c(V, np.complex128)
    M = _c(M, np.complex128)
    g = _c(g, np.complex128)
    tau = _c(tau, np.float64)
    gg = np.empty(NN, np.complex128)
    gt = np.empty(NN, np.float64)
    L = lib().orc_calib_loss_grad_f64(NN, N, _p(V), _p(M), _p(g), _p(tau), _p(gg), _p(gt))
    return L, gg, gt


def calib_apply(V, g, tau, NN, N):
    V = _c(V, np.complex128)
    g = _c(g, np.complex128)
    tau = _c(tau, np.float64)
    out = np.empty_like(V)
    lib().orc_calib_apply_f64(NN, N, _p(V), _p(g), _p(tau), _p(out))
    return out


def num_threads():
    return lib().orc_num_threads()


def tri_index(a, b):
    return a * (a + 1) // 2 + b


def full_to_tri(Vfull):
    """[N][NN][NN] with flat c*NN*NN + a + b*NN  ->  [N][nb] (a >= b)."""
    N, NN, _ = Vfull.shape
    a, b = np.tril_indices(NN)
    # Vfull[c].reshape(NN(b), NN(a)) so element (a,b) sits at [c, b, a]
    return Vfull[:, b, a]
