"""Python front end of libmdtfx (the MI355X-native MDT FX correlator).

Thin ctypes bindings over the C ABI declared in include/mdtfx.h.  PyTorch is
used only for device memory and the current stream: every call takes torch
tensors that already live on the GPU and passes their raw pointers, in the
reference's argument order, to the HIP library.  There is no CPU fallback: a
missing library or a CPU tensor raises immediately.

Mirrors of the reference's GPU call surface (MDT_OneStation_client.cu):
  pfb_window        PFB_Window_min4term_BH   (MDT_device_functions.cu:52-88)
  pps_batch[_i8]    PpS_Batch                (MDT_device_functions.cu:113-161)
  FftPlan           cufftPlan1d/SetStream/ExecC2C (client.cu:622, 630, 1105)
  reorder_foutput   ReorderFOutput           (MDT_device_functions.cu:183-208)
  XEngine           cublasCreate/SetStream/CgemmBatched (client.cu:690-705, 1128)
  history_shift     the (T-1)-spectrum carry (client.cu:1170)
  unpack_u32        SNAP word decode          (client.cu:898-907)
  FxPlan            native fused path (int8 -> triangle visibilities)
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(_HERE)
LIB_PATH = os.path.join(_PKG, "libmdtfx.so")

SUCCESS = 0
ERR_INVALID_VALUE = 1
ERR_NULL_POINTER = 2
ERR_UNSUPPORTED = 3
ERR_LAUNCH = 4
ERR_ALLOC = 5
ERR_DEVICE = 6
ERR_MISALIGNED = 7

WINDOW_BH4 = 0
WINDOW_HAMMING = 1
FFT_FORWARD = -1
FFT_INVERSE = 1
FFT_C2C = 0x29
OP_N, OP_T, OP_C = 0, 1, 2
XPREC_FP32 = 0
XPREC_FP16_SPLIT = 1
XPREC_FP8 = 2
VIS_FULL_COLMAJOR = 0
VIS_TRIANGLE = 1
XIN_AUTO = -1   # X input layout per X-engine precision (default)
XIN_PAIRS = 0   # X input: channel pairs, spectra in 64-byte blocks
XIN_ROWS = 1    # X input: channel-major rows with padding
XIN_BLOCKS = 2  # X input: one channel, spectra in 128-byte blocks

_lib = None


class MdtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mdtfx error {code}: {msg}")
        self.code = code


class _Complex(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float)]


class DumpTimes(ctypes.Structure):
    _fields_ = [("h2d_ms", ctypes.c_float), ("fengine_ms", ctypes.c_float), ("xengine_ms", ctypes.c_float),
                ("d2h_ms", ctypes.c_float), ("total_ms", ctypes.c_float)]


class SynthParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("sigma", ctypes.c_float), ("rho", ctypes.c_float),
                ("max_delay", ctypes.c_int), ("ntones", ctypes.c_int),
                ("tone_freq", ctypes.c_float * 4), ("tone_amp", ctypes.c_float * 4),
                ("t0", ctypes.c_longlong)]


# name -> argtypes (all return int status)
_VP, _I, _F, _SZ = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_size_t
_SIGS = {
    "mdt_version": [ctypes.POINTER(ctypes.c_int)],
    "mdt_pfb_window": [_I, _I, _VP, _I, _VP],
    "mdt_pps_batch": [_I, _I, _VP, _VP, _VP, _I, _I, _VP],
    "mdt_pps_batch_i8": [_I, _I, _VP, _VP, _VP, _I, _I, _VP],
    "mdt_fft_plan_1d": [ctypes.POINTER(_VP), _I, _I, _I],
    "mdt_fft_set_stream": [_VP, _VP],
    "mdt_fft_exec_c2c": [_VP, _VP, _VP, _I],
    "mdt_fft_destroy": [_VP],
    "mdt_reorder_foutput": [_I, _I, _VP, _VP, _I, _VP],
    "mdt_xengine_create": [ctypes.POINTER(_VP)],
    "mdt_xengine_set_stream": [_VP, _VP],
    "mdt_xengine_destroy": [_VP],
    "mdt_cgemm_batched": [_VP, _I, _I, _I, _I, _I, ctypes.POINTER(_Complex), _VP, _I, _VP, _I,
                          ctypes.POINTER(_Complex), _VP, _I, _I],
    "mdt_xengine_herk": [_VP, _I, _I, _F, _VP, _VP, _I],
    "mdt_history_shift": [_VP, _I, _I, _I, _I, _I, _VP],
    "mdt_unpack_u32": [_VP, _VP, _SZ, _VP],
    "mdt_pfb_window_exblackman": [_I, _I, _VP, _VP],
    "mdt_depacketise": [_VP, _I, _I, _I, _I, _I, _I, _I, _I, _VP, _VP, _VP],
    "mdt_csv_filename": [ctypes.c_char_p, _SZ, ctypes.c_longlong],
    "mdt_write_csv": [ctypes.c_char_p, _VP, _I, _I, _I, _I, _I],
    "mdt_write_binary": [ctypes.c_char_p, _VP, _I, _I, _I, _I, _I, _I, ctypes.c_longlong],
    "mdt_fx_plan_create": [ctypes.POINTER(_VP), _I, _I, _I, _I, _I, _I],
    "mdt_fx_set_stream": [_VP, _VP],
    "mdt_fx_set_precision": [_VP, _I],
    "mdt_fx_xinput_bytes": [_VP, ctypes.POINTER(_SZ)],
    "mdt_fx_set_xinput_padding": [_VP, _I, _I],
    "mdt_fx_set_xinput_layout": [_VP, _I],
    "mdt_fx_xinput_strides": [_VP] + [ctypes.POINTER(_SZ)] * 5,
    "mdt_fx_xinput_pitch": [_VP, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)],
    "mdt_fx_get_window": [_VP, _VP],
    "mdt_fx_fengine": [_VP, _VP, _VP],
    "mdt_fx_fengine_f32": [_VP, _VP, _VP],
    "mdt_fx_xengine": [_VP, _VP, _VP, _VP],
    "mdt_fx_dump": [_VP, _VP, _VP, _VP, _VP],
    "mdt_fx_plan_destroy": [_VP],
    "mdt_fx_get_channel_exponents": [_VP, _VP],
    "mdt_fx_xinput_scales": [_VP, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)],
    "mdt_fx_get_split_scale": [_VP, ctypes.POINTER(_F)],
    "mdt_fx_plan_shape": [_VP, _VP, _VP, _VP, _VP, _VP],
    "mdt_stream_create": [ctypes.POINTER(_VP), _VP],
    "mdt_stream_host_stack": [_VP, _I, ctypes.POINTER(_VP)],
    "mdt_stream_host_vis": [_VP, _I, ctypes.POINTER(_VP)],
    "mdt_stream_prime": [_VP, _VP],
    "mdt_stream_set_caller_stream": [_VP, _VP],
    "mdt_stream_submit": [_VP, _I, _VP],
    "mdt_stream_wait": [_VP, _I, _VP],
    "mdt_stream_destroy": [_VP],
    "mdt_pstream_create": [ctypes.POINTER(_VP), _VP, _I, _I, _I],
    "mdt_pstream_host_packets": [_VP, _I, ctypes.POINTER(_VP)],
    "mdt_pstream_host_vis": [_VP, _I, ctypes.POINTER(_VP)],
    "mdt_pstream_set_caller_stream": [_VP, _VP],
    "mdt_pstream_restart": [_VP],
    "mdt_pstream_submit": [_VP, _I, _I, _VP],
    "mdt_pstream_wait": [_VP, _I, _VP, _VP],
    "mdt_pstream_destroy": [_VP],
    "mdt_synth_i8": [ctypes.POINTER(SynthParams), _I, _I, _I, _VP, _VP],
    "mdt_fx_set_gains": [_VP, _VP, _VP],
    "mdt_calib_plan_create": [ctypes.POINTER(_VP), _I, _I],
    "mdt_calib_set_stream": [_VP, _VP],
    "mdt_calib_apply": [_VP, _VP, _VP, _VP, _VP],
    "mdt_calib_loss_grad": [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP],
    "mdt_calib_plan_destroy": [_VP],
    "mdt_calib_loss_grad_noalloc": [_I, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP],
    "mdt_hbm_probe": [_I, _VP, _VP, _SZ, _I, _VP],
}


def lib():
    """Load libmdtfx.so (built in-tree by __graft_entry__.build()); no fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libmdtfx.so not found at {LIB_PATH}: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.mdt_last_error_string.restype = ctypes.c_char_p
        L.mdt_last_error_string.argtypes = []
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != SUCCESS:
        raise MdtError(status, lib().mdt_last_error_string().decode())
    return status


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t, dtype=None):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("mdtfx works on GPU tensors only (no CPU fallback)")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _need(t, n, what, esize=None):
    """the kernels take raw pointers: refuse a buffer smaller than the extent
    they touch, or of the wrong element size, before anything is launched"""
    if esize is not None and t.element_size() != esize:
        raise TypeError(f"{what}: {esize}-byte elements expected, got {t.element_size()}")
    if t.numel() < n:
        raise ValueError(f"{what}: {t.numel()} elements, the call touches {n}")


def version():
    v = ctypes.c_int()
    check(lib().mdt_version(ctypes.byref(v)))
    return v.value


# ------------------------------------------------------------------ a1
def pfb_window_exblackman(N, T, out=None):
    """the client's PFB_Window_ExBlackman call (the BH4 window)"""
    if out is None:
        out = torch.empty((T, N), dtype=torch.float32, device="cuda")
    _need(out, T * N, "Window")
    check(lib().mdt_pfb_window_exblackman(N, T, _ptr(out, torch.float32), _stream()))
    return out


def pfb_window(N, T, kind=WINDOW_BH4, out=None):
    if out is None:
        out = torch.empty((T, N), dtype=torch.float32, device="cuda")
    check(lib().mdt_pfb_window(N, T, _ptr(out, torch.float32), kind, _stream()))
    return out


# ------------------------------------------------------------------ a2
def pps_batch(W, x, NN, Ns, out=None):
    """x: complex64 [(Ns+T-1), NN, N]."""
    T, N = W.shape
    if out is None:
        out = torch.empty((Ns, NN, N), dtype=torch.complex64, device="cuda")
    _need(x, (Ns + T - 1) * NN * N, "input")
    _need(out, Ns * NN * N, "output")
    check(lib().mdt_pps_batch(N, T, _ptr(W, torch.float32), _ptr(x, torch.complex64),
                              _ptr(out, torch.complex64), NN, Ns, _stream()))
    return out


def pps_batch_i8(W, x8, NN, Ns, out=None):
    """x8: int8 [(Ns+T-1), NN, N, 2]."""
    T, N = W.shape
    if out is None:
        out = torch.empty((Ns, NN, N), dtype=torch.complex64, device="cuda")
    _need(x8, (Ns + T - 1) * NN * N * 2, "input")
    _need(out, Ns * NN * N, "output")
    check(lib().mdt_pps_batch_i8(N, T, _ptr(W, torch.float32), _ptr(x8, torch.int8),
                                 _ptr(out, torch.complex64), NN, Ns, _stream()))
    return out


# ------------------------------------------------------------------ a3
class FftPlan:
    """cufftPlan1d(nx, CUFFT_C2C, batch) equivalent."""

    def __init__(self, nx, batch, type=FFT_C2C):
        self._h = ctypes.c_void_p()
        check(lib().mdt_fft_plan_1d(ctypes.byref(self._h), nx, type, batch))
        self.nx, self.batch = nx, batch
        self._bound = None

    def set_stream(self, stream_handle):
        """bind the plan to a HIP stream; unbound plans run on torch's
        current stream at each call"""
        self._bound = stream_handle
        check(lib().mdt_fft_set_stream(self._h, ctypes.c_void_p(stream_handle)))

    def exec_c2c(self, idata, odata, direction=FFT_FORWARD):
        _need(idata, self.nx * self.batch, "idata")
        _need(odata, self.nx * self.batch, "odata")
        if self._bound is None:
            check(lib().mdt_fft_set_stream(self._h, _stream()))
        check(lib().mdt_fft_exec_c2c(self._h, _ptr(idata, torch.complex64),
                                     _ptr(odata, torch.complex64), direction))
        return odata

    def destroy(self):
        if self._h:
            check(lib().mdt_fft_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ------------------------------------------------------------------ a4
def reorder_foutput(F, NN, N, Ns, out=None):
    if out is None:
        out = torch.empty((N, Ns, NN), dtype=F.dtype, device="cuda")
    _need(F, Ns * NN * N, "FOutput", 8)
    _need(out, Ns * NN * N, "XInput", 8)
    check(lib().mdt_reorder_foutput(NN, N, _ptr(F), _ptr(out), Ns, _stream()))
    return out


# ------------------------------------------------------------------ a5
class XEngine:
    """cublasCreate + cublasSetStream + cublasCgemmBatched equivalent."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        check(lib().mdt_xengine_create(ctypes.byref(self._h)))
        self._bound = None

    def set_stream(self, stream_handle):
        """bind the handle to a HIP stream; unbound handles run on torch's
        current stream at each call"""
        self._bound = stream_handle
        check(lib().mdt_xengine_set_stream(self._h, ctypes.c_void_p(stream_handle)))

    def _sync_stream(self):
        if self._bound is None:
            check(lib().mdt_xengine_set_stream(self._h, _stream()))

    def cgemm_batched(self, transa, transb, m, n, k, alpha, Aptrs, lda, Bptrs, ldb, beta,
                      Cptrs, ldc, batch):
        """A/B/C pointer tables: int64 device tensors of data_ptr() values."""
        for t, name in ((Aptrs, "A table"), (Bptrs, "B table"), (Cptrs, "C table")):
            _need(t, batch, name)
        al = _Complex(float(complex(alpha).real), float(complex(alpha).imag))
        be = _Complex(float(complex(beta).real), float(complex(beta).imag))
        self._sync_stream()
        check(lib().mdt_cgemm_batched(self._h, transa, transb, m, n, k, ctypes.byref(al),
                                      _ptr(Aptrs, torch.int64), lda, _ptr(Bptrs, torch.int64), ldb,
                                      ctypes.byref(be), _ptr(Cptrs, torch.int64), ldc, batch))

    def herk(self, NN, Ns, alpha, XIn, XOut, N):
        _need(XIn, N * Ns * NN, "XIn")
        _need(XOut, N * NN * NN, "XOut")
        self._sync_stream()
        check(lib().mdt_xengine_herk(self._h, NN, Ns, float(alpha), _ptr(XIn, torch.complex64),
                                     _ptr(XOut, torch.complex64), N))
        return XOut

    def destroy(self):
        if self._h:
            check(lib().mdt_xengine_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ------------------------------------------------------------------ a6, a7
def history_shift(FInput, NN, N, T, Ns):
    eb = 8 if FInput.dtype == torch.complex64 else 2
    _need(FInput, (Ns + T - 1) * NN * N * (1 if eb == 8 else 2), "FInput", 8 if eb == 8 else 1)
    check(lib().mdt_history_shift(_ptr(FInput), NN, N, T, Ns, eb, _stream()))


def unpack_u32(words, out=None):
    """words: int32 tensor holding the raw uint32 bit patterns."""
    if out is None:
        out = torch.empty(words.shape, dtype=torch.float32, device="cuda")
    _need(words, words.numel(), "words", 4)
    _need(out, words.numel(), "output")
    check(lib().mdt_unpack_u32(_ptr(words), _ptr(out, torch.float32), words.numel(), _stream()))
    return out


# ------------------------------------------------------------------ native path
class FxPlan:
    """One-dump plan of the fused path: int8 [Ns+T-1][NN][N][2] -> X input
    (channel pairs, spectrum blocks or padded rows: xinput_strides /
    xinput_view) -> triangle visibilities [N][NN(NN+1)/2]."""

    def __init__(self, NN, N, T, Ns, window=WINDOW_BH4, xprec=XPREC_FP32):
        self._h = ctypes.c_void_p()
        check(lib().mdt_fx_plan_create(ctypes.byref(self._h), NN, N, T, Ns, window, xprec))
        self.NN, self.N, self.T, self.Ns = NN, N, T, Ns
        self.nb = NN * (NN + 1) // 2
        self.xprec = xprec
        self._bound = None

    def _sync_stream(self):
        """unbound plans launch on torch's current stream at each call"""
        if self._bound is None:
            check(lib().mdt_fx_set_stream(self._h, _stream()))

    def set_precision(self, xprec):
        """X-engine precision; the X input's sample format and pitches follow
        it, so re-allocate the X input and re-run the F-engine after a change."""
        check(lib().mdt_fx_set_precision(self._h, xprec))
        self.xprec = xprec

    def split_scale(self):
        """power-of-two scale s of the FP16 sample formats (x*s is stored)"""
        s = ctypes.c_float()
        check(lib().mdt_fx_get_split_scale(self._h, ctypes.byref(s)))
        return s.value

    def channel_exponents(self):
        """FP8 plans: int32 [N] exponents e_c of the last F-engine run; the
        X-engine quantises e4m3(x*s*2^e_c)."""
        e = torch.empty(self.N, dtype=torch.int32, device="cuda")
        self._sync_stream()
        check(lib().mdt_fx_get_channel_exponents(self._h, _ptr(e, torch.int32)))
        return e

    def set_gains(self, gains=None, delays=None):
        """Per-input complex gains (complex64 [NN]) and delays (float32 [NN],
        samples) applied by the F-engine; None switches the stage off.  The
        tensors must stay alive while the plan uses them."""
        if gains is not None and gains.numel() < self.NN:
            raise ValueError("gains need NN entries")
        if delays is not None and delays.numel() < self.NN:
            raise ValueError("delays need NN entries")
        self._gains = (gains, delays)
        check(lib().mdt_fx_set_gains(self._h, _ptr(gains, torch.complex64) if gains is not None else None,
                                     _ptr(delays, torch.float32) if delays is not None else None))

    def set_stream(self, stream_handle):
        """bind the plan to a HIP stream (None: torch's current stream at each
        call, the default)"""
        self._bound = stream_handle
        if stream_handle is not None:
            check(lib().mdt_fx_set_stream(self._h, ctypes.c_void_p(stream_handle)))

    def window(self):
        W = torch.empty((self.T, self.N), dtype=torch.float32, device="cuda")
        self._sync_stream()
        check(lib().mdt_fx_get_window(self._h, _ptr(W, torch.float32)))
        return W

    def set_xinput_layout(self, layout):
        """XIN_AUTO (default), XIN_PAIRS, XIN_BLOCKS or XIN_ROWS, for the
        buffers filled from now on"""
        check(lib().mdt_fx_set_xinput_layout(self._h, layout))

    def set_xinput_padding(self, row_pad, chan_pad):
        """rows layout: padding (even, samples) after each row / channel"""
        check(lib().mdt_fx_set_xinput_padding(self._h, row_pad, chan_pad))

    def xinput_pitch(self):
        """rows layout: (row_pitch, chan_pitch) in samples"""
        rp, cp = ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().mdt_fx_xinput_pitch(self._h, ctypes.byref(rp), ctypes.byref(cp)))
        return rp.value, cp.value

    def xinput_strides(self):
        """(block, block_pitch, chan_group, chan_pitch, row_pitch): sample
        (k, e, s) at (s // block) * block_pitch + (k // chan_group) * chan_pitch
        + e * row_pitch + (k % chan_group) * block + s % block"""
        v = [ctypes.c_size_t() for _ in range(5)]
        check(lib().mdt_fx_xinput_strides(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def xinput_bytes(self):
        nbytes = ctypes.c_size_t()
        check(lib().mdt_fx_xinput_bytes(self._h, ctypes.byref(nbytes)))
        return nbytes.value

    @property
    def f8block(self):
        """FP8 with block exponents (512 and 1024 channels): 2-byte e4m3
        samples, the exponents after them in the X input"""
        return self.xprec == XPREC_FP8 and self.N in (512, 1024)

    def alloc_xinput(self):
        """X input buffer (flat, mdt_fx_xinput_bytes); xinput_view() gives
        the samples as [N][NN][Ns]."""
        nbytes = self.xinput_bytes()
        if self.f8block:  # 2-byte e4m3 samples and the exponent bytes
            return torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        if self.xprec == XPREC_FP8:  # 4-byte samples: FP16 (re, im)
            return torch.empty(nbytes // 4, dtype=torch.int32, device="cuda")
        return torch.empty(nbytes // 8, dtype=torch.complex64, device="cuda")

    def xinput_scales(self, buf):
        """block-scaled FP8 plans: the E8M0 exponent bytes (eb + 127) of the
        blocks of 16 spectra, uint8 [ceil(Ns/16)][NN] (a view of buf)"""
        off, nblk = ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().mdt_fx_xinput_scales(self._h, ctypes.byref(off), ctypes.byref(nblk)))
        b = buf.view(torch.uint8)
        return b[off.value: off.value + nblk.value * self.NN].view(nblk.value, self.NN)

    def xinput_view(self, buf):
        """The samples as [N][NN][Ns] complex64; for FP8 plans at 2048 and
        4096 channels float16 [N][NN][Ns][2] (re, im) of x*s, at 512 and 1024
        channels the e4m3 codes uint8 [N][NN][Ns][2] (re, im) of x / 2^eb
        (xinput_scales); a view of buf in the rows layout, a copy in the
        others."""
        B, bp, G, cp, rp = self.xinput_strides()
        f8 = self.xprec == XPREC_FP8
        src = buf.view(torch.uint8) if self.f8block else (buf.view(torch.float16) if f8 else buf)
        u = 2 if f8 else 1   # tensor elements per sample
        tail, tstr = ((2,), (1,)) if f8 else ((), ())
        if G == 1 and bp == 0:
            return torch.as_strided(src, (self.N, self.NN, self.Ns) + tail, (u * cp, u * rp, u) + tstr)
        nblk, npair = -(-self.Ns // B), -(-self.N // G)
        v = torch.as_strided(src, (nblk, npair, self.NN, G, B) + tail,
                             (u * bp, u * cp, u * rp, u * B, u) + tstr)
        v = v.permute(1, 3, 2, 0, 4, *range(5, 5 + len(tail)))
        return v.reshape((npair * G, self.NN, nblk * B) + tail)[:self.N, :, :self.Ns]

    def alloc_vis(self):
        return torch.empty((self.N, self.nb), dtype=torch.complex64, device="cuda")

    def _check_sizes(self, x8=None, xin=None, vis=None, acc64=None):
        """The kernels take raw pointers: refuse buffers smaller than the
        extents they read or write, before anything is launched."""
        if x8 is not None and x8.numel() < (self.Ns + self.T - 1) * self.NN * self.N * 2:
            raise ValueError("input buffer smaller than (Ns+T-1) x NN x N int8 (I, Q) samples")
        if xin is not None:
            nbytes = ctypes.c_size_t()
            check(lib().mdt_fx_xinput_bytes(self._h, ctypes.byref(nbytes)))
            if xin.numel() * xin.element_size() < nbytes.value:
                raise ValueError(f"X input buffer smaller than mdt_fx_xinput_bytes = {nbytes.value}")
        if vis is not None and vis.numel() < self.N * self.nb:
            raise ValueError("visibility buffer smaller than N x NN(NN+1)/2")
        if acc64 is not None and acc64.numel() < self.N * self.nb:
            raise ValueError("accumulator smaller than N x NN(NN+1)/2")

    def fengine(self, x8, xin):
        self._check_sizes(x8=x8, xin=xin)
        self._sync_stream()
        check(lib().mdt_fx_fengine(self._h, _ptr(x8, torch.int8), _ptr(xin)))
        return xin

    def fengine_f32(self, fin, xin):
        """F-engine on the reference's float F input: complex64
        [Ns+T-1][NN][N] (FP32 and FP16-split plans)"""
        if fin.numel() < (self.Ns + self.T - 1) * self.NN * self.N:
            raise ValueError("input buffer smaller than (Ns+T-1) x NN x N complex64 samples")
        self._check_sizes(xin=xin)
        self._sync_stream()
        check(lib().mdt_fx_fengine_f32(self._h, _ptr(fin, torch.complex64), _ptr(xin)))
        return xin

    def xengine(self, xin, vis, acc64=None):
        self._check_sizes(xin=xin, vis=vis, acc64=acc64)
        self._sync_stream()
        check(lib().mdt_fx_xengine(self._h, _ptr(xin), _ptr(vis, torch.complex64),
                                   _ptr(acc64, torch.complex128) if acc64 is not None else None))
        return vis

    def dump(self, x8, xin, vis, acc64=None):
        self._check_sizes(x8=x8, xin=xin, vis=vis, acc64=acc64)
        self._sync_stream()
        check(lib().mdt_fx_dump(self._h, _ptr(x8, torch.int8), _ptr(xin),
                                _ptr(vis, torch.complex64),
                                _ptr(acc64, torch.complex128) if acc64 is not None else None))
        return vis

    def destroy(self):
        if self._h:
            check(lib().mdt_fx_plan_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class FxStream:
    """Streaming correlator over consecutive dumps of one FxPlan: two host
    slots, each a pinned int8 input stack [Ns][NN][N][2] and a pinned
    visibility buffer [N][NN(NN+1)/2] complex64 (numpy views).  Fill a slot's
    stack, submit(slot), later wait(slot) and read vis(slot).  The PFB history
    (the previous dump's last T-1 spectra) is carried on the device."""

    def __init__(self, plan):
        import numpy as np
        self._h = ctypes.c_void_p()
        check(lib().mdt_stream_create(ctypes.byref(self._h), plan._h))
        self.plan = plan  # the stream uses the plan until destroyed
        self._stacks, self._vis = [], []
        for slot in range(2):
            p = ctypes.c_void_p()
            check(lib().mdt_stream_host_stack(self._h, slot, ctypes.byref(p)))
            buf = (ctypes.c_int8 * (plan.Ns * plan.NN * plan.N * 2)).from_address(p.value)
            buf._owner = self  # a view keeps the stream (and its pinned memory) alive
            self._stacks.append(np.frombuffer(buf, np.int8).reshape(plan.Ns, plan.NN, plan.N, 2))
            check(lib().mdt_stream_host_vis(self._h, slot, ctypes.byref(p)))
            buf = (ctypes.c_float * (2 * plan.N * plan.nb)).from_address(p.value)
            buf._owner = self
            self._vis.append(np.frombuffer(buf, np.complex64).reshape(plan.N, plan.nb))

    def stack(self, slot):
        """pinned int8 input stack of a slot (valid until destroy())"""
        if not self._h:
            raise MdtError(ERR_NULL_POINTER, "stream destroyed")
        return self._stacks[slot]

    def vis(self, slot):
        """pinned visibilities of a slot (valid until destroy())"""
        if not self._h:
            raise MdtError(ERR_NULL_POINTER, "stream destroyed")
        return self._vis[slot]

    def prime(self, history=None):
        """history: int8 [T-1][NN][N][2] (host), None for zeros"""
        import numpy as np
        if history is None:
            check(lib().mdt_stream_prime(self._h, None))
            return
        h = np.ascontiguousarray(history, dtype=np.int8)
        if h.size < (self.plan.T - 1) * self.plan.NN * self.plan.N * 2:
            raise ValueError("history needs T-1 spectra")
        check(lib().mdt_stream_prime(self._h, ctypes.c_void_p(h.ctypes.data)))

    def submit(self, slot, acc64=None):
        """queue one dump; it first waits for the work queued so far on
        torch's current stream (e.g. zeroing acc64)"""
        if acc64 is not None and acc64.numel() < self.plan.N * self.plan.nb:
            raise ValueError("accumulator smaller than N x NN(NN+1)/2")
        check(lib().mdt_stream_set_caller_stream(self._h, _stream()))
        check(lib().mdt_stream_submit(self._h, slot, _ptr(acc64, torch.complex128) if acc64 is not None else None))

    def wait(self, slot):
        """blocks until the slot's dump is done; returns its phase times (ms)"""
        t = DumpTimes()
        check(lib().mdt_stream_wait(self._h, slot, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in DumpTimes._fields_}

    def destroy(self):
        """frees the pinned slots: views from stack() and vis() must not be
        used afterwards"""
        if self._h:
            self._stacks, self._vis = [], []
            check(lib().mdt_stream_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class PacketStream:
    """Packet streaming correlator (mdt_pstream_*): two host slots, each a
    pinned uint32 packet buffer [max_pkts][words] and pinned visibilities
    [N][NN(NN+1)/2] complex64 (numpy views).  Write a dump's packets into
    packets(slot), submit(slot, npkts), later wait(slot) -> (phase times,
    counts) and read vis(slot).  Dumps are numbered by submit order from 0
    (the first block, which also carries the PFB history)."""

    def __init__(self, plan, words_per_pkt, SperElement, max_pkts):
        import numpy as np
        self._h = ctypes.c_void_p()
        check(lib().mdt_pstream_create(ctypes.byref(self._h), plan._h, words_per_pkt, SperElement, max_pkts))
        self.plan, self.words, self.max_pkts = plan, words_per_pkt, max_pkts
        self._pk, self._vis = [], []
        for slot in range(2):
            p = ctypes.c_void_p()
            check(lib().mdt_pstream_host_packets(self._h, slot, ctypes.byref(p)))
            buf = (ctypes.c_uint32 * (max_pkts * words_per_pkt)).from_address(p.value)
            buf._owner = self
            self._pk.append(np.frombuffer(buf, np.uint32).reshape(max_pkts, words_per_pkt))
            check(lib().mdt_pstream_host_vis(self._h, slot, ctypes.byref(p)))
            buf = (ctypes.c_float * (2 * plan.N * plan.nb)).from_address(p.value)
            buf._owner = self
            self._vis.append(np.frombuffer(buf, np.complex64).reshape(plan.N, plan.nb))

    def packets(self, slot):
        if not self._h:
            raise MdtError(ERR_NULL_POINTER, "stream destroyed")
        return self._pk[slot]

    def vis(self, slot):
        if not self._h:
            raise MdtError(ERR_NULL_POINTER, "stream destroyed")
        return self._vis[slot]

    def submit(self, slot, npkts, acc64=None):
        if acc64 is not None and acc64.numel() < self.plan.N * self.plan.nb:
            raise ValueError("accumulator smaller than N x NN(NN+1)/2")
        check(lib().mdt_pstream_set_caller_stream(self._h, _stream()))
        check(lib().mdt_pstream_submit(self._h, slot, int(npkts),
                                       _ptr(acc64, torch.complex128) if acc64 is not None else None))

    def wait(self, slot):
        """blocks until the slot's dump is done: (phase times in ms, [samples
        to the dump's spectra, to its history, dropped])"""
        t = DumpTimes()
        c = (ctypes.c_ulonglong * 3)()
        check(lib().mdt_pstream_wait(self._h, slot, ctypes.byref(t), c))
        return {k: getattr(t, k) for k, _ in DumpTimes._fields_}, [int(v) for v in c]

    def restart(self):
        check(lib().mdt_pstream_restart(self._h))

    def destroy(self):
        if self._h:
            self._pk, self._vis = [], []
            check(lib().mdt_pstream_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def depacketise(pkts, SperElement, NN, N, Ns, T, first_block, stacks=None, counts=None):
    """SNAP packets (int32/uint32 tensor [npkts][words] on the GPU) -> the
    F-input stacks [2][Ns][NN][N] complex64; counts (uint64 as int64 [3]):
    samples to stack 0, to stack 1, dropped.  Accumulates into the given
    buffers (new zeroed ones when None)."""
    if pkts.dim() != 2 or pkts.element_size() != 4:
        raise ValueError("packets must be a 2-D tensor of 32-bit words")
    if stacks is None:
        stacks = torch.zeros((2, Ns, NN, N), dtype=torch.complex64, device="cuda")
    if counts is None:
        counts = torch.zeros(3, dtype=torch.int64, device="cuda")
    if stacks.numel() < 2 * Ns * NN * N or counts.numel() < 3:
        raise ValueError("stacks or counts too small")
    check(lib().mdt_depacketise(_ptr(pkts), pkts.shape[0], pkts.shape[1], SperElement, NN, N, Ns, T,
                                1 if first_block else 0, _ptr(stacks, torch.complex64),
                                _ptr(counts, torch.int64), _stream()))
    return stacks, counts


def csv_filename(unix_time):
    """MDT_%Y%m%d_%H%M%S.csv in local time (the client's output name)"""
    buf = ctypes.create_string_buffer(64)
    check(lib().mdt_csv_filename(buf, 64, int(unix_time)))
    return buf.value.decode()


def _host_array(a, dtype):
    import numpy as np
    a = np.ascontiguousarray(a)
    if a.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {a.dtype}")
    return a


def write_csv(path, vis, layout, Nelements, Npols, Nchannels, all_pol_pairs=False):
    """CSV in the client's format from host visibilities (numpy complex64):
    layout VIS_FULL_COLMAJOR ([N][NN][NN] column-major, the reference's X
    output) or VIS_TRIANGLE (the native packed triangle)."""
    import numpy as np
    v = _host_array(vis, np.complex64)
    NN = Nelements * Npols
    need = Nchannels * (NN * NN if layout == VIS_FULL_COLMAJOR else NN * (NN + 1) // 2)
    if v.size < need:
        raise ValueError("visibility array smaller than the layout needs")
    check(lib().mdt_write_csv(str(path).encode(), ctypes.c_void_p(v.ctypes.data), layout, Nelements, Npols,
                              Nchannels, 1 if all_pol_pairs else 0))


def write_binary(path, vis, layout, Nelements, Npols, Nchannels, Nspectra, unix_time=0):
    import numpy as np
    v = np.ascontiguousarray(vis)
    dtype = {np.dtype(np.complex64): 0, np.dtype(np.complex128): 1}.get(v.dtype)
    if dtype is None:
        raise TypeError("visibilities must be complex64 or complex128")
    NN = Nelements * Npols
    need = Nchannels * (NN * NN if layout == VIS_FULL_COLMAJOR else NN * (NN + 1) // 2)
    if v.size < need:
        raise ValueError("visibility array smaller than the layout needs")
    check(lib().mdt_write_binary(str(path).encode(), ctypes.c_void_p(v.ctypes.data), layout, dtype, Nelements,
                                 Npols, Nchannels, Nspectra, int(unix_time)))


def read_binary(path):
    """(header dict, visibilities) of a mdt_write_binary file"""
    import numpy as np
    raw = open(path, "rb").read()
    if raw[:8] != b"MDTVIS01":
        raise ValueError("not an MDT visibility dump")
    layout, dtype, Ne, Np, Nc, Ns = np.frombuffer(raw, np.int32, 6, 8)
    t = int(np.frombuffer(raw, np.int64, 1, 32)[0])
    NN = int(Ne) * int(Np)
    per = NN * NN if layout == VIS_FULL_COLMAJOR else NN * (NN + 1) // 2
    data = np.frombuffer(raw, np.complex64 if dtype == 0 else np.complex128, int(Nc) * per, 64)
    hdr = dict(layout=int(layout), dtype=int(dtype), Nelements=int(Ne), Npols=int(Np), Nchannels=int(Nc),
               Nspectra=int(Ns), unix_time=t)
    return hdr, data.reshape(int(Nc), per)


def synth_i8(NN, N, nframes, seed=1, sigma=16.0, rho=0.0, max_delay=0, tones=(), t0=0, out=None):
    """Deterministic int8 (I, Q) samples [nframes][NN][N][2] on the GPU."""
    p = SynthParams()
    p.seed, p.sigma, p.rho, p.max_delay = seed, sigma, rho, max_delay
    p.ntones = len(tones)
    for i, (f, a) in enumerate(tones):
        p.tone_freq[i] = f
        p.tone_amp[i] = a
    p.t0 = t0
    if out is None:
        out = torch.empty((nframes, NN, N, 2), dtype=torch.int8, device="cuda")
    check(lib().mdt_synth_i8(ctypes.byref(p), NN, N, nframes, _ptr(out, torch.int8), _stream()))
    return out


PROBE_READ, PROBE_WRITE, PROBE_COPY = 0, 1, 2


def hbm_probe(kind, src, dst, grid):
    """One streaming pass of the library's HBM probe (read src, write dst, or
    copy src to dst; uint8 CUDA tensors of equal size, multiple of 16 bytes)
    on the current stream: a measurement aid for the kernels' rooflines."""
    t = src if src is not None else dst
    nbytes = t.numel() * t.element_size()
    sp = ctypes.c_void_p(src.data_ptr()) if src is not None else ctypes.c_void_p(0)
    dp = ctypes.c_void_p(dst.data_ptr()) if dst is not None else ctypes.c_void_p(0)
    check(lib().mdt_hbm_probe(kind, sp, dp, nbytes, grid, _stream()))
