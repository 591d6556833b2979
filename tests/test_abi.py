"""CPU tests of the C ABI boundary: every symbol include/mdtfx.h declares is
exported by libmdtfx.so, and argument validation (the limits of
MDT_OneStation_client.h:54-61, null pointers, unsupported operations) fails
with the documented status codes before any GPU work is issued."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mdtfx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(mdt_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


@pytest.fixture(scope="module")
def L():
    import mdtfx
    return mdtfx.lib()


def test_header_declares_the_surface():
    names = declared_functions()
    for required in ["mdt_pfb_window", "mdt_pps_batch", "mdt_pps_batch_i8", "mdt_fft_plan_1d",
                     "mdt_fft_set_stream", "mdt_fft_exec_c2c", "mdt_fft_destroy",
                     "mdt_reorder_foutput", "mdt_xengine_create", "mdt_xengine_set_stream",
                     "mdt_cgemm_batched", "mdt_history_shift", "mdt_unpack_u32",
                     "mdt_fx_plan_create", "mdt_fx_dump", "mdt_last_error_string"]:
        assert required in names


def test_every_declared_symbol_is_exported(L):
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def _err(L):
    return L.mdt_last_error_string().decode()


def test_limits_rejected_before_launch(L):
    import mdtfx
    dummy = ctypes.c_void_p(16)  # never dereferenced: validation fails first
    # Nchannels below MINCHANNELS (client.h:56)
    assert L.mdt_pps_batch(16, 8, dummy, dummy, dummy, 38, 1024, None) == mdtfx.ERR_INVALID_VALUE
    assert "Nchannels" in _err(L)
    # Ntaps above MAXTAPS (client.h:59)
    assert L.mdt_pps_batch(1024, 129, dummy, dummy, dummy, 38, 1024, None) == mdtfx.ERR_INVALID_VALUE
    # Nspectra below MINSPECTRA (client.h:60)
    assert L.mdt_pps_batch(1024, 8, dummy, dummy, dummy, 38, 1, None) == mdtfx.ERR_INVALID_VALUE
    # inputs below MINELEMENTS (client.h:54)
    assert L.mdt_pps_batch(1024, 8, dummy, dummy, dummy, 2, 1024, None) == mdtfx.ERR_INVALID_VALUE
    # null pointers
    assert L.mdt_pps_batch(1024, 8, None, dummy, dummy, 38, 1024, None) == mdtfx.ERR_NULL_POINTER
    assert L.mdt_reorder_foutput(38, 1024, None, dummy, 1024, None) == mdtfx.ERR_NULL_POINTER
    # misaligned complex buffer
    odd = ctypes.c_void_p(4)
    assert L.mdt_reorder_foutput(38, 1024, odd, dummy, 1024, None) == mdtfx.ERR_MISALIGNED
    assert L.mdt_pfb_window(1024, 8, dummy, 7, None) == mdtfx.ERR_INVALID_VALUE


def test_fft_plan_rejections(L):
    import mdtfx
    h = ctypes.c_void_p()
    # lengths outside the client's [32, 16384] that are not powers of two >= 8
    # (1000 is accepted: it needs a GPU to build its tables, tested there)
    for n in (12, 31, 16385, 100000, 4):
        assert L.mdt_fft_plan_1d(ctypes.byref(h), n, mdtfx.FFT_C2C, 4) == mdtfx.ERR_UNSUPPORTED, n
    assert L.mdt_fft_plan_1d(ctypes.byref(h), 1024, 0x2a, 4) == mdtfx.ERR_UNSUPPORTED  # R2C
    assert L.mdt_fft_plan_1d(ctypes.byref(h), 1024, mdtfx.FFT_C2C, 0) == mdtfx.ERR_INVALID_VALUE
    assert L.mdt_fft_exec_c2c(None, None, None, -1) == mdtfx.ERR_NULL_POINTER


def test_cgemm_rejections(L):
    import mdtfx
    h = ctypes.c_void_p()
    assert L.mdt_xengine_create(ctypes.byref(h)) == 0
    one = mdtfx._Complex(1.0, 0.0)
    d = ctypes.c_void_p(256)
    # only OP_N x OP_C (the X-engine call, client.cu:1128) is provided
    assert L.mdt_cgemm_batched(h, mdtfx.OP_N, mdtfx.OP_N, 4, 4, 4, ctypes.byref(one), d, 4, d, 4,
                               ctypes.byref(one), d, 4, 1) == mdtfx.ERR_UNSUPPORTED
    assert L.mdt_cgemm_batched(h, mdtfx.OP_N, mdtfx.OP_C, 8, 4, 4, ctypes.byref(one), d, 4, d, 4,
                               ctypes.byref(one), d, 8, 1) == mdtfx.ERR_INVALID_VALUE  # lda < m
    assert L.mdt_cgemm_batched(h, mdtfx.OP_N, mdtfx.OP_C, 0, 4, 4, ctypes.byref(one), d, 4, d, 4,
                               ctypes.byref(one), d, 4, 1) == 0  # empty: no work
    assert L.mdt_xengine_destroy(h) == 0


def test_fx_plan_rejections(L):
    import mdtfx
    h = ctypes.c_void_p()
    assert L.mdt_fx_plan_create(ctypes.byref(h), 40, 1000, 4, 256, 0, 0) == mdtfx.ERR_UNSUPPORTED
    assert L.mdt_fx_plan_create(ctypes.byref(h), 40, 1024, 4, 256, 0, 9) == mdtfx.ERR_INVALID_VALUE
    assert L.mdt_fx_plan_create(ctypes.byref(h), 5000, 1024, 4, 256, 0, 0) == mdtfx.ERR_INVALID_VALUE
    assert L.mdt_history_shift(ctypes.c_void_p(256), 40, 1024, 4, 256, 4, None) == mdtfx.ERR_INVALID_VALUE


def test_python_front_end_refuses_cpu_tensors():
    import torch
    import mdtfx
    W = torch.zeros((4, 64), dtype=torch.float32)
    x = torch.zeros((7, 3, 64), dtype=torch.complex64)
    with pytest.raises(ValueError):
        mdtfx.pps_batch(W, x, 3, 4, out=torch.zeros((4, 3, 64), dtype=torch.complex64))


def test_python_front_end_checks_buffer_sizes():
    """raw-pointer calls refuse short or wrongly typed buffers before any
    launch (the checks run on the host, so CPU tensors exercise them)"""
    import torch
    import mdtfx
    W = torch.zeros((4, 64), dtype=torch.float32)
    with pytest.raises(ValueError, match="input"):
        mdtfx.pps_batch(W, torch.zeros((6, 3, 64), dtype=torch.complex64), 3, 4,
                        out=torch.zeros((4, 3, 64), dtype=torch.complex64))  # needs Ns+T-1 = 7 spectra
    with pytest.raises(ValueError, match="output"):
        mdtfx.pps_batch_i8(W, torch.zeros((7, 3, 64, 2), dtype=torch.int8), 3, 4,
                           out=torch.zeros((3, 3, 64), dtype=torch.complex64))
    with pytest.raises(TypeError):
        mdtfx.reorder_foutput(torch.zeros((4, 3, 64), dtype=torch.float32), 3, 64, 4,
                              out=torch.zeros((64, 4, 3), dtype=torch.float32))
    with pytest.raises(TypeError):
        mdtfx.unpack_u32(torch.zeros(16, dtype=torch.int16), out=torch.zeros(16))
    with pytest.raises(ValueError):
        mdtfx.history_shift(torch.zeros((5, 3, 64), dtype=torch.complex64), 3, 64, 4, 4)  # needs 7 spectra
