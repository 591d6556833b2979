"""CPU tests of the Python front end's X-input view (FxPlan.xinput_view): for
each layout's strides (block, block_pitch, chan_group, chan_pitch, row_pitch)
a buffer filled through the formula the C header documents,
  (s // block) * block_pitch + (k // chan_group) * chan_pitch
  + e * row_pitch + (k % chan_group) * block + s % block,
must come back as the [N][NN][Ns] samples (complex64, or float16 (re, im)
pairs for the per-channel FP8 format, or e4m3 byte pairs for the
block-scaled one).  No GPU: the strides are given, not queried."""
import types

import numpy as np
import pytest
import torch

import mdtfx


def _fake_plan(N, NN, Ns, xprec, strides, f8block=False):
    return types.SimpleNamespace(N=N, NN=NN, Ns=Ns, xprec=xprec, xinput_strides=lambda: strides, f8block=f8block)


def _layouts(N, NN, Ns, esize):
    B2 = 64 // esize  # pairs: 64 bytes of spectra per channel, two channels side by side
    B1 = 128 // esize  # blocks: 128 bytes of spectra, one channel
    rp = Ns + (Ns & 1) + 8
    return {
        "rows": (1 << 30, 0, 1, NN * rp + 64, rp),
        "pairs": (B2, -(-N // 2) * NN * 2 * B2, 2, NN * 2 * B2, 2 * B2),
        "blocks": (B1, N * NN * B1, 1, NN * B1, B1),
    }


def _size(N, Ns, strides):
    B, bp, G, cp, rp = strides
    return N * cp if bp == 0 else -(-Ns // B) * bp


@pytest.mark.parametrize("N,NN,Ns", [(8, 3, 20), (16, 5, 37), (4, 2, 8), (6, 1, 33)])
@pytest.mark.parametrize("layout", ["rows", "pairs", "blocks"])
@pytest.mark.parametrize("fp8", [False, True])
def test_view_matches_the_documented_formula(N, NN, Ns, layout, fp8):
    esize = 4 if fp8 else 8
    strides = _layouts(N, NN, Ns, esize)[layout]
    B, bp, G, cp, rp = strides
    total = _size(N, Ns, strides)
    k, e, s = np.meshgrid(np.arange(N), np.arange(NN), np.arange(Ns), indexing="ij")
    off = (s // B) * bp + (k // G) * cp + e * rp + (k % G) * B + s % B
    assert off.max() < total and len(np.unique(off)) == off.size  # in range, no two samples share an element
    val = (k * 7 + e * 3 + s) % 997 + 1.0
    if fp8:
        buf = np.zeros((total, 2), np.float16)
        buf[off, 0], buf[off, 1] = val, -val
        plan = _fake_plan(N, NN, Ns, mdtfx.XPREC_FP8, strides)
        v = plan_view(plan, torch.from_numpy(buf.reshape(-1)).view(torch.int32))
        assert tuple(v.shape) == (N, NN, Ns, 2)
        got = v.float().numpy()
        assert np.array_equal(got[..., 0], val) and np.array_equal(got[..., 1], -val)
    else:
        buf = np.zeros(total, np.complex64)
        buf[off] = val - 1j * val
        plan = _fake_plan(N, NN, Ns, mdtfx.XPREC_FP32, strides)
        v = plan_view(plan, torch.from_numpy(buf))
        assert tuple(v.shape) == (N, NN, Ns)
        assert np.array_equal(v.numpy(), (val - 1j * val).astype(np.complex64))


def plan_view(plan, buf):
    return mdtfx.FxPlan.xinput_view(plan, buf)


def test_rows_view_is_a_view():
    """in the rows layout the samples are a strided view of the buffer (no copy)"""
    N, NN, Ns = 4, 3, 10
    strides = _layouts(N, NN, Ns, 8)["rows"]
    buf = torch.zeros(_size(N, Ns, strides), dtype=torch.complex64)
    v = plan_view(_fake_plan(N, NN, Ns, mdtfx.XPREC_FP32, strides), buf)
    v[1, 2, 3] = 5 + 6j
    B, bp, G, cp, rp = strides
    assert buf[1 * cp + 2 * rp + 3] == 5 + 6j


@pytest.mark.parametrize("N,NN,Ns", [(8, 3, 40), (4, 5, 70), (2, 1, 32)])
def test_block_fp8_view(N, NN, Ns):
    """the block-scaled FP8 format: 2-byte (re, im) e4m3 codes in blocks of
    32 spectra (row pitch 32, channel pitch NN*32, block pitch N*NN*32)"""
    strides = (32, N * NN * 32, 1, NN * 32, 32)
    B, bp, G, cp, rp = strides
    total = _size(N, Ns, strides)
    k, e, s = np.meshgrid(np.arange(N), np.arange(NN), np.arange(Ns), indexing="ij")
    off = (s // B) * bp + k * cp + e * rp + s % B
    assert off.max() < total and len(np.unique(off)) == off.size
    re, im = ((k * 7 + e * 3 + s) % 251).astype(np.uint8), ((k + e * 5 + s * 3) % 253).astype(np.uint8)
    buf = np.zeros((total, 2), np.uint8)
    buf[off, 0], buf[off, 1] = re, im
    plan = _fake_plan(N, NN, Ns, mdtfx.XPREC_FP8, strides, f8block=True)
    v = plan_view(plan, torch.from_numpy(buf.reshape(-1)))
    assert tuple(v.shape) == (N, NN, Ns, 2) and v.dtype == torch.uint8
    assert np.array_equal(v[..., 0].numpy(), re) and np.array_equal(v[..., 1].numpy(), im)
