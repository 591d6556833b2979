"""Streaming correlator (the client's rank loop, MDT_OneStation_client.cu:
916-1276, in one process): dumps pushed through the two pinned host slots must
give, bit for bit, what one-dump calls give on the same windows of one
continuous sample stream -- dump d correlates spectra [d*Ns, d*Ns + Ns + T-1)
of it, its first T-1 spectra being the previous dump's last ones (the history
shift of client.cu:1170).  The FP64 accumulator must hold the in-order sum of
the dumps' visibilities.  Each streamed dump is also checked against the FP64
oracle of the whole chain on its own window of the stream, and the stream's
ordering with the caller's work (a just-zeroed accumulator) is exercised."""
import numpy as np
import pytest
import torch

import mdtfx
import oracle as orc

pytestmark = pytest.mark.gpu


def _signal(NN, N, T, Ns, D, seed):
    x = mdtfx.synth_i8(NN, N, T - 1 + D * Ns, seed=seed, sigma=16.0, rho=0.3, max_delay=4,
                       tones=((0.123, 20.0),))
    torch.cuda.synchronize()
    return x, x.cpu().numpy()


def _per_dump(NN, N, T, Ns, xprec, x, d):
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, xprec)
    v = plan.dump(x[d * Ns: d * Ns + Ns + T - 1].contiguous(), plan.alloc_xinput(), plan.alloc_vis())
    torch.cuda.synchronize()
    return v.cpu().numpy()


@pytest.mark.parametrize("xprec,NN,N,T,Ns,slots", [
    (mdtfx.XPREC_FP32, 40, 512, 4, 64, [0, 1, 0, 1, 0]),
    (mdtfx.XPREC_FP16_SPLIT, 40, 512, 8, 48, [0, 0, 1, 1, 0]),   # one slot twice in a row
    (mdtfx.XPREC_FP8, 64, 1024, 4, 32, [1, 0, 1]),
    (mdtfx.XPREC_FP16_SPLIT, 200, 256, 4, 3, [0, 1, 1, 0]),      # Ns = T-1: the history is the whole stack
])
def test_stream_matches_per_dump(xprec, NN, N, T, Ns, slots):
    D = len(slots)
    print(f"stream xprec={xprec} NN={NN} N={N} T={T} Ns={Ns} slots={slots}", flush=True)
    x, xh = _signal(NN, N, T, Ns, D, seed=NN + Ns)
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, xprec)
    st = mdtfx.FxStream(plan)
    acc = torch.zeros((N, plan.nb), dtype=torch.complex128, device="cuda")
    st.prime(xh[:T - 1])
    pending = {}
    got = [None] * D
    for d, slot in enumerate(slots):
        if slot in pending:
            t = st.wait(slot)
            got[pending.pop(slot)] = st.vis(slot).copy()
            assert t["total_ms"] > 0 and t["fengine_ms"] > 0 and t["xengine_ms"] > 0
        st.stack(slot)[...] = xh[T - 1 + d * Ns: T - 1 + (d + 1) * Ns]
        st.submit(slot, acc)
        pending[slot] = d
    for slot, d in pending.items():
        st.wait(slot)
        got[d] = st.vis(slot).copy()
    torch.cuda.synchronize()
    a = acc.cpu().numpy()
    st.destroy()
    want = np.zeros_like(a)
    for d in range(D):
        ref = _per_dump(NN, N, T, Ns, xprec, x, d)
        assert np.array_equal(got[d], ref), f"dump {d}"
        want += ref.astype(np.complex128)
    assert np.array_equal(a, want)


def test_stream_unprimed_history_is_zero():
    NN, N, T, Ns = 16, 256, 4, 32
    x, xh = _signal(NN, N, T, Ns, 1, seed=3)
    plan = mdtfx.FxPlan(NN, N, T, Ns)
    st = mdtfx.FxStream(plan)
    st.stack(1)[...] = xh[T - 1:]
    st.submit(1)
    st.wait(1)
    got = st.vis(1).copy()
    st.destroy()
    z = x.clone()
    z[:T - 1] = 0
    assert np.array_equal(got, _per_dump(NN, N, T, Ns, mdtfx.XPREC_FP32, z, 0))


def test_stream_slot_errors():
    plan = mdtfx.FxPlan(16, 256, 4, 32)
    st = mdtfx.FxStream(plan)
    st.stack(0)[...] = 1
    with pytest.raises(mdtfx.MdtError):
        st.wait(0)  # nothing in flight
    st.submit(0)
    with pytest.raises(mdtfx.MdtError):
        st.submit(0)  # not waited for
    with pytest.raises(mdtfx.MdtError):
        st.prime()  # a dump in flight
    with pytest.raises(mdtfx.MdtError):
        st.submit(2)
    st.wait(0)
    st.destroy()
    with pytest.raises(mdtfx.MdtError):  # the pinned slots are gone
        st.stack(0)
    with pytest.raises(mdtfx.MdtError):  # Ns < T-1
        mdtfx.FxStream(mdtfx.FxPlan(16, 256, 8, 4))


def test_stream_refuses_a_grown_x_input():
    """the stream's X input is sized for the plan at creation: a later layout
    whose X input is larger is refused before anything is queued"""
    NN, N, T, Ns = 16, 256, 4, 32
    x, xh = _signal(NN, N, T, Ns, 1, seed=5)
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP16_SPLIT)  # pairs: no padding
    st = mdtfx.FxStream(plan)
    st.prime(xh[:T - 1])
    st.stack(0)[...] = xh[T - 1:]
    plan.set_xinput_layout(mdtfx.XIN_ROWS)  # padded rows: larger
    with pytest.raises(mdtfx.MdtError):
        st.submit(0)
    plan.set_xinput_layout(mdtfx.XIN_AUTO)
    st.submit(0)
    st.wait(0)
    got = st.vis(0).copy()
    st.destroy()
    assert np.array_equal(got, _per_dump(NN, N, T, Ns, mdtfx.XPREC_FP16_SPLIT, x, 0))


def _autonorm_max(v, v64, NN):
    """max over (channel, a >= b) of |V - V64| / sqrt(V64_aa V64_bb)"""
    a, b = np.tril_indices(NN)
    diag = np.array([orc.tri_index(i, i) for i in range(NN)])
    auto = np.real(v64[:, diag])
    return float(np.max(np.abs(v - v64) / np.sqrt(auto[:, a] * auto[:, b])))


# FP8: e4m3 rounding (<= 2^-4 relative per sample) averages over the Ns
# spectra in noise, 0.35/sqrt(Ns) at ~6 sigma, and repeats coherently for the
# tone's bins: 2 x 2^-4 on top (the bound of tests/test_gpu_fp8.py)
@pytest.mark.parametrize("xprec,NN,N,T,Ns,slots,chans", [
    (mdtfx.XPREC_FP32, 40, 512, 4, 64, [0, 1, 1, 0], None),
    (mdtfx.XPREC_FP16_SPLIT, 64, 1024, 8, 40, [1, 1, 0, 1], None),
    (mdtfx.XPREC_FP16_SPLIT, 200, 256, 4, 3, [0, 1, 0], None),    # Ns = T-1
    (mdtfx.XPREC_FP8, 64, 1024, 4, 32, [0, 1, 0], np.arange(0, 1024, 7)),
])
def test_stream_dumps_vs_fp64_oracle(xprec, NN, N, T, Ns, slots, chans):
    """every streamed dump (primed history, one slot reused) against the FP64
    chain (FIR -> FFT -> correlation, orc_fx_chain_f64) on that dump's window
    of the continuous sample stream; the FP64 accumulator against the sum"""
    D = len(slots)
    print(f"stream vs FP64 xprec={xprec} NN={NN} N={N} T={T} Ns={Ns} slots={slots}", flush=True)
    x, xh = _signal(NN, N, T, Ns, D, seed=7 * NN + Ns)
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, xprec)
    W = plan.window()
    torch.cuda.synchronize()
    W = W.cpu().numpy().astype(np.float64)
    st = mdtfx.FxStream(plan)
    acc = torch.zeros((N, plan.nb), dtype=torch.complex128, device="cuda")
    st.prime(xh[:T - 1])
    got, pending = [None] * D, {}
    for d, slot in enumerate(slots):
        if slot in pending:
            st.wait(slot)
            got[pending.pop(slot)] = st.vis(slot).copy()
        st.stack(slot)[...] = xh[T - 1 + d * Ns: T - 1 + (d + 1) * Ns]
        st.submit(slot, acc)
        pending[slot] = d
    for slot, d in pending.items():
        st.wait(slot)
        got[d] = st.vis(slot).copy()
    torch.cuda.synchronize()
    a = acc.cpu().numpy()
    st.destroy()
    ch = np.arange(N) if chans is None else chans
    bound = 1e-5 if xprec != mdtfx.XPREC_FP8 else 2.0 * 2.0 ** -4 + 0.35 / np.sqrt(Ns)
    total64 = 0
    for d in range(D):
        v64 = orc.fx_chain_f64(W, xh[d * Ns: d * Ns + Ns + T - 1], NN, N, T, Ns, ch)
        err = _autonorm_max(got[d][ch], v64, NN)
        print(f"  dump {d}: max error vs FP64 {err:.3e} (bound {bound:.3e})", flush=True)
        assert err < bound, (d, err)
        total64 = total64 + v64
    # the accumulator holds the sum of the dumps' visibilities (vs FP64 too)
    err = _autonorm_max(a[ch] / D, total64 / D, NN)
    assert err < bound, err


def test_stream_orders_after_caller_stream():
    """the accumulator is zeroed on a non-default torch stream behind ~30 ms
    of queued work, and the dump submitted at once: the dump must wait for the
    zeroing (a dump that ran first would add into NaNs, or be wiped)"""
    NN, N, T, Ns = 40, 512, 4, 64
    x, xh = _signal(NN, N, T, Ns, 1, seed=11)
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP32)
    st = mdtfx.FxStream(plan)  # not primed: zero history
    st.stack(0)[...] = xh[T - 1:]
    acc = torch.full((N, plan.nb), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")
    big = torch.randn(8192, 8192, device="cuda")
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for _ in range(3):
            big = (big @ big) * 1e-4
        acc.zero_()
        st.submit(0, acc)
    st.wait(0)
    torch.cuda.synchronize()
    got = st.vis(0).copy()
    st.destroy()
    z = x.clone()
    z[:T - 1] = 0
    ref = _per_dump(NN, N, T, Ns, mdtfx.XPREC_FP32, z, 0)
    assert np.array_equal(got, ref)
    assert np.array_equal(acc.cpu().numpy(), ref.astype(np.complex128))


def test_plan_follows_torch_current_stream():
    """FxPlan launches on torch's current stream: a dump queued on a
    non-default stream right behind the copy that produces its input (itself
    behind ~30 ms of work) must read the copied samples"""
    NN, N, T, Ns = 64, 1024, 4, 256
    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP16_SPLIT)
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=12, sigma=16.0, rho=0.3, max_delay=4)
    ref = plan.dump(x8, plan.alloc_xinput(), plan.alloc_vis())
    torch.cuda.synchronize()
    ref = ref.cpu().numpy()
    xin, vis = plan.alloc_xinput(), plan.alloc_vis()
    big = torch.randn(8192, 8192, device="cuda")
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        x8b = torch.zeros_like(x8)
        for _ in range(3):
            big = (big @ big) * 1e-4
        x8b.copy_(x8)
        plan.dump(x8b, xin, vis)
    s.synchronize()
    assert np.array_equal(vis.cpu().numpy(), ref)
