"""Streaming correlator (the client's rank loop, MDT_OneStation_client.cu:
916-1276, in one process): dumps pushed through the two pinned host slots must
give, bit for bit, what one-dump calls give on the same windows of one
continuous sample stream -- dump d correlates spectra [d*Ns, d*Ns + Ns + T-1)
of it, its first T-1 spectra being the previous dump's last ones (the history
shift of client.cu:1170).  The FP64 accumulator must hold the in-order sum of
the dumps' visibilities."""
import numpy as np
import pytest
import torch

import mdtfx

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
