"""Packet streaming correlator (mdt_pstream_*): the client's whole per-dump
loop on its own wire data (MDT_OneStation_client.cu:846-1276) -- SNAP packets
of Fix_32_19 word pairs, decoded and scattered on the device into the float F
input behind the PFB history (first block: history from the packets,
client.cu:855-916; later dumps: carried on the device, client.cu:1170), the
float-input F-engine, the X-engine, visibilities and sample counts back to the
host.  Three dumps through the two slots (one reused), packets shuffled, late,
malformed and foreign-station packets among them.  Each dump is compared bit
for bit with the int8 fused path on the same samples (the words carry int8
values exactly) and with the FP64 chain oracle; the counts with the numbers
of samples each dump's packets carry."""
import numpy as np
import pytest
import torch

import mdtfx
import oracle as orc

pytestmark = pytest.mark.gpu

SNAP_INPUTS = (10, 10, 10, 8)
SNAP_FIRST = (0, 10, 20, 30)


def _words(v):
    """int values -> Fix_32_19 words (two's complement, 2^-19 units)"""
    return ((v.astype(np.int64) << 19) & 0xFFFFFFFF).astype(np.uint32)


def _packets(x8, cnts, S, words, station=0):
    """packets of every SNAP of one station for the given counters; sample i
    of counter c is element-sample t = c*S + i of the int8 frames x8
    [frames][NN][N][2] (t = frame*N + channel)"""
    N = x8.shape[2]
    out = []
    for sn in range(4):
        nin, e0 = SNAP_INPUTS[sn], SNAP_FIRST[sn]
        t = cnts[:, None] * S + np.arange(S)[None, :]                 # [cnt][S]
        f, n = t // N, t % N
        v = x8[f[:, :, None], (e0 + np.arange(nin))[None, None, :], n[:, :, None]]  # [cnt][S][nin][2]
        p = np.zeros((len(cnts), words), np.uint32)
        p[:, 0], p[:, 1], p[:, 2] = station, sn, cnts
        p[:, 4:4 + 2 * S * nin] = _words(v.reshape(len(cnts), -1))
        out.append(p)
    return np.concatenate(out)


@pytest.mark.parametrize("xprec", [mdtfx.XPREC_FP32, mdtfx.XPREC_FP16_SPLIT])
def test_packet_stream_three_dumps(xprec):
    NN, N, T, Ns, D, S = 38, 512, 4, 16, 3, 16
    words = 4 + 2 * S * 10
    begins, per_dump = (T - 1) * N, Ns * N
    print(f"packet stream xprec={xprec} NN={NN} N={N} T={T} Ns={Ns} dumps={D}", flush=True)
    nfr = T - 1 + D * Ns
    x8g = mdtfx.synth_i8(NN, N, nfr, seed=77, sigma=16.0, rho=0.3, max_delay=3, tones=((0.21, 10.0),))
    torch.cuda.synchronize()
    x8 = x8g.cpu().numpy()
    rng = np.random.default_rng(78)
    # counters of each dump's samples: t = u + begins in [d*per_dump, (d+1)*per_dump) + begins,
    # dump 0 from t = 0 (its history)
    bounds = [0] + [(d + 1) * per_dump + begins for d in range(D)]
    dump_pk = []
    for d in range(D):
        cnts = np.arange(bounds[d] // S, bounds[d + 1] // S)
        pk = _packets(x8, cnts, S, words)
        if d == 1:  # late packets of dump 0, a bad SNAP index, a station beyond NN
            late = _packets(x8, np.arange(bounds[1] // S - 3, bounds[1] // S), S, words)[:3]
            bad = pk[:1].copy()
            bad[0, 1] = 5
            foreign = _packets(x8, cnts[:1], S, words, station=1)[:1]
            pk = np.concatenate([pk, late, bad, foreign])
        dump_pk.append(pk[rng.permutation(len(pk))])
    max_pkts = max(len(p) for p in dump_pk)

    plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, xprec)
    W = plan.window()
    torch.cuda.synchronize()
    W = W.cpu().numpy().astype(np.float64)
    ps = mdtfx.PacketStream(plan, words, S, max_pkts)
    acc = torch.zeros((N, plan.nb), dtype=torch.complex128, device="cuda")
    got, counts, pending = [None] * D, [None] * D, {}
    for d, slot in enumerate([0, 1, 0]):
        if slot in pending:
            t, c = ps.wait(slot)
            dd = pending.pop(slot)
            got[dd], counts[dd] = ps.vis(slot).copy(), c
            assert t["total_ms"] > 0 and t["fengine_ms"] > 0
        ps.packets(slot)[:len(dump_pk[d])] = dump_pk[d]
        ps.submit(slot, len(dump_pk[d]), acc)
        pending[slot] = d
    for slot, d in pending.items():
        _, c = ps.wait(slot)
        got[d], counts[d] = ps.vis(slot).copy(), c
    torch.cuda.synchronize()
    a = acc.cpu().numpy()
    ps.destroy()

    full = 4 * S * (per_dump // S)  # per (packet, sample): 4 SNAPs
    assert counts[0] == [full, 4 * begins, 0], counts[0]
    assert counts[1] == [full, 0, 5 * S], counts[1]
    assert counts[2] == [full, 0, 0], counts[2]
    ref_plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, xprec)
    a_, b_ = np.tril_indices(NN)
    diag = np.array([orc.tri_index(i, i) for i in range(NN)])
    total = np.zeros_like(a)
    for d in range(D):
        win = x8g[d * Ns: d * Ns + Ns + T - 1].contiguous()
        ref = ref_plan.dump(win, ref_plan.alloc_xinput(), ref_plan.alloc_vis())
        torch.cuda.synchronize()
        ref = ref.cpu().numpy()
        assert np.array_equal(got[d], ref), f"dump {d}: packet stream != int8 fused path"
        v64 = orc.fx_chain_f64(W, x8[d * Ns: d * Ns + Ns + T - 1], NN, N, T, Ns, np.arange(N))
        auto = np.real(v64[:, diag])
        err = float(np.max(np.abs(got[d] - v64) / np.sqrt(auto[:, a_] * auto[:, b_])))
        print(f"  dump {d}: vs FP64 {err:.3e}, counts {counts[d]}", flush=True)
        assert err < 1e-5, (d, err)
        total += ref.astype(np.complex128)
    assert np.array_equal(a, total)


def test_fengine_f32_equals_int8_and_dropin():
    """the float-input F-engine on int8-valued floats gives the int8 F-engine's
    X input bit for bit (N = 1024: persistent int8 kernel against the generic
    float kernel), FP32 and FP16 split; FP8 plans refuse float input"""
    NN, N, T, Ns = 64, 1024, 8, 40
    x8 = mdtfx.synth_i8(NN, N, Ns + T - 1, seed=79, sigma=20.0, rho=0.2, max_delay=4)
    fin = torch.view_as_complex(x8.float().contiguous())
    for xprec in (mdtfx.XPREC_FP32, mdtfx.XPREC_FP16_SPLIT):
        plan = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_HAMMING, xprec)
        xa = plan.fengine(x8, plan.alloc_xinput())
        xb = plan.fengine_f32(fin, plan.alloc_xinput())
        torch.cuda.synchronize()
        va = np.ascontiguousarray(plan.xinput_view(xa).cpu().numpy())
        vb = np.ascontiguousarray(plan.xinput_view(xb).cpu().numpy())
        assert np.array_equal(va.view(np.uint64), vb.view(np.uint64)), xprec
    p8 = mdtfx.FxPlan(NN, N, T, Ns, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP8)
    with pytest.raises(mdtfx.MdtError):
        p8.fengine_f32(fin, p8.alloc_xinput())
    with pytest.raises(mdtfx.MdtError):
        mdtfx.PacketStream(p8, 324, 16, 8)
