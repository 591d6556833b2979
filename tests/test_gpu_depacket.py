"""GPU depacketiser (the client's rank-1 loop) against the oracle's
restatement (MDT_OneStation_client.cu:846-911, 945-1005): bit-exact stacks
and sample counts for shuffled packet streams in both modes, and malformed
packets dropped, counted and never written."""
import numpy as np
import pytest
import torch

import mdtfx
import oracle as orc

pytestmark = pytest.mark.gpu

SNAP_INPUTS = (10, 10, 10, 8)


def _packets(rng, stations, cnts, S, words):
    pk = []
    for st in range(stations):
        for sn in range(4):
            for c in cnts:
                p = np.zeros(words, np.uint32)
                p[:4] = (st, sn, c, 0)
                n = 2 * S * SNAP_INPUTS[sn]
                p[4:4 + n] = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
                pk.append(p)
    pk = np.array(pk)
    pk[rng.permutation(len(pk))[:3], 4] = 0x80000000  # the +4096.0 quirk, a few times
    return pk[rng.permutation(len(pk))]


def _gpu(pk, S, NN, N, Ns, T, first):
    st, cn = mdtfx.depacketise(torch.from_numpy(pk.view(np.int32)).cuda(), S, NN, N, Ns, T, first)
    torch.cuda.synchronize()
    return st.cpu().numpy(), cn.cpu().numpy()


@pytest.mark.parametrize("first", [True, False])
# (1, 128, 4, 8, 16): a first-block history of T-1 = 7 spectra is longer than a
# 4-spectrum stack; the reference would write past its stacks, both sides skip
@pytest.mark.parametrize("stations,N,Ns,T,S", [(2, 64, 8, 4, 8), (1, 128, 4, 8, 16), (3, 32, 6, 5, 4)])
def test_depacketise_matches_oracle(first, stations, N, Ns, T, S):
    rng = np.random.default_rng(stations * 100 + N + S + first)
    NN = 38 * stations
    words = 4 + 2 * S * 10
    begins = (T - 1) * N
    if first:  # history (T-1)*N samples, then one stack of Ns*N
        cnts = range((begins + Ns * N) // S)
    else:      # two stacks' worth of steady-state samples
        cnts = range(begins // S, begins // S + 2 * Ns * N // S)
    pk = _packets(rng, stations, cnts, S, words)
    ref, ns = orc.depacketise(pk, S, NN, N, Ns, T, first_block=first)
    got, cnt = _gpu(pk, S, NN, N, Ns, T, first)
    assert np.array_equal(got.view(np.uint32), ref.astype(np.complex64).view(np.uint32))
    assert list(cnt) == list(ns)


def test_malformed_packets_dropped():
    rng = np.random.default_rng(7)
    stations, N, Ns, T, S = 1, 64, 4, 4, 8
    NN, words, begins = 38, 4 + 2 * S * 10, 3 * 64
    cnts = range(begins // S, begins // S + Ns * N // S)
    good = _packets(rng, stations, cnts, S, words)
    bad = np.zeros((4, words), np.uint32)
    bad[:, 4:] = 0x00080000  # 1.0
    bad[0, :3] = (0, 4, begins // S)          # SNAP 4 does not exist
    bad[1, :3] = (1, 0, begins // S)          # station 1: elements 38.. outside NN = 38
    bad[2, :3] = (0, 0, 0)                    # steady state, before the first sample
    bad[3, :3] = (0xFFFFFFFF, 1, begins // S)  # absurd station
    short = np.zeros((1, words - 2), np.uint32)  # payload larger than the packet
    ref, ns = orc.depacketise(good, S, NN, N, Ns, T, first_block=False)
    st, cn = mdtfx.depacketise(torch.from_numpy(good.view(np.int32)).cuda(), S, NN, N, Ns, T, False)
    st, cn = mdtfx.depacketise(torch.from_numpy(bad.view(np.int32)).cuda(), S, NN, N, Ns, T, False, st, cn)
    st, cn = mdtfx.depacketise(torch.from_numpy(short.view(np.int32)).cuda(), S, NN, N, Ns, T, False, st, cn)
    torch.cuda.synchronize()
    got, cnt = st.cpu().numpy(), cn.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), ref.astype(np.complex64).view(np.uint32))
    assert list(cnt[:2]) == list(ns[:2]) and ns[2] == 0
    assert cnt[2] == 4 * S + S  # four bad headers and one short packet, S samples each
    # the oracle applies the same rule to the same malformed packets
    _, ns_bad = orc.depacketise(bad, S, NN, N, Ns, T, first_block=False)
    assert list(ns_bad) == [0, 0, 4 * S]
    with pytest.raises(mdtfx.MdtError):
        mdtfx.depacketise(torch.from_numpy(good.view(np.int32)).cuda(), 0, NN, N, Ns, T, False)


def test_depacketise_throughput():
    """one station, 2048 channels, 512 spectra per stack (1.1e6 packets' worth of samples)"""
    rng = np.random.default_rng(3)
    stations, N, Ns, T, S = 1, 2048, 512, 4, 16
    NN, words, begins = 38, 4 + 2 * S * 10, 3 * 2048
    npk_per = Ns * N // S
    pk = np.zeros((4 * npk_per, words), np.uint32)
    for sn in range(4):
        blk = pk[sn * npk_per:(sn + 1) * npk_per]
        blk[:, 1] = sn
        blk[:, 2] = np.arange(begins // S, begins // S + npk_per)
        blk[:, 4:] = rng.integers(0, 2 ** 32, (npk_per, words - 4), dtype=np.uint64).astype(np.uint32)
    d = torch.from_numpy(pk.view(np.int32)).cuda()
    stacks = torch.zeros((2, Ns, NN, N), dtype=torch.complex64, device="cuda")
    counts = torch.zeros(3, dtype=torch.int64, device="cuda")
    mdtfx.depacketise(d, S, NN, N, Ns, T, False, stacks, counts)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        mdtfx.depacketise(d, S, NN, N, Ns, T, False, stacks, counts)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    print(f"depacketiser: {pk.shape[0]} packets ({pk.nbytes / 1e6:.0f} MB) in {ms:.3f} ms = "
          f"{pk.nbytes / ms / 1e6:.1f} GB/s of packets, {38 * Ns * N / ms / 1e3:.0f} Msamples/s")
    assert counts.cpu().numpy()[2] == 0
