"""The client's per-dump path end to end on the GPU, two dumps: SNAP packets ->
GPU depacketiser -> F input = history + stack (client.cu:916, 1022, 1048) ->
the drop-in calls PpS_Batch, cufftExecC2C, ReorderFOutput and
cublasCgemmBatched (client.cu:1083-1128) -> visibilities -> the client's CSV
(client.cu:1239-1266), with the history carry between the dumps
(client.cu:1170).  The first dump is the first block (history in stack 1,
the spectra in stack 0); the second arrives in steady state and lands in
stack 1.  Checked against the oracle: its depacketiser on the same packets,
then its FP64 FIR, DFT, corner turn and correlation."""
import numpy as np
import pytest
import torch

import mdtfx
import oracle as orc

pytestmark = pytest.mark.gpu

SNAP_INPUTS = (10, 10, 10, 8)


def _packets(rng, cnts, S, words, amp=0.05):
    """one station, 4 SNAPs: Fix_32_19 words of Gaussian samples"""
    pk = []
    for sn in range(4):
        for c in cnts:
            p = np.zeros(words, np.uint32)
            p[:4] = (0, sn, c, 0)
            n = 2 * S * SNAP_INPUTS[sn]
            v = np.rint(rng.standard_normal(n) * amp * 2 ** 19).astype(np.int64)
            p[4:4 + n] = (v & 0xFFFFFFFF).astype(np.uint32)
            pk.append(p)
    pk = np.array(pk)
    return pk[rng.permutation(len(pk))]


def _autonorm_err(V, V64):
    d = np.real(np.einsum("caa->ca", V64))
    return float(np.max(np.abs(V - V64) / np.sqrt(d[:, :, None] * d[:, None, :])))


def _gpu_dump(W, F_in, NN, N, Ns, fft, xe):
    F = mdtfx.pps_batch(W, F_in, NN, Ns)
    fft.exec_c2c(F, F, mdtfx.FFT_FORWARD)
    XIn = mdtfx.reorder_foutput(F, NN, N, Ns)
    XOut = torch.full((N, NN, NN), complex(np.nan, np.nan), dtype=torch.complex64, device="cuda")
    Aptr = torch.tensor([XIn.data_ptr() + c * NN * Ns * 8 for c in range(N)], dtype=torch.int64, device="cuda")
    Cptr = torch.tensor([XOut.data_ptr() + c * NN * NN * 8 for c in range(N)], dtype=torch.int64, device="cuda")
    xe.cgemm_batched(mdtfx.OP_N, mdtfx.OP_C, NN, NN, Ns, 1.0 / Ns, Aptr, NN, Aptr, NN, 0.0, Cptr, NN, N)
    torch.cuda.synchronize()
    return XOut.cpu().numpy()


def _oracle_dump(W, F_in, NN, N, Ns):
    F = orc.pps_f64(W.astype(np.float64), F_in.astype(np.complex128), NN, Ns)
    F = orc.fft_f64(F.reshape(-1, N)).reshape(Ns, NN, N)
    XIn = orc.reorder(F, NN, N, Ns)
    return orc.xengine_full(XIn, NN, Ns, N)


def test_packets_to_visibilities_and_csv(tmp_path):
    Nel, Npol, N, Ns, T, S = 19, 2, 64, 16, 4, 8
    NN = Nel * Npol  # one station: 38 inputs
    words = 4 + 2 * S * 10
    begins = (T - 1) * N
    print(f"pipeline NN={NN} N={N} Ns={Ns} T={T} S={S}", flush=True)
    rng = np.random.default_rng(2026)
    c1 = range((begins + Ns * N) // S)                               # first block
    c2 = range((begins + Ns * N) // S, (begins + 2 * Ns * N) // S)    # next dump, steady state
    pk1, pk2 = _packets(rng, c1, S, words), _packets(rng, c2, S, words)

    W = mdtfx.pfb_window(N, T, mdtfx.WINDOW_BH4)
    fft = mdtfx.FftPlan(N, Ns * NN)
    xe = mdtfx.XEngine()
    stream = torch.cuda.current_stream().cuda_stream
    fft.set_stream(stream)
    xe.set_stream(stream)

    # dump 1: first block
    st1, cn1 = mdtfx.depacketise(torch.from_numpy(pk1.view(np.int32)).cuda(), S, NN, N, Ns, T, True)
    F_in = torch.cat([st1[1, :T - 1], st1[0]]).contiguous()  # [Ns+T-1][NN][N]
    V1 = _gpu_dump(W, F_in, NN, N, Ns, fft, xe)
    # dump 2: history shift, then the steady-state stack (stack 1)
    st2, cn2 = mdtfx.depacketise(torch.from_numpy(pk2.view(np.int32)).cuda(), S, NN, N, Ns, T, False)
    mdtfx.history_shift(F_in, NN, N, T, Ns)
    F_in[T - 1:] = st2[1]
    V2 = _gpu_dump(W, F_in, NN, N, Ns, fft, xe)

    # oracle on the same packets
    r1, n1 = orc.depacketise(pk1, S, NN, N, Ns, T, first_block=True)
    r2, n2 = orc.depacketise(pk2, S, NN, N, Ns, T, first_block=False)
    assert np.array_equal(st1.cpu().numpy(), r1.astype(np.complex64))
    assert np.array_equal(st2.cpu().numpy(), r2.astype(np.complex64))
    assert list(cn1.cpu().numpy()) == list(n1) and list(cn2.cpu().numpy()) == list(n2)
    assert n1[0] == 4 * Ns * N and n1[1] == 4 * begins and n2[1] == 4 * Ns * N  # S per packet, 4 SNAPs
    Wh = W.cpu().numpy()
    f1 = np.concatenate([r1[1, :T - 1], r1[0]])
    f2 = np.concatenate([f1[Ns:Ns + T - 1], r2[1]])
    for V, f in ((V1, f1), (V2, f2)):
        V64 = _oracle_dump(Wh, f, NN, N, Ns)
        assert np.all(np.isfinite(V))
        err = _autonorm_err(V, V64)
        print(f"max normalised error vs FP64: {err:.3e}")
        assert err < 1e-5

    # the client's CSV of dump 2: rows, order and values
    path = tmp_path / mdtfx.csv_filename(1733270399)
    mdtfx.write_csv(path, V2, mdtfx.VIS_FULL_COLMAJOR, Nel, Npol, N)
    rows = [r.split(",") for r in path.read_text().splitlines()]
    pairs = [(j, p, k, q) for j in range(Nel) for p in range(Npol) for k in range(j, Nel) for q in range(p, Npol)]
    assert len(rows) == len(pairs) * N
    flat = V2.reshape(-1)
    for idx in rng.integers(0, len(rows), 200):
        j, p, k, q, z = (int(v) for v in rows[idx][:5])
        assert (j, p, k, q) == pairs[idx // N] and z == idx % N
        v = flat[((z + N // 2) % N) * NN * NN + (j * Npol + p) + (k * Npol + q) * NN]
        assert rows[idx][5] == "%.6f" % float(v.real) and rows[idx][6] == "%.6f" % float(v.imag)
