"""The MDT station's real-time shape on the drop-in calls (40 inputs, 4-tap
PFB, 2048 channels, 4096 spectra per dump, 8 dumps; int8 samples sigma 16
with a correlated common component, rho 0.3): per dump the reference's call
sequence PpS_Batch -> cufftExecC2C -> ReorderFOutput -> cublasCgemmBatched
(MDT_OneStation_client.cu:1083-1128) on the float F input, the history carry
(client.cu:1170) between dumps.  Dumps 0 and 7 against the FP64 oracle chain;
every dump against the native fused FP32 path on the same samples."""
import numpy as np
import pytest
import torch

import mdtfx
import oracle as orc

pytestmark = pytest.mark.gpu

NN, N, T, NS, D = 40, 2048, 4, 4096, 8


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _tri_autonorm_err(v, v64):
    a, b = np.tril_indices(NN)
    diag = np.array([orc.tri_index(i, i) for i in range(NN)])
    auto = np.real(v64[:, diag])
    return float(np.max(np.abs(v - v64) / np.sqrt(auto[:, a] * auto[:, b])))


def test_station_shape_dropin_eight_dumps():
    print(f"station NN={NN} N={N} T={T} Ns={NS} dumps={D}", flush=True)
    x8 = mdtfx.synth_i8(NN, N, D * NS + T - 1, seed=2019, sigma=16.0, rho=0.3, max_delay=4)
    W = mdtfx.pfb_window(N, T, mdtfx.WINDOW_BH4)
    stream = torch.cuda.current_stream().cuda_stream
    fft = mdtfx.FftPlan(N, NS * NN)
    fft.set_stream(stream)
    xe = mdtfx.XEngine()
    xe.set_stream(stream)
    plan = mdtfx.FxPlan(NN, N, T, NS, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP32)
    plan.set_stream(stream)
    xin = plan.alloc_xinput()
    vis = plan.alloc_vis()

    # the reference's d_FInput: complex64 [Ns+T-1][NN][N], first block = the
    # first Ns+T-1 spectra, then the history carry and Ns new spectra per dump
    f_in = torch.view_as_complex(x8[:NS + T - 1].float().contiguous())
    F = torch.empty((NS, NN, N), dtype=torch.complex64, device="cuda")
    XOut = torch.empty((N, NN, NN), dtype=torch.complex64, device="cuda")
    Cptr = torch.tensor([XOut.data_ptr() + c * NN * NN * 8 for c in range(N)], dtype=torch.int64, device="cuda")
    a, b = np.tril_indices(NN)
    chans = np.array([0, 1, 511, 1024, 1500, 2047])
    for d in range(D):
        if d > 0:
            mdtfx.history_shift(f_in, NN, N, T, NS)
            f_in[T - 1:] = torch.view_as_complex(x8[T - 1 + d * NS: T - 1 + (d + 1) * NS].float())
        mdtfx.pps_batch(W, f_in, NN, NS, F)
        fft.exec_c2c(F, F, mdtfx.FFT_FORWARD)
        XIn = mdtfx.reorder_foutput(F, NN, N, NS)
        Aptr = torch.tensor([XIn.data_ptr() + c * NN * NS * 8 for c in range(N)], dtype=torch.int64, device="cuda")
        xe.cgemm_batched(mdtfx.OP_N, mdtfx.OP_C, NN, NN, NS, 1.0 / NS, Aptr, NN, Aptr, NN, 0.0, Cptr, NN, N)
        dump = x8[d * NS: d * NS + NS + T - 1].contiguous()
        plan.dump(dump, xin, vis)
        torch.cuda.synchronize()
        Vd = XOut.cpu().numpy()[:, b, a]  # column-major [c][col b][row a] -> triangle (a >= b)
        Vn = vis.cpu().numpy()
        e_nat = _tri_autonorm_err(Vd[chans].astype(np.complex128), Vn[chans].astype(np.complex128))
        msg = f"dump {d}: drop-in vs native {e_nat:.3e}"
        assert e_nat < 1e-5, msg
        if d in (0, D - 1):
            v64 = orc.fx_chain_f64(W.cpu().numpy().astype(np.float64), dump.cpu().numpy(), NN, N, T, NS, chans)
            e64 = _tri_autonorm_err(Vd[chans], v64)
            msg += f", drop-in vs FP64 {e64:.3e}"
            assert e64 < 1e-5, msg
        if d == 0:
            # x = sqrt(1 - rho) n + sqrt(rho) c(t - d_e): the normalised cross
            # power of every baseline is ~rho; Cauchy-Schwarz holds everywhere
            diag = np.array([orc.tri_index(i, i) for i in range(NN)])
            auto = np.real(Vd[:, diag]).astype(np.float64)
            coef = np.abs(Vd) / np.sqrt(auto[:, a] * auto[:, b])
            assert np.all(coef <= 1.0 + 1e-6)
            rho = float(coef[:, a != b].mean())
            msg += f", mean normalised cross power {rho:.4f}"
            assert abs(rho - 0.3) < 0.01, msg
        print(msg, flush=True)


def test_deployed_shape_dropin_and_native():
    """The operating point the reference ships in its configuration
    (MDT_SystemParams.conf:49-66: 19 elements x 2 pols = 38 inputs, 1024
    channels, 8-tap PFB, 1024 spectra per dump): two dumps through the drop-in
    sequence with the history carry, and the native FP32 and FP16-split paths
    on the same windows, every channel against the FP64 oracle chain."""
    NN2, N2, T2, NS2, D2 = 38, 1024, 8, 1024, 2
    print(f"deployed NN={NN2} N={N2} T={T2} Ns={NS2} dumps={D2}", flush=True)
    x8 = mdtfx.synth_i8(NN2, N2, D2 * NS2 + T2 - 1, seed=1907, sigma=16.0, rho=0.3, max_delay=4)
    W = mdtfx.pfb_window(N2, T2, mdtfx.WINDOW_BH4)
    W64 = W.cpu().numpy().astype(np.float64)
    stream = torch.cuda.current_stream().cuda_stream
    fft = mdtfx.FftPlan(N2, NS2 * NN2)
    fft.set_stream(stream)
    xe = mdtfx.XEngine()
    xe.set_stream(stream)
    f_in = torch.view_as_complex(x8[:NS2 + T2 - 1].float().contiguous())
    F = torch.empty((NS2, NN2, N2), dtype=torch.complex64, device="cuda")
    XIn = torch.empty((N2, NS2, NN2), dtype=torch.complex64, device="cuda")
    XOut = torch.empty((N2, NN2, NN2), dtype=torch.complex64, device="cuda")
    Aptr = torch.tensor([XIn.data_ptr() + c * NN2 * NS2 * 8 for c in range(N2)], dtype=torch.int64, device="cuda")
    Cptr = torch.tensor([XOut.data_ptr() + c * NN2 * NN2 * 8 for c in range(N2)], dtype=torch.int64, device="cuda")
    a, b = np.tril_indices(NN2)
    diag = np.array([orc.tri_index(i, i) for i in range(NN2)])
    chans = np.arange(N2)

    def err(v, v64):
        auto = np.real(v64[:, diag])
        return float(np.max(np.abs(v - v64) / np.sqrt(auto[:, a] * auto[:, b])))

    plans = {p: mdtfx.FxPlan(NN2, N2, T2, NS2, mdtfx.WINDOW_BH4, p) for p in (mdtfx.XPREC_FP32, mdtfx.XPREC_FP16_SPLIT)}
    for d in range(D2):
        if d > 0:
            mdtfx.history_shift(f_in, NN2, N2, T2, NS2)
            f_in[T2 - 1:] = torch.view_as_complex(x8[T2 - 1 + d * NS2: T2 - 1 + (d + 1) * NS2].float())
        mdtfx.pps_batch(W, f_in, NN2, NS2, F)
        fft.exec_c2c(F, F, mdtfx.FFT_FORWARD)
        mdtfx.reorder_foutput(F, NN2, N2, NS2, XIn)
        xe.cgemm_batched(mdtfx.OP_N, mdtfx.OP_C, NN2, NN2, NS2, 1.0 / NS2, Aptr, NN2, Aptr, NN2, 0.0, Cptr, NN2, N2)
        dump = x8[d * NS2: d * NS2 + NS2 + T2 - 1].contiguous()
        torch.cuda.synchronize()
        v64 = orc.fx_chain_f64(W64, dump.cpu().numpy(), NN2, N2, T2, NS2, chans)
        e = {"drop-in": err(XOut.cpu().numpy()[:, b, a].astype(np.complex128), v64)}
        for p, plan in plans.items():
            vis = plan.dump(dump, plan.alloc_xinput(), plan.alloc_vis())
            e["fp32" if p == mdtfx.XPREC_FP32 else "fp16-split"] = err(_host(vis).astype(np.complex128), v64)
        print(f"dump {d}: max normalised error vs FP64 " + ", ".join(f"{k} {v:.3e}" for k, v in e.items()), flush=True)
        assert max(e.values()) < 1e-5, e
