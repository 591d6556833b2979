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
