"""FX correlator benchmark on one MI355X.

Workload (the scaled array): 256 inputs, 8-tap PFB, 4096-point FFT, 2048
spectra per dump, int8 samples sigma=20 with a common component (rho=0.3)
delayed per input by 0-8 samples; FP16 hi+lo split X-engine with FP32
accumulation and the FP64 cross-dump accumulator.  A step is one dump: the
fused F-engine (PFB FIR + FFT + corner turn) and the X-engine over samples
already resident in HBM.  The same dumps are timed with the FP32-MFMA
X-engine for the metric's comparison.

Prints ONE JSON line.  value = Msamples/s/input of the FP16-split path =
Nspectra * Nchannels / (seconds per dump) / 1e6.  Beside it, each skippable:
the host-inclusive streaming rates, int8 stacks and the reference's SNAP
packets (--no-stream), the CPU port and the FP64 error (--no-cpu), the
512-input FP8 wide-band shape (--no-fp8), the reference's drop-in call
sequence at the MDT station's shape with the FFT at lengths that are not
powers of two (--no-station; --station-only runs that leg alone), the
gain/delay calibration shape (--no-calib) and the HBM read / write / copy
rates this device delivers to a streaming kernel, with the F-engine's time
against the floor they set (--no-hbm-probe).  --fengine-shapes runs only the
F-engine across channel and tap counts (both of its kernels) and prints that
table instead.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mdt_fpga-gpu_pipeline_amd"))

METRIC = "FX Msamples/s/input at 256 inputs, FP16-split vs FP32 MFMA; max rel err vs FP64"
NN, N, T, NS = 256, 4096, 8, 2048
SEED, SIGMA, RHO, MAXDELAY = 2026, 20.0, 0.3, 8
HBM_PEAK_GBS = 8000.0
FP16_PEAK_TFLOPS = 2500.0
FP32_PEAK_TFLOPS = 157.3
FP8_PEAK_TFLOPS = 5000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_path(torch, mdtfx, plan, x8, xin, vis, acc, steps, warmup):
    """Run warmup + timed dumps; returns wall ms/dump and per-kernel mean ms
    from HIP events recorded on the stream the kernels run on."""
    for _ in range(warmup):
        plan.dump(x8, xin, vis, acc)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        ev[k][0].record()
        plan.fengine(x8, xin)
        ev[k][1].record()
        plan.xengine(xin, vis, acc)
        ev[k][2].record()
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    f_ms = sum(e[0].elapsed_time(e[1]) for e in ev) / steps
    x_ms = sum(e[1].elapsed_time(e[2]) for e in ev) / steps
    return wall_ms, f_ms, x_ms


def fp8_leg(np, torch, mdtfx, steps, warmup, stream, err_channels, with_err):
    """The wide-band shape (512 inputs, 4-tap PFB, 1024-pt FFT, 8192 spectra
    per dump, noise plus a tone ~40 dB above it) on the FP8 e4m3 X-engine:
    the F-engine writes e4m3 samples with one E8M0 exponent per (input, 16
    spectra), the MFMA's block scales; full dump resident in HBM."""
    NN4, N4, T4, NS4 = 512, 1024, 4, 8192
    tone_bin, tone_amp, sigma = 300, 12.0, 4.0
    log(f"bench: FP8 path, wide-band NN={NN4} N={N4} T={T4} Ns={NS4}")
    plan = mdtfx.FxPlan(NN4, N4, T4, NS4, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP8)
    plan.set_stream(stream)
    x8 = mdtfx.synth_i8(NN4, N4, NS4 + T4 - 1, seed=SEED, sigma=sigma, rho=RHO, max_delay=MAXDELAY,
                        tones=((tone_bin / N4, tone_amp),))
    xin = plan.alloc_xinput()
    vis = plan.alloc_vis()
    torch.cuda.synchronize()
    wall, f_ms, x_ms = time_path(torch, mdtfx, plan, x8, xin, vis, None, steps, warmup)
    nb4 = NN4 * (NN4 + 1) // 2
    f_bytes = NN4 * (NS4 + T4 - 1) * N4 * 2 + T4 * N4 * 4 + NN4 * NS4 * N4 * 2 + NN4 * (NS4 // 16)
    x_flops = 8.0 * N4 * NS4 * nb4
    out = {
        "workload": "wide-band: 512 inputs, 4-tap PFB, 1024-pt FFT, 8192 spectra/dump, FP8 e4m3 "
                    "X-engine, E8M0 block exponents per (input, 16 spectra), FP32 accumulate",
        "data": f"synthetic int8 sigma={sigma}, rho={RHO}, tone at bin {tone_bin} amplitude {tone_amp}, seed {SEED}",
        "value": NS4 * N4 / (wall * 1e-3) / 1e6,
        "unit": "Msamples/s/input",
        "ms_per_step": wall,
        "steps": steps,
        "kernels": {
            "fengine_e4m3out": {"ms": f_ms, "algorithmic_bytes": f_bytes,
                                "GB/s": f_bytes / (f_ms * 1e-3) / 1e9,
                                "frac_hbm": f_bytes / (f_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
            "xengine_fp8": {"ms": x_ms, "useful_flops": x_flops,
                            "useful_TFLOP/s": x_flops / (x_ms * 1e-3) / 1e12,
                            "frac_fp8_peak": x_flops / (x_ms * 1e-3) / 1e12 / FP8_PEAK_TFLOPS},
        },
    }
    if with_err:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc

        W = plan.window().cpu().numpy().astype(np.float64)
        chans = np.unique(np.concatenate([[tone_bin], np.linspace(0, N4 - 1, err_channels).astype(int)]))
        t0 = time.perf_counter()
        v64 = orc.fx_chain_f64(W, x8.cpu().numpy(), NN4, N4, T4, NS4, chans)
        t64 = time.perf_counter() - t0
        v8 = vis.cpu().numpy()[chans]
        a, b = np.tril_indices(NN4)
        diag = np.array([orc.tri_index(i, i) for i in range(NN4)])
        auto = np.real(v64[:, diag])
        e = np.max(np.abs(v8 - v64) / np.sqrt(auto[:, a] * auto[:, b]), axis=1)
        it = int(np.nonzero(chans == tone_bin)[0][0])
        noise = [i for i in range(len(chans)) if i != it]
        out["max_rel_err_vs_fp64"] = {
            "tone_channel": float(e[it]), "noise_channels_max": float(e[noise].max()),
            "tone_dB_above_noise": float(10 * np.log10(auto[it].mean() / np.median(auto[noise].mean(axis=1)))),
            "channels_checked": [int(c) for c in chans], "oracle_seconds": round(t64, 1),
            "definition": "max over (a>=b) of |V - V64| / sqrt(V64_aa V64_bb); e4m3 rounding is "
                          "coherent for a bin-centred tone, so its channel carries up to 2 x 2^-4"}
    del plan, x8, xin, vis
    torch.cuda.empty_cache()
    return out


def stream_leg(torch, mdtfx, x8, dumps, err_channels=0):
    """The same scaled-array dumps through the streaming correlator: int8 stacks from
    the two pinned host slots over PCIe (the history carried on the device),
    F+X, visibilities back to pinned host memory, copies and compute on
    separate streams.  Reports the host-inclusive rate beside the
    device-resident value.  The host stacks are filled once; every dump
    copies a full stack.  With err_channels: the last dump's visibilities
    against the FP64 oracle on its window (the carried history, the stack)."""
    plan = mdtfx.FxPlan(NN, N, T, NS, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP16_SPLIT)
    st = mdtfx.FxStream(plan)
    for slot in range(2):
        torch.from_numpy(st.stack(slot)).copy_(x8[T - 1:])
    st.prime(x8[:T - 1].cpu().numpy())
    acc = torch.zeros((N, plan.nb), dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()

    def run(n):
        times = []
        for d in range(n):
            if d >= 2:
                times.append(st.wait(d % 2))
            st.submit(d % 2, acc)
        for d in range(max(0, n - 2), n):
            times.append(st.wait(d % 2))
        return times

    run(2)
    t0 = time.perf_counter()
    times = run(dumps)
    wall = (time.perf_counter() - t0) * 1e3 / dumps
    vis_last = st.vis((dumps - 1) % 2).copy()
    st.destroy()
    mean = {k: sum(t[k] for t in times) / len(times) for k in times[0]}
    stack_bytes = NS * NN * N * 2
    out = {"value": NS * N / (wall * 1e-3) / 1e6, "unit": "Msamples/s/input", "ms_per_dump": wall,
           "dumps": dumps, "phase_ms": mean, "h2d_GB/s": stack_bytes / (mean["h2d_ms"] * 1e-3) / 1e9,
           "stack_bytes": stack_bytes}
    if err_channels:
        import numpy as np
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc

        # every dump after the first carries the stack's last T-1 spectra
        xh = x8.cpu().numpy()
        window = np.concatenate([xh[-(T - 1):], xh[T - 1:]])
        chans = np.unique(np.linspace(0, N - 1, err_channels).astype(int))
        t0 = time.perf_counter()
        v64 = orc.fx_chain_f64(plan.window().cpu().numpy().astype(np.float64), window, NN, N, T, NS, chans)
        out["max_rel_err_vs_fp64"] = {
            "last_dump_fp16_split": _autonorm_err(np, orc, vis_last[chans], v64, NN),
            "channels_checked": len(chans), "oracle_seconds": round(time.perf_counter() - t0, 1),
            "definition": "max over (channel, a>=b) of |V - V64| / sqrt(V64_aa V64_bb)"}
    return out


def station_leg(np, torch, mdtfx, steps, warmup, err_channels=0):
    """The MDT station shape (40 inputs, 4-tap PFB, 2048 channels, 4096
    spectra per dump) on the drop-in calls, issued in the reference's rank-2
    order per dump (MDT_OneStation_client.cu:1170 history carry, then 1083
    PpS_Batch on the float F input, 1105 C2C FFT in place, 1114
    ReorderFOutput, 1128 cublasCgemmBatched over the full NN x NN with
    alpha = 1/Ns), every call bracketed by HIP events on the stream it runs
    on; beside it the native fused path on the same samples.  Per-call
    rooflines use each call's minimum bytes (or the GEMM's flops); FP64 error
    of the drop-in and native visibilities on err_channels channels."""
    NN2, N2, T2, NS2 = 40, 2048, 4, 4096
    log(f"bench: drop-in station leg NN={NN2} N={N2} T={T2} Ns={NS2}")
    x8 = mdtfx.synth_i8(NN2, N2, NS2 + T2 - 1, seed=SEED, sigma=16.0, rho=RHO, max_delay=4)
    W = mdtfx.pfb_window(N2, T2, mdtfx.WINDOW_BH4)
    stream = torch.cuda.current_stream().cuda_stream
    fft = mdtfx.FftPlan(N2, NS2 * NN2)
    fft.set_stream(stream)
    xe = mdtfx.XEngine()
    xe.set_stream(stream)
    f_in = torch.view_as_complex(x8.float().contiguous())  # [Ns+T-1][NN][N] complex64
    F = torch.empty((NS2, NN2, N2), dtype=torch.complex64, device="cuda")
    XIn = torch.empty((N2, NS2, NN2), dtype=torch.complex64, device="cuda")
    XOut = torch.empty((N2, NN2, NN2), dtype=torch.complex64, device="cuda")
    Aptr = torch.tensor([XIn.data_ptr() + c * NN2 * NS2 * 8 for c in range(N2)], dtype=torch.int64, device="cuda")
    Cptr = torch.tensor([XOut.data_ptr() + c * NN2 * NN2 * 8 for c in range(N2)], dtype=torch.int64, device="cuda")
    pts = NS2 * NN2 * N2
    calls = [  # name, call, algorithmic bytes, flops
        ("history_shift", lambda: mdtfx.history_shift(f_in, NN2, N2, T2, NS2), 2 * (T2 - 1) * NN2 * N2 * 8, 0),
        ("pps_batch", lambda: mdtfx.pps_batch(W, f_in, NN2, NS2, F),
         (NS2 + T2 - 1) * NN2 * N2 * 8 + T2 * N2 * 4 + pts * 8, 0),
        ("fft_c2c", lambda: fft.exec_c2c(F, F, mdtfx.FFT_FORWARD), 16 * pts, 0),
        ("reorder_foutput", lambda: mdtfx.reorder_foutput(F, NN2, N2, NS2, XIn), 16 * pts, 0),
        ("cgemm_batched", lambda: xe.cgemm_batched(mdtfx.OP_N, mdtfx.OP_C, NN2, NN2, NS2, 1.0 / NS2, Aptr, NN2,
                                                   Aptr, NN2, 0.0, Cptr, NN2, N2),
         pts * 8 + N2 * NN2 * NN2 * 8, 8.0 * N2 * NS2 * NN2 * NN2),
    ]
    for _ in range(warmup):
        for c in calls:
            c[1]()
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(len(calls) + 1)] for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        ev[k][0].record()
        for i, c in enumerate(calls):
            c[1]()
            ev[k][i + 1].record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    per_call = {}
    for i, (name, _, nbytes, flops) in enumerate(calls):
        ms = sum(e[i].elapsed_time(e[i + 1]) for e in ev) / steps
        r = {"ms": ms, "algorithmic_bytes": nbytes, "GB/s": nbytes / (ms * 1e-3) / 1e9,
             "frac_hbm": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        if flops:
            r.update({"flops": flops, "TFLOP/s": flops / (ms * 1e-3) / 1e12,
                      "frac_fp32_mfma": flops / (ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS})
        per_call[name] = r
    native, native_vis = {}, {}
    for label, xprec in (("fp32", mdtfx.XPREC_FP32), ("fp16_split", mdtfx.XPREC_FP16_SPLIT)):
        plan = mdtfx.FxPlan(NN2, N2, T2, NS2, mdtfx.WINDOW_BH4, xprec)
        plan.set_stream(stream)
        xin = plan.alloc_xinput()
        vis = plan.alloc_vis()
        w_ms, f_ms, x_ms = time_path(torch, mdtfx, plan, x8, xin, vis, None, steps, warmup)
        native[label] = {"value": NS2 * N2 / (w_ms * 1e-3) / 1e6, "ms_per_dump": w_ms, "fengine_ms": f_ms,
                         "xengine_ms": x_ms}
        native_vis[label] = vis.cpu().numpy()
        del plan, xin, vis
    out = {"workload": "MDT station: 40 inputs, 4-tap PFB (BH4), 2048-pt FFT, 4096 spectra/dump; the "
                       "reference's four drop-in calls + history carry per dump on the float F input",
           "value": NS2 * N2 / (wall * 1e-3) / 1e6, "unit": "Msamples/s/input", "ms_per_dump": wall,
           "steps": steps, "calls": per_call, "native_fused": native}
    if err_channels:
        # the drop-in dump read the F input after its history carry: its first
        # T-1 spectra are the last T-1 of the buffer
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc

        xh = x8.cpu().numpy()
        xd = xh.copy()
        xd[:T2 - 1] = xh[NS2:NS2 + T2 - 1]
        chans = np.unique(np.linspace(0, N2 - 1, err_channels).astype(int))
        Wh = W.cpu().numpy().astype(np.float64)
        t0 = time.perf_counter()
        v64d = orc.fx_chain_f64(Wh, xd, NN2, N2, T2, NS2, chans)
        v64n = orc.fx_chain_f64(Wh, xh, NN2, N2, T2, NS2, chans)
        t64 = time.perf_counter() - t0
        a, b = np.tril_indices(NN2)
        vd = XOut.cpu().numpy()[chans][:, b, a]  # column-major [c][col b][row a] -> triangle
        out["max_rel_err_vs_fp64"] = {
            "dropin_cgemm": _autonorm_err(np, orc, vd, v64d, NN2),
            "native_fp32": _autonorm_err(np, orc, native_vis["fp32"][chans], v64n, NN2),
            "native_fp16_split": _autonorm_err(np, orc, native_vis["fp16_split"][chans], v64n, NN2),
            "channels_checked": len(chans), "oracle_seconds": round(t64, 1),
            "definition": "max over (channel, a>=b) of |V - V64| / sqrt(V64_aa V64_bb)"}
    del fft, xe, f_in, F, XIn, XOut, x8
    torch.cuda.empty_cache()
    # the drop-in FFT at channel counts that are not powers of two, same
    # number of rows: mixed radix (1000 = 2^3 5^3) and Bluestein (1009, prime)
    out["fft_non_pow2"] = {}
    rows = NS2 * NN2
    for n, how in ((1000, "mixed radix 8 x 5 x 5 x 5"), (1009, "Bluestein, 2048-point convolution")):
        fp = mdtfx.FftPlan(n, rows)
        fp.set_stream(stream)
        buf = torch.randn((rows, n), dtype=torch.complex64, device="cuda")
        for _ in range(warmup):
            fp.exec_c2c(buf, buf, mdtfx.FFT_FORWARD)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fp.exec_c2c(buf, buf, mdtfx.FFT_FORWARD)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        nbytes = 16 * rows * n
        out["fft_non_pow2"][str(n)] = {"algorithm": how, "rows": rows, "ms": ms, "algorithmic_bytes": nbytes,
                                       "GB/s": nbytes / (ms * 1e-3) / 1e9,
                                       "frac_hbm": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        fp.destroy()
        del buf
    torch.cuda.empty_cache()
    return out


def _autonorm_err(np, orc, v, v64, nn):
    a, b = np.tril_indices(nn)
    diag = np.array([orc.tri_index(i, i) for i in range(nn)])
    auto = np.real(v64[:, diag])
    return float(np.max(np.abs(v - v64) / np.sqrt(auto[:, a] * auto[:, b])))


def calib_leg(np, torch, mdtfx, steps, warmup, with_err):
    """The calibration shape: 64 inputs, 4-tap PFB, 1024-pt FFT, 512 spectra;
    per-input gains |g| ~ U(0.8, 1.2), phase ~ U(-pi, pi), delays ~ U(-4, 4)
    samples applied in the F-engine; FP32 X-engine; the loss sum |V' - M|^2
    and its hand-written adjoint (mdt_calib_loss_grad) on the raw
    visibilities.  Each kernel timed with HIP events on its stream."""
    from mdtfx.calib import CalibPlan

    NN5, N5, T5, NS5 = 64, 1024, 4, 512
    log(f"bench: calibration leg NN={NN5} N={N5} T={T5} Ns={NS5}")
    rng = np.random.default_rng(55)
    plan = mdtfx.FxPlan(NN5, N5, T5, NS5, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP32)
    x8 = mdtfx.synth_i8(NN5, N5, NS5 + T5 - 1, seed=SEED, sigma=16.0, rho=0.5, max_delay=4)
    xin, V = plan.alloc_xinput(), plan.alloc_vis()
    plan.dump(x8, xin, V)
    g_true = (rng.uniform(0.8, 1.2, NN5) * np.exp(1j * rng.uniform(-np.pi, np.pi, NN5))).astype(np.complex64)
    tau_true = rng.uniform(-4, 4, NN5).astype(np.float32)
    g = torch.from_numpy((g_true * (1 + 0.05 * (rng.standard_normal(NN5) + 1j * rng.standard_normal(NN5))))
                         .astype(np.complex64)).cuda()
    tau = torch.from_numpy((tau_true + 0.1 * rng.standard_normal(NN5)).astype(np.float32)).cuda()
    cal = CalibPlan(NN5, N5)
    M = cal.apply(V, torch.from_numpy(g_true).cuda(), torch.from_numpy(tau_true).cuda())
    plan.set_gains(g, tau)
    vis = plan.alloc_vis()
    for _ in range(warmup):
        plan.fengine(x8, xin)
        plan.xengine(xin, vis)
        cal.loss_grad(V, M, g, tau)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        ev[k][0].record()
        plan.fengine(x8, xin)
        ev[k][1].record()
        plan.xengine(xin, vis)
        ev[k][2].record()
        loss, gg, gt = cal.loss_grad(V, M, g, tau)
        ev[k][3].record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    ms = [sum(e[i].elapsed_time(e[i + 1]) for e in ev) / steps for i in range(3)]
    nb5 = NN5 * (NN5 + 1) // 2
    f_bytes = NN5 * (NS5 + T5 - 1) * N5 * 2 + T5 * N5 * 4 + NN5 * NS5 * N5 * 8
    x_flops = 8.0 * N5 * NS5 * nb5
    c_bytes = 2 * 2 * N5 * nb5 * 8  # V and M, read by the row and the column sweeps
    out = {"workload": "calibration: 64 inputs, 4-tap PFB, 1024-pt FFT, 512 spectra; gains and delays in the "
                       "F-engine, FP32 X-engine, FP32 loss with the hand-written adjoint",
           "value": NS5 * N5 / (wall * 1e-3) / 1e6, "unit": "Msamples/s/input", "ms_per_step": wall,
           "steps": steps,
           "kernels": {
               "fengine_gained": {"ms": ms[0], "algorithmic_bytes": f_bytes, "GB/s": f_bytes / (ms[0] * 1e-3) / 1e9,
                                  "frac_hbm": f_bytes / (ms[0] * 1e-3) / 1e9 / HBM_PEAK_GBS},
               "xengine_fp32": {"ms": ms[1], "useful_flops": x_flops,
                                "useful_TFLOP/s": x_flops / (ms[1] * 1e-3) / 1e12,
                                "frac_fp32_mfma": x_flops / (ms[1] * 1e-3) / 1e12 / FP32_PEAK_TFLOPS},
               "calib_loss_grad": {"ms": ms[2], "algorithmic_bytes": c_bytes,
                                   "GB/s": c_bytes / (ms[2] * 1e-3) / 1e9,
                                   "frac_hbm": c_bytes / (ms[2] * 1e-3) / 1e9 / HBM_PEAK_GBS}}}
    if with_err:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc

        V64, M64 = V.cpu().numpy().astype(np.complex128), M.cpu().numpy().astype(np.complex128)
        L64, gg64, gt64 = orc.calib_loss_grad(V64, M64, g.cpu().numpy().astype(np.complex128),
                                              tau.cpu().numpy().astype(np.float64), NN5, N5)
        out["adjoint_vs_fp64"] = {
            "loss_rel": abs(float(loss.cpu()) - L64) / L64,
            "grad_gain_rel": float(np.max(np.abs(gg.cpu().numpy() - gg64)) / np.max(np.abs(gg64))),
            "grad_delay_rel": float(np.max(np.abs(gt.cpu().numpy() - gt64)) / np.max(np.abs(gt64))),
            "definition": "max |GPU - FP64 analytic| / max |FP64| (orc_calib_loss_grad_f64 on the same V, M, g, tau)"}
    del plan, cal, x8, xin, V, M, vis
    torch.cuda.empty_cache()
    return out


def packet_stream_leg(np, torch, mdtfx, dumps, err_channels=0):
    """The client's loop on its own wire data at the shipped station's wire
    shape (one station: 38 inputs from 4 SNAPs, 2048 channels, 4-tap PFB,
    4096 spectra per dump; 32 samples per element per packet): each dump's
    1 048 576 SNAP packets (Fix_32_19 word pairs, 2.7 GB) from pinned host
    memory over PCIe, decoded on the device into the float F input behind
    the history, F-engine on the float input, FP32 X-engine, visibilities
    back.  Packet payloads are built once from int8 samples; each dump only
    renumbers the packet counters.  With err_channels: the last dump's
    visibilities against the FP64 oracle on its window (the Fix_32_19 words
    decode to the int8 samples exactly)."""
    NN6, N6, T6, NS6, S = 38, 2048, 4, 4096, 32
    words = 4 + 2 * S * 10
    log(f"bench: packet stream leg NN={NN6} N={N6} T={T6} Ns={NS6} S={S}")
    x8 = mdtfx.synth_i8(NN6, N6, NS6, seed=SEED, sigma=16.0, rho=RHO, max_delay=4).cpu().numpy()
    per_dump, begins = NS6 * N6, (T6 - 1) * N6
    ncnt = per_dump // S
    pk = np.zeros((4 * ncnt, words), np.uint32)
    X = x8.reshape(NS6, NN6, N6 // S, S, 2)
    for sn, (nin, e0) in enumerate(zip((10, 10, 10, 8), (0, 10, 20, 30))):
        v = X[:, e0:e0 + nin].transpose(0, 2, 3, 1, 4).reshape(ncnt, S * nin * 2)
        rows = pk[sn * ncnt:(sn + 1) * ncnt]
        rows[:, 1] = sn
        rows[:, 4:4 + 2 * S * nin] = ((v.astype(np.int32) << 19).view(np.uint32))
    del X
    if not err_channels:
        del x8
    plan = mdtfx.FxPlan(NN6, N6, T6, NS6, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP32)
    ps = mdtfx.PacketStream(plan, words, S, len(pk))
    for slot in range(2):
        ps.packets(slot)[:] = pk
    cnt = np.tile(np.arange(ncnt, dtype=np.uint32), 4)
    del pk
    torch.cuda.synchronize()

    def run(n, d0):
        times = []
        for d in range(d0, d0 + n):
            b = d % 2
            if d >= d0 + 2:
                times.append(ps.wait(b))
            ps.packets(b)[:, 2] = cnt + np.uint32((d * per_dump + begins) // S)
            ps.submit(b, ps.max_pkts)
        for d in range(max(d0, d0 + n - 2), d0 + n):
            times.append(ps.wait(d % 2))
        return times

    run(2, 0)
    t0 = time.perf_counter()
    times = run(dumps, 2)
    wall = (time.perf_counter() - t0) * 1e3 / dumps
    vis_last = ps.vis((2 + dumps - 1) % 2).copy()
    ps.destroy()
    mean = {k: sum(t[0][k] for t in times) / len(times) for k in times[0][0]}
    pkt_bytes = 4 * ncnt * words * 4
    out = {"workload": "MDT station wire format: 38 inputs (4 SNAPs), 4-tap PFB, 2048-pt FFT, 4096 spectra/dump, "
                       "32 samples per element per packet, FP32 X-engine",
           "value": NS6 * N6 / (wall * 1e-3) / 1e6, "unit": "Msamples/s/input", "ms_per_dump": wall,
           "dumps": dumps, "phase_ms": mean, "packet_bytes": pkt_bytes,
           "h2d_GB/s": pkt_bytes / (mean["h2d_ms"] * 1e-3) / 1e9,
           "counts_last_dump": times[-1][1]}
    if err_channels:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc

        # dumps after the first carry the previous payload's last T-1 spectra
        window = np.concatenate([x8[-(T6 - 1):], x8])
        chans = np.unique(np.linspace(0, N6 - 1, err_channels).astype(int))
        t0 = time.perf_counter()
        v64 = orc.fx_chain_f64(plan.window().cpu().numpy().astype(np.float64), window, NN6, N6, T6, NS6, chans)
        out["max_rel_err_vs_fp64"] = {
            "last_dump_fp32": _autonorm_err(np, orc, vis_last[chans], v64, NN6),
            "channels_checked": len(chans), "oracle_seconds": round(time.perf_counter() - t0, 1),
            "definition": "max over (channel, a>=b) of |V - V64| / sqrt(V64_aa V64_bb)"}
        del x8
    del plan
    torch.cuda.empty_cache()
    return out


def load_traffic():
    """HBM bytes per launch from the rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    runs, scaled by the copy-kernel calibration (profiles/traffic_fx_dumps.json)."""
    p = os.path.join(ROOT, "profiles", "traffic_fx_dumps.json")
    if not os.path.exists(p):
        return {}
    with open(p) as f:
        return json.load(f)


def _cpu_model():
    """lscpu's model name, with family / model / stepping and AVX-512
    presence (virtualised hosts may report only a generic name)"""
    import subprocess
    fields = {}
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                fields[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if "Model name" not in fields:
        try:
            for line in open("/proc/cpuinfo"):
                if ":" in line:
                    k, v = line.split(":", 1)
                    k = {"model name": "Model name", "cpu family": "CPU family", "model": "Model",
                         "stepping": "Stepping", "flags": "Flags"}.get(k.strip())
                    if k and k not in fields:
                        fields[k] = v.strip()
        except OSError:
            pass
    name = fields.get("Model name", "unknown")
    fms = "/".join(fields.get(k, "?") for k in ("CPU family", "Model", "Stepping"))
    avx512 = "avx512f" in fields.get("Flags", "").split()
    return f"{name} (family/model/stepping {fms}, AVX-512 {'yes' if avx512 else 'no'})"


def cpu_baseline_and_error(np, torch, plan, x8, vis16, vis32, sample_spectra, err_channels):
    """The oracle's two legs on the host: (1) the FP32 CPU port of the chain
    timed on a bounded sample (all inputs and channels, the first
    sample_spectra spectra of the dump); (2) the FP64 oracle on err_channels
    channels of the full dump, for the error of the GPU visibilities."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    W = plan.window().cpu().numpy()
    x_host = x8.cpu().numpy()
    L, build = orc.lib_native()  # this host's own -march=native build
    nthreads = orc.num_threads()
    sub = np.ascontiguousarray(x_host[: sample_spectra + T - 1])
    orc.fx_chain_f32_with(L, W, np.ascontiguousarray(x_host[: 4 + T - 1]), NN, N, T, 4, nthreads)  # warm-up
    runs = []
    for _ in range(5):
        t0 = time.perf_counter()
        orc.fx_chain_f32_with(L, W, sub, NN, N, T, sample_spectra, nthreads)
        runs.append(time.perf_counter() - t0)
    dt = float(np.median(runs))
    # one core, on a sample 1/16 the size, median of 3
    s1 = max(2, sample_spectra // 16)
    sub1 = np.ascontiguousarray(x_host[: s1 + T - 1])
    runs1 = []
    for _ in range(3):
        t0 = time.perf_counter()
        orc.fx_chain_f32_with(L, W, sub1, NN, N, T, s1, 1)
        runs1.append(time.perf_counter() - t0)
    dt1 = float(np.median(runs1))
    cpu = {
        "value": sample_spectra * N / dt / 1e6,
        "unit": "Msamples/s/input",
        "cores": nthreads,
        "kind": "port",
        "sample": f"FP32 CPU restatement of the chain (fmaf FIR, radix-2 FFT, corner turn, "
                  f"FP32 triangle correlation) on all {NN} inputs x {N} channels, first "
                  f"{sample_spectra} of {NS} spectra of the same dump; median of 5 runs of {dt:.2f} s "
                  f"after one warm-up; build: {build}",
        "runs_s": [round(r, 3) for r in runs],
        "single_core": {"value": s1 * N / dt1 / 1e6, "sample_spectra": s1, "seconds_median_of_3": round(dt1, 2)},
        "host": {"cpu_model": _cpu_model(), "os_cpu_count": os.cpu_count(), "omp_threads": nthreads},
    }
    chans = np.unique(np.linspace(0, N - 1, err_channels).astype(int))
    t0 = time.perf_counter()
    v64 = orc.fx_chain_f64(W.astype(np.float64), x_host, NN, N, T, NS, chans)
    t64 = time.perf_counter() - t0
    a, b = np.tril_indices(NN)
    diag = np.array([orc.tri_index(i, i) for i in range(NN)])
    auto = np.real(v64[:, diag])
    norm = np.sqrt(auto[:, a] * auto[:, b])
    big = np.abs(v64) > 1e-3 * norm

    def errs(v):
        d = np.abs(v[chans] - v64)
        return {"autocorr_normalised_max": float(np.max(d / norm)),
                "elementwise_rel_max_where_|V|>1e-3_norm": float(np.max(d[big] / np.abs(v64[big]))),
                "per_channel_max": [float(x) for x in np.max(d / norm, axis=1)]}

    err = {"fp16_split": errs(vis16), "fp32_mfma": errs(vis32), "channels_checked": [int(c) for c in chans],
           "definition": "max over (channel, a>=b) of |V - V64| / sqrt(V64_aa V64_bb); FP64 oracle "
                         "chain from the same int8 samples", "oracle_seconds": round(t64, 1)}
    return cpu, err



def hbm_probe_leg(torch, mdtfx, nbytes=2684354560, grids=(1024, 2048, 4096, 8192, 16384), reps=10):
    """What this device's HBM delivers to a streaming kernel with 16-byte lanes
    (the library's probe kernels): reads alone, streaming writes alone, and a
    copy (read + write), best over a few grid sizes, each timed with HIP
    events over `reps` passes of `nbytes`."""
    log("bench: HBM probe (read, write, copy ceilings)")
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    best = {}
    for kind, name, moved in ((mdtfx.PROBE_READ, "read", 1), (mdtfx.PROBE_WRITE, "write", 1),
                              (mdtfx.PROBE_COPY, "copy", 2)):
        rates = {}
        for g in grids:
            mdtfx.hbm_probe(kind, src, dst, g)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                mdtfx.hbm_probe(kind, src, dst, g)
            e1.record()
            torch.cuda.synchronize()
            rates[g] = moved * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
        g = max(rates, key=rates.get)
        best[name] = {"GB/s": rates[g], "grid": g, "frac_of_8TBs": rates[g] / HBM_PEAK_GBS}
    del src, dst
    return {"bytes_per_pass": nbytes, "unit": "GB/s", **best,
            "definition": "best over grids of 256-thread workgroups; read and copy count every byte moved"}

# (inputs, channels, taps, spectra): about 1 GB of FP32 X input each
FE_SHAPES = [(128, 64, 4, 16384), (64, 256, 8, 8192), (64, 512, 8, 4096), (64, 1024, 4, 2048),
             (64, 1024, 16, 2048), (64, 2048, 6, 1024), (64, 4096, 8, 512), (64, 4096, 16, 512),
             (64, 8192, 4, 256), (64, 16384, 4, 128), (64, 16384, 16, 128)]


def fengine_shapes_leg(torch, mdtfx, reps=5):
    """The F-engine alone (int8 in, FP32 or FP16 hi/lo X input out) across
    channel and tap counts, each precision's default X-input layout, timed
    with HIP events over `reps` launches after two untimed ones.  The library
    runs the persistent kernel at 256-4096 channels (any tap count 4-128) and
    the generic fused kernel elsewhere; "kernel" restates that rule."""
    rows = []
    for nn, n, t, ns in FE_SHAPES:
        log(f"bench: F-engine shape NN={nn} N={n} T={t} Ns={ns}")
        x8 = mdtfx.synth_i8(nn, n, ns + t - 1, seed=5, sigma=16.0)
        alg = nn * (ns + t - 1) * n * 2 + t * n * 4 + nn * ns * n * 8
        row = {"inputs": nn, "channels": n, "taps": t, "spectra": ns,
               "kernel": "persistent" if 256 <= n <= 4096 else "generic", "algorithmic_bytes": alg}
        for name, xp in (("fp32", mdtfx.XPREC_FP32), ("fp16_split", mdtfx.XPREC_FP16_SPLIT)):
            plan = mdtfx.FxPlan(nn, n, t, ns, mdtfx.WINDOW_BH4, xp)
            xin = plan.alloc_xinput()
            for _ in range(2):
                plan.fengine(x8, xin)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                plan.fengine(x8, xin)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            row[name] = {"ms": ms, "GB/s": alg / ms / 1e6, "frac_hbm": alg / ms / 1e6 / HBM_PEAK_GBS}
            del plan, xin
        rows.append(row)
        del x8
        torch.cuda.empty_cache()
    return {"unit": "GB/s of algorithmic bytes (int8 in + window + 8-byte samples out)", "shapes": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-spectra", type=int, default=512)
    ap.add_argument("--err-channels", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and FP64 error legs")
    ap.add_argument("--no-fp8", action="store_true", help="skip the wide-band FP8 leg")
    ap.add_argument("--no-stream", action="store_true", help="skip the host-inclusive streaming leg")
    ap.add_argument("--no-station", action="store_true", help="skip the drop-in station leg")
    ap.add_argument("--no-calib", action="store_true", help="skip the calibration leg")
    ap.add_argument("--station-only", action="store_true",
                    help="only run the drop-in station leg and print its JSON (for per-kernel profiles)")
    ap.add_argument("--no-hbm-probe", action="store_true", help="skip the HBM read/write/copy ceiling leg")
    ap.add_argument("--fengine-shapes", action="store_true",
                    help="only time the F-engine across channel and tap counts and print that table")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="only run a plain device copy of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures a single MI355X; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding is design-only here)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import mdtfx

    torch.cuda.set_device(0)
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 3}))
        return
    if args.station_only:
        print(json.dumps(station_leg(np, torch, mdtfx, args.steps, args.warmup, 0)), flush=True)
        return
    if args.fengine_shapes:
        print(json.dumps(fengine_shapes_leg(torch, mdtfx)), flush=True)
        return
    stream = torch.cuda.current_stream().cuda_stream
    log(f"bench: scaled array NN={NN} N={N} T={T} Ns={NS} steps={args.steps} warmup={args.warmup}")
    plan = mdtfx.FxPlan(NN, N, T, NS, mdtfx.WINDOW_BH4, mdtfx.XPREC_FP16_SPLIT)
    plan.set_stream(stream)
    x8 = mdtfx.synth_i8(NN, N, NS + T - 1, seed=SEED, sigma=SIGMA, rho=RHO, max_delay=MAXDELAY)
    xin = plan.alloc_xinput()
    vis16 = plan.alloc_vis()
    vis32 = plan.alloc_vis()
    acc = torch.zeros((N, plan.nb), dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()

    log("bench: FP16-split path")
    w16, f16_ms, x16_ms = time_path(torch, mdtfx, plan, x8, xin, vis16, acc, args.steps, args.warmup)
    plan.set_precision(mdtfx.XPREC_FP32)
    del xin  # each precision's own X-input layout (FP32: padded rows, a little larger)
    torch.cuda.empty_cache()
    xin = plan.alloc_xinput()
    log("bench: FP32-MFMA path")
    w32, f32_ms, x32_ms = time_path(torch, mdtfx, plan, x8, xin, vis32, acc, args.steps, args.warmup)
    plan.set_precision(mdtfx.XPREC_FP16_SPLIT)
    del xin
    torch.cuda.empty_cache()
    torch.cuda.synchronize()

    per_dump = NS * N
    value16 = per_dump / (w16 * 1e-3) / 1e6
    value32 = per_dump / (w32 * 1e-3) / 1e6
    nb = NN * (NN + 1) // 2
    f_bytes = NN * (NS + T - 1) * N * 2 + T * N * 4 + NN * NS * N * 8
    x_flops = 8.0 * N * NS * nb
    kernels = {
        "fengine": {"ms": f16_ms, "algorithmic_bytes": f_bytes, "GB/s": f_bytes / (f16_ms * 1e-3) / 1e9},
        "xengine_fp16split": {"ms": x16_ms, "useful_flops": x_flops,
                              "useful_TFLOP/s": x_flops / (x16_ms * 1e-3) / 1e12,
                              "issued_TFLOP/s": 3 * x_flops / (x16_ms * 1e-3) / 1e12},
        "xengine_fp32": {"ms": x32_ms, "useful_flops": x_flops,
                         "useful_TFLOP/s": x_flops / (x32_ms * 1e-3) / 1e12},
        "fengine_in_fp32_path_ms": f32_ms,
    }
    traffic = load_traffic()
    if f16_ms >= x16_ms:
        ach = f_bytes / (f16_ms * 1e-3) / 1e9
        roof = {"kernel": "fengine_persistent_kernel<4096, 8, 1, false, 1>", "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic.get("fengine")}
    else:
        ach = x_flops / (x16_ms * 1e-3) / 1e12
        roof = {"kernel": "xeng_f16split_chan_kernel", "bound": "mfma", "achieved": ach, "peak": FP16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": ach / FP16_PEAK_TFLOPS, "traffic": traffic.get("xengine_fp16split")}

    out = {
        "metric": METRIC,
        "value": value16,
        "unit": "Msamples/s/input",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": w16,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16-split",
        "data": f"synthetic int8 (I,Q): sigma={SIGMA}, common component rho={RHO} delayed 0-{MAXDELAY} "
                f"samples per input, seed {SEED}",
        "config": {"workload": "scaled array: 256 inputs, 8-tap PFB (BH4-windowed sinc), 4096-pt FFT, "
                               "2048 spectra/dump, FP16 hi+lo split X-engine, FP32 accumulate, FP64 cross-dump "
                               "accumulator",
                   "inputs": NN, "taps": T, "channels": N, "spectra_per_dump": NS, "baselines": nb},
        "fp32_mfma": {"value": value32, "ms_per_step": w32},
        "speedup_fp16split_vs_fp32": value16 / value32,
        "kernels": kernels,
        "roofline": roof,
    }
    if not args.no_stream:
        log("bench: streaming leg (host stacks over PCIe)")
        out["stream_host_inclusive"] = stream_leg(torch, mdtfx, x8, max(4, min(args.steps, 8)),
                                                  0 if args.no_cpu else 8)
    if not args.no_cpu:
        log("bench: CPU baseline + FP64 error legs")
        cpu, err = cpu_baseline_and_error(np, torch, plan, x8, vis16.cpu().numpy(), vis32.cpu().numpy(),
                                          args.cpu_spectra, args.err_channels)
        out["cpu_baseline"] = cpu
        out["max_rel_err_vs_fp64"] = err
    else:
        out["cpu_baseline"] = None
    if not args.no_fp8:
        out["fp8_wideband"] = fp8_leg(np, torch, mdtfx, min(args.steps, 10), min(args.warmup, 2), stream,
                                4, not args.no_cpu)
    if not args.no_station:
        out["station_dropin"] = station_leg(np, torch, mdtfx, min(args.steps, 10), min(args.warmup, 2),
                                            0 if args.no_cpu else args.err_channels)
    if not args.no_calib:
        out["calibration"] = calib_leg(np, torch, mdtfx, min(args.steps, 10), min(args.warmup, 2), not args.no_cpu)
    if not args.no_stream:
        out["packet_stream_host_inclusive"] = packet_stream_leg(np, torch, mdtfx, 4, 0 if args.no_cpu else 16)
    if not args.no_hbm_probe:
        probe = hbm_probe_leg(torch, mdtfx)
        # the dominant kernel against what this device measured: its reads at
        # the read rate plus its writes at the write rate
        fe = out["kernels"]["fengine"]
        rd = NN * (NS + T - 1) * N * 2 + T * N * 4
        wr = fe["algorithmic_bytes"] - rd
        floor_ms = (rd / probe["read"]["GB/s"] + wr / probe["write"]["GB/s"]) / 1e6
        probe["fengine_floor_ms"] = floor_ms
        probe["fengine_frac_of_measured_floor"] = floor_ms / fe["ms"]
        out["hbm_measured_ceiling"] = probe
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
