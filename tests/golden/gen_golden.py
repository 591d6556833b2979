"""Writes tests/golden/golden.npz: FP64 golden vectors for the oracle pins.

Every expected value here comes from a library routine or a closed form that
is independent of the oracle's code:
  * window: numpy.sinc x scipy.signal.windows.blackmanharris / hamming
    (periodic, sym=False) -- the same windowed-sinc that
    MDT_device_functions.cu:52-88 evaluates;
  * fft: numpy.fft.fft (pocketfft, FP64), the contract of the cufftExecC2C
    forward call at MDT_OneStation_client.cu:1105, at powers of two and at
    lengths that are not (1000, 1536, 3000, 4099);
  * xengine: numpy.einsum in FP64 of alpha * A A^H per channel, the contract of
    cublasCgemmBatched(OP_N, OP_C) at MDT_OneStation_client.cu:1128;
  * chain: numpy FP64 FIR (einsum) -> numpy.fft -> einsum on a small int8
    input, the whole four-call chain.
Run:  python tests/golden/gen_golden.py
"""
import os

import numpy as np
import scipy.signal.windows as sw

HERE = os.path.dirname(os.path.abspath(__file__))


def window_np(N, T, kind):
    TN = N * T
    i = np.arange(TN, dtype=np.float64)
    sinc = np.sinc((i - TN / 2) / N)
    if kind == "bh4":
        w = sw.blackmanharris(TN, sym=False)
    else:
        w = sw.hamming(TN, sym=False)
    return (sinc * w).reshape(T, N)


def main():
    rng = np.random.default_rng(20240611)
    out = {}
    for N, T in [(64, 4), (1024, 8), (2048, 4)]:
        out[f"window_bh4_N{N}_T{T}"] = window_np(N, T, "bh4")
        out[f"window_hamming_N{N}_T{T}"] = window_np(N, T, "hamming")
    for N in [8, 32, 1024, 2048, 4096]:
        x = rng.standard_normal((1, N)) + 1j * rng.standard_normal((1, N))
        out[f"fft_in_N{N}"] = x
        out[f"fft_out_N{N}"] = np.fft.fft(x, axis=-1)
        out[f"ifft_out_N{N}"] = np.fft.ifft(x, axis=-1) * N
    # X-engine: XIn [N][Ns][NN] complex, V[c][b][a] = alpha sum_s x_a conj(x_b)
    Nc, Ns, NN = 4, 7, 5
    x = (rng.standard_normal((Nc, Ns, NN)) + 1j * rng.standard_normal((Nc, Ns, NN))).astype(np.complex64)
    out["xeng_in"] = x
    out["xeng_out"] = np.einsum("csa,csb->cba", x.astype(np.complex128), np.conj(x.astype(np.complex128))) / Ns
    # whole chain on int8 input [Ns+T-1][NN][N][2]
    NN, N, T, Ns = 3, 64, 4, 8
    x8 = rng.integers(-128, 128, size=(Ns + T - 1, NN, N, 2), dtype=np.int8)
    W = window_np(N, T, "hamming")
    xc = x8[..., 0].astype(np.float64) + 1j * x8[..., 1].astype(np.float64)
    y = np.zeros((Ns, NN, N), np.complex128)
    for i in range(T):
        y += W[i][None, None, :] * xc[i:i + Ns]
    Y = np.fft.fft(y, axis=-1)                  # [Ns][NN][N]
    V = np.einsum("sac,sbc->cab", Y, np.conj(Y)) / Ns   # [N][a][b]
    a, b = np.tril_indices(NN)
    out["chain_x8"] = x8
    out["chain_W"] = W
    out["chain_vis"] = V[:, a, b]
    out["chain_shape"] = np.array([NN, N, T, Ns])
    # FFT lengths that are not powers of two (mixed radix 2^a 3^b 5^c; the
    # prime 4099 takes the chirp-z path), from their own generator so the
    # vectors above stay as they were
    rng2 = np.random.default_rng(20261204)
    for N in [1000, 1536, 3000, 4099]:
        x = rng2.standard_normal((1, N)) + 1j * rng2.standard_normal((1, N))
        out[f"fft_in_N{N}"] = x
        out[f"fft_out_N{N}"] = np.fft.fft(x, axis=-1)
        out[f"ifft_out_N{N}"] = np.fft.ifft(x, axis=-1) * N
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"), sum(v.nbytes for v in out.values()), "bytes raw")


if __name__ == "__main__":
    main()
