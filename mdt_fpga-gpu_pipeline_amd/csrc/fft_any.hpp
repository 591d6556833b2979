// fft_any.hpp -- the batched C2C FFT for lengths that are not powers of two
// (cufftPlan1d accepts any Nchannels the client allows, 32..16384:
// MDT_OneStation_client.h:56-57, MDT_OneStation_client.cu:286, 622).
//   * 7-smooth lengths (N = 2^a 3^b 5^c 7^d): mixed-radix Stockham passes of
//     radix 8/4/2, 3, 5 and 7 in LDS.  Up to 8192 points, persistent
//     workgroups that load the next rows while the current ones are
//     transformed, twiddles from a per-pass table in LDS; longer, one
//     workgroup per transform, the first pass reading and the last writing
//     global memory;
//   * any other length: Bluestein's chirp-z algorithm, the length-N DFT as a
//     circular convolution of length M >= 2N-1 (M a power of two) done with
//     the power-of-two Stockham core in LDS (M = 512..4096: persistent
//     workgroups loading the next rows during the transforms); for N > 8192
//     the M = 32768 convolution is split into its even and odd halves, two
//     16384-point transforms each way.
// Tables (twiddles, chirps and the transformed chirp filter) are evaluated in
// FP64 and rounded once to FP32 at plan creation.
#pragma once

#include "fft_core.hpp"

namespace mdt {

enum { FFT_ANY_MIXED = 1, FFT_ANY_BLUESTEIN = 2, FFT_ANY_BLUESTEIN_SPLIT = 3 };
constexpr int kMaxPasses = 16;

// one mixed-radix pass: radix, stage span NS (product of the earlier
// radices), offset of its entries in the per-pass twiddle table, and
// ceil(2^32 / NS) for j / NS (NS > 1)
struct MixedPassInfo {
  int R, NS, twoff;
  unsigned magic;
};

struct FftAny {
  int kind = 0;
  int N = 0;
  // mixed radix
  int npass = 0;
  int rad[kMaxPasses] = {};
  int TPB = 0, FPB = 0;
  size_t lds_bytes = 0;
  // Bluestein
  int M = 0;
  float2 *tw = nullptr;                   // W_N^k (mixed), W_M^k (Bluestein) or W_{M/2}^k (split)
  float2 *chirp[2] = {nullptr, nullptr};  // [forward, inverse]: exp(-+ pi i (n^2 mod 2N) / N), n < N
  float2 *filt[2] = {nullptr, nullptr};   // [forward, inverse]: FFT_M(conj chirp, wrapped) / M; split: even then odd
  float2 *twm = nullptr;                  // split: W_M^n, n < M/2
  // mixed radix in persistent workgroups (N <= 8192, tables in LDS)
  bool stream = false;
  MixedPassInfo ps[kMaxPasses] = {};
  float2 *twp = nullptr;  // per-pass twiddles, [pass][r - 1][jm], ntw = N - R_0 entries
  int ntw = 0, sTPB = 0, sFPB = 0;
  size_t stream_lds = 0;
  long sgrid = 0;  // resident workgroups
};

// 7-smooth?  Fills the radix sequence (8s first, then 4 or 2, 3s, 5s, 7s).
bool fft_factor(int N, int *npass, int *rad);
int fft_any_create(int N, FftAny *p);
int fft_any_exec(const FftAny &p, const float2 *in, float2 *out, long batch, bool inv, hipStream_t st);
void fft_any_destroy(FftAny *p);

}  // namespace mdt
