// fengine.hpp -- launch interface of the fused F-engine.
#pragma once

#include "common.hpp"

namespace mdt {

struct FArgs {
  const char2 *in;      // [Ns+T-1][NN][N] int8 (I, Q)
  const float2 *fin;    // or the same frames as float2 (the reference's d_FInput); nullptr: int8
  const float *W;       // [T][N] FP32 window
  const float *Wt;      // persistent kernel: the window as [m][i/4][t][4] (arrange_taps)
  const float2 *tw;     // W_N^k twiddles
  float2 *out;          // X input: (k, e, s) at xg.chan(k) + xg.at(e, s) (samples)
  int NN, Ns, T;
  XGeom xg;             // X input geometry (rows or channel pairs)
  const float2 *gain;   // optional [NN] complex gains
  const float *delay;   // optional [NN] delays in samples
  int split;            // 0: FP32 (re, im); 1: FP16 (hi_re, hi_im, lo_re, lo_im) of x*split_scale;
                        // 2: FP16 (re, im) of x*split_scale in 4-byte elements + channel power;
                        // 3: e4m3 (re, im) in 2-byte elements + block exponents (scales)
  float split_scale;    // exact power of two (without gains)
  const float *dscale;  // with gains: the power-of-two scale for |g| x the bound, on the device
  float *power;         // split 2: per-channel sum of |x*split_scale|^2 [N]
  float *power_part;    // split 2: per-workgroup partial sums [power_slots][N], added in order into power
  int power_slots;      // split 2: rows of power_part (the persistent grid is capped to it)
  unsigned char *scales;  // split 3: E8M0 exponent byte per (16-spectrum block, input), [nblk16][NN]
  int nblk16;             // split 3: ceil(Ns / 16)
};

// split 3 (FP8 e4m3 with block scales): the exponent eb of a block whose
// largest |re| or |im| is maxabs: the smallest integer with maxabs <= 448 2^eb
// (448 = 0.875 2^9 is e4m3's largest finite value), clamped to E8M0's
// [-127, 127]; samples are stored as e4m3(x / 2^eb), the byte eb + 127.
__host__ __device__ inline int fp8_block_exponent(float maxabs) {
  if (!(maxabs > 0.f)) return -127;
  int E = 0;
  const float m = frexpf(maxabs, &E);  // maxabs = m 2^E, m in [0.5, 1)
  const int eb = E - (m <= 0.875f ? 9 : 8);
  return eb < -127 ? -127 : (eb > 127 ? 127 : eb);
}

// The FP16 sample formats store x*s, s the largest power of two with
// s*bound <= 2^15 (below the FP16 maximum 65504), bound a rigorous bound of
// |x|.  Exact: with bound = m 2^E, m in [0.5, 1), s = 2^(15-E), or 2^(16-E)
// when m = 0.5.
__host__ __device__ inline float f16_prescale(double bound) {
  int E = 0;
  const double m = frexp(bound, &E);
  return ldexpf(1.0f, 15 - E + (m == 0.5 ? 1 : 0));
}

// FP16 (re, im) of a sample scaled by an exact power of two, as 4 bytes
__device__ __forceinline__ unsigned half_sample(float2 v, float sc) {
  const _Float16 r = (_Float16)(v.x * sc), i = (_Float16)(v.y * sc);
  return (unsigned)__builtin_bit_cast(unsigned short, r) |
         ((unsigned)__builtin_bit_cast(unsigned short, i) << 16);
}

// FP16 hi/lo split of one complex sample scaled by an exact power of two:
// hi = fp16(x*sc), lo = fp16(x*sc - hi); 8 bytes (hi_re, hi_im, lo_re, lo_im).
__device__ __forceinline__ float2 split_sample(float2 v, float sc) {
  const float xr = v.x * sc, xi = v.y * sc;
  const _Float16 hr = (_Float16)xr, hi_ = (_Float16)xi;
  const _Float16 lr = (_Float16)(xr - (float)hr), li = (_Float16)(xi - (float)hi_);
  const unsigned w0 = (unsigned)__builtin_bit_cast(unsigned short, hr) |
                      ((unsigned)__builtin_bit_cast(unsigned short, hi_) << 16);
  const unsigned w1 = (unsigned)__builtin_bit_cast(unsigned short, lr) |
                      ((unsigned)__builtin_bit_cast(unsigned short, li) << 16);
  return make_float2(__uint_as_float(w0), __uint_as_float(w1));
}

// int8 input (a.in) or float2 input (a.fin, the generic kernel only)
int launch_fengine(const FArgs &a, int N, hipStream_t st);
// shapes served by the persistent kernel (N in [256, 4096], any tap count;
// 4 and 8 taps with their own code, the others in groups of four)
bool persistent_shape(int N, int T);
// floats of the arranged window (taps zero-padded to a multiple of 4)
int arranged_taps_floats(int N, int T);
// FP8 sample formats by shape: e4m3 with block scales where an F-engine item
// spans at least 16 spectra (N = 512, 1024), per-channel FP16 + power above
bool fp8_block_shape(int N);
// the window [T][N] re-arranged for the persistent kernel: Wt[((m*G4 + i/4)*TPB + t)*4 + i%4]
// = W[i*N + t + m*TPB], G4 = ceil(T/4), TPB = N / (points per thread), taps i >= T zero
int arrange_taps(int N, int T, const float *W, float *Wt, hipStream_t st);

}  // namespace mdt
