// calib.hpp -- per-input delay phasor shared by the F-engine gain stage and
// the calibration kernels.
#pragma once

#include "common.hpp"

namespace mdt {

// exp(-2 pi i k tau / N) with the argument reduced exactly: p + e = k*tau (e
// the fma rounding error of the product), p reduced modulo N without rounding
// (N a power of two, |p| < 2^24), so the sincospi argument 2(p' + e)/N in
// [-1, 1] carries about one ulp of absolute error.
__device__ __forceinline__ float2 delay_phasor(int k, float tau, int N) {
  const float kf = (float)k;
  const float p = kf * tau;
  const float e = __builtin_fmaf(kf, tau, -p);
  const float nf = (float)N;
  const float pr = p - nf * rintf(p / nf);
  const float arg = -2.0f * (pr + e) / nf;
  float s, c;
  sincospif(arg, &s, &c);
  return make_float2(c, s);
}

__device__ __forceinline__ int signed_bin(int c, int N) { return c < N / 2 ? c : c - N; }

}  // namespace mdt
