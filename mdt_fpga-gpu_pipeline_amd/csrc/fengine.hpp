// fengine.hpp -- launch interface of the fused F-engine.
#pragma once

#include "common.hpp"

namespace mdt {

struct FArgs {
  const char2 *in;      // [Ns+T-1][NN][N] int8 (I, Q)
  const float *W;       // [T][N] FP32 window
  const float2 *tw;     // W_N^k twiddles
  float2 *out;          // [N][NN][Ns] complex FP32
  int NN, Ns, T;
  const float2 *gain;   // optional [NN] complex gains
  const float *delay;   // optional [NN] delays in samples
};

int launch_fengine(const FArgs &a, int N, hipStream_t st);

}  // namespace mdt
