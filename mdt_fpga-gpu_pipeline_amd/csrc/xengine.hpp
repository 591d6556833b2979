// xengine.hpp -- launch interface of the MFMA X-engine kernels.
#pragma once

#include "common.hpp"

namespace mdt {

// operand layouts
enum { XL_KM = 0,   // element (row, s) at base[row + s*ld]: rows contiguous (reference XInput)
       XL_MK = 1 }; // element (row, s) at base[row*ld + s]: samples contiguous (native)
// output forms
enum { XO_FULL = 0,  // column-major C[a + b*ldc] = alpha V + beta C (cuBLAS form)
       XO_TRI = 1 }; // packed lower triangle C[bl(a,b)], bl = a(a+1)/2 + b, a >= b

struct XArgs {
  // per-batch operands: pointer tables (drop-in) or base + batch stride
  const float2 *const *Atab;
  const float2 *const *Btab;
  float2 *const *Ctab;
  const float2 *Abase;
  const float2 *Bbase;
  float2 *Cbase;
  long Abatch, Bbatch, Cbatch;
  long ldI, ldJ;  // operand leading dimensions (I = B side, J = A side)
  int ldc;
  int m, n, k, batch;
  float2 alpha, beta;
  int herm;   // A == B: Hermitian, triangle tiles only
  int tiles;  // set by xeng_launch
  int nbJ;
  double *acc64;  // XO_TRI only: FP64 complex accumulator, += V (nullable)
  int prec;       // MDT_XPREC_*
  float in_scale; // FP16 split: exact power-of-two pre-scale of the FP32 samples
};

int xeng_launch(const XArgs &g, int layout, int out, hipStream_t st);
int xeng_launch_f16split(const XArgs &g, hipStream_t st);
int xeng_launch_f8(const XArgs &g, const float *power, hipStream_t st);
int launch_chan_exponents(int N, const float *power, float count, int *e, hipStream_t st);

}  // namespace mdt
