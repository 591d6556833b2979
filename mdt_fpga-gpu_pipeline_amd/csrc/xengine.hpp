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
  const float *scale_dev = nullptr;  // the same scale on the device, as the F-engine used it (nullable)
  // FP8 block-scaled input: E8M0 exponent byte per (16-spectrum block, input), [nblk16][m]
  const unsigned char *f8scales = nullptr;
  int nblk16 = 0;
  // native X input geometry (XGeom, common.hpp) beyond the row pitch: batch
  // b's operand starts at base + (b >> xlgG) * Abatch + (b mod 2^xlgG) *
  // 2^xlgB, sample (row, s) at row * ld + (s >> xlgB) * xbp + (s mod
  // 2^xlgB).  The defaults give base + b * Abatch + row * ld + s.
  int xlgB = 30, xlgG = 0;
  long xbp = 0;
};

// offset of batch b's operand, and the spectrum part of a sample's offset
__device__ __forceinline__ long xin_batch(const XArgs &g, int b) {
  return (long)(b >> g.xlgG) * g.Abatch + ((long)(b & ((1 << g.xlgG) - 1)) << g.xlgB);
}
__device__ __forceinline__ long xin_spec(int lgB, long bp, int s) {
  return (long)(s >> lgB) * bp + (s & ((1 << lgB) - 1));
}

// Triangle epilogue of one 32x32 MFMA block: register r of the lane holds
// the block element (a = a0 + 8(r/4) + 4 kh + r%4, b) whose value is
// alpha' * (dre[r], -dim[r]) (D[i][j] = V[j][i] = conj of V[i][j]).  Stores
// V[a][b] at cb + a(a+1)/2 + b where a >= b and a, b < NN, and adds it into
// the FP64 accumulator (nullable).  The accumulator is read in groups of G
// loads issued before the first is used -- 16/G memory latencies per block
// instead of 16; elements outside the triangle load the channel's first
// entry (in range, unused) so that the loads stay branch-free.
typedef float xf32x16 __attribute__((ext_vector_type(16)));
template <int G>
__device__ __forceinline__ void tri_block_out(float2 *__restrict__ C, double2 *__restrict__ acc, long cb, int NN,
                                              int a0, int kh, int b, float scale, const xf32x16 &dre,
                                              const xf32x16 &dim) {
  static_assert(16 % G == 0, "groups must split the 16 registers");
#pragma unroll
  for (int g0 = 0; g0 < 16; g0 += G) {
    long off[G];
    bool ok[G];
#pragma unroll
    for (int i = 0; i < G; i++) {
      const int r = g0 + i;
      const int a = a0 + 8 * (r >> 2) + 4 * kh + (r & 3);
      ok[i] = a < NN && b < NN && a >= b;
      off[i] = cb + (ok[i] ? (long)a * (a + 1) / 2 + b : 0);
    }
    if (acc) {
      double2 old[G];
#pragma unroll
      for (int i = 0; i < G; i++) old[i] = acc[off[i]];
#pragma unroll
      for (int i = 0; i < G; i++)
        if (ok[i]) {
          const float vr = scale * dre[g0 + i], vi = -scale * dim[g0 + i];
          C[off[i]] = make_float2(vr, vi);
          acc[off[i]] = make_double2(old[i].x + (double)vr, old[i].y + (double)vi);
        }
    } else {
#pragma unroll
      for (int i = 0; i < G; i++)
        if (ok[i]) C[off[i]] = make_float2(scale * dre[g0 + i], -scale * dim[g0 + i]);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

int xeng_launch(const XArgs &g, int layout, int out, hipStream_t st);
int xeng_launch_f16split(const XArgs &g, hipStream_t st);
int xeng_launch_f8(const XArgs &g, const float *power, hipStream_t st);
int xeng_launch_f8_block(const XArgs &g, hipStream_t st);
int launch_chan_exponents(int N, const float *power, float count, int *e, hipStream_t st);

}  // namespace mdt
