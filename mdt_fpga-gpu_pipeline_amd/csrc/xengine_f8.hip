// xengine_f8.hip -- the FP8 (OCP e4m3) X-engine of the native path on
// v_mfma_scale_f32_32x32x64_f8f6f4 with FP32 accumulation.
//
// Per-channel power scaling: the F-engine writes FP16 samples x*sc (sc a
// global power of two) and the per-channel power P_c = sum |x*sc|^2.  Each
// workgroup maps its channel's RMS to about 32 (e4m3 max 448) with an exact
// power of two 2^e_c, quantises on the way into LDS with the one-instruction
// scaled conversion (v_cvt_scalef32_pk_fp8_f16 divides by its scale, so the
// scale passed is 2^-e_c), and rescales the accumulators by
// alpha / (sc 2^e_c)^2.  A strong tone and the noise-only channels each keep
// the full e4m3 dynamic range (the 40 dB case).
//
// Complex scheme as in the other X-engines: K runs over interleaved (re, im)
// bytes, the I operand is fed as (-im, re) for the imaginary part (a byte
// permute and a sign-bit flip), D[i][j] = V[j][i].  Any K ordering is valid as
// long as both operands use the same one, which the shared LDS layout ensures.
// Workgroup tile 128 x 128 complex (4 waves x 2 x 2 MFMA 32x32 tiles), one
// 64-wide K step (32 spectra) per chunk, the next chunk prefetched into
// registers.  Lower-triangle tiles only; packed triangle output; optional
// FP64 accumulation.
#include "common.hpp"
#include "xengine.hpp"

namespace mdt {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef short short2_t __attribute__((ext_vector_type(2)));

constexpr int Q8T = 128;                // workgroup tile edge (complex rows)
constexpr int Q8KC = 32;                // spectra per K chunk (one 32x32x64 step)
constexpr int Q8ROWB = Q8KC * 2 + 16;   // LDS bytes per row: 64 data + 16 pad
constexpr int Q8PLANE = Q8T * Q8ROWB;   // bytes per operand

// channel exponent e: RMS of x*sc times 2^e close to 32
__device__ __forceinline__ int chan_exponent(float power, float count) {
  if (!(power > 0.f)) return 0;
  const float rms = sqrtf(power / count);
  int e = (int)rintf(log2f(32.0f / rms));
  return e < -30 ? -30 : (e > 30 ? 30 : e);
}

__global__ void chan_exponent_kernel(int N, const float *power, float count, int *e) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < N) e[c] = chan_exponent(power[c], count);
}

struct Prefetch8 {
  uint4 v[4];     // 4 x 4 samples (FP16 re, im)
  unsigned ok;    // bit 4q + i: sample i of row q valid
};

// 128 rows x 32 samples of 4 bytes: 8 lanes x 16 bytes per row, branch-free
__device__ __forceinline__ void fetch8(Prefetch8 &p, const unsigned *__restrict__ base, long ld,
                                       int row0, int nrows, int s0, int K, int lgB, long bp) {
  const int seg = threadIdx.x & 7, rq = threadIdx.x >> 3;
  const int s = s0 + 4 * seg;
  const long so = xin_spec(lgB, bp, s < K ? s : 0);  // 4 samples never straddle a block
  unsigned okm = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int row = row0 + rq + 32 * q;
    const bool r_ok = row < nrows;
    p.v[q] = *(const uint4 *)(base + (long)(r_ok ? row : 0) * ld + so);
#pragma unroll
    for (int i = 0; i < 4; i++) okm |= ((r_ok && s + i < K) ? 1u : 0u) << (4 * q + i);
  }
  p.ok = okm;
}

// one FP16 (re, im) pair -> two e4m3 bytes of (re, im) / inv_scale, clamped
// first to +-448 inv_scale (e4m3's largest finite value) so that an outlier
// saturates instead of becoming NaN
__device__ __forceinline__ unsigned q2(unsigned h2, float inv_scale, half2_t lim) {
  short2_t old = {0, 0};
  half2_t h = __builtin_bit_cast(half2_t, h2);
  h = __builtin_elementwise_max(__builtin_elementwise_min(h, lim), -lim);
  short2_t r = __builtin_amdgcn_cvt_scalef32_pk_fp8_f16(old, h, inv_scale, false);
  return (unsigned)(unsigned short)r[0];
}

// FP16 pairs -> e4m3 bytes (re, im per sample) into LDS [row][k]
__device__ __forceinline__ void stage8(const Prefetch8 &p, char *lds, float inv_scale) {
  const int seg = threadIdx.x & 7, rq = threadIdx.x >> 3;
  const _Float16 l = (_Float16)fminf(448.0f * inv_scale, 65504.0f);
  const half2_t lim = {l, l};
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int row = rq + 32 * q;
    unsigned w[4] = {p.v[q].x, p.v[q].y, p.v[q].z, p.v[q].w};
#pragma unroll
    for (int i = 0; i < 4; i++) w[i] = ((p.ok >> (4 * q + i)) & 1u) ? w[i] : 0u;
    const unsigned d0 = q2(w[0], inv_scale, lim) | (q2(w[1], inv_scale, lim) << 16);
    const unsigned d1 = q2(w[2], inv_scale, lim) | (q2(w[3], inv_scale, lim) << 16);
    *(uint2 *)(lds + row * Q8ROWB + seg * 8) = make_uint2(d0, d1);
  }
}

// (re, im) e4m3 byte pairs -> (-im, re)
__device__ __forceinline__ i32x8 rot_conj8(i32x8 v) {
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const unsigned u = (unsigned)v[q];
    v[q] = (int)(__builtin_amdgcn_perm(u, u, 0x02030001u) ^ 0x00800080u);
  }
  return v;
}

__device__ __forceinline__ i32x8 ld_op(const char *p) {
  const int4 a = *(const int4 *)p, b = *(const int4 *)(p + 16);
  i32x8 r;
  r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w;
  r[4] = b.x; r[5] = b.y; r[6] = b.z; r[7] = b.w;
  return r;
}


__global__ __launch_bounds__(256, 2) void xeng_f8_kernel(XArgs g, const float *power, float count) {
  __shared__ __attribute__((aligned(16))) char lds[2 * Q8PLANE];  // [op I|J][row][k]

  const long nwg = gridDim.x;
  const long w = blockIdx.x;
  const long logical = (nwg % 8 == 0) ? (w % 8) * (nwg / 8) + w / 8 : w;
  const int batch = (int)(logical / g.tiles);
  const int tile = (int)(logical % g.tiles);
  int P = 0;
  while ((P + 1) * (P + 2) / 2 <= tile) P++;
  const int Q = tile - P * (P + 1) / 2;
  const int bi = P, bj = Q;
  const bool same = (bi == bj);
  const unsigned *base = (const unsigned *)g.Abase + xin_batch(g, batch);
  const int NN = g.m, K = g.k;

  const int ec = chan_exponent(power[batch], count);
  const float inv_scale = exp2f((float)-ec);  // the conversion divides by this
  const float s_total = (g.scale_dev != nullptr ? *g.scale_dev : g.in_scale) * exp2f((float)ec);
  const float oscale = g.alpha.x / (s_total * s_total);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wi = wave >> 1, wj = wave & 1;
  const int l32 = lane & 31, kh = lane >> 5;
  bool act[2][2];
  bool any = false;
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int ra = bi * 4 + wi * 2 + t, rb = bj * 4 + wj * 2 + u;
      act[t][u] = (ra >= rb) && (ra * 32 < NN) && (rb * 32 < NN);
      any |= act[t][u];
    }

  f32x16 accR[2][2], accI[2][2];
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        accR[t][u][r] = 0.f;
        accI[t][u][r] = 0.f;
      }

  char *ldsI = lds;
  char *ldsJ = same ? lds : lds + Q8PLANE;
  Prefetch8 pI, pJ;
  fetch8(pI, base, g.ldI, bi * Q8T, NN, 0, K, g.xlgB, g.xbp);
  if (!same) fetch8(pJ, base, g.ldJ, bj * Q8T, NN, 0, K, g.xlgB, g.xbp);

  for (int s0 = 0; s0 < K; s0 += Q8KC) {
    __syncthreads();
    stage8(pI, lds, inv_scale);
    if (!same) stage8(pJ, lds + Q8PLANE, inv_scale);
    __syncthreads();
    if (s0 + Q8KC < K) {
      fetch8(pI, base, g.ldI, bi * Q8T, NN, s0 + Q8KC, K, g.xlgB, g.xbp);
      if (!same) fetch8(pJ, base, g.ldJ, bj * Q8T, NN, s0 + Q8KC, K, g.xlgB, g.xbp);
    }
    if (any) {
      const int koff = kh * 32;  // 32 bytes = 16 samples per half-wave
      i32x8 I[2], Pr[2], J[2];
#pragma unroll
      for (int t = 0; t < 2; t++) {
        I[t] = ld_op(ldsI + (wi * 64 + t * 32 + l32) * Q8ROWB + koff);
        Pr[t] = rot_conj8(I[t]);
      }
#pragma unroll
      for (int u = 0; u < 2; u++) J[u] = ld_op(ldsJ + (wj * 64 + u * 32 + l32) * Q8ROWB + koff);
#pragma unroll
      for (int t = 0; t < 2; t++)
#pragma unroll
        for (int u = 0; u < 2; u++) {
          if (!act[t][u]) continue;
          accR[t][u] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(I[t], J[u], accR[t][u], 0, 0, 0,
                                                                      127, 0, 127);
          accI[t][u] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(Pr[t], J[u], accI[t][u], 0, 0,
                                                                      0, 127, 0, 127);
        }
    }
  }
  if (!any) return;

#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int u = 0; u < 2; u++) {
      if (!act[t][u]) continue;
      tri_block_out<8>(g.Cbase, (double2 *)g.acc64, (long)batch * g.Cbatch, NN, bi * Q8T + wi * 64 + t * 32, kh,
                    bj * Q8T + wj * 64 + u * 32 + l32, oscale, accR[t][u], accI[t][u]);
      __builtin_amdgcn_sched_barrier(0);
    }
}

// ------------------------------------------------------------------ block-scaled e4m3 input
// The F-engine writes e4m3 (re, im) samples x / 2^eb, eb one exponent per
// (input, block of 16 spectra) shared by all channels (E8M0 bytes eb + 127,
// [block][input]), in blocks of 32 spectra: a tile row's K chunk is one
// 64-byte run and the 128 rows of a tile are 8 KB in a row.  No conversion on
// the way into LDS; lane l of an MFMA holds 16 spectra (32 bytes) of row
// l mod 32, exactly one block, so the block's exponent is the lane's scale
// operand: D += sum (qa 2^ea)(qb 2^eb), the samples themselves.
struct PrefetchB {
  uint4 v[2];  // rows rq and rq + 64 of the tile: 8 samples (16 bytes) each
  int nv[2];   // valid samples of each (0 for rows past NN, fewer in the last chunk)
};

__device__ __forceinline__ int f8_scale(const XArgs &g, int row, int blk) {
  return (row < g.m && blk < g.nblk16) ? (int)g.f8scales[(long)blk * g.m + row] : 127;
}

// one operand's 128 rows x 32 spectra (64 bytes a row, rows 64 bytes apart)
__device__ __forceinline__ void fetch8b(PrefetchB &p, const char *__restrict__ base, const XArgs &g, int row0,
                                        int s0) {
  const int seg = threadIdx.x & 3, rq = threadIdx.x >> 2;  // 4 lanes x 16 bytes per row
  const long so = ((long)(s0 >> 5) * g.xbp + seg * 8) * 2;  // s0: a multiple of 32
  const int nvs = min(8, max(0, g.k - s0 - seg * 8));
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int row = row0 + rq + 64 * h;
    const bool ok = row < g.m;
    p.v[h] = *(const uint4 *)(base + (long)(ok ? row : 0) * g.ldI * 2 + so);
    p.nv[h] = ok ? nvs : 0;
  }
}

// the lane's scale bytes for a chunk: I rows t = 0, 1, then J rows u = 0, 1
__device__ __forceinline__ void fetch_scales(int (&sc)[4], const XArgs &g, int rowI, int rowJ, int s0, int kh) {
  const int blk = (s0 >> 4) + kh;
  sc[0] = f8_scale(g, rowI, blk);
  sc[1] = f8_scale(g, rowI + 32, blk);
  sc[2] = f8_scale(g, rowJ, blk);
  sc[3] = f8_scale(g, rowJ + 32, blk);
}

// 16 bytes of e4m3 pairs into LDS, samples past nv zeroed
__device__ __forceinline__ void stage8b(const PrefetchB &p, char *lds) {
  const int seg = threadIdx.x & 3, rq = threadIdx.x >> 2;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    unsigned w[4] = {p.v[h].x, p.v[h].y, p.v[h].z, p.v[h].w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const unsigned lo = 2 * i < p.nv[h] ? 0x0000ffffu : 0u, hi = 2 * i + 1 < p.nv[h] ? 0xffff0000u : 0u;
      w[i] &= lo | hi;
    }
    *(uint4 *)(lds + (rq + 64 * h) * Q8ROWB + seg * 16) = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

__global__ __launch_bounds__(256, 2) void xeng_f8b_kernel(XArgs g) {
  __shared__ __attribute__((aligned(16))) char lds[2 * Q8PLANE];  // [op I|J][row][k]

  const long nwg = gridDim.x;
  const long w = blockIdx.x;
  const long logical = (nwg % 8 == 0) ? (w % 8) * (nwg / 8) + w / 8 : w;
  const int batch = (int)(logical / g.tiles);
  const int tile = (int)(logical % g.tiles);
  int P = 0;
  while ((P + 1) * (P + 2) / 2 <= tile) P++;
  const int Q = tile - P * (P + 1) / 2;
  const int bi = P, bj = Q;
  const bool same = (bi == bj);
  const char *base = (const char *)g.Abase + xin_batch(g, batch) * 2;  // 2-byte samples
  const int NN = g.m, K = g.k;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wi = wave >> 1, wj = wave & 1;
  const int l32 = lane & 31, kh = lane >> 5;
  bool act[2][2];
  bool any = false;
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int ra = bi * 4 + wi * 2 + t, rb = bj * 4 + wj * 2 + u;
      act[t][u] = (ra >= rb) && (ra * 32 < NN) && (rb * 32 < NN);
      any |= act[t][u];
    }

  f32x16 accR[2][2], accI[2][2];
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        accR[t][u][r] = 0.f;
        accI[t][u][r] = 0.f;
      }

  char *ldsI = lds;
  char *ldsJ = same ? lds : lds + Q8PLANE;
  const int rowI = bi * Q8T + wi * 64 + l32, rowJ = bj * Q8T + wj * 64 + l32;  // the lane's MFMA rows (t, u = 0)
  PrefetchB pI, pJ;
  int scn[4];
  fetch8b(pI, base, g, bi * Q8T, 0);
  if (!same) fetch8b(pJ, base, g, bj * Q8T, 0);
  fetch_scales(scn, g, rowI, rowJ, 0, kh);

  for (int s0 = 0; s0 < K; s0 += Q8KC) {
    __syncthreads();
    stage8b(pI, lds);
    if (!same) stage8b(pJ, lds + Q8PLANE);
    const int scI[2] = {scn[0], scn[1]}, scJ[2] = {scn[2], scn[3]};
    __syncthreads();
    if (s0 + Q8KC < K) {
      fetch8b(pI, base, g, bi * Q8T, s0 + Q8KC);
      if (!same) fetch8b(pJ, base, g, bj * Q8T, s0 + Q8KC);
      fetch_scales(scn, g, rowI, rowJ, s0 + Q8KC, kh);
    }
    if (any) {
      const int koff = kh * 32;  // 32 bytes = 16 samples per half-wave: one block
      i32x8 I[2], Pr[2], J[2];
#pragma unroll
      for (int t = 0; t < 2; t++) {
        I[t] = ld_op(ldsI + (wi * 64 + t * 32 + l32) * Q8ROWB + koff);
        Pr[t] = rot_conj8(I[t]);
      }
#pragma unroll
      for (int u = 0; u < 2; u++) J[u] = ld_op(ldsJ + (wj * 64 + u * 32 + l32) * Q8ROWB + koff);
#pragma unroll
      for (int t = 0; t < 2; t++)
#pragma unroll
        for (int u = 0; u < 2; u++) {
          if (!act[t][u]) continue;
          accR[t][u] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(I[t], J[u], accR[t][u], 0, 0, 0, scI[t], 0,
                                                                      scJ[u]);
          accI[t][u] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(Pr[t], J[u], accI[t][u], 0, 0, 0, scI[t], 0,
                                                                      scJ[u]);
        }
    }
  }
  if (!any) return;

  const float oscale = g.alpha.x;  // the samples are x itself
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int u = 0; u < 2; u++) {
      if (!act[t][u]) continue;
      tri_block_out<8>(g.Cbase, (double2 *)g.acc64, (long)batch * g.Cbatch, NN, bi * Q8T + wi * 64 + t * 32, kh,
                       bj * Q8T + wj * 64 + u * 32 + l32, oscale, accR[t][u], accI[t][u]);
      __builtin_amdgcn_sched_barrier(0);
    }
}

int xeng_launch_f8_block(const XArgs &g0, hipStream_t st) {
  XArgs g = g0;
  const int nb = (g.m + Q8T - 1) / Q8T;
  g.tiles = nb * (nb + 1) / 2;
  const long total = (long)g.tiles * g.batch;
  MDT_TRY(check_grid(total, 256, "X-engine"));
  if (g.f8scales == nullptr) return fail(MDT_ERR_INVALID_VALUE, "FP8 block-scaled input needs its scales");
  hipLaunchKernelGGL(xeng_f8b_kernel, dim3((unsigned)total), dim3(256), 0, st, g);
  return check_launch("xeng_f8b_kernel");
}

int xeng_launch_f8(const XArgs &g0, const float *power, hipStream_t st) {
  XArgs g = g0;
  const int nb = (g.m + Q8T - 1) / Q8T;
  g.tiles = nb * (nb + 1) / 2;
  long total = (long)g.tiles * g.batch;
  MDT_TRY(check_grid(total, 256, "X-engine"));
  if (!(g.in_scale > 0.f)) return fail(MDT_ERR_INVALID_VALUE, "FP8 path needs a positive input scale");
  const float count = (float)g.m * (float)g.k;
  hipLaunchKernelGGL(xeng_f8_kernel, dim3((unsigned)total), dim3(256), 0, st, g, power, count);
  return check_launch("xeng_f8_kernel");
}

int launch_chan_exponents(int N, const float *power, float count, int *e, hipStream_t st) {
  hipLaunchKernelGGL(chan_exponent_kernel, dim3((N + 255) / 256), dim3(256), 0, st, N, power, count, e);
  return check_launch("chan_exponent_kernel");
}

}  // namespace mdt
