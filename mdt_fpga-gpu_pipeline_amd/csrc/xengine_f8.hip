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

// Block-scaled operand of one 32x32x64 step from a 64-byte row of 32 spectra
// (re, im e4m3 pairs).  The scaled MFMA splits K into two blocks of 32: block
// h is registers 4h..4h+3 (16 bytes) of both lane halves, and lane 32h + row
// supplies its scale.  (Measured: with lane half 0 holding spectra 0-15 of a
// step as 32 contiguous bytes, its second 16 bytes took the other half's
// scale.)  So lane half kh loads spectra 8kh..8kh+7 (bytes 16kh) into
// registers 0-3 and spectra 16 + 8kh.. (bytes 32 + 16kh) into 4-7: K block 0
// is spectra 0-15, K block 1 spectra 16-31, each one 16-spectrum block whose
// exponent lane half 0, resp. 1, supplies.  Both operands use the same order,
// so the products pair the same spectra.
__device__ __forceinline__ i32x8 ld_op_blk(const char *row64, int kh) {
  const int4 a = *(const int4 *)(row64 + 16 * kh), b = *(const int4 *)(row64 + 32 + 16 * kh);
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
// the way into LDS; each of the MFMA's two K blocks holds one 16-spectrum
// block of a row (ld_op_blk), so that block's exponent is the scale operand
// of the lane that supplies it: D += sum (qa 2^ea)(qb 2^eb), the samples
// themselves.
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
      i32x8 I[2], Pr[2], J[2];
#pragma unroll
      for (int t = 0; t < 2; t++) {
        I[t] = ld_op_blk(ldsI + (wi * 64 + t * 32 + l32) * Q8ROWB, kh);
        Pr[t] = rot_conj8(I[t]);
      }
#pragma unroll
      for (int u = 0; u < 2; u++) J[u] = ld_op_blk(ldsJ + (wj * 64 + u * 32 + l32) * Q8ROWB, kh);
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

// ---------------------------------------------------------------------------
// Whole-channel block-scaled kernel for 128 < NN <= 512 (the 512-input
// wide-band array): the channel's triangle of 32 x 32 blocks (136 at NN =
// 512) is split over four 8-wave workgroups that run side by side on one XCD
// and stream the same K chunks, so the channel comes from HBM about once and
// its re-reads from L2:
//   kind 0, 1: the triangles of rows [0, 256) and [256, 512), 36 blocks each,
//              in the FP16-split whole-channel kernel's arrangement (rows pair
//              up (7, 0), (6, 1), (5, 2), (4, 3), one pair per SIMD);
//   kind 2, 3: the rectangles rows [256, 512) x columns [0, 128) and
//              [128, 256), 32 blocks each, wave w owning block row w and the
//              four column blocks.
// At NN <= 256 only kind 0 runs.  Two waves per SIMD with 256 registers each;
// K chunks of 64 spectra (two 32x32x64 steps) for all of a workgroup's rows
// (256, or 384 for a rectangle: its 256 I rows and 128 J rows) are
// double-buffered in LDS beside their scale bytes, one barrier per chunk, and
// prefetched two chunks ahead in registers.  Each row's four E8M0 bytes of a
// chunk are one LDS word, byte 2 kh + ks for the lane half kh and K step ks.
constexpr int F8C_THREADS = 512;
constexpr int F8C_KC = 64;                  // spectra per chunk
constexpr int F8C_ROWB = F8C_KC * 2 + 16;   // LDS bytes per row: 128 data + 16 pad
constexpr int F8C_MAXR = 384;               // LDS rows of a rectangle
constexpr int F8C_BUF = F8C_MAXR * F8C_ROWB + F8C_MAXR * 4;  // data, then one scale word per row
constexpr int F8C_A = 5, F8C_B = 4, F8C_R = 4;  // blocks of an A, B and rectangle wave

template <int NR>
struct F8cPrefetch {
  static constexpr int L = NR * 8 / F8C_THREADS;  // 16-byte pieces per thread per chunk
  uint4 v[L];
  unsigned ok;  // bit i: piece i valid
  unsigned sc;  // the scale word of row threadIdx.x (threads < NR)
};

// the channel row of LDS row r: r0 + r below 256, c0 + r - 256 above (rectangles' J rows)
__device__ __forceinline__ int f8c_chrow(int r, int r0, int c0) { return r < 256 ? r0 + r : c0 + r - 256; }

// global -> registers: chunk c's two 32-spectrum blocks of NR rows (four 16-byte
// pieces a row per block; 16 consecutive rows of a block are 1 KB in a row).
// Piece i of a thread lies in block i / (L / 2) (NR * 4 pieces a block, a
// multiple of the 512 threads), so the block offset is uniform and the lane's
// part a 32-bit offset; branch-free: invalid pieces read the block's first
// bytes (or block 0's past the last block) and are zeroed when staged.  Scale
// bytes past the rows or blocks read as 127 (scale 1).
template <int NR>
__device__ __forceinline__ void f8c_fetch(F8cPrefetch<NR> &p, const char *__restrict__ base, const XArgs &g,
                                          int r0, int c0, int c) {
  constexpr int L = F8cPrefetch<NR>::L;
  static_assert((NR * 4) % F8C_THREADS == 0 && L % 2 == 0, "pieces of a block must divide among the threads");
  const int nblk32 = (g.k + 31) >> 5;
  const int tid = threadIdx.x;
  unsigned ok = 0;
#pragma unroll
  for (int i = 0; i < L; i++) {
    const int blk = i / (L / 2);
    const int rem = tid + (i - blk * (L / 2)) * F8C_THREADS;  // piece within the block
    const int row = f8c_chrow(rem >> 2, r0, c0), seg = rem & 3;
    const int b32 = 2 * c + blk;
    const bool bv = b32 < nblk32;
    const char *cb = base + (long)(bv ? b32 : 0) * g.xbp * 2;
    const bool v = bv && row < g.m;
    p.v[i] = *(const uint4 *)(cb + (v ? row * (int)g.ldI * 2 + seg * 16 : 0));
    ok |= (v ? 1u : 0u) << i;
  }
  p.ok = ok;
  // the scale word of LDS row tid (threads past NR load row NR - 1's and do
  // not stage it); clamped indices, so the loads are unconditional
  const int srow = f8c_chrow(tid < NR ? tid : NR - 1, r0, c0);
  const bool rv = srow < g.m;
  unsigned sc = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {  // byte j = 2 kh + ks: 16-spectrum block 4c + 2 ks + kh
    const int b16 = 4 * c + 2 * (j & 1) + (j >> 1);
    const bool bv = b16 < g.nblk16;
    const unsigned char *sb = g.f8scales + (long)(bv ? b16 : 0) * g.m;
    const unsigned sv = sb[rv ? srow : 0];
    sc |= ((bv && rv) ? sv : 127u) << (8 * j);
  }
  p.sc = sc;
}

template <int NR>
__device__ __forceinline__ void f8c_stage(const F8cPrefetch<NR> &p, char *buf) {
  constexpr int L = F8cPrefetch<NR>::L;
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < L; i++) {
    const int blk = i / (L / 2);
    const int rem = tid + (i - blk * (L / 2)) * F8C_THREADS;
    const unsigned m = 0u - ((p.ok >> i) & 1u);  // all ones for a valid piece (a value mask, not a select)
    const uint4 x = p.v[i];
    *(uint4 *)(buf + (rem >> 2) * F8C_ROWB + blk * 64 + (rem & 3) * 16) = make_uint4(x.x & m, x.y & m, x.z & m, x.w & m);
  }
  if (tid < NR) ((unsigned *)(buf + F8C_MAXR * F8C_ROWB))[tid] = p.sc;
}

// operand and scale of LDS row `row` at K step ks for lane half kh
__device__ __forceinline__ i32x8 f8c_op(const char *buf, int row, int ks, int kh) {
  return ld_op_blk(buf + row * F8C_ROWB + ks * 64, kh);
}
__device__ __forceinline__ int f8c_sc(const char *buf, int row, int ks, int kh) {
  return (int)((((const unsigned *)(buf + F8C_MAXR * F8C_ROWB))[row] >> (16 * kh + 8 * ks)) & 0xffu);
}

__device__ __forceinline__ void f8c_block(i32x8 I, i32x8 Pr, i32x8 J, int sI, int sJ, f32x16 &aR, f32x16 &aI) {
  aR = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(I, J, aR, 0, 0, 0, sI, 0, sJ);
  aI = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(Pr, J, aI, 0, 0, 0, sI, 0, sJ);
}

// no LDS read moves across this point (everything else may): keeps each
// operand's load one block ahead of its MFMAs instead of all of a chunk's
// loads hoisted to its top, which would need more registers than a wave has
__device__ __forceinline__ void f8c_fence_lds_reads() {
  __builtin_amdgcn_sched_barrier(0x0002 | 0x0004 | 0x0008 | 0x0010 | 0x0020 | 0x0040 | 0x0200);
}

// A wave of a triangle: LDS block row r, blocks (r, 0..4); the next block's
// J operand is read from LDS while the current block's MFMAs run
__device__ __forceinline__ void f8c_chunk_a(const char *buf, int r, int l32, int kh, f32x16 (&accR)[F8C_A],
                                            f32x16 (&accI)[F8C_A]) {
#pragma unroll
  for (int ks = 0; ks < 2; ks++) {
    const int irow = r * 32 + l32;
    const i32x8 I = f8c_op(buf, irow, ks, kh);
    const int sI = f8c_sc(buf, irow, ks, kh);
    i32x8 J = f8c_op(buf, l32, ks, kh);
    int sJ = f8c_sc(buf, l32, ks, kh);
    const i32x8 Pr = rot_conj8(I);
#pragma unroll
    for (int j = 0; j < F8C_A; j++) {
      i32x8 Jn = J;
      int sJn = sJ;
      if (j + 1 < F8C_A) {
        Jn = f8c_op(buf, (j + 1) * 32 + l32, ks, kh);
        sJn = f8c_sc(buf, (j + 1) * 32 + l32, ks, kh);
      }
      f8c_block(I, Pr, J, sI, sJ, accR[j], accI[j]);
      f8c_fence_lds_reads();
      J = Jn;
      sJ = sJn;
    }
  }
}

// B waves of a triangle, wave 4 + W: blocks (r[b], c[b]) as constants;
// consecutive blocks of one row share its I operand, a diagonal block uses
// its I operand as J
template <int W> struct F8cBw;
template <> struct F8cBw<0> { static constexpr int r[F8C_B] = {7, 7, 7, 0}, c[F8C_B] = {5, 6, 7, 0}; };
template <> struct F8cBw<1> { static constexpr int r[F8C_B] = {6, 6, 1, 1}, c[F8C_B] = {5, 6, 0, 1}; };
template <> struct F8cBw<2> { static constexpr int r[F8C_B] = {5, 2, 2, 2}, c[F8C_B] = {5, 0, 1, 2}; };
template <> struct F8cBw<3> { static constexpr int r[F8C_B] = {3, 3, 3, 3}, c[F8C_B] = {0, 1, 2, 3}; };

template <int W, int B>
__device__ __forceinline__ void f8c_b_step(const char *buf, int l32, int ks, int kh, i32x8 &I, i32x8 &Pr, int &sI,
                                           f32x16 (&accR)[F8C_B], f32x16 (&accI)[F8C_B]) {
  if constexpr (B < F8C_B) {
    constexpr int R = F8cBw<W>::r[B], Cc = F8cBw<W>::c[B];
    if constexpr (B == 0 || R != F8cBw<W>::r[B > 0 ? B - 1 : 0]) {
      const int irow = R * 32 + l32;
      I = f8c_op(buf, irow, ks, kh);
      Pr = rot_conj8(I);
      sI = f8c_sc(buf, irow, ks, kh);
    }
    if constexpr (Cc == R) {
      f8c_block(I, Pr, I, sI, sI, accR[B], accI[B]);
    } else {
      const int jrow = Cc * 32 + l32;
      f8c_block(I, Pr, f8c_op(buf, jrow, ks, kh), sI, f8c_sc(buf, jrow, ks, kh), accR[B], accI[B]);
    }
    f8c_b_step<W, B + 1>(buf, l32, ks, kh, I, Pr, sI, accR, accI);
  }
}

template <int W>
__device__ __forceinline__ void f8c_chunk_b(const char *buf, int l32, int kh, f32x16 (&accR)[F8C_B],
                                            f32x16 (&accI)[F8C_B]) {
#pragma unroll
  for (int ks = 0; ks < 2; ks++) {
    i32x8 I, Pr;
    int sI;
    f8c_b_step<W, 0>(buf, l32, ks, kh, I, Pr, sI, accR, accI);
  }
}

// a rectangle's wave w: I block row w (LDS rows 32 w..), J column blocks 0..3 (LDS rows 256 + 32 j..)
__device__ __forceinline__ void f8c_chunk_r(const char *buf, int w, int l32, int kh, f32x16 (&accR)[F8C_R],
                                            f32x16 (&accI)[F8C_R]) {
#pragma unroll
  for (int ks = 0; ks < 2; ks++) {
    const int irow = w * 32 + l32;
    const i32x8 I = f8c_op(buf, irow, ks, kh);
    const i32x8 Pr = rot_conj8(I);
    const int sI = f8c_sc(buf, irow, ks, kh);
#pragma unroll
    for (int j = 0; j < F8C_R; j++) {
      const int jrow = 256 + j * 32 + l32;
      f8c_block(I, Pr, f8c_op(buf, jrow, ks, kh), sI, f8c_sc(buf, jrow, ks, kh), accR[j], accI[j]);
    }
  }
}

// the chunk loop, branch-free: at the top of iteration c (after its barrier)
// chunk c + 1, fetched during iteration c - 1, is staged into buffer
// (c + 1) % 2 (last read for chunk c - 1, before the barrier), chunk c + 2 is
// fetched into registers, then chunk c is multiplied out of buffer c % 2.
// Past the last chunk the fetch reads clamped addresses, masked to zeros, and
// the staging fills a buffer that is not read again.  The MFMAs end the loop
// body, so each operand's LDS read stays next to its use.
template <int NR, typename Chunk>
__device__ __forceinline__ void f8c_loop(char *lds, const char *base, const XArgs &g, int r0, int c0, Chunk chunk) {
  const int nchunk = (g.k + F8C_KC - 1) / F8C_KC;
  F8cPrefetch<NR> pa;
  f8c_fetch<NR>(pa, base, g, r0, c0, 0);
  f8c_stage<NR>(pa, lds);
  f8c_fetch<NR>(pa, base, g, r0, c0, 1);
  for (int c = 0; c < nchunk; c++) {
    __syncthreads();  // chunk c staged in buffer c % 2; buffer (c + 1) % 2 free
    f8c_stage<NR>(pa, lds + ((c + 1) & 1) * F8C_BUF);
    f8c_fetch<NR>(pa, base, g, r0, c0, c + 2);
    chunk(lds + (c & 1) * F8C_BUF);
  }
}

template <int NB>
__device__ __forceinline__ void f8c_zero(f32x16 (&accR)[NB], f32x16 (&accI)[NB]) {
#pragma unroll
  for (int b = 0; b < NB; b++)
#pragma unroll
    for (int q = 0; q < 16; q++) {
      accR[b][q] = 0.f;
      accI[b][q] = 0.f;
    }
}

// block (a0 block row, b0 block column) in channel rows: D[i][j] = V[j][i]
template <int NB>
__device__ __forceinline__ void f8c_out(const XArgs &g, int batch, const int (&ra)[NB], const int (&rb)[NB], int l32,
                                        int kh, const f32x16 (&accR)[NB], const f32x16 (&accI)[NB]) {
#pragma unroll
  for (int b = 0; b < NB; b++) {
    tri_block_out<8>(g.Cbase, (double2 *)g.acc64, (long)batch * g.Cbatch, g.m, ra[b], kh, rb[b] + l32, g.alpha.x,
                     accR[b], accI[b]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int W>
__device__ __forceinline__ void f8c_run_b(char *lds, const char *base, const XArgs &g, int batch, int r0, int l32,
                                          int kh) {
  f32x16 accR[F8C_B], accI[F8C_B];
  f8c_zero(accR, accI);
  f8c_loop<256>(lds, base, g, r0, 0, [&](const char *buf) { f8c_chunk_b<W>(buf, l32, kh, accR, accI); });
  int ra[F8C_B], rb[F8C_B];
#pragma unroll
  for (int b = 0; b < F8C_B; b++) {
    ra[b] = r0 + 32 * F8cBw<W>::r[b];
    rb[b] = r0 + 32 * F8cBw<W>::c[b];
  }
  f8c_out(g, batch, ra, rb, l32, kh, accR, accI);
}

__global__ __launch_bounds__(F8C_THREADS, 2) void xeng_f8b_chan_kernel(XArgs g, int kinds) {
  __shared__ __attribute__((aligned(16))) char lds[2 * F8C_BUF];
  const long nwg = gridDim.x;
  const long w = blockIdx.x;
  // XCD-aware: a channel's workgroups are consecutive logical indices on one XCD
  const long logical = (nwg % 8 == 0) ? (w % 8) * (nwg / 8) + w / 8 : w;
  const int batch = (int)(logical / kinds);
  const int kind = (int)(logical % kinds);
  const char *base = (const char *)g.Abase + xin_batch(g, batch) * 2;  // 2-byte samples
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int l32 = lane & 31, kh = lane >> 5;

  if (kind >= 2) {  // rectangle: rows [256, 512) x columns [c0, c0 + 128)
    const int c0 = 128 * (kind - 2);
    f32x16 accR[F8C_R], accI[F8C_R];
    f8c_zero(accR, accI);
    f8c_loop<F8C_MAXR>(lds, base, g, 256, c0, [&](const char *buf) { f8c_chunk_r(buf, wave, l32, kh, accR, accI); });
    int ra[F8C_R], rb[F8C_R];
#pragma unroll
    for (int b = 0; b < F8C_R; b++) {
      ra[b] = 256 + 32 * wave;
      rb[b] = c0 + 32 * b;
    }
    f8c_out(g, batch, ra, rb, l32, kh, accR, accI);
    return;
  }
  const int r0 = 256 * kind;  // a triangle of rows [r0, r0 + 256)
  if (wave < 4) {
    const int r = 7 - wave;
    f32x16 accR[F8C_A], accI[F8C_A];
    f8c_zero(accR, accI);
    f8c_loop<256>(lds, base, g, r0, 0, [&](const char *buf) { f8c_chunk_a(buf, r, l32, kh, accR, accI); });
    int ra[F8C_A], rb[F8C_A];
#pragma unroll
    for (int b = 0; b < F8C_A; b++) {
      ra[b] = r0 + 32 * r;
      rb[b] = r0 + 32 * b;
    }
    f8c_out(g, batch, ra, rb, l32, kh, accR, accI);
  } else {
    switch (wave) {
      case 4: f8c_run_b<0>(lds, base, g, batch, r0, l32, kh); break;
      case 5: f8c_run_b<1>(lds, base, g, batch, r0, l32, kh); break;
      case 6: f8c_run_b<2>(lds, base, g, batch, r0, l32, kh); break;
      default: f8c_run_b<3>(lds, base, g, batch, r0, l32, kh); break;
    }
  }
}

int xeng_launch_f8_block(const XArgs &g0, hipStream_t st) {
  XArgs g = g0;
  if (g.f8scales == nullptr) return fail(MDT_ERR_INVALID_VALUE, "FP8 block-scaled input needs its scales");
  if (g.m > 128 && g.m <= 512 && g.ldI == g.ldJ && g.Abase == g.Bbase) {
    const int kinds = g.m > 256 ? 4 : 1;
    const long total = (long)kinds * g.batch;
    MDT_TRY(check_grid(total, F8C_THREADS, "X-engine"));
    hipLaunchKernelGGL(xeng_f8b_chan_kernel, dim3((unsigned)total), dim3(F8C_THREADS), 0, st, g, kinds);
    return check_launch("xeng_f8b_chan_kernel");
  }
  const int nb = (g.m + Q8T - 1) / Q8T;
  g.tiles = nb * (nb + 1) / 2;
  const long total = (long)g.tiles * g.batch;
  MDT_TRY(check_grid(total, 256, "X-engine"));
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
