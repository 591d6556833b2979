// xengine_f16.hip -- the split-precision X-engine of the native path on
// v_mfma_f32_32x32x16_f16: each FP32 sample x (pre-scaled by an exact power
// of two) is split into hi = fp16(x) and lo = fp16(x - hi) by the F-engine's
// epilogue, and
//   x_i * x_j ~= hi_i hi_j + hi_i lo_j + lo_i hi_j        (FP32 accumulate)
// which keeps ~22 significant bits per operand; the dropped lo_i lo_j term is
// below 2^-22 relative.  Same complex scheme as the FP32 kernel: K runs over
// interleaved (re, im) halves, the I operand is fed as (-im, re) for the
// imaginary part, D[i][j] = V[j][i].
//
// Workgroup tile 128 x 128 complex (4 waves of 64 x 64 = 2 x 2 MFMA tiles),
// K streamed in chunks of 16 samples; the next chunk is prefetched into
// registers while the current one is multiplied out of LDS.  Only tiles on or
// below the diagonal are visited and 32 x 32 MFMA tiles strictly above it are
// skipped.  Output: packed lower triangle, optional FP64 accumulation.
#include "common.hpp"
#include "xengine.hpp"

namespace mdt {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int HT = 128;               // workgroup tile edge (complex rows)
constexpr int HKC = 16;               // samples per K chunk
constexpr int HROWB = HKC * 4 + 16;   // bytes per LDS row per plane (64 data + 16 pad)
constexpr int HPLANE = HT * HROWB;    // bytes per plane

__device__ __forceinline__ unsigned pack_h2(_Float16 a, _Float16 b) {
  return (unsigned)__builtin_bit_cast(unsigned short, a) |
         ((unsigned)__builtin_bit_cast(unsigned short, b) << 16);
}


// (re, im) half pairs -> (-im, re)
__device__ __forceinline__ half8 rot_conj(half8 v) {
  u32x4 u = __builtin_bit_cast(u32x4, v);
#pragma unroll
  for (int q = 0; q < 4; q++) u[q] = __builtin_amdgcn_alignbit(u[q], u[q], 16) ^ 0x8000u;
  return __builtin_bit_cast(half8, u);
}

struct Prefetch {
  float4 v[4];
  unsigned ok;  // bit 2q: first sample of row q valid, bit 2q+1: second
};

// global -> registers: 128 rows x 16 samples (8 lanes x 16 B per row).
// Branch-free so the four loads are in flight together: out-of-range rows and
// samples read a clamped, valid address and are zeroed afterwards.  Rows are
// 16-byte aligned (even row pitch, even channel pitch).
__device__ __forceinline__ void fetch(Prefetch &p, const float2 *__restrict__ base, long ld,
                                      int row0, int nrows, int s0, int K, int lgB, long bp) {
  const int seg = threadIdx.x & 7, rq = threadIdx.x >> 3;
  const int s = s0 + 2 * seg;
  const bool s_ok0 = s < K, s_ok1 = s + 1 < K;
  const long so = xin_spec(lgB, bp, s_ok0 ? s : 0);  // pairs (s, s+1) never straddle a block
  unsigned ok = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int row = row0 + rq + 32 * q;
    const bool r_ok = row < nrows;
    p.v[q] = *(const float4 *)(base + (long)(r_ok ? row : 0) * ld + so);
    ok |= ((r_ok && s_ok0) ? 1u : 0u) << (2 * q);
    ok |= ((r_ok && s_ok1) ? 1u : 0u) << (2 * q + 1);
  }
  p.ok = ok;  // masks applied when staging, so the loads stay in flight
}

// registers -> LDS hi/lo planes, [row][k] with k = 2*s + (re|im).  The X
// input already holds each sample as (hi_re, hi_im | lo_re, lo_im) halves (the
// F-engine's split epilogue), so staging only routes dwords to the planes.
__device__ __forceinline__ void stage(const Prefetch &p, char *lds) {
  const int seg = threadIdx.x & 7, rq = threadIdx.x >> 3;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int row = rq + 32 * q;
    const float4 v = p.v[q];
    const unsigned m0 = (p.ok >> (2 * q)) & 1u ? 0xFFFFFFFFu : 0u;
    const unsigned m1 = (p.ok >> (2 * q + 1)) & 1u ? 0xFFFFFFFFu : 0u;
    *(uint2 *)(lds + row * HROWB + seg * 8) =
        make_uint2(__float_as_uint(v.x) & m0, __float_as_uint(v.z) & m1);
    *(uint2 *)(lds + HPLANE + row * HROWB + seg * 8) =
        make_uint2(__float_as_uint(v.y) & m0, __float_as_uint(v.w) & m1);
  }
}


__global__ __launch_bounds__(256, 2) void xeng_f16split_kernel(XArgs g, float oscale) {
  if (g.scale_dev != nullptr) {  // the scale the F-engine used (gains change it)
    const float sc = *g.scale_dev;
    oscale = g.alpha.x / (sc * sc);
  }
  __shared__ __attribute__((aligned(16))) char lds[2 * 2 * HPLANE];  // [op I|J][hi|lo][row][k]

  const long nwg = gridDim.x;
  const long w = blockIdx.x;
  const long logical = (nwg % 8 == 0) ? (w % 8) * (nwg / 8) + w / 8 : w;
  const int batch = (int)(logical / g.tiles);
  const int tile = (int)(logical % g.tiles);
  int P = 0;
  while ((P + 1) * (P + 2) / 2 <= tile) P++;
  const int Q = tile - P * (P + 1) / 2;
  const int bi = P, bj = Q;  // I rows <-> a, J rows <-> b, a >= b
  const bool same = (bi == bj);
  const float2 *base = g.Abase + xin_batch(g, batch);
  const int NN = g.m, K = g.k;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // NN <= 64: the one tile's three 32 x 32 blocks go to waves 0, 1, 2 (all
  // waves address the tile's first 64 rows) instead of all to wave 0 (at 40
  // inputs the X-engine 1.28 -> 1.22 ms; each block's products unchanged)
  const bool small = NN <= 64;
  const int wi = small ? 0 : wave >> 1, wj = small ? 0 : wave & 1;
  const int l32 = lane & 31, kh = lane >> 5;
  // 32-row block indices of this wave's 2 x 2 MFMA tiles
  bool act[2][2];
  bool any = false;
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int ra = bi * 4 + wi * 2 + t, rb = bj * 4 + wj * 2 + u;
      act[t][u] = (ra >= rb) && (ra * 32 < NN) && (rb * 32 < NN) && (!small || wave == t + u);
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
  char *ldsJ = same ? lds : lds + 2 * HPLANE;
  Prefetch pI, pJ;
  fetch(pI, base, g.ldI, bi * HT, NN, 0, K, g.xlgB, g.xbp);
  if (!same) fetch(pJ, base, g.ldJ, bj * HT, NN, 0, K, g.xlgB, g.xbp);

  for (int s0 = 0; s0 < K; s0 += HKC) {
    __syncthreads();
    stage(pI, lds);
    if (!same) stage(pJ, lds + 2 * HPLANE);
    __syncthreads();
    if (s0 + HKC < K) {  // prefetch the next chunk while this one is multiplied
      fetch(pI, base, g.ldI, bi * HT, NN, s0 + HKC, K, g.xlgB, g.xbp);
      if (!same) fetch(pJ, base, g.ldJ, bj * HT, NN, s0 + HKC, K, g.xlgB, g.xbp);
    }
    if (any) {
#pragma unroll
      for (int ks = 0; ks < HKC / 8; ks++) {
        const int koff = ks * 32 + kh * 16;  // bytes: 8 samples = 16 halves per k-step
        half8 Ih[2], Il[2], Jh[2], Jl[2], Ph[2], Pl[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
          const int row = wi * 64 + t * 32 + l32;
          Ih[t] = *(const half8 *)(ldsI + row * HROWB + koff);
          Il[t] = *(const half8 *)(ldsI + HPLANE + row * HROWB + koff);
          Ph[t] = rot_conj(Ih[t]);
          Pl[t] = rot_conj(Il[t]);
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
          const int row = wj * 64 + u * 32 + l32;
          Jh[u] = *(const half8 *)(ldsJ + row * HROWB + koff);
          Jl[u] = *(const half8 *)(ldsJ + HPLANE + row * HROWB + koff);
        }
        // three split products, each swept over the 2 x 2 tiles so that
        // dependent MFMAs on one accumulator are 8 instructions apart
#pragma unroll
        for (int p = 0; p < 3; p++) {
#pragma unroll
          for (int t = 0; t < 2; t++)
#pragma unroll
            for (int u = 0; u < 2; u++) {
              if (!act[t][u]) continue;
              const half8 a = (p == 2) ? Il[t] : Ih[t];
              const half8 ar = (p == 2) ? Pl[t] : Ph[t];
              const half8 bb = (p == 1) ? Jl[u] : Jh[u];
              accR[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, bb, accR[t][u], 0, 0, 0);
              accI[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ar, bb, accI[t][u], 0, 0, 0);
            }
        }
      }
    }
  }
  if (!any) return;

  // D[i][j] = V[b = j][a = i]; V[a][b] = conj(D) * alpha / sc^2
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int u = 0; u < 2; u++) {
      if (!act[t][u]) continue;
      tri_block_out<8>(g.Cbase, (double2 *)g.acc64, (long)batch * g.Cbatch, NN, bi * HT + wi * 64 + t * 32, kh,
                    bj * HT + wj * 64 + u * 32 + l32, oscale, accR[t][u], accI[t][u]);
      __builtin_amdgcn_sched_barrier(0);
    }
}

// ---------------------------------------------------------------------------
// Whole-channel kernel for 128 < NN <= 256 (the 256-input scaled array): one workgroup per
// channel computes every 32 x 32 block of the lower triangle (36 blocks for
// NN = 256), so each sample is read from HBM exactly once (128 x 128 tiles
// read a 256-input channel twice).  8 waves, two per SIMD with 256 registers
// each: waves 0-3 own 5 blocks and waves 4-7 own 4, so each SIMD carries 9.
// K chunks of 16 spectra for all 256 rows are double-buffered in LDS (one
// barrier per chunk) and prefetched two chunks ahead in registers, so about
// 64 KB of loads per CU are in flight while the MFMAs run.  Consecutive blocks
// of a wave mostly share their I rows; the I operand is reloaded only when
// the row changes.
constexpr int WC_NB = 8;                  // 32-row blocks (NN <= 256)
constexpr int WC_WAVES = 8, WC_THREADS = WC_WAVES * 64;
constexpr int WC_KC = 16;                 // samples per K chunk
constexpr int WC_ROWB = WC_KC * 4 + 16;   // bytes per row per plane (64 data + 16 pad)
constexpr int WC_PLANE = WC_NB * 32 * WC_ROWB;
constexpr int WC_BUF = 2 * WC_PLANE;      // hi + lo planes
constexpr int WC_SEGS = WC_KC / 2;        // 16-byte segments (2 samples) per row per chunk
constexpr int WC_LOADS = WC_NB * 32 * WC_SEGS / WC_THREADS;  // per thread per chunk
static_assert(WC_NB * 32 * WC_SEGS % WC_THREADS == 0, "chunk loads must divide among the threads");

// Block assignment: rows pair up (7, 0), (6, 1), (5, 2), (4, 3), 9 blocks a
// pair, one pair per SIMD.  Waves 0-3 ("A", one per SIMD) take the five
// blocks (r, 0..4) of row r = 7 - wave: one I row, J rows known at compile
// time, straight-line code with no per-block branch.  Waves 4-7 ("B") take
// the remaining four blocks of their pair (WcBw below), each wave's blocks
// as constants of its own code path.
constexpr int WC_A = 5, WC_B = 4;

struct WcPrefetch {
  float4 v[WC_LOADS];
  unsigned ok0, ok1;  // bit i: first / second sample of load i valid
};

__device__ __forceinline__ void wc_fetch(WcPrefetch &p, const float2 *__restrict__ base, long ld, int lgB,
                                         long bp, int NN, int s0, int K) {
  unsigned ok0 = 0, ok1 = 0;
#pragma unroll
  for (int i = 0; i < WC_LOADS; i++) {
    const int c = threadIdx.x + i * WC_THREADS;
    const int row = c / WC_SEGS, seg = c % WC_SEGS;
    const int s = s0 + 2 * seg;
    const bool r_ok = row < NN;
    const int sc = s < K ? s : 0;
    p.v[i] = *(const float4 *)(base + (long)(r_ok ? row : 0) * ld + xin_spec(lgB, bp, sc));
    ok0 |= ((r_ok && s < K) ? 1u : 0u) << i;
    ok1 |= ((r_ok && s + 1 < K) ? 1u : 0u) << i;
  }
  p.ok0 = ok0;
  p.ok1 = ok1;
}

__device__ __forceinline__ void wc_stage(const WcPrefetch &p, char *lds) {
#pragma unroll
  for (int i = 0; i < WC_LOADS; i++) {
    const int c = threadIdx.x + i * WC_THREADS;
    const int row = c / WC_SEGS, seg = c % WC_SEGS;
    const float4 v = p.v[i];
    const unsigned m0 = (p.ok0 >> i) & 1u ? 0xFFFFFFFFu : 0u;
    const unsigned m1 = (p.ok1 >> i) & 1u ? 0xFFFFFFFFu : 0u;
    *(uint2 *)(lds + row * WC_ROWB + seg * 8) = make_uint2(__float_as_uint(v.x) & m0, __float_as_uint(v.z) & m1);
    *(uint2 *)(lds + WC_PLANE + row * WC_ROWB + seg * 8) =
        make_uint2(__float_as_uint(v.y) & m0, __float_as_uint(v.w) & m1);
  }
}

// the six products of one block and one K-step: real and imaginary parts of
// hi*hi + hi*lo + lo*hi
__device__ __forceinline__ void wc_block(half8 Ih, half8 Il, half8 Ph, half8 Pl, half8 Jh, half8 Jl, f32x16 &aR,
                                         f32x16 &aI) {
  aR = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ih, Jh, aR, 0, 0, 0);
  aI = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ph, Jh, aI, 0, 0, 0);
  aR = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ih, Jl, aR, 0, 0, 0);
  aI = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ph, Jl, aI, 0, 0, 0);
  aR = __builtin_amdgcn_mfma_f32_32x32x16_f16(Il, Jh, aR, 0, 0, 0);
  aI = __builtin_amdgcn_mfma_f32_32x32x16_f16(Pl, Jh, aI, 0, 0, 0);
}

// Diagonal 32 x 32 blocks as three 16 x 16 blocks: rows 0-15 x columns 0-15,
// rows 16-31 x columns 0-15 and rows 16-31 x columns 16-31 (the fourth lies
// above the diagonal), on v_mfma_f32_16x16x32_f16, which takes a chunk's 16
// samples in one step: 3/4 of the block's MFMA work.  Lane l supplies row
// l % 16 and samples 4 (l / 16) .. +3 of the chunk; result register r holds
// D[4 (l / 16) + r][l % 16].
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void wc_block16(half8 Ih, half8 Il, half8 Ph, half8 Pl, half8 Jh, half8 Jl, f32x4 &aR,
                                           f32x4 &aI) {
  aR = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ih, Jh, aR, 0, 0, 0);
  aI = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ph, Jh, aI, 0, 0, 0);
  aR = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ih, Jl, aR, 0, 0, 0);
  aI = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ph, Jl, aI, 0, 0, 0);
  aR = __builtin_amdgcn_mfma_f32_16x16x32_f16(Il, Jh, aR, 0, 0, 0);
  aI = __builtin_amdgcn_mfma_f32_16x16x32_f16(Pl, Jh, aI, 0, 0, 0);
}
// one chunk of diagonal block R (rows 32 R .. 32 R + 31)
__device__ __forceinline__ void wc_diag16(const char *bufp, int R, int lane, f32x4 (&dR)[3], f32x4 (&dI)[3]) {
  const int row0 = R * 32 + (lane & 15), row1 = row0 + 16;
  const int koff = (lane >> 4) * 16;
  const half8 A0h = *(const half8 *)(bufp + row0 * WC_ROWB + koff);
  const half8 A0l = *(const half8 *)(bufp + WC_PLANE + row0 * WC_ROWB + koff);
  const half8 A1h = *(const half8 *)(bufp + row1 * WC_ROWB + koff);
  const half8 A1l = *(const half8 *)(bufp + WC_PLANE + row1 * WC_ROWB + koff);
  const half8 P0h = rot_conj(A0h), P0l = rot_conj(A0l), P1h = rot_conj(A1h), P1l = rot_conj(A1l);
  wc_block16(A0h, A0l, P0h, P0l, A0h, A0l, dR[0], dI[0]);
  wc_block16(A1h, A1l, P1h, P1l, A0h, A0l, dR[1], dI[1]);
  wc_block16(A1h, A1l, P1h, P1l, A1h, A1l, dR[2], dI[2]);
}
// V of the three 16 x 16 blocks of diagonal block R (as tri_block_out)
__device__ __forceinline__ void wc_diag16_out(const XArgs &g, int batch, float oscale, int R, int lane,
                                              const f32x4 (&dR)[3], const f32x4 (&dI)[3]) {
  float2 *__restrict__ C = g.Cbase;
  double2 *__restrict__ acc = (double2 *)g.acc64;
  const long cb = (long)batch * g.Cbatch;
  const int NN = g.m;
#pragma unroll
  for (int sb = 0; sb < 3; sb++) {
    const int a0 = R * 32 + (sb == 0 ? 0 : 16) + 4 * (lane >> 4);
    const int b = R * 32 + (sb == 2 ? 16 : 0) + (lane & 15);
    long off[4];
    bool ok[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int a = a0 + r;
      ok[r] = a < NN && b < NN && a >= b;
      off[r] = cb + (ok[r] ? (long)a * (a + 1) / 2 + b : 0);
    }
    if (acc) {
      double2 old[4];
#pragma unroll
      for (int r = 0; r < 4; r++) old[r] = acc[off[r]];
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (ok[r]) {
          const float vr = oscale * dR[sb][r], vi = -oscale * dI[sb][r];
          C[off[r]] = make_float2(vr, vi);
          acc[off[r]] = make_double2(old[r].x + (double)vr, old[r].y + (double)vi);
        }
    } else {
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (ok[r]) C[off[r]] = make_float2(oscale * dR[sb][r], -oscale * dI[sb][r]);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// one chunk (2 K-steps) of an A wave: I row `irow`, J rows 32j + l32, j = 0..4;
// DIAG (wave 3, row 4): block (4, 4) is diagonal, three 16 x 16 blocks
template <bool DIAG>
__device__ __forceinline__ void wc_chunk_a(const char *bufp, int irow, int l32, int kh, int lane,
                                           f32x16 (&accR)[WC_A], f32x16 (&accI)[WC_A], f32x4 (&dR)[3],
                                           f32x4 (&dI)[3]) {
#pragma unroll
  for (int ks = 0; ks < WC_KC / 8; ks++) {
    const int koff = ks * 32 + kh * 16;
    const half8 Ih = *(const half8 *)(bufp + irow * WC_ROWB + koff);
    const half8 Il = *(const half8 *)(bufp + WC_PLANE + irow * WC_ROWB + koff);
    const half8 Ph = rot_conj(Ih), Pl = rot_conj(Il);
#pragma unroll
    for (int j = 0; j < (DIAG ? WC_A - 1 : WC_A); j++) {
      const int jrow = j * 32 + l32;
      const half8 Jh = *(const half8 *)(bufp + jrow * WC_ROWB + koff);
      const half8 Jl = *(const half8 *)(bufp + WC_PLANE + jrow * WC_ROWB + koff);
      wc_block(Ih, Il, Ph, Pl, Jh, Jl, accR[j], accI[j]);
    }
  }
  if constexpr (DIAG) wc_diag16(bufp, WC_A - 1, lane, dR, dI);
}

// A wave: the five blocks (r, 0..4) of row r over all chunks, then their
// epilogue; r = 4 (DIAG) has the diagonal block (4, 4)
template <bool DIAG>
__device__ __forceinline__ void wc_run_a(char *lds, const float2 *base, const XArgs &g, int batch, float oscale,
                                         int NN, int K, int r, int l32, int kh, int lane) {
  constexpr int NO = DIAG ? WC_A - 1 : WC_A;
  int ra[NO], rb[NO];
  f32x16 accR[WC_A], accI[WC_A];
  f32x4 dR[3], dI[3];
#pragma unroll
  for (int b = 0; b < WC_A; b++) {
    if (b < NO) {
      ra[b] = r;
      rb[b] = b;
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
      accR[b][q] = 0.f;
      accI[b][q] = 0.f;
    }
  }
#pragma unroll
  for (int sb = 0; sb < 3; sb++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      dR[sb][q] = 0.f;
      dI[sb][q] = 0.f;
    }
  const int irow = r * 32 + l32;
  wc_loop(lds, base, g.ldI, g.xlgB, g.xbp, NN, K,
          [&](const char *bufp) { wc_chunk_a<DIAG>(bufp, irow, l32, kh, lane, accR, accI, dR, dI); });
  f32x16 oR[NO], oI[NO];
#pragma unroll
  for (int b = 0; b < NO; b++) {
    oR[b] = accR[b];
    oI[b] = accI[b];
  }
  wc_epilogue(g, batch, oscale, ra, rb, l32, kh, oR, oI);
  if constexpr (DIAG) wc_diag16_out(g, batch, oscale, WC_A - 1, lane, dR, dI);
}

// The B waves' blocks (row, column) as constants, wave 4 + W: consecutive
// blocks of one row share its I operand (loaded and rotated once), and a
// diagonal block's J operand is its I.  Per K-step a B wave reads 8 operand
// rows from LDS instead of 16 (X-engine 7.22-7.25 -> 7.01-7.03 ms at 256
// inputs, at the power limit); the MFMAs and their order are those of one
// block after another, as in the A waves.
template <int W> struct WcBw;
template <> struct WcBw<0> { static constexpr int r[WC_B] = {7, 7, 7, 0}, c[WC_B] = {5, 6, 7, 0}; };
template <> struct WcBw<1> { static constexpr int r[WC_B] = {6, 6, 1, 1}, c[WC_B] = {5, 6, 0, 1}; };
template <> struct WcBw<2> { static constexpr int r[WC_B] = {5, 2, 2, 2}, c[WC_B] = {5, 0, 1, 2}; };
template <> struct WcBw<3> { static constexpr int r[WC_B] = {3, 3, 3, 3}, c[WC_B] = {0, 1, 2, 3}; };

template <int W, int B>
__device__ __forceinline__ void wc_b_step(const char *bufp, int l32, int koff, half8 &Ih, half8 &Il, half8 &Ph,
                                          half8 &Pl, f32x16 (&accR)[WC_B], f32x16 (&accI)[WC_B]) {
  if constexpr (B < WC_B) {
    constexpr int R = WcBw<W>::r[B], Cc = WcBw<W>::c[B];
    if constexpr (B == 0 || R != WcBw<W>::r[B > 0 ? B - 1 : 0]) {
      const int irow = R * 32 + l32;
      Ih = *(const half8 *)(bufp + irow * WC_ROWB + koff);
      Il = *(const half8 *)(bufp + WC_PLANE + irow * WC_ROWB + koff);
      Ph = rot_conj(Ih);
      Pl = rot_conj(Il);
    }
    if constexpr (Cc == R) {
      // diagonal: three 16 x 16 blocks per chunk (wc_chunk_b)
    } else {
      const int jrow = Cc * 32 + l32;
      const half8 Jh = *(const half8 *)(bufp + jrow * WC_ROWB + koff);
      const half8 Jl = *(const half8 *)(bufp + WC_PLANE + jrow * WC_ROWB + koff);
      wc_block(Ih, Il, Ph, Pl, Jh, Jl, accR[B], accI[B]);
    }
    wc_b_step<W, B + 1>(bufp, l32, koff, Ih, Il, Ph, Pl, accR, accI);
  }
}

// one chunk of B wave 4 + W
template <int W>
constexpr int wc_b_ndiag() {
  int n = 0;
  for (int b = 0; b < WC_B; b++) n += WcBw<W>::r[b] == WcBw<W>::c[b];
  return n;
}
template <int W>
constexpr int wc_b_diag_row(int d) {  // row of wave 4 + W's d-th diagonal block
  for (int b = 0; b < WC_B; b++)
    if (WcBw<W>::r[b] == WcBw<W>::c[b] && d-- == 0) return WcBw<W>::r[b];
  return -1;
}
template <int W>
__device__ __forceinline__ void wc_chunk_b(const char *bufp, int l32, int kh, int lane, f32x16 (&accR)[WC_B],
                                           f32x16 (&accI)[WC_B], f32x4 (&dR)[2][3], f32x4 (&dI)[2][3]) {
#pragma unroll
  for (int ks = 0; ks < WC_KC / 8; ks++) {
    half8 Ih, Il, Ph, Pl;
    wc_b_step<W, 0>(bufp, l32, ks * 32 + kh * 16, Ih, Il, Ph, Pl, accR, accI);
  }
#pragma unroll
  for (int d = 0; d < wc_b_ndiag<W>(); d++) wc_diag16(bufp, wc_b_diag_row<W>(d), lane, dR[d], dI[d]);
}

// V for blocks (ra[b], rb[b]): D[i][j] = V[j][i], so accumulator register r of
// lane l holds V[32 ra + 8(r/4) + 4 kh + r%4][32 rb + l32]
template <int NB>
__device__ __forceinline__ void wc_epilogue(const XArgs &g, int batch, float oscale, const int (&ra)[NB],
                                            const int (&rb)[NB], int l32, int kh, const f32x16 (&accR)[NB],
                                            const f32x16 (&accI)[NB]) {
#pragma unroll
  for (int b = 0; b < NB; b++) {
    tri_block_out<8>(g.Cbase, (double2 *)g.acc64, (long)batch * g.Cbatch, g.m, ra[b] * 32, kh, rb[b] * 32 + l32, oscale,
                  accR[b], accI[b]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// the chunk loop of one wave kind: two-deep register prefetch (pa holds chunk
// c+1, pb chunk c+2 while chunk c is multiplied out of LDS), LDS double buffer
template <typename Chunk>
__device__ __forceinline__ void wc_loop(char *lds, const float2 *base, long ld, int lgB, long bp, int NN, int K,
                                        Chunk chunk) {
  const int nchunk = (K + WC_KC - 1) / WC_KC;
  WcPrefetch pa, pb;
  wc_fetch(pa, base, ld, lgB, bp, NN, 0, K);
  wc_stage(pa, lds);
  if (nchunk > 1) wc_fetch(pa, base, ld, lgB, bp, NN, WC_KC, K);
  if (nchunk > 2) wc_fetch(pb, base, ld, lgB, bp, NN, 2 * WC_KC, K);
  for (int c = 0; c < nchunk; c += 2) {
    __syncthreads();  // chunk c staged in buffer 0; buffer 1 is free
    chunk(lds);
    if (c + 1 < nchunk) wc_stage(pa, lds + WC_BUF);
    if (c + 3 < nchunk) wc_fetch(pa, base, ld, lgB, bp, NN, (c + 3) * WC_KC, K);
    if (c + 1 < nchunk) {
      __syncthreads();  // chunk c+1 staged in buffer 1; buffer 0 is free
      chunk(lds + WC_BUF);
      if (c + 2 < nchunk) wc_stage(pb, lds);
      if (c + 4 < nchunk) wc_fetch(pb, base, ld, lgB, bp, NN, (c + 4) * WC_KC, K);
    }
  }
}

// B wave 4 + W: its four blocks over all chunks, then their epilogue
template <int W>
__device__ __forceinline__ void wc_run_b(char *lds, const float2 *base, const XArgs &g, int batch, float oscale,
                                         int NN, int K, int l32, int kh, int lane) {
  constexpr int ND = wc_b_ndiag<W>(), NO = WC_B - ND;  // diagonal and other blocks
  int ra[NO], rb[NO];
  f32x16 accR[WC_B], accI[WC_B];  // diagonal entries unused
  f32x4 dR[2][3], dI[2][3];
  {
    int o = 0;
#pragma unroll
    for (int b = 0; b < WC_B; b++) {
      if (WcBw<W>::r[b] != WcBw<W>::c[b]) {
        ra[o] = WcBw<W>::r[b];
        rb[o] = WcBw<W>::c[b];
        o++;
      }
#pragma unroll
      for (int q = 0; q < 16; q++) {
        accR[b][q] = 0.f;
        accI[b][q] = 0.f;
      }
    }
  }
#pragma unroll
  for (int d = 0; d < 2; d++)
#pragma unroll
    for (int sb = 0; sb < 3; sb++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        dR[d][sb][q] = 0.f;
        dI[d][sb][q] = 0.f;
      }
  wc_loop(lds, base, g.ldI, g.xlgB, g.xbp, NN, K,
          [&](const char *bufp) { wc_chunk_b<W>(bufp, l32, kh, lane, accR, accI, dR, dI); });
  f32x16 oR[NO], oI[NO];
  {
    int o = 0;
#pragma unroll
    for (int b = 0; b < WC_B; b++)
      if (WcBw<W>::r[b] != WcBw<W>::c[b]) {
        oR[o] = accR[b];
        oI[o] = accI[b];
        o++;
      }
  }
  wc_epilogue(g, batch, oscale, ra, rb, l32, kh, oR, oI);
#pragma unroll
  for (int d = 0; d < ND; d++) wc_diag16_out(g, batch, oscale, wc_b_diag_row<W>(d), lane, dR[d], dI[d]);
}

__global__ __launch_bounds__(WC_THREADS, 2) void xeng_f16split_chan_kernel(XArgs g, float oscale) {
  if (g.scale_dev != nullptr) {  // the scale the F-engine used (gains change it)
    const float sc = *g.scale_dev;
    oscale = g.alpha.x / (sc * sc);
  }
  __shared__ __attribute__((aligned(16))) char lds[2 * WC_BUF];

  const long nwg = gridDim.x;
  const long w = blockIdx.x;
  const int batch = (int)((nwg % 8 == 0) ? (w % 8) * (nwg / 8) + w / 8 : w);
  const float2 *base = g.Abase + xin_batch(g, batch);
  const int NN = g.m, K = g.k;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int l32 = lane & 31, kh = lane >> 5;

  // blocks past NN (NN < 256) are multiplied out of zero rows and not stored
  if (wave < 4) {
    if (wave == 3)
      wc_run_a<true>(lds, base, g, batch, oscale, NN, K, 4, l32, kh, lane);
    else
      wc_run_a<false>(lds, base, g, batch, oscale, NN, K, 7 - wave, l32, kh, lane);
  } else {
    switch (wave) {
      case 4: wc_run_b<0>(lds, base, g, batch, oscale, NN, K, l32, kh, lane); break;
      case 5: wc_run_b<1>(lds, base, g, batch, oscale, NN, K, l32, kh, lane); break;
      case 6: wc_run_b<2>(lds, base, g, batch, oscale, NN, K, l32, kh, lane); break;
      default: wc_run_b<3>(lds, base, g, batch, oscale, NN, K, l32, kh, lane); break;
    }
  }
}

// ---------------------------------------------------------------------------
// Small arrays (NN <= 48, the MDT station's 40 inputs): 16 x 16 blocks on
// v_mfma_f32_16x16x32_f16 (40 inputs pad to 48; the 128-input tile kernel's
// 32 x 32 blocks cover 64 x 64 for three of its four waves).  At this size the
// products take a fraction of the time of reading the X input once, so the
// kernel is built to stream: one 128-thread workgroup per channel, 32-sample
// chunks of all NN rows loaded as 16-byte pieces (16 lanes read a row's
// 256 contiguous bytes) into registers one chunk ahead, then routed into hi
// and lo planes [row][k] in LDS (k = 2 s + re/im; a 144-byte row pitch puts
// the 16 rows of an operand read on distinct banks).  Block roles as in the
// FP32 small-array kernel (xengine.hip), transposed for the triangle: with I
// and J the A and B block rows and I >= J, role 0 = (0,0), (1,0), (2,0),
// role 1 = (1,1), (2,1), (2,2), split once at the top.  Operand layout of the
// 16 x 16 x 32 MFMA: lane l supplies A[l%16][8(l/16) .. +7] and
// B[8(l/16) .. +7][l%16], i.e. samples 4(l/16) .. +3 of row l%16 as (re, im)
// halves; D register r of lane l is D[4(l/16) + r][l%16].  The three split
// products of a block go to one accumulator, interleaved over the role's
// blocks so that dependent MFMAs are six apart.
constexpr int XS16 = 16;                           // block edge
constexpr int HS_KC = 32;                          // samples per chunk
constexpr int HS_NT = 128;                         // threads
constexpr int HS_MAXN = 48;
constexpr int HS_ROWB = HS_KC * 4 + 16;            // bytes per LDS row per plane (128 data + 16 pad)
constexpr int HS_PLANE = HS_MAXN * HS_ROWB;
constexpr int HS_V4 = HS_MAXN * HS_KC / 2 / HS_NT;  // 16-byte pieces per thread per chunk (6)

// piece q: samples s0 + 2(q%16), +1 of row q/16 as (hi0, lo0, hi1, lo1)
// dwords; rows >= NN and samples >= K read as zero
__device__ __forceinline__ void hs_fetch(float4 (&v)[HS_V4], const float2 *__restrict__ base, long ld, int lgB,
                                         long bp, int NN, int s0, int K) {
#pragma unroll
  for (int i = 0; i < HS_V4; i++) {
    const int q = threadIdx.x + i * HS_NT;
    const int row = q >> 4, s = s0 + 2 * (q & 15);
    if (row < NN && s < K) {
      float4 x = *(const float4 *)(base + row * ld + xin_spec(lgB, bp, s));
      if (s + 1 >= K) x.z = x.w = 0.f;
      v[i] = x;
    } else {
      v[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
}
__device__ __forceinline__ void hs_stage(const float4 (&v)[HS_V4], char *lds, int NN) {
#pragma unroll
  for (int i = 0; i < HS_V4; i++) {
    const int q = threadIdx.x + i * HS_NT;
    const int row = q >> 4, s = 2 * (q & 15);
    if (row < NN) {
      *(uint2 *)(lds + row * HS_ROWB + s * 4) = make_uint2(__float_as_uint(v[i].x), __float_as_uint(v[i].z));
      *(uint2 *)(lds + HS_PLANE + row * HS_ROWB + s * 4) =
          make_uint2(__float_as_uint(v[i].y), __float_as_uint(v[i].w));
    }
  }
}

template <int ROLE>
__device__ __forceinline__ constexpr int hs_blk_I(int q) {
  return ROLE == 0 ? q : (q == 0 ? 1 : 2);
}
template <int ROLE>
__device__ __forceinline__ constexpr int hs_blk_J(int q) {
  return ROLE == 0 ? 0 : (q == 2 ? 2 : 1);
}

template <int ROLE, int NB>
__device__ __forceinline__ void hs_role(const XArgs &g, int batch, const float2 *__restrict__ base, char *lds,
                                        float oscale) {
  constexpr int NQ = ROLE == 0 ? NB : (NB == 3 ? 3 : NB - 1);
  const int NN = g.m, K = g.k;
  const int lane = threadIdx.x & 63, l16 = lane & 15, kq = lane >> 4;
  f32x4 accR[3], accI[3];
#pragma unroll
  for (int q = 0; q < 3; q++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      accR[q][r] = 0.f;
      accI[q][r] = 0.f;
    }
  float4 v[HS_V4];
  hs_fetch(v, base, g.ldI, g.xlgB, g.xbp, NN, 0, K);
  for (int s0 = 0; s0 < K; s0 += HS_KC) {
    __syncthreads();
    hs_stage(v, lds, NN);
    __syncthreads();
    if (s0 + HS_KC < K) hs_fetch(v, base, g.ldI, g.xlgB, g.xbp, NN, s0 + HS_KC, K);
#pragma unroll
    for (int ks = 0; ks < HS_KC / 16; ks++) {
      const int off = ks * 64 + kq * 16;
      half8 H[NB], L[NB], Hr[NB], Lr[NB];
#pragma unroll
      for (int gr = 0; gr < NB; gr++) {
        H[gr] = *(const half8 *)(lds + (XS16 * gr + l16) * HS_ROWB + off);
        L[gr] = *(const half8 *)(lds + HS_PLANE + (XS16 * gr + l16) * HS_ROWB + off);
        Hr[gr] = rot_conj(H[gr]);
        Lr[gr] = rot_conj(L[gr]);
      }
#pragma unroll
      for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          const int I = hs_blk_I<ROLE>(q), J = hs_blk_J<ROLE>(q);
          const half8 a = p == 2 ? L[I] : H[I], ar = p == 2 ? Lr[I] : Hr[I], b = p == 1 ? L[J] : H[J];
          accR[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, accR[q], 0, 0, 0);
          accI[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ar, b, accI[q], 0, 0, 0);
        }
    }
  }
  // V[a][b] = conj(D[a][b]) * oscale, a = 16 I + 4 kq + r, b = 16 J + l16
  const long cb = (long)batch * g.Cbatch;
  double2 *acc = (double2 *)g.acc64;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int I = hs_blk_I<ROLE>(q), J = hs_blk_J<ROLE>(q);
    const int b = XS16 * J + l16;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int a = XS16 * I + 4 * kq + r;
      if (a < NN && b < NN && a >= b) {
        const long o = cb + (long)a * (a + 1) / 2 + b;
        const float vr = oscale * accR[q][r], vim = -oscale * accI[q][r];
        g.Cbase[o] = make_float2(vr, vim);
        if (acc) {
          double2 tt = acc[o];
          tt.x += (double)vr;
          tt.y += (double)vim;
          acc[o] = tt;
        }
      }
    }
  }
}

template <int NB>
__global__ __launch_bounds__(HS_NT) void xeng_f16split_small_kernel(XArgs g, float oscale) {
  if (g.scale_dev != nullptr) {  // the scale the F-engine used (gains change it)
    const float sc = *g.scale_dev;
    oscale = g.alpha.x / (sc * sc);
  }
  __shared__ __attribute__((aligned(16))) char lds[2 * HS_PLANE];  // [hi|lo][row][k]
  const long nwg = gridDim.x;
  const long w = blockIdx.x;
  const int batch = (int)((nwg % 8 == 0) ? (w % 8) * (nwg / 8) + w / 8 : w);
  const float2 *base = g.Abase + xin_batch(g, batch);
  // rows NN..47 stay zero for the whole launch
  for (int i = threadIdx.x; i < 2 * HS_PLANE / 16; i += HS_NT) ((uint4 *)lds)[i] = make_uint4(0u, 0u, 0u, 0u);
  if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0)
    hs_role<0, NB>(g, batch, base, lds, oscale);
  else
    hs_role<1, NB>(g, batch, base, lds, oscale);
}

int xeng_launch_f16split(const XArgs &g0, hipStream_t st) {
  XArgs g = g0;
  if (g.m <= HS_MAXN && g.ldI == g.ldJ && g.Abase == g.Bbase && (g.ldI & 1) == 0 && (g.xbp & 1) == 0 &&
      g.xlgB >= 1) {
    MDT_TRY(check_grid(g.batch, HS_NT, "X-engine"));
    const float sc = g.in_scale;
    if (!(sc > 0.f)) return fail(MDT_ERR_INVALID_VALUE, "FP16 split needs a positive input scale");
    const float oscale = g.alpha.x / (sc * sc);
    const int nb = (g.m + XS16 - 1) / XS16;
    if (nb == 1)
      hipLaunchKernelGGL(xeng_f16split_small_kernel<1>, dim3((unsigned)g.batch), dim3(HS_NT), 0, st, g, oscale);
    else if (nb == 2)
      hipLaunchKernelGGL(xeng_f16split_small_kernel<2>, dim3((unsigned)g.batch), dim3(HS_NT), 0, st, g, oscale);
    else
      hipLaunchKernelGGL(xeng_f16split_small_kernel<3>, dim3((unsigned)g.batch), dim3(HS_NT), 0, st, g, oscale);
    return check_launch("xeng_f16split_small_kernel");
  }
  if (g.m > 128 && g.m <= WC_NB * 32 && g.ldI == g.ldJ && g.Abase == g.Bbase) {
    MDT_TRY(check_grid(g.batch, WC_THREADS, "X-engine"));
    const float sc = g.in_scale;
    if (!(sc > 0.f)) return fail(MDT_ERR_INVALID_VALUE, "FP16 split needs a positive input scale");
    hipLaunchKernelGGL(xeng_f16split_chan_kernel, dim3((unsigned)g.batch), dim3(WC_THREADS), 0, st, g,
                       g.alpha.x / (sc * sc));
    return check_launch("xeng_f16split_chan_kernel");
  }
  const int nb = (g.m + HT - 1) / HT;
  g.tiles = nb * (nb + 1) / 2;
  long total = (long)g.tiles * g.batch;
  MDT_TRY(check_grid(total, 256, "X-engine"));
  // power-of-two input scale (exact) chosen by the caller; products scale by sc^2
  const float sc = g.in_scale;
  if (!(sc > 0.f)) return fail(MDT_ERR_INVALID_VALUE, "FP16 split needs a positive input scale");
  const float oscale = g.alpha.x / (sc * sc);
  hipLaunchKernelGGL(xeng_f16split_kernel, dim3((unsigned)total), dim3(256), 0, st, g, oscale);
  return check_launch("xeng_f16split_kernel");
}

}  // namespace mdt
