// xengine.hip -- a5: the X-engine, drop-in for cublasCreate / cublasSetStream
// / cublasCgemmBatched(OP_N, OP_C, NN, NN, Ns, 1/Ns, A[], NN, A[], NN, 0, C[],
// NN, N) (MDT_OneStation_client.cu:656-705, 1128), on FP32 MFMA.
//
// Complex products on the real v_mfma_f32_32x32x2f32: one MFMA consumes one
// complex sample as its K = 2 (re, im) pair.  With I-operand rows x_i and
// J-operand columns x_j,
//   D_re[i][j] = sum_s (re_i re_j + im_i im_j)          = Re(x_j conj(x_i))
//   D_im[i][j] = sum_s ((-im_i) re_j + re_i im_j)       = Im(x_j conj(x_i))
// i.e. D[i][j] = V[j][i] with V[p][q] = sum_s x_p conj(x_q); the I operand is
// fed as (-im, re) for the second product.  No 3M/Gauss trick: the four real
// products keep FP32 accuracy.
//
// A workgroup computes a 64 x 64 complex tile with 4 waves (2 x 2 MFMA tiles
// of 32 x 32), streaming K in chunks of KC samples through LDS stored as
// planar [s][re|im][row] so each MFMA operand read is 32 consecutive floats.
// Hermitian products (A == B) visit only tiles on or below the diagonal.
// Workgroups of one batch entry are placed on the same XCD so their shared
// operand rows hit the same L2.
#include <new>

#include "common.hpp"
#include "xengine.hpp"

namespace mdt {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int XT = 64;        // tile edge (complex rows)
constexpr int XKC = 32;       // K samples per LDS chunk
constexpr int XRL = XT + 1;   // padded plane row length (floats)
constexpr int XPLANE = 2 * XRL;  // floats per sample: re row + im row

// Load a 64-row x KC-sample operand chunk into LDS [s][comp][row].
// KM: element (row, s) at base[row + s*ld] (column-major, rows contiguous).
// MK: element (row, s) at base[row*ld + s] (samples contiguous, 16-byte
//     aligned rows: even ld).
// Branch-free: out-of-range rows/samples read a clamped valid address and are
// zeroed, so all loads of a chunk are in flight together.
template <int LAYOUT>
__device__ __forceinline__ void load_chunk(float *lds, const float2 *__restrict__ base, long ld,
                                           int row0, int nrows, int s0, int K, int lgB, long bp) {
  const int tid = threadIdx.x;
  if constexpr (LAYOUT == XL_KM) {
    const int r = tid & 63, sq = tid >> 6;  // 64 rows x 4 sample lanes
    const bool r_ok = row0 + r < nrows;
    const long rr = r_ok ? row0 + r : 0;
    float2 v[XKC / 4];
#pragma unroll
    for (int q = 0; q < XKC / 4; q++) {
      const int s = s0 + sq + 4 * q;
      const bool ok = r_ok && s < K;
      const float2 x = base[rr + (long)(s < K ? s : 0) * ld];
      v[q] = ok ? x : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int q = 0; q < XKC / 4; q++) {
      const int s = sq + 4 * q;
      lds[s * XPLANE + r] = v[q].x;
      lds[s * XPLANE + XRL + r] = v[q].y;
    }
  } else {
    const int sp = tid & 15, rq = tid >> 4;  // 16 sample pairs x 16 rows
    const int s = 2 * sp;
    const bool s_ok0 = s0 + s < K, s_ok1 = s0 + s + 1 < K;
    const long so = xin_spec(lgB, bp, s_ok0 ? s0 + s : 0);  // pairs (s, s+1) never straddle a block
    float4 v[XT / 16];
#pragma unroll
    for (int q = 0; q < XT / 16; q++) {
      const int grow = row0 + rq + 16 * q;
      const bool r_ok = grow < nrows;
      const float4 x = *(const float4 *)(base + (long)(r_ok ? grow : 0) * ld + so);
      const bool ok0 = r_ok && s_ok0, ok1 = r_ok && s_ok1;
      v[q] = make_float4(ok0 ? x.x : 0.f, ok0 ? x.y : 0.f, ok1 ? x.z : 0.f, ok1 ? x.w : 0.f);
    }
#pragma unroll
    for (int q = 0; q < XT / 16; q++) {
      const int r = rq + 16 * q;
      lds[s * XPLANE + r] = v[q].x;
      lds[s * XPLANE + XRL + r] = v[q].y;
      lds[(s + 1) * XPLANE + r] = v[q].z;
      lds[(s + 1) * XPLANE + XRL + r] = v[q].w;
    }
  }
}

__device__ __forceinline__ long tri_bl(int a, int b) { return (long)a * (a + 1) / 2 + b; }

// grid.x = tiles_per_batch * batch (XCD-remapped); block = 256 (4 waves)
template <int LAYOUT, int OUT>
__global__ __launch_bounds__(256) void xeng_f32_kernel(XArgs g) {
  __shared__ float ldsI[XKC * XPLANE];
  __shared__ float ldsJ[XKC * XPLANE];

  // XCD-aware remap: hardware places workgroup w on XCD w % 8; give each XCD a
  // contiguous range of logical tiles so one batch entry stays on one XCD.
  const long nwg = gridDim.x;
  const long w = blockIdx.x;
  long logical = w;
  if (nwg % 8 == 0) logical = (w % 8) * (nwg / 8) + w / 8;
  const int batch = (int)(logical / g.tiles);
  int tile = (int)(logical % g.tiles);
  int P, Q;  // block indices, P >= Q in Hermitian mode
  if (g.herm) {
    P = 0;
    while ((P + 1) * (P + 2) / 2 <= tile) P++;
    Q = tile - P * (P + 1) / 2;
  } else {
    P = tile / g.nbJ;
    Q = tile % g.nbJ;
  }
  // I-operand block and J-operand block
  int bi, bj;
  if (OUT == XO_TRI) {  // i <-> a (row of triangle), j <-> b, a >= b
    bi = P;
    bj = Q;
  } else if (g.herm) {  // j <-> a, i <-> b, a >= b
    bi = Q;
    bj = P;
  } else {
    bi = P;  // i <-> B rows (n), j <-> A rows (m)
    bj = Q;
  }
  const float2 *baseI, *baseJ;
  if (OUT == XO_TRI) {
    baseI = g.Abase + xin_batch(g, batch);
    baseJ = baseI;
  } else {
    baseJ = g.Atab ? g.Atab[batch] : g.Abase + (long)batch * g.Abatch;
    baseI = g.Btab ? g.Btab[batch] : g.Bbase + (long)batch * g.Bbatch;
  }
  const int nI = (OUT == XO_TRI) ? g.m : g.n;  // rows available in the I operand
  const int nJ = g.m;
  const bool same = g.herm && (bi == bj);

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wi = wave >> 1, wj = wave & 1;
  // sub-tile strictly on the excluded side of the diagonal: skip the MFMAs
  bool skip = false;
  if (same) skip = (OUT == XO_TRI) ? (wi < wj) : (wj < wi);

  f32x16 accR, accI;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    accR[r] = 0.f;
    accI[r] = 0.f;
  }
  const int lrow = lane & 31;
  const int comp = lane >> 5;  // 0: re, 1: im
  const int rowI = wi * 32 + lrow, rowJ = wj * 32 + lrow;
  const float *srcJ = same ? ldsI : ldsJ;

  for (int s0 = 0; s0 < g.k; s0 += XKC) {
    __syncthreads();
    load_chunk<LAYOUT>(ldsI, baseI, g.ldI, bi * XT, nI, s0, g.k, g.xlgB, g.xbp);
    if (!same) load_chunk<LAYOUT>(ldsJ, baseJ, g.ldJ, bj * XT, nJ, s0, g.k, g.xlgB, g.xbp);
    __syncthreads();
    if (!skip) {
#pragma unroll 8
      for (int s = 0; s < XKC; s++) {
        const float *pI = ldsI + s * XPLANE;
        const float *pJ = srcJ + s * XPLANE;
        float vi = pI[comp * XRL + rowI];
        float vit = pI[(1 - comp) * XRL + rowI];
        vit = comp ? vit : -vit;  // (-im, re)
        float vj = pJ[comp * XRL + rowJ];
        accR = __builtin_amdgcn_mfma_f32_32x32x2f32(vi, vj, accR, 0, 0, 0);
        accI = __builtin_amdgcn_mfma_f32_32x32x2f32(vit, vj, accI, 0, 0, 0);
      }
    }
  }
  if (skip) return;

  // D[i][j] = V[j][i]; lane holds column j = lane & 31, rows
  // i = 8*(r/4) + 4*(lane/32) + r%4.
  const int jg = bj * XT + wj * 32 + lrow;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int ig = bi * XT + wi * 32 + 8 * (r >> 2) + 4 * comp + (r & 3);
    const float dre = accR[r], dim = accI[r];
    if constexpr (OUT == XO_TRI) {
      // a = ig, b = jg, V[a][b] = conj(D)
      const int a = ig, b = jg;
      if (a < g.m && b < g.m && a >= b) {
        const long o = (long)batch * g.Cbatch + tri_bl(a, b);
        const float vr = g.alpha.x * dre, vim = -g.alpha.x * dim;
        g.Cbase[o] = make_float2(vr, vim);
        if (g.acc64) {
          double2 *acc = (double2 *)g.acc64;
          double2 t = acc[o];
          t.x += (double)vr;
          t.y += (double)vim;
          acc[o] = t;
        }
      }
    } else {
      // a = jg (row of C), b = ig (column of C), V[a][b] = D
      const int a = jg, b = ig;
      if (a >= g.m || b >= g.n) continue;
      if (g.herm && a < b) continue;
      float2 *C = g.Ctab ? g.Ctab[batch] : g.Cbase + (long)batch * g.Cbatch;
      const float2 al = g.alpha, be = g.beta;
      {
        float2 c0 = (be.x != 0.f || be.y != 0.f) ? C[a + (long)b * g.ldc] : make_float2(0.f, 0.f);
        float2 o = make_float2(al.x * dre - al.y * dim + be.x * c0.x - be.y * c0.y,
                               al.x * dim + al.y * dre + be.x * c0.y + be.y * c0.x);
        C[a + (long)b * g.ldc] = o;
      }
      if (g.herm && a != b) {  // mirror: V[b][a] = conj(V[a][b])
        float2 c0 = (be.x != 0.f || be.y != 0.f) ? C[b + (long)a * g.ldc] : make_float2(0.f, 0.f);
        float2 o = make_float2(al.x * dre + al.y * dim + be.x * c0.x - be.y * c0.y,
                               -al.x * dim + al.y * dre + be.x * c0.y + be.y * c0.x);
        C[b + (long)a * g.ldc] = o;
      }
    }
  }
}

// Hermitian products of small arrays (NN <= 48, the MDT station's 40
// inputs): the drop-in CGEMM with the reference's X-input layout (NN even,
// one sample's NN rows contiguous, ld == NN: a chunk of 32 samples is one
// contiguous span of 32·NN complex values) and the native triangle in any
// X-input layout (a chunk is 32 samples of each row).  16 x 16 blocks on
// v_mfma_f32_16x16x4f32 pad 40 inputs to 48 rather than 64.  Operand layout:
// lane l supplies A[l%16][l/16] and B[l/16][l%16], K = 4 being two complex
// samples (k = 2t + re/im); D register r of lane l is D[4(l/16) + r][l%16].
// Products and epilogues as in xeng_f32_kernel.
//
// One 128-thread workgroup per batch entry.  The nb(nb+1)/2 blocks go to
// the two waves with fixed roles, split once at the top so that each wave's
// loop carries no per-block branch (a per-block `if` put every MFMA pair
// behind its own exec mask, each waiting for its LDS reads: 1.16 against
// 0.90 ms at the station's shape).  In the full form, with (I, J) the A
// and B block rows and I <= J: role 0 = (0,0), (0,1), (0,2), role 1 =
// (1,1), (1,2), (2,2); the triangle uses the transposed blocks.  Each block
// has two accumulator sets (alternate sample pairs, summed at the end), so a
// wave runs six to twelve independent MFMA chains.  The chunk sits in LDS as
// complex [sample][row] with rows NN..47 zero (a 16-byte piece of the
// drop-in layout is one 16-byte LDS store; the native layout's pieces are
// two samples of one row, written as two 8-byte stores, with an odd row
// pitch so that the 16 lanes of a row hit distinct banks); each wave reads
// its three 16-row operand groups per sample pair one step ahead of the
// MFMAs that use them.  The next chunk is loaded into registers while the
// current one is multiplied.  128 VGPRs + 48 accumulation registers: four
// waves per SIMD.
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int XS = 16;                       // block edge
constexpr int XS_MAXN = 48;                  // largest NN
constexpr int XS_KC = 32;                    // samples per chunk
constexpr int XS_NT = 128;                   // threads per workgroup
constexpr int XS_V4 = XS_KC * XS_MAXN / 2 / XS_NT;  // 16-byte pieces per thread per chunk (6)
template <int LAYOUT>
struct XsRow {
  static constexpr int RL = LAYOUT == XL_KM ? XS_MAXN : XS_MAXN + 1;  // LDS row pitch (complex values)
};

// AL16: the operand base is 16-byte aligned, so a piece is one 16-byte
// load; otherwise two 8-byte loads (table entries are checked per workgroup,
// not on the host)
template <bool AL16>
__device__ __forceinline__ void xs_fetch(float4 (&v)[XS_V4], const float2 *base, int NN, int s0, int K) {
  const long nval = (long)(K - s0 < XS_KC ? K - s0 : XS_KC) * NN / 2;  // valid pieces (NN even)
  const float2 *src = base + (long)s0 * NN;
#pragma unroll
  for (int i = 0; i < XS_V4; i++) {
    const int q = threadIdx.x + i * XS_NT;
    if (q < nval) {
      if constexpr (AL16) {
        v[i] = ((const float4 *)src)[q];
      } else {
        const float2 x0 = src[2 * q], x1 = src[2 * q + 1];
        v[i] = make_float4(x0.x, x0.y, x1.x, x1.y);
      }
    } else {
      v[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
}
// pieces beyond the chunk's valid samples were zeroed by the fetch, so every
// sample of the chunk is written (rows NN..47 stay zero from the start)
__device__ __forceinline__ void xs_stage(const float4 (&v)[XS_V4], float2 *lds, int NN) {
  const int npiece = XS_KC * NN / 2;
#pragma unroll
  for (int i = 0; i < XS_V4; i++) {
    const int q = threadIdx.x + i * XS_NT;
    if (q < npiece) {
      const int e = 2 * q, s = e / NN, r = e - s * NN;  // r, r + 1 in the same sample (NN even)
      *(float4 *)(lds + s * XsRow<XL_KM>::RL + r) = v[i];
    }
  }
}

// The native X-input layouts (element (row e, sample s) at
// e*ld + xin_spec(lgB, bp, s); ld, bp and the block length even, so a pair
// of samples (s even) is one aligned 16-byte piece inside one block): piece q
// is samples 2(q%16), 2(q%16)+1 of row q/16; samples past K are zeroed (the
// pair's second sample lies in the row's or the block's padding then)
__device__ __forceinline__ void xs_fetch_mk(float4 (&v)[XS_V4], const float2 *base, long ld, int lgB, long bp,
                                            int NN, int s0, int K) {
#pragma unroll
  for (int i = 0; i < XS_V4; i++) {
    const int q = threadIdx.x + i * XS_NT;
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
__device__ __forceinline__ void xs_stage_mk(const float4 (&v)[XS_V4], float2 *lds, int NN) {
  constexpr int RL = XsRow<XL_MK>::RL;
#pragma unroll
  for (int i = 0; i < XS_V4; i++) {
    const int q = threadIdx.x + i * XS_NT;
    const int r = q >> 4, s = 2 * (q & 15);
    if (r < NN) {
      lds[s * RL + r] = make_float2(v[i].x, v[i].y);
      lds[(s + 1) * RL + r] = make_float2(v[i].z, v[i].w);
    }
  }
}

// blocks of one role: slot q -> (I, J), the A and B operand block rows
template <int OUT, int ROLE>
__device__ __forceinline__ constexpr int xs_blk_I(int q) {
  const int lo = ROLE == 0 ? 0 : (q == 2 ? 2 : 1), hi = ROLE == 0 ? q : (q == 0 ? 1 : 2);
  return OUT == XO_FULL ? lo : hi;
}
template <int OUT, int ROLE>
__device__ __forceinline__ constexpr int xs_blk_J(int q) {
  const int lo = ROLE == 0 ? 0 : (q == 2 ? 2 : 1), hi = ROLE == 0 ? q : (q == 0 ? 1 : 2);
  return OUT == XO_FULL ? hi : lo;
}

template <int LAYOUT, int OUT, int ROLE, int NB, bool AL16>
__device__ __forceinline__ void xs_role(const XArgs &g, int batch, const float2 *base, float2 *lds) {
  constexpr int RL = XsRow<LAYOUT>::RL;
  constexpr int NQ = ROLE == 0 ? NB : (NB == 3 ? 3 : NB - 1);  // blocks of this role
  const int NN = g.m, K = g.k;
  const int lane = threadIdx.x & 63;
  const int l16 = lane & 15, kq = lane >> 4;
  const int t = kq >> 1, comp = kq & 1;
  f32x4 accR[3][2], accI[3][2];
#pragma unroll
  for (int q = 0; q < 3; q++)
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        accR[q][h][r] = 0.f;
        accI[q][h][r] = 0.f;
      }
  // x[h][gr]: sample s + 2h + t of row 16 gr + l16 (complex)
  auto rd = [&](int s, float2 (&x)[2][NB]) {
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int gr = 0; gr < NB; gr++) x[h][gr] = lds[(s + 2 * h + t) * RL + XS * gr + l16];
  };
  auto mm = [&](const float2 (&x)[2][NB]) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      float vv[NB], vt[NB];  // this lane's component, and (-im, re)
#pragma unroll
      for (int gr = 0; gr < NB; gr++) {
        vv[gr] = comp ? x[h][gr].y : x[h][gr].x;
        vt[gr] = comp ? x[h][gr].x : -x[h][gr].y;
      }
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const int I = xs_blk_I<OUT, ROLE>(q), J = xs_blk_J<OUT, ROLE>(q);
        accR[q][h] = __builtin_amdgcn_mfma_f32_16x16x4f32(vv[I], vv[J], accR[q][h], 0, 0, 0);
        accI[q][h] = __builtin_amdgcn_mfma_f32_16x16x4f32(vt[I], vv[J], accI[q][h], 0, 0, 0);
      }
    }
  };

  float4 v[XS_V4];
  if constexpr (LAYOUT == XL_KM)
    xs_fetch<AL16>(v, base, NN, 0, K);
  else
    xs_fetch_mk(v, base, g.ldI, g.xlgB, g.xbp, NN, 0, K);
  for (int s0 = 0; s0 < K; s0 += XS_KC) {
    __syncthreads();
    if constexpr (LAYOUT == XL_KM)
      xs_stage(v, lds, NN);
    else
      xs_stage_mk(v, lds, NN);
    __syncthreads();
    if (s0 + XS_KC < K) {
      if constexpr (LAYOUT == XL_KM)
        xs_fetch<AL16>(v, base, NN, s0 + XS_KC, K);
      else
        xs_fetch_mk(v, base, g.ldI, g.xlgB, g.xbp, NN, s0 + XS_KC, K);
    }
    float2 x0[2][NB], x1[2][NB];
    rd(0, x0);
#pragma unroll 1
    for (int s = 0; s < XS_KC; s += 8) {
      rd(s + 4, x1);
      mm(x0);
      if (s + 8 < XS_KC) rd(s + 8, x0);
      mm(x1);
    }
  }

  float2 *C = g.Ctab ? g.Ctab[batch] : g.Cbase + (long)batch * g.Cbatch;
  const float2 al = g.alpha, be = g.beta;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int I = xs_blk_I<OUT, ROLE>(q), J = xs_blk_J<OUT, ROLE>(q);
    const int jg = J * XS + l16;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int ig = I * XS + 4 * kq + r;
      const float dre = accR[q][0][r] + accR[q][1][r], dim = accI[q][0][r] + accI[q][1][r];
      if constexpr (OUT == XO_TRI) {
        const int a = ig, b = jg;  // V[a][b] = conj(D)
        if (a < NN && b < NN && a >= b) {
          const long o = (long)batch * g.Cbatch + tri_bl(a, b);
          const float vr = al.x * dre, vim = -al.x * dim;
          g.Cbase[o] = make_float2(vr, vim);
          if (g.acc64) {
            double2 *acc = (double2 *)g.acc64;
            double2 tt = acc[o];
            tt.x += (double)vr;
            tt.y += (double)vim;
            acc[o] = tt;
          }
        }
        continue;
      }
      const int a = jg, b = ig;  // V[a][b] = D, a >= b; mirror V[b][a] = conj(D)
      if (a >= NN || b >= NN || a < b) continue;
      {
        const float2 c0 = (be.x != 0.f || be.y != 0.f) ? C[a + (long)b * g.ldc] : make_float2(0.f, 0.f);
        C[a + (long)b * g.ldc] = make_float2(al.x * dre - al.y * dim + be.x * c0.x - be.y * c0.y,
                                             al.x * dim + al.y * dre + be.x * c0.y + be.y * c0.x);
      }
      if (a != b) {
        const float2 c0 = (be.x != 0.f || be.y != 0.f) ? C[b + (long)a * g.ldc] : make_float2(0.f, 0.f);
        C[b + (long)a * g.ldc] = make_float2(al.x * dre + al.y * dim + be.x * c0.x - be.y * c0.y,
                                             -al.x * dim + al.y * dre + be.x * c0.y + be.y * c0.x);
      }
    }
  }
}

template <int LAYOUT, int OUT, int NB>
__global__ __launch_bounds__(XS_NT, 4) void xeng_f32_small_kernel(XArgs g) {
  constexpr int RL = XsRow<LAYOUT>::RL;
  __shared__ float2 lds[XS_KC * RL];
  const long nwg = gridDim.x;
  const long w = blockIdx.x;
  const long logical = (nwg % 8 == 0) ? (w % 8) * (nwg / 8) + w / 8 : w;
  const int batch = (int)logical;
  const float2 *base;
  if constexpr (OUT == XO_TRI)
    base = g.Abase + xin_batch(g, batch);
  else
    base = g.Btab ? g.Btab[batch] : g.Bbase + (long)batch * g.Bbatch;
  // rows NN..47 of every sample are zero for the whole launch
  for (int i = threadIdx.x; i < XS_KC * RL; i += XS_NT) lds[i] = make_float2(0.f, 0.f);
  const bool role0 = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0;
  if (LAYOUT == XL_MK || ((uintptr_t)base & 15) == 0) {
    if (role0)
      xs_role<LAYOUT, OUT, 0, NB, true>(g, batch, base, lds);
    else
      xs_role<LAYOUT, OUT, 1, NB, true>(g, batch, base, lds);
  } else {
    if (role0)
      xs_role<LAYOUT, OUT, 0, NB, false>(g, batch, base, lds);
    else
      xs_role<LAYOUT, OUT, 1, NB, false>(g, batch, base, lds);
  }
}

template <int LAYOUT, int OUT>
static int xs_launch(const XArgs &g, hipStream_t st) {
  MDT_TRY(check_grid(g.batch, XS_NT, "X-engine"));
  const int nb = (g.m + XS - 1) / XS;
  if (nb == 1)
    hipLaunchKernelGGL((xeng_f32_small_kernel<LAYOUT, OUT, 1>), dim3((unsigned)g.batch), dim3(XS_NT), 0, st, g);
  else if (nb == 2)
    hipLaunchKernelGGL((xeng_f32_small_kernel<LAYOUT, OUT, 2>), dim3((unsigned)g.batch), dim3(XS_NT), 0, st, g);
  else
    hipLaunchKernelGGL((xeng_f32_small_kernel<LAYOUT, OUT, 3>), dim3((unsigned)g.batch), dim3(XS_NT), 0, st, g);
  return check_launch("xeng_f32_small_kernel");
}

int xeng_launch(const XArgs &g0, int layout, int out, hipStream_t st) {
  XArgs g = g0;
  if (layout == XL_KM && out == XO_FULL && g.herm && g.m <= XS_MAXN && (g.m & 1) == 0 && g.ldI == g.m) {
    return xs_launch<XL_KM, XO_FULL>(g, st);
  }
  // the native triangle, every X-input layout (even pitches: 16-byte aligned pieces)
  if (layout == XL_MK && out == XO_TRI && g.herm && g.m <= XS_MAXN && (g.ldI & 1) == 0 && (g.xbp & 1) == 0 &&
      g.xlgB >= 1) {
    return xs_launch<XL_MK, XO_TRI>(g, st);
  }
  const int nbM = (g.m + XT - 1) / XT;
  const int nbN = ((out == XO_TRI ? g.m : g.n) + XT - 1) / XT;
  if (g.herm) {
    g.tiles = nbM * (nbM + 1) / 2;
    g.nbJ = nbM;
  } else {
    g.tiles = nbN * nbM;  // P over I blocks (n), Q over J blocks (m)
    g.nbJ = nbM;
  }
  long total = (long)g.tiles * g.batch;
  MDT_TRY(check_grid(total, 256, "X-engine"));
  dim3 grid((unsigned)total), block(256);
  if (layout == XL_KM && out == XO_FULL)
    hipLaunchKernelGGL((xeng_f32_kernel<XL_KM, XO_FULL>), grid, block, 0, st, g);
  else if (layout == XL_KM && out == XO_TRI)
    hipLaunchKernelGGL((xeng_f32_kernel<XL_KM, XO_TRI>), grid, block, 0, st, g);
  else if (layout == XL_MK && out == XO_TRI)
    hipLaunchKernelGGL((xeng_f32_kernel<XL_MK, XO_TRI>), grid, block, 0, st, g);
  else
    hipLaunchKernelGGL((xeng_f32_kernel<XL_MK, XO_FULL>), grid, block, 0, st, g);
  return check_launch("xeng_f32_kernel");
}

}  // namespace mdt

struct mdt_xengine_s {
  hipStream_t stream = 0;
};

using namespace mdt;

extern "C" {

int mdt_xengine_create(mdt_xengine *handle) {
  MDT_REQUIRE_PTR(handle);
  *handle = new (std::nothrow) mdt_xengine_s;
  if (!*handle) return fail(MDT_ERR_ALLOC, "X-engine handle");
  return MDT_SUCCESS;
}

int mdt_xengine_set_stream(mdt_xengine handle, hipStream_t stream) {
  MDT_REQUIRE_PTR(handle);
  handle->stream = stream;
  return MDT_SUCCESS;
}

int mdt_xengine_destroy(mdt_xengine handle) {
  delete handle;
  return MDT_SUCCESS;
}

int mdt_cgemm_batched(mdt_xengine handle, int transa, int transb, int m, int n, int k,
                      const mdt_complex *alpha, const mdt_complex *const Aarray[], int lda,
                      const mdt_complex *const Barray[], int ldb, const mdt_complex *beta,
                      mdt_complex *const Carray[], int ldc, int batchCount) {
  MDT_REQUIRE_PTR(handle);
  MDT_REQUIRE_PTR(alpha);
  MDT_REQUIRE_PTR(beta);
  MDT_REQUIRE_PTR(Aarray);
  MDT_REQUIRE_PTR(Barray);
  MDT_REQUIRE_PTR(Carray);
  if (transa != MDT_OP_N || transb != MDT_OP_C)
    return fail(MDT_ERR_UNSUPPORTED, "only transa=N, transb=C (the X-engine call) is provided");
  if (m < 0 || n < 0 || k < 0 || batchCount < 0)
    return fail(MDT_ERR_INVALID_VALUE, "negative dimension");
  if (lda < (m > 1 ? m : 1) || ldb < (n > 1 ? n : 1) || ldc < (m > 1 ? m : 1))
    return fail(MDT_ERR_INVALID_VALUE, "leading dimension too small (lda>=m, ldb>=n, ldc>=m)");
  if (m == 0 || n == 0 || batchCount == 0) return MDT_SUCCESS;
  XArgs g{};
  g.m = m;
  g.n = n;
  g.k = k;
  g.batch = batchCount;
  g.alpha = make_float2(alpha->x, alpha->y);
  g.beta = make_float2(beta->x, beta->y);
  g.Atab = (const float2 *const *)Aarray;
  g.Btab = (const float2 *const *)Barray;
  g.Ctab = (float2 *const *)Carray;
  g.ldJ = lda;  // J operand = A (rows a), column-major: element (row, s) at row + s*lda
  g.ldI = ldb;
  g.ldc = ldc;
  g.herm = ((const void *)Aarray == (const void *)Barray) && m == n && lda == ldb;
  return xeng_launch(g, XL_KM, XO_FULL, handle->stream);
}

int mdt_xengine_herk(mdt_xengine handle, int NpolsxNelements, int Nspectra, float alpha,
                     const mdt_complex *XInput, mdt_complex *XOutput, int Nchannels) {
  MDT_REQUIRE_PTR(handle);
  MDT_REQUIRE_PTR(XInput);
  MDT_REQUIRE_PTR(XOutput);
  MDT_TRY(check_inputs(NpolsxNelements));
  MDT_TRY(check_spectra(Nspectra));
  MDT_TRY(check_channels(Nchannels));
  XArgs g{};
  const int NN = NpolsxNelements;
  g.m = g.n = NN;
  g.k = Nspectra;
  g.batch = Nchannels;
  g.alpha = make_float2(alpha, 0.f);
  g.beta = make_float2(0.f, 0.f);
  g.Abase = g.Bbase = (const float2 *)XInput;
  g.Abatch = g.Bbatch = (long)NN * Nspectra;
  g.ldI = g.ldJ = NN;
  g.Cbase = (float2 *)XOutput;
  g.Cbatch = (long)NN * NN;
  g.ldc = NN;
  g.herm = 1;
  return xeng_launch(g, XL_KM, XO_FULL, handle->stream);
}

}  // extern "C"

