// fft_core.hpp -- batched C2C FFT building block shared by the standalone
// FFT (drop-in for cufftExecC2C, MDT_OneStation_client.cu:1105) and the fused
// F-engine.
//
// Algorithm: Stockham autosort, radix-P passes plus one remainder pass.  One
// transform of N points is held by TPB = N/P threads, P points each; thread t
// owns elements t + m*TPB (m = 0..P-1) on entry AND owns X[t + m*TPB] on exit,
// so the first pass reads straight from registers (or global memory, fully
// coalesced) and the last pass writes straight to registers.  Passes in between
// exchange through LDS.  For the pass of radix R at stage NS (product of the
// earlier radices), butterfly j reads v[r] = x[j + r*N/R], multiplies by
// W_{NS*R}^{(j mod NS)*r}, does an R-point DFT and stores v[r] at
// (j/NS)*NS*R + (j mod NS) + r*NS.  Twiddles W_N^k come from a table evaluated
// in FP64 and rounded once to FP32.
#pragma once

#include "common.hpp"

namespace mdt {

constexpr int ipow_c(int b, int e) { return e == 0 ? 1 : b * ipow_c(b, e - 1); }
constexpr int ilog2_c(int v) { return v <= 1 ? 0 : 1 + ilog2_c(v / 2); }
constexpr int full_passes(int n, int p) {
  int l = 0;
  while (n > 1 && n % p == 0) {
    n /= p;
    l++;
  }
  return l;
}

// points per thread for an N-point transform: 16 at N = 1024, so that one
// transform is one wave (64 threads) and its LDS exchanges need no workgroup
// barrier, and at N = 2048 (three passes, two exchanges instead of three);
// 4 at N = 256, for the same reason (a transform is one whole wave, so the
// persistent F-engine's sub-blocks stay wave-uniform); 8 otherwise up to
// 4096 (a 512-point transform is one wave already; at 4096 the F-engine's
// FP16 hi/lo output measured twice as slow with 16); 16 above
constexpr int fft_points(int N) {
  return (N == 1024 || N == 2048) ? 16 : (N == 256 ? 4 : (N <= 4096 ? 8 : 16));
}

// exchange synchronisation between passes: a transform held by one wave (or
// part of one) only orders its own LDS accesses (a wave's LDS operations
// execute in order; the wavefront-scope fence keeps the compiler from moving
// one lane's accesses past another's); larger transforms span waves and
// need the workgroup barrier.  RAW: a bare s_barrier after this wave's LDS
// operations have completed (lgkmcnt(0)), which, unlike __syncthreads() and
// its release fence, does not wait for the global loads in flight (used where
// the next rows are being loaded during the transform)
template <int TPB, bool RAW = false>
__device__ __forceinline__ void fft_exchange_sync() {
  if constexpr (TPB <= 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else if constexpr (RAW) {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt, expcnt unchanged
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  } else {
    __syncthreads();
  }
}

// launch shape of the standalone N-point transforms: FPB transforms of TPB
// threads (P points each) per workgroup
template <int N>
struct FftCfg {
  static constexpr int P = fft_points(N);
  static constexpr int TPB = N / P;
  static constexpr int FPB = TPB >= 256 ? 1 : 256 / TPB;  // transforms per block
  static constexpr int THREADS = TPB * FPB;
};

// W_16^k constants, k = 0..7: cos(2 pi k / 16), sin(2 pi k / 16)
__device__ constexpr float kCos16[8] = {1.0f, 0.92387953251128674f, 0.70710678118654752f,
                                        0.38268343236508978f, 0.0f, -0.38268343236508978f,
                                        -0.70710678118654752f, -0.92387953251128674f};
__device__ constexpr float kSin16[8] = {0.0f, 0.38268343236508978f, 0.70710678118654752f,
                                        0.92387953251128674f, 1.0f, 0.92387953251128674f,
                                        0.70710678118654752f, 0.38268343236508978f};

// z * W_R^k, W_R = exp(-2 pi i / R) (forward) or exp(+2 pi i / R) (inverse);
// k < R/2, R <= 16.  The quarter-turn case is an exact swap.
template <int R, int K, bool INV>
__device__ __forceinline__ float2 rot(float2 z) {
  constexpr int idx = K * (16 / R);  // angle 2 pi idx / 16
  if constexpr (idx == 0) {
    return z;
  } else if constexpr (idx == 4) {
    return INV ? make_float2(-z.y, z.x) : make_float2(z.y, -z.x);
  } else {
    const float c = kCos16[idx], s = kSin16[idx];
    if constexpr (INV)
      return make_float2(__builtin_fmaf(z.x, c, -(z.y * s)), __builtin_fmaf(z.y, c, z.x * s));
    else
      return make_float2(__builtin_fmaf(z.x, c, z.y * s), __builtin_fmaf(z.y, c, -(z.x * s)));
  }
}

// In-register R-point DFT (natural order in and out), radix-2 split.
template <int R, bool INV>
struct Dft {
  template <int K>
  __device__ __forceinline__ static void combine(float2 *v, const float2 *e, const float2 *o) {
    if constexpr (K < R / 2) {
      float2 t = rot<R, K, INV>(o[K]);
      v[K] = cadd(e[K], t);
      v[K + R / 2] = csub(e[K], t);
      combine<K + 1>(v, e, o);
    }
  }
  __device__ __forceinline__ static void run(float2 *v) {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int k = 0; k < R / 2; k++) {
      e[k] = v[2 * k];
      o[k] = v[2 * k + 1];
    }
    Dft<R / 2, INV>::run(e);
    Dft<R / 2, INV>::run(o);
    combine<0>(v, e, o);
  }
};
template <bool INV>
struct Dft<2, INV> {
  __device__ __forceinline__ static void run(float2 *v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};
template <bool INV>
struct Dft<1, INV> {
  __device__ __forceinline__ static void run(float2 *) {}
};

// LDS index padding: one float2 of padding every 16 entries breaks the
// power-of-two strides of the Stockham scatter across banks.
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
template <int N>
constexpr int lds_entries() {
  return N + N / 16;
}

// Twiddle sources.  GlobalTw: the W_N^k table (global memory, L1/L2 cached).
// PassTw: per-pass table in LDS, entry [pass][r][j mod NS] holding
// W_{NS*R}^{(j mod NS)*r}: for a given r, lanes with consecutive j read
// consecutive words (no bank conflicts); radix-P passes hold P slots of NS
// entries, the remainder pass R.
// get(jm, r, q): the twiddle of input r of the thread's q-th butterfly of
// the pass, j mod NS = jm.
struct GlobalTw {
  const float2 *__restrict__ tw;
  template <int N, int NS, int R>
  __device__ __forceinline__ float2 get(int jm, int r, int) const {
    return tw[jm * r * (N / (NS * R))];
  }
};
template <int P>
constexpr int pass_tw_offset(int pass) {  // entries before pass `pass` (pass >= 1)
  int off = 0, ns = P;
  for (int p = 1; p < pass; p++) {
    off += ns * P;
    ns *= P;
  }
  return off;
}
template <int N, int P>
constexpr int pass_tw_entries() {  // radix-P passes hold P slots of NS, the remainder pass RL
  constexpr int L = full_passes(N, P);
  constexpr int RL = N / ipow_c(P, L);
  return RL > 1 ? pass_tw_offset<P>(L) + ipow_c(P, L) * RL : pass_tw_offset<P>(L);
}
// fills the per-pass table below from the W_N^k table: pass p >= 1 (stage
// NS = P^p, radix R) holds W_{NS*R}^{jm*r} = W_N^{jm*r*N/(NS*R)} at
// [offset(p) + r*NS + jm]; entry idx < COUNT (the table's size: the whole
// table, or the passes before the last where the last pass's twiddles are
// held in registers) is written by thread idx mod nthreads
template <int N, int P, int COUNT = pass_tw_entries<N, P>()>
__device__ __forceinline__ void fill_pass_tw(float2 *tab, const float2 *tw, int tid, int nthreads) {
  static_assert(COUNT >= 0 && COUNT <= pass_tw_entries<N, P>(), "table size");
  constexpr int L = full_passes(N, P);
  constexpr int RL = N / ipow_c(P, L);
  constexpr int NPASS = L + (RL > 1 ? 1 : 0);
  for (int idx = tid; idx < COUNT; idx += nthreads) {
    int pass = 1, off = 0, ns = P;
    while (pass < NPASS - 1 && idx >= off + ns * P) {
      off += ns * P;
      ns *= P;
      pass++;
    }
    const int R = (pass < L) ? P : RL;
    const int rel = idx - off, r = rel / ns, jm = rel % ns;
    tab[idx] = tw[jm * r * (N / (ns * R))];
  }
}
template <int P>
struct PassTw {
  const float2 *tab;  // LDS
  template <int N, int NS, int R>
  __device__ __forceinline__ float2 get(int jm, int r, int) const {
    constexpr int pass = ilog2_c(NS) / ilog2_c(P);
    return tab[pass_tw_offset<P>(pass) + r * NS + jm];
  }
};

// Stockham scatter addresses as one base + compile-time offsets: with
// lpad(i) = i + (i >> 4), lpad(base + r*NS) splits into wbase(j) + woff(r)
// whenever the carries of (base + r*NS) >> 4 do not depend on j:
//   NS % 16 == 0:    lpad(base) + r*NS*17/16
//   NS == 1, R = 8:  8j + (j >> 1) + r        (8j + r, r < 8, never carries into bit 4 beyond j/2)
//   NS == 1, R = 16: 17j + r
//   NS == 8, R = 8:  (j/8)*68 + j%8 + 8r + (r >> 1)
template <int NS, int R>
constexpr bool split_scatter() {
  return NS % 16 == 0 || (NS == 1 && (R == 8 || R == 16)) || (NS == 8 && R == 8);
}
template <int NS, int R>
__device__ __forceinline__ int scatter_base(int j) {
  const int jm = j & (NS - 1);
  const int base = (j / NS) * NS * R + jm;
  if constexpr (NS % 16 == 0) return lpad(base);
  else if constexpr (NS == 1 && R == 8) return 8 * j + (j >> 1);
  else if constexpr (NS == 1 && R == 16) return 17 * j;
  else return (j / 8) * 68 + jm;  // NS == 8, R == 8
}
template <int NS, int R>
constexpr int scatter_off(int r) {
  if (NS % 16 == 0) return r * NS + (r * NS) / 16;
  if (NS == 1) return r;
  return 8 * r + (r >> 1);  // NS == 8, R == 8
}

template <int N, int P, int R, int NS, bool LAST, bool INV, typename TW>
__device__ __forceinline__ void fft_pass(float2 (&x)[P], int t, float2 *lds, const TW &tw) {
  constexpr int TPB = N / P;
  constexpr int Q = P / R;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int j = t + q * TPB;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = x[q + r * Q];
    const int jm = j & (NS - 1);
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) {
        float2 w = tw.template get<N, NS, R>(jm, r, q);
        if constexpr (INV) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    Dft<R, INV>::run(v);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; r++) x[q + r * Q] = v[r];
    } else if constexpr (split_scatter<NS, R>()) {
      float2 *wp = lds + scatter_base<NS, R>(j);
#pragma unroll
      for (int r = 0; r < R; r++) wp[scatter_off<NS, R>(r)] = v[r];
    } else {
      const int base = (j / NS) * NS * R + jm;
#pragma unroll
      for (int r = 0; r < R; r++) lds[lpad(base + r * NS)] = v[r];
    }
  }
}

template <int N, int P, bool INV, int PASS, bool RAW = false, typename TW>
__device__ __forceinline__ void fft_stage(float2 (&x)[P], int t, float2 *lds, const TW &tw) {
  constexpr int L = full_passes(N, P);
  constexpr int RL = N / ipow_c(P, L);
  constexpr int NPASS = L + (RL > 1 ? 1 : 0);
  constexpr int R = (PASS < L) ? P : RL;
  constexpr int NS = ipow_c(P, PASS);
  constexpr int TPB = N / P;
  if constexpr (PASS > 0) {
    fft_exchange_sync<TPB, RAW>();
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (TPB % 16 == 0) {  // lpad(t + m*TPB) = lpad(t) + m*TPB*17/16
      const float2 *rp = lds + lpad(t);
#pragma unroll
      for (int m = 0; m < P; m++) x[m] = rp[m * (TPB + TPB / 16)];
    } else {
#pragma unroll
      for (int m = 0; m < P; m++) x[m] = lds[lpad(t + m * TPB)];
    }
    fft_exchange_sync<TPB, RAW>();
  }
  fft_pass<N, P, R, NS, PASS == NPASS - 1, INV>(x, t, lds, tw);
  if constexpr (PASS + 1 < NPASS) fft_stage<N, P, INV, PASS + 1, RAW>(x, t, lds, tw);
}

// NX independent transforms advanced together: each pass does the NX sets of
// butterflies, one pair of barriers per exchange round for all of them.
template <int N, int P, bool INV, int PASS, int NX, typename TW>
__device__ __forceinline__ void fft_stage_multi(float2 (&x)[NX][P], int t, float2 *lds,
                                                const TW &tw) {
  constexpr int L = full_passes(N, P);
  constexpr int RL = N / ipow_c(P, L);
  constexpr int NPASS = L + (RL > 1 ? 1 : 0);
  constexpr int R = (PASS < L) ? P : RL;
  constexpr int NS = ipow_c(P, PASS);
  constexpr int TPB = N / P;
  if constexpr (PASS > 0) {
    fft_exchange_sync<TPB>();
#pragma unroll
    for (int v = 0; v < NX; v++)
#pragma unroll
      for (int m = 0; m < P; m++) x[v][m] = lds[v * lds_entries<N>() + lpad(t + m * TPB)];
    fft_exchange_sync<TPB>();
  }
#pragma unroll
  for (int v = 0; v < NX; v++)
    fft_pass<N, P, R, NS, PASS == NPASS - 1, INV>(x[v], t, lds + v * lds_entries<N>(), tw);
  if constexpr (PASS + 1 < NPASS) fft_stage_multi<N, P, INV, PASS + 1, NX>(x, t, lds, tw);
}
template <int N, int P, bool INV, int NX, typename TW>
__device__ __forceinline__ void fft_block_multi(float2 (&x)[NX][P], int t, float2 *lds,
                                                const TW &tw) {
  fft_stage_multi<N, P, INV, 0, NX>(x, t, lds, tw);
}

// Two independent transforms advanced together, their exchange buffers at two
// separate places (l0, l1: lds_entries<N>() float2 each): one pair of
// barriers per exchange round for both, and the LDS reads of the second
// overlap the butterflies of the first.  Same arithmetic as fft_block_tw.
template <int N, int P, bool INV, int PASS, typename TW>
__device__ __forceinline__ void fft_stage_pair(float2 (&x)[2][P], int t, float2 *l0, float2 *l1,
                                               const TW &tw) {
  constexpr int L = full_passes(N, P);
  constexpr int RL = N / ipow_c(P, L);
  constexpr int NPASS = L + (RL > 1 ? 1 : 0);
  constexpr int R = (PASS < L) ? P : RL;
  constexpr int NS = ipow_c(P, PASS);
  constexpr int TPB = N / P;
  static_assert(TPB % 16 == 0, "strided exchange reads");
  if constexpr (PASS > 0) {
    fft_exchange_sync<TPB>();
    const float2 *r0 = l0 + lpad(t), *r1 = l1 + lpad(t);
#pragma unroll
    for (int m = 0; m < P; m++) x[0][m] = r0[m * (TPB + TPB / 16)];
#pragma unroll
    for (int m = 0; m < P; m++) x[1][m] = r1[m * (TPB + TPB / 16)];
    fft_exchange_sync<TPB>();
  }
  fft_pass<N, P, R, NS, PASS == NPASS - 1, INV>(x[0], t, l0, tw);
  fft_pass<N, P, R, NS, PASS == NPASS - 1, INV>(x[1], t, l1, tw);
  if constexpr (PASS + 1 < NPASS) fft_stage_pair<N, P, INV, PASS + 1>(x, t, l0, l1, tw);
}
template <int N, int P, bool INV, typename TW>
__device__ __forceinline__ void fft_block_pair(float2 (&x)[2][P], int t, float2 *l0, float2 *l1,
                                               const TW &tw) {
  fft_stage_pair<N, P, INV, 0>(x, t, l0, l1, tw);
}

// Twiddles of the passes before the last from the per-pass LDS table (as
// PassTw), those of the last radix-P pass from registers: there NS = N/P =
// the threads per transform, so thread t's butterfly uses W_N^{t r}, the
// same every transform.
template <int N, int P>
struct LastRegTw {
  const float2 *tab;  // LDS, passes 1 .. L-2
  float2 w[P];        // W_N^{t r}, r = 0 .. P-1
  template <int N_, int NS, int R>
  __device__ __forceinline__ float2 get(int jm, int r, int) const {
    if constexpr (NS * P == N && R == P) {
      return w[r];
    } else {
      constexpr int pass = ilog2_c(NS) / ilog2_c(P);
      return tab[pass_tw_offset<P>(pass) + r * NS + jm];
    }
  }
};

// Twiddles of the radix-P passes from the per-pass LDS table (as PassTw),
// those of the remainder pass (radix RL = N / P^L, the last) from
// registers: there NS = N/RL and thread t's butterflies j = t + q*TPB,
// q < P/RL, use W_N^{j r}, the same every transform.  For transforms whose
// whole table does not fit in LDS beside the exchange buffer.
template <int N, int P>
struct RemRegTw {
  static constexpr int L = full_passes(N, P);
  static constexpr int RL = N / ipow_c(P, L);
  static_assert(RL > 1, "a remainder pass");
  static constexpr int NW = (P / RL) * (RL - 1);
  static constexpr int ENTRIES = pass_tw_offset<P>(L);  // LDS table: passes 1 .. L-1
  const float2 *tab;
  float2 w[NW];  // [q][r - 1]: W_N^{(t + q TPB) r}
  __device__ __forceinline__ void load(const float2 *__restrict__ tw, int t) {
#pragma unroll
    for (int q = 0; q < P / RL; q++)
#pragma unroll
      for (int r = 1; r < RL; r++) w[q * (RL - 1) + r - 1] = tw[((t + q * (N / P)) & (N / RL - 1)) * r];
  }
  template <int N_, int NS, int R>
  __device__ __forceinline__ float2 get(int jm, int r, int q) const {
    if constexpr (NS * RL == N && R == RL) {
      return w[q * (RL - 1) + r - 1];
    } else {
      constexpr int pass = ilog2_c(NS) / ilog2_c(P);
      return tab[pass_tw_offset<P>(pass) + r * NS + jm];
    }
  }
};

// Transform one N-point vector held as x[m] = element t + m*(N/P).  All
// threads of the block must call this the same number of times (it contains
// __syncthreads when the transform spans more than one wave).  lds:
// lds_entries<N>() float2 private to this transform.
template <int N, int P, bool INV, bool RAW = false, typename TW>
__device__ __forceinline__ void fft_block_tw(float2 (&x)[P], int t, float2 *lds, const TW &tw) {
  static_assert(N % P == 0 && N >= P, "N must be a multiple of P");
  fft_stage<N, P, INV, 0, RAW>(x, t, lds, tw);
}
template <int N, int P, bool INV>
__device__ __forceinline__ void fft_block(float2 (&x)[P], int t, float2 *lds,
                                          const float2 *__restrict__ tw) {
  fft_block_tw<N, P, INV>(x, t, lds, GlobalTw{tw});
}

// Twiddle table W_N^k = exp(-2 pi i k / N), k < N, FP64 evaluated, FP32 stored.
int make_twiddles(int N, float2 **dev_table);

}  // namespace mdt
