// fft_any.hip -- a3 for the lengths that are not powers of two: the drop-in
// cufftPlan1d / cufftExecC2C (MDT_OneStation_client.cu:622, 1105) accepts any
// Nchannels in [32, 16384] (MDT_OneStation_client.h:56-57).  Unnormalised,
// natural order, forward sign -1, in place or out of place (see fft_any.hpp).
#include <atomic>
#include <cmath>
#include <complex>
#include <mutex>
#include <vector>

#include "fft_any.hpp"

namespace mdt {

// ------------------------------------------------------------------ mixed radix
// cos / sin(2 pi j / R), j < R, for the odd radices
__device__ constexpr float kC3[3] = {1.0f, -0.5f, -0.5f};
__device__ constexpr float kS3[3] = {0.0f, 0.86602540378443865f, -0.86602540378443865f};
__device__ constexpr float kC5[5] = {1.0f, 0.30901699437494742f, -0.80901699437494742f, -0.80901699437494742f,
                                     0.30901699437494742f};
__device__ constexpr float kS5[5] = {0.0f, 0.95105651629515357f, 0.58778525229247313f, -0.58778525229247313f,
                                     -0.95105651629515357f};
__device__ constexpr float kC7[7] = {1.0f, 0.62348980185873353f, -0.22252093395631440f, -0.90096886790241913f,
                                     -0.90096886790241913f, -0.22252093395631440f, 0.62348980185873353f};
__device__ constexpr float kS7[7] = {0.0f, 0.78183148246802981f, 0.97492791218182361f, 0.43388373911755812f,
                                     -0.43388373911755812f, -0.97492791218182361f, -0.78183148246802981f};

template <int R>
struct OddTab;
template <>
struct OddTab<3> {
  __device__ static float c(int j) { return kC3[j]; }
  __device__ static float s(int j) { return kS3[j]; }
};
template <>
struct OddTab<5> {
  __device__ static float c(int j) { return kC5[j]; }
  __device__ static float s(int j) { return kS5[j]; }
};
template <>
struct OddTab<7> {
  __device__ static float c(int j) { return kC7[j]; }
  __device__ static float s(int j) { return kS7[j]; }
};

// R-point DFT, R odd, natural order in and out.  Inputs r and R-r pair up:
// a_m = x_m + x_{R-m}, b_m = x_m - x_{R-m}; then with theta = 2 pi m k / R
//   y_k = x_0 + sum_m a_m cos(theta) - i sum_m b_m sin(theta),  y_{R-k} = the same with + i
// (forward; the inverse swaps the signs of the i terms).
template <int R, bool INV>
__device__ __forceinline__ void dft_odd(float2 *v) {
  constexpr int H = (R - 1) / 2;
  float2 a[H], b[H];
#pragma unroll
  for (int m = 1; m <= H; m++) {
    a[m - 1] = cadd(v[m], v[R - m]);
    b[m - 1] = csub(v[m], v[R - m]);
  }
  float2 y[R];
  y[0] = v[0];
#pragma unroll
  for (int m = 0; m < H; m++) y[0] = cadd(y[0], a[m]);
#pragma unroll
  for (int k = 1; k <= H; k++) {
    float2 A = v[0], B = make_float2(0.f, 0.f);
#pragma unroll
    for (int m = 1; m <= H; m++) {
      const int j = (m * k) % R;
      const float c = OddTab<R>::c(j), s = OddTab<R>::s(j);
      A = make_float2(__builtin_fmaf(a[m - 1].x, c, A.x), __builtin_fmaf(a[m - 1].y, c, A.y));
      B = make_float2(__builtin_fmaf(b[m - 1].x, s, B.x), __builtin_fmaf(b[m - 1].y, s, B.y));
    }
    // -i B = (B.y, -B.x)
    const float2 mi = make_float2(A.x + B.y, A.y - B.x), pi = make_float2(A.x - B.y, A.y + B.x);
    y[k] = INV ? pi : mi;
    y[R - k] = INV ? mi : pi;
  }
#pragma unroll
  for (int r = 0; r < R; r++) v[r] = y[r];
}

template <int R, bool INV>
__device__ __forceinline__ void dft_any(float2 *v) {
  if constexpr (R == 3 || R == 5 || R == 7)
    dft_odd<R, INV>(v);
  else
    Dft<R, INV>::run(v);
}

// One Stockham pass of radix R at stage NS (the product of the earlier
// radices) over an N-point vector: butterfly j < N/R reads v_r = x[j + r N/R],
// multiplies by W_{NS R}^{(j mod NS) r} = W_N^{(j mod NS) r N/(NS R)}, does an
// R-point DFT and writes y[(j / NS) NS R + j mod NS + r NS].  The thread
// handles butterflies t, t + TPB, ... (at most QM: TPB >= N/16).  Sources and
// destinations are global memory (the first pass reads the input row, the
// last writes the output row, both 64-lane contiguous) or the transform's LDS
// buffer (padded by lpad).  A pass that reads and writes LDS waits at a
// barrier between its reads and its writes; a pass that writes LDS waits at
// one after its writes.
template <int R, bool INV>
__device__ __attribute__((noinline)) void mixed_pass(const float2 *gsrc, float2 *gdst, float2 *lds,
                                           bool from_lds, bool to_lds, bool valid, int N, int NS, int t, int TPB,
                                           const float2 *__restrict__ tw) {
  constexpr int QM = (16 + R - 1) / R;
  const int nb = N / R;
  float2 v[QM][R];
#pragma unroll
  for (int q = 0; q < QM; q++) {
    const int j = t + q * TPB;
    if (j < nb) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = j + r * nb;
        v[q][r] = from_lds ? lds[lpad(i)] : (valid ? gsrc[i] : make_float2(0.f, 0.f));
      }
    }
  }
  if (from_lds && to_lds) __syncthreads();
  const int tstep = N / (NS * R);
#pragma unroll
  for (int q = 0; q < QM; q++) {
    const int j = t + q * TPB;
    if (j < nb) {
      const int jm = j % NS;
      if (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) {
          float2 w = tw[jm * r * tstep];
          if (INV) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      dft_any<R, INV>(v[q]);
      const int base = (j - jm) * R + jm;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int o = base + r * NS;
        if (to_lds)
          lds[lpad(o)] = v[q][r];
        else if (valid)
          store_stream(gdst + o, v[q][r]);
      }
    }
  }
  if (to_lds) __syncthreads();
}

struct MixedArgs {
  const float2 *in;
  float2 *out;
  long batch;
  const float2 *tw;
  int N, npass, TPB, FPB, lds_stride;
  int rad[kMaxPasses];
};

template <bool INV>
__global__ __launch_bounds__(1024) void fft_mixed_kernel(MixedArgs a) {
  extern __shared__ float2 lds_dyn[];
  const int f = threadIdx.x / a.TPB, t = threadIdx.x % a.TPB;
  const long row = (long)blockIdx.x * a.FPB + f;
  const bool valid = row < a.batch;
  float2 *lds = lds_dyn + f * a.lds_stride;
  const float2 *src = a.in + (valid ? row : 0) * a.N;
  float2 *dst = a.out + (valid ? row : 0) * a.N;
  int NS = 1;
  for (int p = 0; p < a.npass; p++) {
    const bool from_lds = p > 0, to_lds = p + 1 < a.npass;
    switch (a.rad[p]) {
      case 2: mixed_pass<2, INV>(src, dst, lds, from_lds, to_lds, valid, a.N, NS, t, a.TPB, a.tw); break;
      case 3: mixed_pass<3, INV>(src, dst, lds, from_lds, to_lds, valid, a.N, NS, t, a.TPB, a.tw); break;
      case 4: mixed_pass<4, INV>(src, dst, lds, from_lds, to_lds, valid, a.N, NS, t, a.TPB, a.tw); break;
      case 5: mixed_pass<5, INV>(src, dst, lds, from_lds, to_lds, valid, a.N, NS, t, a.TPB, a.tw); break;
      case 7: mixed_pass<7, INV>(src, dst, lds, from_lds, to_lds, valid, a.N, NS, t, a.TPB, a.tw); break;
      default: mixed_pass<8, INV>(src, dst, lds, from_lds, to_lds, valid, a.N, NS, t, a.TPB, a.tw); break;
    }
    NS *= a.rad[p];
  }
}

// ------------------------------------------------------------ mixed, short
// Below kStreamMinN points the one-shot form again, compiled per length: the
// radix sequence, stage spans, butterfly counts and twiddle strides are
// constants, the passes inlined, so j mod NS and the table indices cost no
// run-time division and no call.  Same passes, twiddles and order as
// fft_mixed_kernel (same bits).
struct MixedSeq {
  int n, rad[kMaxPasses];
};
constexpr MixedSeq mixed_seq(int N) {  // fft_factor's sequence: 8s, a 4 or 2, 3s, 5s, 7s
  MixedSeq s{0, {}};
  int n = N;
  while (n % 8 == 0) s.rad[s.n++] = 8, n /= 8;
  if (n % 4 == 0) s.rad[s.n++] = 4, n /= 4;
  if (n % 2 == 0) s.rad[s.n++] = 2, n /= 2;
  for (int r = 3; r <= 7; r += 2)
    while (n % r == 0) s.rad[s.n++] = r, n /= r;
  return s;
}
template <int N, int PASS>
constexpr int seq_ns() {
  int ns = 1;
  for (int p = 0; p < PASS; p++) ns *= mixed_seq(N).rad[p];
  return ns;
}

template <int N, bool INV, int PASS>
__device__ __forceinline__ void ms_short_pass(const float2 *src, float2 *dst, float2 *lds, bool valid, int t,
                                              const float2 *__restrict__ tw) {
  constexpr MixedSeq S = mixed_seq(N);
  constexpr int R = S.rad[PASS], NS = seq_ns<N, PASS>(), TPB = (N + 7) / 8, nb = N / R;
  constexpr int QM = (nb + TPB - 1) / TPB;
  constexpr bool FROM_LDS = PASS > 0, TO_LDS = PASS + 1 < S.n;
  float2 v[QM][R];
#pragma unroll
  for (int q = 0; q < QM; q++) {
    const int j = t + q * TPB;
    if (j < nb) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = j + r * nb;
        v[q][r] = FROM_LDS ? lds[lpad(i)] : (valid ? src[i] : make_float2(0.f, 0.f));
      }
    }
  }
  if (FROM_LDS && TO_LDS) __syncthreads();
#pragma unroll
  for (int q = 0; q < QM; q++) {
    const int j = t + q * TPB;
    if (j < nb) {
      const int jm = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) {
          float2 w = tw[jm * r * (N / (NS * R))];
          if (INV) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      dft_any<R, INV>(v[q]);
      const int base = (j - jm) * R + jm;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int o = base + r * NS;
        if (TO_LDS)
          lds[lpad(o)] = v[q][r];
        else if (valid)
          store_stream(dst + o, v[q][r]);
      }
    }
  }
  if (TO_LDS) __syncthreads();
  if constexpr (PASS + 1 < S.n) ms_short_pass<N, INV, PASS + 1>(src, dst, lds, valid, t, tw);
}

template <int N, bool INV>
__global__ __launch_bounds__(256) void fft_mixed_short_kernel(const float2 *in, float2 *out, long batch,
                                                              const float2 *__restrict__ tw) {
  constexpr int TPB = (N + 7) / 8, FPB = 256 / TPB, STRIDE = N + (N >> 4) + 1;
  static_assert(mixed_seq(N).n >= 2, "at least two passes");
  __shared__ float2 lds[FPB * STRIDE];
  const int f = threadIdx.x / TPB, t = threadIdx.x % TPB;
  const long row = (long)blockIdx.x * FPB + f;
  const bool valid = row < batch;
  ms_short_pass<N, INV, 0>(in + (valid ? row : 0) * N, out + (valid ? row : 0) * N, lds + f * STRIDE, valid, t,
                           tw);
}

template <int N>
static int launch_mixed_short(const float2 *tw, const float2 *in, float2 *out, long batch, bool inv,
                              hipStream_t st) {
  constexpr int TPB = (N + 7) / 8, FPB = 256 / TPB;
  const long blocks = (batch + FPB - 1) / FPB;
  MDT_TRY(check_grid(blocks, TPB * FPB, "FFT batch"));
  if (inv)
    hipLaunchKernelGGL((fft_mixed_short_kernel<N, true>), dim3((unsigned)blocks), dim3(TPB * FPB), 0, st, in, out,
                       batch, tw);
  else
    hipLaunchKernelGGL((fft_mixed_short_kernel<N, false>), dim3((unsigned)blocks), dim3(TPB * FPB), 0, st, in, out,
                       batch, tw);
  return check_launch("fft_mixed_short_kernel");
}

// the compiled short lengths: the 7-smooth lengths in [32, 143] that are not
// powers of two and where the compiled form measured faster (1.6e8-sample
// batches: 1.04-1.89x, e.g. 35 points 1.43 -> 0.76 ms, 63 2.11 -> 1.12, 135
// 1.27 -> 0.88); at 72, 80, 100 and 112 points the run-time kernel was
// faster (0.81-0.96x) and is kept.  -1 when N is not one of them.
static int mixed_short_dispatch(int N, const float2 *tw, const float2 *in, float2 *out, long batch, bool inv,
                                hipStream_t st) {
  switch (N) {
#define MDT_SHORT(n) \
  case n: return launch_mixed_short<n>(tw, in, out, batch, inv, st);
    MDT_SHORT(35) MDT_SHORT(36) MDT_SHORT(40) MDT_SHORT(42) MDT_SHORT(45) MDT_SHORT(48) MDT_SHORT(49)
    MDT_SHORT(50) MDT_SHORT(54) MDT_SHORT(56) MDT_SHORT(60) MDT_SHORT(63) MDT_SHORT(70) MDT_SHORT(75)
    MDT_SHORT(81) MDT_SHORT(84) MDT_SHORT(90) MDT_SHORT(96) MDT_SHORT(98) MDT_SHORT(105) MDT_SHORT(108)
    MDT_SHORT(120) MDT_SHORT(125) MDT_SHORT(126) MDT_SHORT(135) MDT_SHORT(140)
#undef MDT_SHORT
    default: return -1;
  }
}

// ------------------------------------------------------------ mixed, streamed
// The same passes in persistent workgroups (N <= 8192 where the tables fit in
// LDS): each workgroup loops over groups of FPB rows, and the next group's
// rows are loaded into registers while the current group is transformed.
// The twiddles of every pass are gathered once per plan into a per-pass
// table, [pass][r - 1][jm] = W_N^{jm r N/(NS R)} (the same FP32 values as the
// one-shot kernel's W_N table), and copied into LDS once per workgroup, so
// that the passes issue no global loads of their own: vmcnt then counts only
// the prefetch and the stores.  j / NS by a multiply-high with
// ceil(2^32 / NS), exact for j NS < 2^32.  Same butterflies, twiddles and
// order of operations as fft_mixed_kernel, so the same bits.
struct MixedStreamArgs {
  const float2 *in;
  float2 *out;
  long batch;
  const float2 *twp;  // per-pass twiddle table, ntw entries
  int N, npass, TPB, FPB, lds_stride, ntw;
  MixedPassInfo ps[kMaxPasses];
};

constexpr int kStreamPts = 8;  // points a thread loads per row: TPB = ceil(N / 8)

// one pass from LDS, to LDS or (last) to the output row; QM butterflies a
// thread (N / R <= QM TPB since 8 TPB >= N)
template <int R, bool INV>
__device__ __forceinline__ void ms_pass(float2 *lds, float2 *gdst, bool last, bool valid, int N,
                                        const MixedPassInfo &ps, int t, int TPB, const float2 *twl) {
  constexpr int QM = (kStreamPts + R - 1) / R;
  const int nb = N / R;
  float2 v[QM][R];
#pragma unroll
  for (int q = 0; q < QM; q++) {
    const int j = t + q * TPB;
    if (j < nb) {
#pragma unroll
      for (int r = 0; r < R; r++) v[q][r] = lds[lpad(j + r * nb)];
    }
  }
  __syncthreads();  // every read of the buffer done before it is overwritten
#pragma unroll
  for (int q = 0; q < QM; q++) {
    const int j = t + q * TPB;
    if (j < nb) {
      int jd = j, jm = 0;
      if (ps.NS > 1) {
        jd = (int)__umulhi((unsigned)j, ps.magic);
        jm = j - jd * ps.NS;
        const float2 *tw = twl + ps.twoff + jm;
#pragma unroll
        for (int r = 1; r < R; r++) {
          float2 w = tw[(r - 1) * ps.NS];
          if (INV) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      dft_any<R, INV>(v[q]);
      const int base = jd * ps.NS * R + jm;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int o = base + r * ps.NS;
        if (!last)
          lds[lpad(o)] = v[q][r];
        else if (valid)
          store_stream(gdst + o, v[q][r]);
      }
    }
  }
  if (!last) __syncthreads();
}

template <bool INV>
__global__ __launch_bounds__(1024) void fft_mixed_stream_kernel(MixedStreamArgs a) {
  extern __shared__ float2 lds_dyn[];
  float2 *twl = lds_dyn;
  const int f = threadIdx.x / a.TPB, t = threadIdx.x % a.TPB;
  float2 *lds = lds_dyn + a.ntw + f * a.lds_stride;
  for (int i = threadIdx.x; i < a.ntw; i += blockDim.x) twl[i] = a.twp[i];
  const int N = a.N, TPB = a.TPB;
  const long ngroups = (a.batch + a.FPB - 1) / a.FPB;
  // the first pass straight from the loaded points when they are its one
  // butterfly's inputs (radix 8 first, N = 8 TPB: x[t + r TPB])
  const bool reg0 = a.ps[0].R == 8 && N == 8 * TPB;
  float2 x[kStreamPts];
  auto load = [&](long g) {
    const long row0 = g * a.FPB + f;
    const long row = row0 < a.batch ? row0 : a.batch - 1;  // clamped, never stored
    const float2 *src = a.in + row * N;
#pragma unroll
    for (int i = 0; i < kStreamPts; i++) {
      const int n = t + i * TPB;
      x[i] = src[n < N ? n : N - 1];
    }
  };
  long g = blockIdx.x;
  if (g < ngroups) load(g);
  __syncthreads();  // the twiddle table in LDS
  for (; g < ngroups; g += gridDim.x) {
    const long row = g * a.FPB + f;
    const bool valid = row < a.batch;
    float2 *dst = a.out + (valid ? row : 0) * N;
    int p0 = 0;
    if (reg0) {  // NS = 1: no twiddles; butterfly t writes x'[8 t + r]
      dft_any<8, INV>(x);
#pragma unroll
      for (int r = 0; r < 8; r++) lds[lpad(8 * t + r)] = x[r];
      p0 = 1;
    } else {
#pragma unroll
      for (int i = 0; i < kStreamPts; i++) {
        const int n = t + i * TPB;
        if (n < N) lds[lpad(n)] = x[i];
      }
    }
    __syncthreads();
    if (g + gridDim.x < ngroups) load(g + gridDim.x);
    for (int p = p0; p < a.npass; p++) {
      const bool last = p + 1 == a.npass;
      const MixedPassInfo &ps = a.ps[p];
      switch (ps.R) {
        case 2: ms_pass<2, INV>(lds, dst, last, valid, N, ps, t, TPB, twl); break;
        case 3: ms_pass<3, INV>(lds, dst, last, valid, N, ps, t, TPB, twl); break;
        case 4: ms_pass<4, INV>(lds, dst, last, valid, N, ps, t, TPB, twl); break;
        case 5: ms_pass<5, INV>(lds, dst, last, valid, N, ps, t, TPB, twl); break;
        case 7: ms_pass<7, INV>(lds, dst, last, valid, N, ps, t, TPB, twl); break;
        default: ms_pass<8, INV>(lds, dst, last, valid, N, ps, t, TPB, twl); break;
      }
    }
    __syncthreads();  // the last pass's reads done before the next group's writes
  }
}

// twp[ps.twoff + (r - 1) NS + jm] = tw[jm r N / (NS R)] for every pass with NS > 1
__global__ void mixed_twiddle_kernel(MixedStreamArgs a, const float2 *__restrict__ tw, float2 *twp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.ntw) return;
  int p = 1;  // pass 0 has NS = 1 and no entries
  while (p + 1 < a.npass && i >= a.ps[p + 1].twoff) p++;
  const MixedPassInfo &ps = a.ps[p];
  const int rel = i - ps.twoff, r = rel / ps.NS + 1, jm = rel % ps.NS;
  twp[i] = tw[jm * r * (a.N / (ps.NS * ps.R))];
}

// ------------------------------------------------------------------ Bluestein
// v as a value the compiler cannot see through.  The transforms and table
// reads of one kernel repeat the same addresses (twiddles, chirps, LDS
// offsets): laundered per phase, they are recomputed or reloaded where used
// instead of being kept live across the transforms in between (without this
// the 16384-point kernels spill).
template <typename V>
__device__ __forceinline__ V fresh(V v) {
  if constexpr (sizeof(V) == 8)
    asm volatile("" : "+s"(v));
  else
    asm volatile("" : "+v"(v));
  return v;
}

// X_k = c_k sum_n (x_n c_n) conj(c_{k-n}), c_n = exp(-+ pi i n^2 / N) (nk =
// (n^2 + k^2 - (k-n)^2) / 2): a circular convolution of a_n = x_n c_n (zero
// past N) with b_m = conj(c_m) (wrapped, b_{M-m} = b_m) of length M >= 2N-1,
// done as IFFT_M(FFT_M(a) FFT_M(b)) with FFT_M(b) / M precomputed (filt).
// The thread holds points t + m TPB of the M-point vector throughout: the
// power-of-two core returns X[t + m TPB] where it took x[t + m TPB].
template <int M>
__global__ __launch_bounds__(FftCfg<M>::THREADS) void fft_bluestein_kernel(const float2 *in, float2 *out, long batch,
                                                                          int N, const float2 *__restrict__ chirp,
                                                                          const float2 *__restrict__ filt,
                                                                          const float2 *__restrict__ twm) {
  using C = FftCfg<M>;
  __shared__ float2 lds[C::FPB * lds_entries<M>()];
  const int t = threadIdx.x % C::TPB;
  const int f = threadIdx.x / C::TPB;
  const long row = (long)blockIdx.x * C::FPB + f;
  const bool valid = row < batch;
  const float2 *src = in + (valid ? row : 0) * N;
  float2 x[C::P];
#pragma unroll
  for (int m = 0; m < C::P; m++) {
    const int n = t + m * C::TPB;
    x[m] = (valid && n < N) ? cmul(src[n], chirp[n]) : make_float2(0.f, 0.f);
  }
  float2 *l = lds + f * lds_entries<M>();
  fft_block<M, C::P, false>(x, t, l, twm);
#pragma unroll
  for (int m = 0; m < C::P; m++) x[m] = cmul(x[m], filt[t + m * C::TPB]);
  fft_block<M, C::P, true>(x, fresh(t), l, fresh(twm));
  if (valid) {
    float2 *dst = out + row * N;
    const float2 *ck = fresh(chirp);
#pragma unroll
    for (int m = 0; m < C::P; m++) {
      const int k = t + m * C::TPB;
      if (k < N) store_stream(dst + k, cmul(x[m], ck[k]));
    }
  }
}

// Streaming form for 512 <= M <= 4096 (N in [257, 2048]; a transform spans
// whole waves, so the row and its buffer descriptor are wave-uniform), the
// fft_stream_kernel pattern:
// persistent workgroups load their next rows into registers while
// transforming the current ones; twiddles (per-pass table), the filter and
// the chirp sit in LDS, filled once per workgroup, so nothing inside the
// transforms waits for the rows in flight.  Rows are read and written with
// buffer instructions bounded by the row's N points (reads past it return
// zero, writes past it are dropped), so every step issues the same loads
// and stores and the compiler's waits count exactly.  Same tables, products
// and transforms as fft_bluestein_kernel, so the same bits.
constexpr int kBlueStreamMinM = 512, kBlueStreamMaxM = 4096;
typedef unsigned blue_u2 __attribute__((ext_vector_type(2)));

template <int M>
__global__ __launch_bounds__(FftCfg<M>::THREADS) void fft_bluestein_stream_kernel(
    const float2 *in, float2 *out, long batch, int N, const float2 *__restrict__ chirp,
    const float2 *__restrict__ filt, const float2 *__restrict__ twm) {
  using C = FftCfg<M>;
  constexpr int P = C::P, TPB = C::TPB, H = P / 2;  // points t + m TPB < N <= M/2: m < H
  static_assert(TPB % 64 == 0, "one row per wave at most");
  __shared__ float2 lds[C::FPB * lds_entries<M>()];
  __shared__ float2 twl[pass_tw_entries<M, P>()];
  __shared__ float2 fl[M];
  __shared__ float2 cl[M / 2];
  fill_pass_tw<M, P>(twl, twm, threadIdx.x, C::THREADS);
  for (int i = threadIdx.x; i < M; i += C::THREADS) fl[i] = filt[i];
  for (int i = threadIdx.x; i < N; i += C::THREADS) cl[i] = chirp[i];
  __syncthreads();
  const int t = threadIdx.x % TPB;
  const int f = __builtin_amdgcn_readfirstlane(threadIdx.x / TPB);  // wave-uniform (TPB % 64 == 0)
  float2 *l = lds + f * lds_entries<M>();
  const long ngroups = (batch + C::FPB - 1) / C::FPB;
  const long G = gridDim.x;
  auto row_of = [&](long grp) {
    const long row = grp * C::FPB + f;
    return row < batch ? row : batch - 1;  // past the batch: row batch - 1 again, same result
  };
  auto load = [&](float2 (&xv)[H], long grp) {
    grp = grp < ngroups ? grp : grp - G;  // past the last group: the current one again (discarded)
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc((void *)(in + row_of(grp) * N), (short)0, N * 8, 0x00020000);
#pragma unroll
    for (int m = 0; m < H; m++)
      xv[m] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (t + m * TPB) * 8, 0, 0));
  };
  auto step = [&](const float2 (&xv)[H], long grp) {
    float2 x[P];
#pragma unroll
    for (int m = 0; m < P; m++) {
      if (m < H) {  // (compile time) branch-free: the product, then +0 past N
        const int n = t + m * TPB;
        const float2 v = cmul(xv[m < H ? m : 0], cl[n < N ? n : 0]);
        x[m] = n < N ? v : make_float2(0.f, 0.f);
      } else {
        x[m] = make_float2(0.f, 0.f);
      }
    }
    fft_block_tw<M, P, false, true>(x, t, l, PassTw<P>{twl});
#pragma unroll
    for (int m = 0; m < P; m++) x[m] = cmul(x[m], fl[t + m * TPB]);
    fft_block_tw<M, P, true, true>(x, t, l, PassTw<P>{twl});
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc((void *)(out + row_of(grp) * N), (short)0, N * 8, 0x00020000);
#pragma unroll
    for (int m = 0; m < H; m++) {
      const int k = t + m * TPB;
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(blue_u2, cmul(x[m], cl[k < N ? k : 0])), r, k * 8,
                                            0, 0);
    }
  };
  float2 xa[H], xb[H];
  long g = blockIdx.x;
  if (g < ngroups) load(xa, g);
  for (; g < ngroups; g += 2 * G) {  // uniform: every thread runs the same barriers
    load(xb, g + G);
    step(xa, g);
    if (g + G >= ngroups) break;
    load(xa, g + 2 * G);
    step(xb, g + G);
  }
}

// N > 8192: M = 32768 = 2H.  With a_n = 0 for n >= N and N <= H,
//   A_{2k} = FFT_H(a)_k,  A_{2k+1} = FFT_H(a_n W_M^n)_k,
//   c_k = IFFT_H(C_even)_k + W_M^{-k} IFFT_H(C_odd)_k  (k < H),
// C = A FFT_M(b) / M.  Two H-point transforms each way; the even half's
// result stays in registers while the odd half runs.
constexpr int kSplitH = 16384;
__global__ __launch_bounds__(FftCfg<kSplitH>::THREADS) void fft_bluestein_split_kernel(
    const float2 *in, float2 *out, long batch, int N, const float2 *__restrict__ chirp,
    const float2 *__restrict__ filt, const float2 *__restrict__ twh, const float2 *__restrict__ twm) {
  using C = FftCfg<kSplitH>;
  static_assert(C::FPB == 1, "one transform per workgroup");
  __shared__ float2 lds[lds_entries<kSplitH>()];
  const int t = threadIdx.x;
  const long row = blockIdx.x;
  const float2 *src = in + row * N;
  float2 e[C::P], x[C::P];
#pragma unroll
  for (int m = 0; m < C::P; m++) {
    const int n = t + m * C::TPB;
    e[m] = n < N ? cmul(src[n], chirp[n]) : make_float2(0.f, 0.f);
  }
  fft_block<kSplitH, C::P, false>(e, t, lds, twh);
#pragma unroll
  for (int m = 0; m < C::P; m++) e[m] = cmul(e[m], filt[t + m * C::TPB]);
  fft_block<kSplitH, C::P, true>(e, fresh(t), lds, fresh(twh));
  {
    const int t1 = fresh(t);
    const float2 *s1 = fresh(src), *c1 = fresh(chirp), *w1 = fresh(twm);
#pragma unroll
    for (int m = 0; m < C::P; m++) {
      const int n = t1 + m * C::TPB;
      x[m] = n < N ? cmul(cmul(s1[n], c1[n]), w1[n]) : make_float2(0.f, 0.f);
    }
  }
  fft_block<kSplitH, C::P, false>(x, fresh(t), lds, fresh(twh));
  {
    const int t1 = fresh(t);
    const float2 *f1 = fresh(filt) + kSplitH;
#pragma unroll
    for (int m = 0; m < C::P; m++) x[m] = cmul(x[m], f1[t1 + m * C::TPB]);
  }
  fft_block<kSplitH, C::P, true>(x, fresh(t), lds, fresh(twh));
  float2 *dst = out + row * N;
  const int t1 = fresh(t);
  const float2 *c1 = fresh(chirp), *w1 = fresh(twm);
#pragma unroll
  for (int m = 0; m < C::P; m++) {
    const int k = t1 + m * C::TPB;
    if (k < N) {
      const float2 w = w1[k];
      const float2 o = cmul(x[m], make_float2(w.x, -w.y));  // W_M^{-k}
      store_stream(dst + k, cmul(cadd(e[m], o), c1[k]));
    }
  }
  (void)batch;
}

// ------------------------------------------------------------------ host side
bool fft_factor(int N, int *npass, int *rad) {
  int n = N, p = 0;
  while (n % 8 == 0 && p < kMaxPasses) rad[p++] = 8, n /= 8;
  if (n % 4 == 0 && p < kMaxPasses) rad[p++] = 4, n /= 4;
  if (n % 2 == 0 && p < kMaxPasses) rad[p++] = 2, n /= 2;
  for (int r : {3, 5, 7})
    while (n % r == 0 && p < kMaxPasses) rad[p++] = r, n /= r;
  *npass = p;
  return n == 1;
}

// in-place radix-2 FFT in FP64 (host), sign -1
static void host_fft(std::vector<std::complex<double>> &a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; i++) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; k++) {
        const double ang = -2.0 * M_PI * (double)k / (double)len;
        const std::complex<double> w(std::cos(ang), std::sin(ang));
        const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
  }
}

// exp(i pi q / N) for an integer q in [0, 2N), argument reduced to [-1, 1)
static std::complex<double> expi_pi(long q, int N) {
  double r = (double)q / (double)N;
  if (r >= 1.0) r -= 2.0;
  return std::complex<double>(std::cos(M_PI * r), std::sin(M_PI * r));
}

static int upload(const std::vector<float2> &h, float2 **d) {
  if (hipMalloc(d, sizeof(float2) * h.size()) != hipSuccess) return fail(MDT_ERR_ALLOC, "FFT table");
  if (hipMemcpy(*d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail(MDT_ERR_DEVICE, "FFT table upload");
  return MDT_SUCCESS;
}

static float2 to_f2(std::complex<double> z) { return make_float2((float)z.real(), (float)z.imag()); }

// The dynamic-LDS limit belongs to a kernel, not to a plan: it is set once,
// to all of a gfx950 CU's LDS (160 KiB, more than any plan uses), so that a
// plan created later never lowers it below what an earlier one needs.
static int allow_full_lds() {
  static std::once_flag once;
  static bool ok = false;
  std::call_once(once, [] {
    const void *ks[] = {(const void *)fft_mixed_kernel<false>, (const void *)fft_mixed_kernel<true>,
                        (const void *)fft_mixed_stream_kernel<false>, (const void *)fft_mixed_stream_kernel<true>};
    ok = true;
    for (const void *k : ks)
      ok = ok && hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) == hipSuccess;
  });
  return ok ? MDT_SUCCESS : fail(MDT_ERR_DEVICE, "LDS limit of the mixed-radix FFT kernels");
}

// The streamed mixed-radix plan where it fits: ceil(N / 8) threads a row
// (at most 1024), the per-pass twiddle table and FPB padded rows in LDS.
// Otherwise, and below kStreamMinN points, the plan keeps the one-shot
// kernel.  Measured on 1.6e8-sample batches (HIP events): streamed 0.56-0.78
// ms against 1.20-1.60 ms one-shot at 1000-8000 points, 0.77 against 1.35
// ms at 243, 0.53 against 0.82 at 192, 0.93-1.10 against 1.04-1.30 at 144,
// 150, 180, 189 (168 and 175 within 5 % either way); at 35-140 points the
// one-shot kernel was as fast or faster (its many small workgroups).
constexpr int kStreamMinN = 144;

static int mixed_stream_create(FftAny *p) {
  const int N = p->N;
  const int tpb = (N + kStreamPts - 1) / kStreamPts;
  if (N < kStreamMinN || tpb > 1024) return MDT_SUCCESS;
  const int fpb = tpb >= 256 ? 1 : 256 / tpb;
  const int stride = N + (N >> 4) + 1;
  int ns = 1, off = 0;
  for (int i = 0; i < p->npass; i++) {
    MixedPassInfo &q = p->ps[i];
    q.R = p->rad[i];
    q.NS = ns;
    q.twoff = off;
    q.magic = ns > 1 ? (unsigned)(((1ULL << 32) + (unsigned long long)ns - 1) / (unsigned long long)ns) : 0u;
    if (ns > 1) off += ns * (q.R - 1);
    ns *= q.R;
  }
  const size_t lds = sizeof(float2) * ((size_t)off + (size_t)fpb * stride);
  if (lds > 160 * 1024) return MDT_SUCCESS;
  // j / NS by the multiply-high is exact while j NS < 2^32 (j < N)
  if ((unsigned long long)N * (unsigned long long)ns >= (1ULL << 32)) return MDT_SUCCESS;
  // resident workgroups: the smaller occupancy of the two directions
  int dev = 0, cus = 0, per = 0, per_inv = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                               hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fft_mixed_stream_kernel<false>, tpb * fpb, lds) !=
          hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_inv, fft_mixed_stream_kernel<true>, tpb * fpb, lds) !=
          hipSuccess)
    return fail(MDT_ERR_DEVICE, "occupancy of the %d-point FFT", N);
  per = per_inv < per ? per_inv : per;
  if (cus <= 0 || per <= 0) return MDT_SUCCESS;
  if (off > 0) {
    if (hipMalloc(&p->twp, sizeof(float2) * (size_t)off) != hipSuccess)
      return fail(MDT_ERR_ALLOC, "per-pass twiddle table");
    MixedStreamArgs a{};
    a.N = N;
    a.npass = p->npass;
    a.ntw = off;
    for (int i = 0; i < p->npass; i++) a.ps[i] = p->ps[i];
    hipLaunchKernelGGL(mixed_twiddle_kernel, dim3((off + 255) / 256), dim3(256), 0, 0, a, p->tw, p->twp);
    MDT_TRY(check_launch("mixed_twiddle_kernel"));
    if (hipStreamSynchronize(0) != hipSuccess) return fail(MDT_ERR_DEVICE, "per-pass twiddle table build");
  }
  p->ntw = off;
  p->sTPB = tpb;
  p->sFPB = fpb;
  p->stream_lds = lds;
  p->sgrid = (long)cus * per;
  p->stream = true;
  return MDT_SUCCESS;
}

int fft_any_create(int N, FftAny *p) {
  *p = FftAny{};
  p->N = N;
  if (fft_factor(N, &p->npass, p->rad)) {
    p->kind = FFT_ANY_MIXED;
    p->TPB = (N + 7) / 8 <= 1024 ? (N + 7) / 8 : (N + 15) / 16;
    p->FPB = p->TPB >= 256 ? 1 : 256 / p->TPB;
    const int stride = N + (N >> 4) + 1;  // lpad(N - 1) < stride
    p->lds_bytes = sizeof(float2) * (size_t)stride * p->FPB;
    if (p->lds_bytes > 160 * 1024) return fail(MDT_ERR_UNSUPPORTED, "%d-point FFT: LDS of %zu bytes", N, p->lds_bytes);
    MDT_TRY(allow_full_lds());
    MDT_TRY(make_twiddles(N, &p->tw));
    return mixed_stream_create(p);
  }
  int M = 1;
  while (M < 2 * N - 1) M <<= 1;
  p->M = M;
  p->kind = M <= kSplitH ? FFT_ANY_BLUESTEIN : FFT_ANY_BLUESTEIN_SPLIT;
  int st = make_twiddles(p->kind == FFT_ANY_BLUESTEIN ? M : kSplitH, &p->tw);
  for (int inv = 0; inv < 2 && st == MDT_SUCCESS; inv++) {
    const double sg = inv ? 1.0 : -1.0;  // the transform's sign
    std::vector<std::complex<double>> c(N), b(M, 0.0);
    std::vector<float2> hc(N);
    for (long n = 0; n < N; n++) {
      const std::complex<double> e = expi_pi((n * n) % (2L * N), N);
      c[n] = sg < 0 ? std::conj(e) : e;
      hc[n] = to_f2(c[n]);
    }
    for (long m = 0; m < N; m++) {
      b[m] = std::conj(c[m]);
      if (m > 0) b[M - m] = std::conj(c[m]);
    }
    host_fft(b);
    std::vector<float2> hf(M);
    if (p->kind == FFT_ANY_BLUESTEIN) {
      for (int k = 0; k < M; k++) hf[k] = to_f2(b[k] / (double)M);
    } else {  // even bins, then odd bins
      for (int k = 0; k < kSplitH; k++) {
        hf[k] = to_f2(b[2 * k] / (double)M);
        hf[kSplitH + k] = to_f2(b[2 * k + 1] / (double)M);
      }
    }
    st = upload(hc, &p->chirp[inv]);
    if (st == MDT_SUCCESS) st = upload(hf, &p->filt[inv]);
  }
  if (st == MDT_SUCCESS && p->kind == FFT_ANY_BLUESTEIN_SPLIT) {
    std::vector<float2> h(kSplitH);
    for (long n = 0; n < kSplitH; n++) h[n] = to_f2(std::conj(expi_pi(n, kSplitH)));  // exp(-2 pi i n / M)
    st = upload(h, &p->twm);
  }
  return st;
}

void fft_any_destroy(FftAny *p) {
  float2 *ptrs[] = {p->tw, p->chirp[0], p->chirp[1], p->filt[0], p->filt[1], p->twm, p->twp};
  for (float2 *q : ptrs)
    if (q) (void)hipFree(q);
  *p = FftAny{};
}

template <int M>
static int launch_blue(const FftAny &p, const float2 *in, float2 *out, long batch, bool inv, hipStream_t st) {
  using C = FftCfg<M>;
  if constexpr (M >= kBlueStreamMinM && M <= kBlueStreamMaxM) {
    // resident workgroups (CUs x occupancy), queried once per device
    static std::atomic<int> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return fail(MDT_ERR_DEVICE, "current device");
    int grid = cached[dev].load(std::memory_order_relaxed);
    if (grid == 0) {
      int cus = 0, per = 0;
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fft_bluestein_stream_kernel<M>, C::THREADS, 0) !=
              hipSuccess ||
          cus <= 0 || per <= 0)
        return fail(MDT_ERR_DEVICE, "occupancy of the %d-point Bluestein transform", M);
      grid = cus * per;
      cached[dev].store(grid, std::memory_order_relaxed);
    }
    const long groups = (batch + C::FPB - 1) / C::FPB;
    const long g = groups < grid ? groups : grid;
    hipLaunchKernelGGL(fft_bluestein_stream_kernel<M>, dim3((unsigned)g), dim3(C::THREADS), 0, st, in, out, batch,
                       p.N, p.chirp[inv], p.filt[inv], p.tw);
    return check_launch("fft_bluestein_stream_kernel");
  }
  const long blocks = (batch + C::FPB - 1) / C::FPB;
  MDT_TRY(check_grid(blocks, C::THREADS, "FFT batch"));
  hipLaunchKernelGGL(fft_bluestein_kernel<M>, dim3((unsigned)blocks), dim3(C::THREADS), 0, st, in, out, batch, p.N,
                     p.chirp[inv], p.filt[inv], p.tw);
  return check_launch("fft_bluestein_kernel");
}

int fft_any_exec(const FftAny &p, const float2 *in, float2 *out, long batch, bool inv, hipStream_t st) {
  if (p.kind == FFT_ANY_MIXED && p.stream) {
    MixedStreamArgs a{};
    a.in = in;
    a.out = out;
    a.batch = batch;
    a.twp = p.twp;
    a.N = p.N;
    a.npass = p.npass;
    a.TPB = p.sTPB;
    a.FPB = p.sFPB;
    a.lds_stride = p.N + (p.N >> 4) + 1;
    a.ntw = p.ntw;
    for (int i = 0; i < p.npass; i++) a.ps[i] = p.ps[i];
    const long groups = (batch + p.sFPB - 1) / p.sFPB;
    const long grid = groups < p.sgrid ? groups : p.sgrid;
    const int threads = p.sTPB * p.sFPB;
    if (inv)
      hipLaunchKernelGGL(fft_mixed_stream_kernel<true>, dim3((unsigned)grid), dim3(threads), p.stream_lds, st, a);
    else
      hipLaunchKernelGGL(fft_mixed_stream_kernel<false>, dim3((unsigned)grid), dim3(threads), p.stream_lds, st, a);
    return check_launch("fft_mixed_stream_kernel");
  }
  if (p.kind == FFT_ANY_MIXED) {
    const int sh = mixed_short_dispatch(p.N, p.tw, in, out, batch, inv, st);
    if (sh >= 0) return sh;
    MixedArgs a{};
    a.in = in;
    a.out = out;
    a.batch = batch;
    a.tw = p.tw;
    a.N = p.N;
    a.npass = p.npass;
    a.TPB = p.TPB;
    a.FPB = p.FPB;
    a.lds_stride = p.N + (p.N >> 4) + 1;
    for (int i = 0; i < p.npass; i++) a.rad[i] = p.rad[i];
    const long blocks = (batch + p.FPB - 1) / p.FPB;
    const int threads = p.TPB * p.FPB;
    MDT_TRY(check_grid(blocks, threads, "FFT batch"));
    if (inv)
      hipLaunchKernelGGL(fft_mixed_kernel<true>, dim3((unsigned)blocks), dim3(threads), p.lds_bytes, st, a);
    else
      hipLaunchKernelGGL(fft_mixed_kernel<false>, dim3((unsigned)blocks), dim3(threads), p.lds_bytes, st, a);
    return check_launch("fft_mixed_kernel");
  }
  const int d = inv ? 1 : 0;
  if (p.kind == FFT_ANY_BLUESTEIN_SPLIT) {
    MDT_TRY(check_grid(batch, FftCfg<kSplitH>::THREADS, "FFT batch"));
    hipLaunchKernelGGL(fft_bluestein_split_kernel, dim3((unsigned)batch), dim3(FftCfg<kSplitH>::THREADS), 0, st, in,
                       out, batch, p.N, p.chirp[d], p.filt[d], p.tw, p.twm);
    return check_launch("fft_bluestein_split_kernel");
  }
  switch (p.M) {
    case 64: return launch_blue<64>(p, in, out, batch, d, st);
    case 128: return launch_blue<128>(p, in, out, batch, d, st);
    case 256: return launch_blue<256>(p, in, out, batch, d, st);
    case 512: return launch_blue<512>(p, in, out, batch, d, st);
    case 1024: return launch_blue<1024>(p, in, out, batch, d, st);
    case 2048: return launch_blue<2048>(p, in, out, batch, d, st);
    case 4096: return launch_blue<4096>(p, in, out, batch, d, st);
    case 8192: return launch_blue<8192>(p, in, out, batch, d, st);
    case 16384: return launch_blue<16384>(p, in, out, batch, d, st);
    default: return fail(MDT_ERR_UNSUPPORTED, "Bluestein length %d", p.M);
  }
}

}  // namespace mdt
