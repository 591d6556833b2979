// fengine.hip -- the native fused F-engine: int8 (I, Q) samples -> polyphase
// FIR (PpS_Batch, MDT_device_functions.cu:113-161) -> C2C FFT
// (cufftExecC2C, MDT_OneStation_client.cu:1105) -> corner turn
// (ReorderFOutput, MDT_device_functions.cu:183-208) in one kernel, one HBM
// read of the samples and one HBM write of the channelised output.
//
// Workgroup = one input e (or EB inputs for short transforms) x SB consecutive
// spectra.  Thread t owns FFT points n = t + m*TPB.  For every frame f of the
// SB + T - 1 needed, the thread loads its P int8 samples once and
// fma-accumulates them into the SB outputs that use the frame, in tap order
// (bit-identical to mdt_pps_batch_i8).  Each of the SB spectra is then
// transformed with the shared Stockham core (bit-identical to
// mdt_fft_exec_c2c), and the SB x N results are corner-turned through LDS so
// that each store instruction writes 64-byte runs of SB consecutive spectra of
// one (channel, input) row of the channel-major output [N][NN][Ns].
#include "calib.hpp"
#include "fengine.hpp"
#include "fft_core.hpp"

namespace mdt {

// persistent-kernel configuration (spectra per item, transforms per exchange,
// corner-turn method)
constexpr int FE_SB = 4;
constexpr int FE_NX = 2;
constexpr int FE_WMODE = 0;
constexpr bool FE_WPERSIST = true;
constexpr int FE_MR = 1;

template <int N>
struct FCfg {
  static constexpr int P = fft_points(N);
  static constexpr int TPB = N / P;
  static constexpr int EB = TPB >= 64 ? 1 : 64 / TPB;  // inputs per block
  static constexpr int THREADS = TPB * EB;
  static constexpr int SB = N <= 4096 ? 8 : (N == 8192 ? 2 : 1);  // spectra per block
  static constexpr int FFT_LDS = lds_entries<N>();
  static constexpr int ROW = SB + 1;  // padded staging row (float2)
  static constexpr int STAGE_LDS = TPB * ROW;
  static constexpr int SLICE = FFT_LDS > STAGE_LDS ? FFT_LDS : STAGE_LDS;
};

template <int N, int T>
__global__ __launch_bounds__(FCfg<N>::THREADS) void fengine_kernel(FArgs a) {
  using C = FCfg<N>;
  constexpr int P = C::P, TPB = C::TPB, SB = C::SB;
  __shared__ float2 lds_all[C::EB * C::SLICE];

  // XCD-aware: consecutive spectrum blocks of one input on one XCD (shared
  // FIR halo frames, adjacent output segments merge in the same L2).
  const long nwg = gridDim.x;
  const long w = blockIdx.x;
  const long logical = (nwg % 8 == 0) ? (w % 8) * (nwg / 8) + w / 8 : w;
  const int nsb = (a.Ns + SB - 1) / SB;
  const int sb = (int)(logical % nsb);
  const int eg = (int)(logical / nsb);
  const int t = threadIdx.x % TPB;
  const int ey = threadIdx.x / TPB;
  const int e = eg * C::EB + ey;
  const bool valid_e = e < a.NN;
  const int s0 = sb * SB;
  const int scount = min(SB, a.Ns - s0);
  float2 *lds = lds_all + ey * C::SLICE;
  const int Tn = (T > 0) ? T : a.T;

  float2 acc[SB][P];
#pragma unroll
  for (int sl = 0; sl < SB; sl++)
#pragma unroll
    for (int m = 0; m < P; m++) acc[sl][m] = make_float2(0.f, 0.f);

  const long spec = (long)a.NN * N;
  const char2 *src = a.in + (long)s0 * spec + (long)(valid_e ? e : 0) * N + t;
  const float *Wt = a.W + t;

  if constexpr (T > 0) {
    // frame f feeds outputs sl = f - i for taps i = 0..T-1; f ascending means
    // each output accumulates its taps in order i = 0..T-1.
#pragma unroll
    for (int f = 0; f < SB + T - 1; f++) {
      if (f < scount + T - 1 && valid_e) {
        float2 x[P];
#pragma unroll
        for (int m = 0; m < P; m++) {
          char2 v = src[(long)f * spec + m * TPB];
          x[m] = make_float2((float)v.x, (float)v.y);
        }
#pragma unroll
        for (int sl = 0; sl < SB; sl++) {
          const int i = f - sl;
          if (i >= 0 && i < T && sl < scount) {
#pragma unroll
            for (int m = 0; m < P; m++) {
              const float wv = Wt[(long)i * N + m * TPB];
              acc[sl][m].x = __builtin_fmaf(wv, x[m].x, acc[sl][m].x);
              acc[sl][m].y = __builtin_fmaf(wv, x[m].y, acc[sl][m].y);
            }
          }
        }
      }
    }
  } else {
#pragma unroll
    for (int sl = 0; sl < SB; sl++) {
      if (sl < scount && valid_e) {
        for (int i = 0; i < Tn; i++) {
#pragma unroll
          for (int m = 0; m < P; m++) {
            char2 v = src[(long)(sl + i) * spec + m * TPB];
            const float wv = Wt[(long)i * N + m * TPB];
            acc[sl][m].x = __builtin_fmaf(wv, (float)v.x, acc[sl][m].x);
            acc[sl][m].y = __builtin_fmaf(wv, (float)v.y, acc[sl][m].y);
          }
        }
      }
    }
  }

  // FFT of each spectrum (block-uniform trip count: scount is uniform)
#pragma unroll
  for (int sl = 0; sl < SB; sl++)
    if (sl < scount) fft_block<N, P, false>(acc[sl], t, lds, a.tw);

  // optional per-input complex gain and delay (calibration), per channel k:
  // X'[k] = g_e exp(-2 pi i kk tau_e / N) X[k], kk the signed bin
  if (a.gain != nullptr && valid_e) {
    const float2 g = a.gain[e];
    const float tau = a.delay ? a.delay[e] : 0.0f;
#pragma unroll
    for (int m = 0; m < P; m++) {
      const int k = t + m * TPB;
      const float2 ph = cmul(g, delay_phasor(signed_bin(k, N), tau, N));
#pragma unroll
      for (int sl = 0; sl < SB; sl++) acc[sl][m] = cmul(acc[sl][m], ph);
    }
  }

  // corner turn: out[(k*NN + e)*Ns + s0 + sl], staged per m through LDS
  if constexpr (SB == 1) {
    if (valid_e) {
#pragma unroll
      for (int m = 0; m < P; m++)
        a.out[(long)(t + m * TPB) * a.cp + (long)e * a.rp + s0] =
            a.split ? split_sample(acc[0][m], a.split_scale) : acc[0][m];
    }
  } else {
    constexpr int CH = SB / 2;  // 16-byte chunks per (k, e) row segment
    const bool vec = ((a.rp & 1) == 0) && ((a.cp & 1) == 0);
#pragma unroll
    for (int m = 0; m < P; m++) {
      __syncthreads();
#pragma unroll
      for (int sl = 0; sl < SB; sl++)
        lds[t * C::ROW + sl] = a.split ? split_sample(acc[sl][m], a.split_scale) : acc[sl][m];
      __syncthreads();
      if (valid_e) {
#pragma unroll
        for (int v = 0; v < CH; v++) {
          const int q = t + v * TPB;
          const int tt = q / CH, pr = q % CH;
          const int k = tt + m * TPB;
          const int sl = 2 * pr;
          float2 *dst = a.out + (long)k * a.cp + (long)e * a.rp + s0 + sl;
          const float2 v0 = lds[tt * C::ROW + sl], v1 = lds[tt * C::ROW + sl + 1];
          if (vec && sl + 1 < scount) {
            *(float4 *)dst = make_float4(v0.x, v0.y, v1.x, v1.y);
          } else {
            if (sl < scount) dst[0] = v0;
            if (sl + 1 < scount) dst[1] = v1;
          }
        }
      }
    }
  }
}

// Persistent variant for N in [512, 4096] (8 points per thread) and T in
// {4, 8}: one workgroup per slot of the CU, looping over (input, spectrum
// block) items.  Each thread keeps its T x 8 window taps in registers for the
// whole kernel (the window is read from memory once per workgroup, not once per
// spectrum), and the workgroup stages the per-pass twiddle table in LDS once.
// Same arithmetic in the same order as fengine_kernel, so the output is
// bit-identical.
__device__ __forceinline__ float swap_pair(float v) {  // lane l <-> l ^ 1 (DPP quad_perm 1,0,3,2)
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}

// NX: transforms per LDS exchange round.  WMODE: corner turn through LDS
// staging (0), lane-pair DPP swaps with SB = 4 (1), or each lane storing its
// own SB-spectrum segment (2).  WPERSIST: taps stay in registers for the
// whole kernel (true) or are reloaded (L2-resident) for each item and are dead
// outside the FIR, which frees registers for larger SB.
template <int N, int T, int SB, int NX, int WMODE, bool WPERSIST, int MR>
__global__ __launch_bounds__(N / 8, 1) void fengine_persistent_kernel(FArgs a, long nitems) {
  constexpr bool DPPW = (WMODE == 1);
  static_assert(!DPPW || SB == 4, "the lane-pair corner turn writes 4 spectra per segment");
  static_assert(SB % NX == 0, "spectra per block must be a multiple of NX");
  constexpr int P = 8, TPB = N / P;
  constexpr int L = full_passes(N, P);
  constexpr int RL = N / ipow_c(P, L);
  constexpr int NPASS = L + (RL > 1 ? 1 : 0);
  constexpr int TWN = pass_tw_entries<N, P>();
  constexpr int ROW = SB + 1;
  constexpr int FBUF = NX * lds_entries<N>();
  constexpr int SBUF = MR * TPB * ROW;  // staging for MR channel rounds
  constexpr int BUF = (WMODE != 0 || FBUF >= SBUF) ? FBUF : SBUF;
  __shared__ float2 twl[TWN > 0 ? TWN : 1];
  __shared__ float2 buf[BUF];
  const int t = threadIdx.x;

  // per-pass twiddle table: entry [pass][jm][r] = W_{NS*R}^{jm*r}
  for (int idx = t; idx < TWN; idx += TPB) {
    int pass = 1, off = 0, ns = P;
    while (pass < NPASS - 1 && idx >= off + ns * 8) {
      off += ns * 8;
      ns *= P;
      pass++;
    }
    const int R = (pass < L) ? P : RL;
    const int rel = idx - off, jm = rel / 8, r = rel % 8;
    twl[idx] = (r < R) ? a.tw[jm * r * (N / (ns * R))] : make_float2(0.f, 0.f);
  }
  float w[T][P];
  if constexpr (WPERSIST) {
#pragma unroll
    for (int i = 0; i < T; i++)
#pragma unroll
      for (int m = 0; m < P; m++) w[i][m] = a.W[(long)i * N + t + m * TPB];
  }
  __syncthreads();
  const PassTw<P> tws{twl};

  const int nsb = (a.Ns + SB - 1) / SB;
  const long G = gridDim.x;
  const long b = blockIdx.x;
  const long spec = (long)a.NN * N;
  float pw[P];  // split 2: this thread's channels' power over all its items
#pragma unroll
  for (int m = 0; m < P; m++) pw[m] = 0.f;
  for (long base = 0; base < nitems; base += G) {
    // XCD-aware: workgroup b runs on XCD b % 8; consecutive logical items (the
    // spectrum blocks of one input) go to the same XCD in each round.
    const long logical = (G % 8 == 0) ? base + (b % 8) * (G / 8) + b / 8 : base + b;
    if (logical >= nitems) break;
    const int sb = (int)(logical % nsb);
    const int e = (int)(logical / nsb);
    const int s0 = sb * SB;
    const int scount = min(SB, a.Ns - s0);
    const char2 *src = a.in + (long)s0 * spec + (long)e * N + t;

    if constexpr (!WPERSIST) {
#pragma unroll
      for (int i = 0; i < T; i++)
#pragma unroll
        for (int m = 0; m < P; m++) w[i][m] = a.W[(long)i * N + t + m * TPB];
    }
    float2 acc[SB][P];
#pragma unroll
    for (int sl = 0; sl < SB; sl++)
#pragma unroll
      for (int m = 0; m < P; m++) acc[sl][m] = make_float2(0.f, 0.f);
#pragma unroll
    for (int f = 0; f < SB + T - 1; f++) {
      if (f < scount + T - 1) {
        float2 x[P];
#pragma unroll
        for (int m = 0; m < P; m++) {
          char2 v = src[(long)f * spec + m * TPB];
          x[m] = make_float2((float)v.x, (float)v.y);
        }
#pragma unroll
        for (int sl = 0; sl < SB; sl++) {
          const int i = f - sl;
          if (i >= 0 && i < T && sl < scount) {
#pragma unroll
            for (int m = 0; m < P; m++) {
              acc[sl][m].x = __builtin_fmaf(w[i][m], x[m].x, acc[sl][m].x);
              acc[sl][m].y = __builtin_fmaf(w[i][m], x[m].y, acc[sl][m].y);
            }
          }
        }
      }
    }
    // transforms in pairs (block-uniform trip count)
#pragma unroll
    for (int sp = 0; sp < SB; sp += NX) {
      if (sp < scount) {
        float2 (&pair)[NX][P] = *reinterpret_cast<float2(*)[NX][P]>(&acc[sp][0]);
        fft_block_multi<N, P, false, NX>(pair, t, buf, tws);
      }
    }

    if (a.gain != nullptr) {
      const float2 g = a.gain[e];
      const float tau = a.delay ? a.delay[e] : 0.0f;
#pragma unroll
      for (int m = 0; m < P; m++) {
        const int k = t + m * TPB;
        const float2 ph = cmul(g, delay_phasor(signed_bin(k, N), tau, N));
#pragma unroll
        for (int sl = 0; sl < SB; sl++) acc[sl][m] = cmul(acc[sl][m], ph);
      }
    }

    if (a.split == 2) {
#pragma unroll
      for (int m = 0; m < P; m++)
#pragma unroll
        for (int sl = 0; sl < SB; sl++)
          if (sl < scount) {
            const float xr = acc[sl][m].x * a.split_scale, xi = acc[sl][m].y * a.split_scale;
            pw[m] = __builtin_fmaf(xr, xr, __builtin_fmaf(xi, xi, pw[m]));
          }
    }

    if constexpr (WMODE == 2) {
      // each lane stores its own channels' SB-spectrum segments (16-byte
      // stores), no LDS staging and no barriers
      const bool full = (scount == SB) && ((a.rp & 1) == 0) && ((a.cp & 1) == 0);
#pragma unroll
      for (int m = 0; m < P; m++) {
        const int k = t + m * TPB;
        float2 *dst = a.out + (long)k * a.cp + (long)e * a.rp + s0;
        float2 o[SB];
#pragma unroll
        for (int sl = 0; sl < SB; sl++) o[sl] = a.split ? split_sample(acc[sl][m], a.split_scale) : acc[sl][m];
        if (full) {
#pragma unroll
          for (int sl = 0; sl < SB; sl += 2)
            *(float4 *)(dst + sl) = make_float4(o[sl].x, o[sl].y, o[sl + 1].x, o[sl + 1].y);
        } else {
#pragma unroll
          for (int sl = 0; sl < SB; sl++)
            if (sl < scount) dst[sl] = o[sl];
        }
      }
    } else if constexpr (DPPW) {
      // corner turn without LDS: lanes 2g and 2g+1 hold channels k and k+1 for
      // 4 spectra (two 16-byte chunks each); one DPP swap gives each lane pair
      // the two halves of one channel's 32-byte segment for each of 2 stores.
      const int odd = t & 1;
      const bool full = (scount == SB) && ((a.rp & 1) == 0) && ((a.cp & 1) == 0);
#pragma unroll
      for (int m = 0; m < P; m++) {
        const int k = t + m * TPB;
        float2 o[SB];
#pragma unroll
        for (int sl = 0; sl < SB; sl++) o[sl] = a.split ? split_sample(acc[sl][m], a.split_scale) : acc[sl][m];
        if (full) {
          const float4 c0 = make_float4(o[0].x, o[0].y, o[1].x, o[1].y);
          const float4 c1 = make_float4(o[2].x, o[2].y, o[3].x, o[3].y);
          const float4 snd = odd ? c0 : c1;
          const float4 rcv =
              make_float4(swap_pair(snd.x), swap_pair(snd.y), swap_pair(snd.z), swap_pair(snd.w));
          float4 *d1 = (float4 *)(a.out + (long)(k - odd) * a.cp + (long)e * a.rp + s0 + 2 * odd);
          float4 *d2 = (float4 *)(a.out + (long)(k + 1 - odd) * a.cp + (long)e * a.rp + s0 + 2 * odd);
          *d1 = odd ? rcv : c0;
          *d2 = odd ? c1 : rcv;
        } else {
          float2 *dst = a.out + (long)k * a.cp + (long)e * a.rp + s0;
#pragma unroll
          for (int sl = 0; sl < SB; sl++)
            if (sl < scount) dst[sl] = o[sl];
        }
      }
    } else {
      // corner turn staged through LDS, MR channel rounds per barrier pair:
      // each store instruction writes SB consecutive spectra of a (channel,
      // input) row per SB/2 lanes
      constexpr int CH = SB / 2;
      const bool vec = ((a.rp & 1) == 0) && ((a.cp & 1) == 0);
#pragma unroll
      for (int m0 = 0; m0 < P; m0 += MR) {
        __syncthreads();
#pragma unroll
        for (int mr = 0; mr < MR; mr++)
#pragma unroll
          for (int sl = 0; sl < SB; sl++)
            buf[mr * TPB * ROW + t * ROW + sl] =
                (a.split == 1) ? split_sample(acc[sl][m0 + mr], a.split_scale) : acc[sl][m0 + mr];
        __syncthreads();
#pragma unroll
        for (int mr = 0; mr < MR; mr++)
#pragma unroll
          for (int v = 0; v < CH; v++) {
            const int q = t + v * TPB;
            const int tt = q / CH, pr = q % CH;
            const int k = tt + (m0 + mr) * TPB;
            const int sl = 2 * pr;
            const long o = (long)k * a.cp + (long)e * a.rp + s0 + sl;
            float2 *dst = a.out + o;
            const float2 *src_ = buf + mr * TPB * ROW + tt * ROW;
            const float2 v0 = src_[sl], v1 = src_[sl + 1];
            if (a.split == 2) {  // 4-byte FP16 samples
              unsigned *d4 = (unsigned *)a.out + o;
              if (vec && sl + 1 < scount) {
                *(uint2 *)d4 = make_uint2(half_sample(v0, a.split_scale), half_sample(v1, a.split_scale));
              } else {
                if (sl < scount) d4[0] = half_sample(v0, a.split_scale);
                if (sl + 1 < scount) d4[1] = half_sample(v1, a.split_scale);
              }
            } else if (vec && sl + 1 < scount) {
              *(float4 *)dst = make_float4(v0.x, v0.y, v1.x, v1.y);
            } else {
              if (sl < scount) dst[0] = v0;
              if (sl + 1 < scount) dst[1] = v1;
            }
          }
      }
      __syncthreads();  // the staging buffer is the next item's FFT buffer
    }
  }
  if (a.split == 2) {
#pragma unroll
    for (int m = 0; m < P; m++) atomicAdd(&a.power[t + m * TPB], pw[m]);
  }
}

template <int N, int T, int SB, int NX, int WMODE, bool WPERSIST, int MR>
static int launch_persistent(const FArgs &a, hipStream_t st) {
  const long nitems = (long)a.NN * ((a.Ns + SB - 1) / SB);
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fengine_persistent_kernel<N, T, SB, NX, WMODE, WPERSIST, MR>, N / 8,
                                                   0) != hipSuccess)
    return fail(MDT_ERR_DEVICE, "device attributes for the F-engine grid");
  long grid = (long)cus * (per_cu > 0 ? per_cu : 1);
  if (grid > nitems) grid = nitems;
  hipLaunchKernelGGL((fengine_persistent_kernel<N, T, SB, NX, WMODE, WPERSIST, MR>), dim3((unsigned)grid), dim3(N / 8), 0, st, a,
                     nitems);
  return check_launch("fengine_persistent_kernel");
}

template <int N>
static int launch_fengine_n(const FArgs &a, hipStream_t st) {
  using C = FCfg<N>;
  if constexpr (N >= 512 && N <= 4096) {
    if (a.split == 2 && FE_WMODE != 0) return fail(MDT_ERR_UNSUPPORTED, "FP8 needs the LDS corner turn");
    if (a.T == 4) return launch_persistent<N, 4, FE_SB, FE_NX, FE_WMODE, FE_WPERSIST, FE_MR>(a, st);
    if (a.T == 8) return launch_persistent<N, 8, FE_SB, FE_NX, FE_WMODE, FE_WPERSIST, FE_MR>(a, st);
  }
  if (a.split == 2)
    return fail(MDT_ERR_UNSUPPORTED, "FP8 X-engine input needs Nchannels in [512, 4096] and Ntaps 4 or 8");
  const long nsb = (a.Ns + C::SB - 1) / C::SB;
  const long neg = (a.NN + C::EB - 1) / C::EB;
  const long total = nsb * neg;
  MDT_TRY(check_grid(total, C::THREADS, "F-engine"));
  dim3 grid((unsigned)total), block(C::THREADS);
  switch (a.T) {
    case 4: hipLaunchKernelGGL((fengine_kernel<N, 4>), grid, block, 0, st, a); break;
    case 8: hipLaunchKernelGGL((fengine_kernel<N, 8>), grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL((fengine_kernel<N, 0>), grid, block, 0, st, a); break;
  }
  return check_launch("fengine_kernel");
}

int launch_fengine(const FArgs &a, int N, hipStream_t st) {
  switch (N) {
    case 32: return launch_fengine_n<32>(a, st);
    case 64: return launch_fengine_n<64>(a, st);
    case 128: return launch_fengine_n<128>(a, st);
    case 256: return launch_fengine_n<256>(a, st);
    case 512: return launch_fengine_n<512>(a, st);
    case 1024: return launch_fengine_n<1024>(a, st);
    case 2048: return launch_fengine_n<2048>(a, st);
    case 4096: return launch_fengine_n<4096>(a, st);
    case 8192: return launch_fengine_n<8192>(a, st);
    case 16384: return launch_fengine_n<16384>(a, st);
    default: return fail(MDT_ERR_UNSUPPORTED, "Nchannels=%d: the fused F-engine needs a power of two", N);
  }
}

}  // namespace mdt
