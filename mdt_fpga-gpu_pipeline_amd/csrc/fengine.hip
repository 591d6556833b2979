// fengine.hip -- the native fused F-engine: int8 (I, Q) samples -> polyphase
// FIR (PpS_Batch, MDT_device_functions.cu:113-161) -> C2C FFT
// (cufftExecC2C, MDT_OneStation_client.cu:1105) -> corner turn
// (ReorderFOutput, MDT_device_functions.cu:183-208) in one kernel, one HBM
// read of the samples and one HBM write of the channelised output.
//
// Workgroup = one input e (or EB inputs for short transforms) x SB consecutive
// spectra.  Thread t owns FFT points n = t + m*TPB.  For every frame f of the
// SB + T - 1 needed, the thread loads its P samples once (int8, or the
// reference's float2) and fma-accumulates them into the SB outputs that use
// the frame, in tap order (bit-identical to mdt_pps_batch[_i8]).  Each of the SB spectra is then
// transformed with the shared Stockham core (bit-identical to
// mdt_fft_exec_c2c), and the SB x N results are corner-turned through LDS
// into the X input (rows or channel-pair layout, XGeom in common.hpp): each
// store instruction writes runs of consecutive spectra of one (channel, input).
#include <atomic>

#include "calib.hpp"
#include "fengine.hpp"
#include "fft_core.hpp"

namespace mdt {

template <int N>
struct FCfg {
  static constexpr int P = fft_points(N);
  static constexpr int TPB = N / P;
  static constexpr int EB = TPB >= 64 ? 1 : 64 / TPB;  // inputs per block
  static constexpr int THREADS = TPB * EB;
  // spectra per block: 8 x 8 points, 4 x 16 points (N = 1024) in registers
  static constexpr int SB = N <= 4096 ? (P == 16 ? 4 : 8) : (N == 8192 ? 2 : 1);
  static constexpr int FFT_LDS = lds_entries<N>();
  static constexpr int ROW = SB + 1;  // padded staging row (float2)
  static constexpr int STAGE_LDS = TPB * ROW;
  static constexpr int SLICE = FFT_LDS > STAGE_LDS ? FFT_LDS : STAGE_LDS;
};

__device__ __forceinline__ float2 fe_sample(char2 v) { return make_float2((float)v.x, (float)v.y); }
__device__ __forceinline__ float2 fe_sample(float2 v) { return v; }

// In: char2 (int8 I, Q) or float2 (the reference's float F input)
template <int N, int T, typename In>
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
  const float sc = a.gain != nullptr ? *a.dscale : a.split_scale;

  float2 acc[SB][P];
#pragma unroll
  for (int sl = 0; sl < SB; sl++)
#pragma unroll
    for (int m = 0; m < P; m++) acc[sl][m] = make_float2(0.f, 0.f);

  const long spec = (long)a.NN * N;
  const In *in0;
  if constexpr (sizeof(In) == 8)
    in0 = a.fin;
  else
    in0 = a.in;
  const In *src = in0 + (long)s0 * spec + (long)(valid_e ? e : 0) * N + t;
  const float *Wt = a.W + t;

  if constexpr (T > 0) {
    // frame f feeds outputs sl = f - i for taps i = 0..T-1; f ascending means
    // each output accumulates its taps in order i = 0..T-1.
#pragma unroll
    for (int f = 0; f < SB + T - 1; f++) {
      if (f < scount + T - 1 && valid_e) {
        float2 x[P];
#pragma unroll
        for (int m = 0; m < P; m++) x[m] = fe_sample(src[(long)f * spec + m * TPB]);
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
            const float2 v = fe_sample(src[(long)(sl + i) * spec + m * TPB]);
            const float wv = Wt[(long)i * N + m * TPB];
            acc[sl][m].x = __builtin_fmaf(wv, v.x, acc[sl][m].x);
            acc[sl][m].y = __builtin_fmaf(wv, v.y, acc[sl][m].y);
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
        a.out[a.xg.chan(t + m * TPB) + a.xg.at(e, s0)] =
            a.split ? split_sample(acc[0][m], sc) : acc[0][m];
    }
  } else {
    constexpr int CH = SB / 2;  // 16-byte chunks per (k, e) row segment
    const bool vec = ((a.xg.rp | a.xg.cp | a.xg.bp) & 1) == 0;  // 16-byte aligned pairs (B is even)
#pragma unroll
    for (int m = 0; m < P; m++) {
      __syncthreads();
#pragma unroll
      for (int sl = 0; sl < SB; sl++)
        lds[t * C::ROW + sl] = a.split ? split_sample(acc[sl][m], sc) : acc[sl][m];
      __syncthreads();
      if (valid_e) {
#pragma unroll
        for (int v = 0; v < CH; v++) {
          const int q = t + v * TPB;
          const int tt = q / CH, pr = q % CH;
          const int k = tt + m * TPB;
          const int sl = 2 * pr;
          float2 *dst = a.out + a.xg.chan(k) + a.xg.at(e, s0 + sl);
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

// Persistent kernel for N in [256, 4096] (P = fft_points(N) points per thread:
// 16 at N = 1024 and 4 at N = 256, so that a transform is one wave and its
// exchanges need no workgroup barrier; 8 otherwise; TPB = N/P threads per
// transform, at least one wave, so that a sub-block is wave-uniform) and T in
// {4, 8}.  A workgroup is Q sub-blocks of TPB
// threads; an item is one input and 4Q consecutive spectra, sub-block q
// taking spectra 4q..4q+3.  Each sub-block runs the FIR and the four
// transforms with its own exchange buffer (the twiddle table is shared), then
// the whole workgroup corner-turns through LDS so that every (channel, input)
// row gets 4Q consecutive samples in one piece: 64 B and more per row instead
// of 32 B.  In the channel-pair layout the rows of channels 2j and 2j+1 of an
// input are adjacent within each block of spectra, so consecutive lanes
// cover both and at N = 4096 every piece is 128 B (two channels x 8 spectra
// of 8 bytes), which is what scattered writes need to run near the HBM write
// rate.  Per thread: 128-VGPR budget, taps read
// per item from a lane-contiguous copy (one 16-byte load per 4 taps), the four
// spectra transformed one at a time in place in the FIR accumulators.  The
// arithmetic and its order are those of fengine_kernel and of the drop-in
// mdt_pps_batch_i8 + mdt_fft_exec_c2c, so every path agrees bit for bit.
template <int N, int SPLIT, bool GAIN>
struct PCfg {
  static constexpr int P = fft_points(N), TPB = N / P, SB = 4;
  // sub-blocks per workgroup: up to 1024 threads (512 with the gain stage,
  // whose phasors need the registers of two waves per SIMD, and with 16
  // points per thread)
  static constexpr int QMAX = ((GAIN || P == 16) ? 512 : 1024) / TPB;
  static constexpr int Q = QMAX < 1 ? 1 : (QMAX < 8 ? QMAX : 8);
  static constexpr int THREADS = Q * TPB;
  static constexpr int SPAN = SB * Q;                 // spectra per item
  static constexpr int SPC = SPLIT == 2 ? 4 : (SPLIT == 3 ? 8 : 2);  // samples per 16-byte output chunk
  static constexpr int CPR = SPAN / SPC;              // chunks per row
  // Staging rows (float2).  Rows of 8 spectra of 8 bytes (N = 4096): no pad,
  // each 16-byte chunk stored at chunk ^ ((row >> 1) & 3), so that both the
  // writers' and the readers' 16-byte accesses are free of bank conflicts;
  // otherwise one float2 of pad per row.
  static constexpr bool SWZ = SPAN == 8 && SPC == 2;
  static constexpr int ROWW = SWZ ? SPAN : SPAN + 1;
  static constexpr int FFT_LDS = lds_entries<N>();    // per sub-block
  // channel rounds per staging pass: as many as fit in the exchange buffers
  // (counted with padded rows)
  static constexpr int MR0 = (Q * FFT_LDS) / (TPB * (SPAN + 1));
  static constexpr int MR = MR0 >= 4 ? 4 : (MR0 >= 2 ? 2 : 1);
  static constexpr int STAGE = MR * TPB * ROWW;
  static constexpr int BUF = Q * FFT_LDS > STAGE ? Q * FFT_LDS : STAGE;
  static constexpr int TWN = pass_tw_entries<N, P>();
};

// float2 index of staging column col (spectrum within the item) of row row
template <class C>
__device__ __forceinline__ int stage_idx(int row, int col) {
  if constexpr (C::SWZ)
    return row * C::ROWW + ((((col >> 1) ^ (row >> 1)) & 3) << 1) + (col & 1);
  else
    return row * C::ROWW + col;
}

// Wt[((m*G4 + i/4)*TPB + t)*4 + i%4] = W[i*N + t + m*TPB], G4 = ceil(T/4)
// groups of four taps, taps i >= T zero
__global__ void arrange_taps_kernel(int N, int T, const float *__restrict__ W, float *__restrict__ Wt) {
  const int G4 = (T + 3) / 4;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N * 4 * G4) return;
  const int TPB = N / fft_points(N);
  const int ii = idx & 3;
  int r = idx >> 2;
  const int t = r % TPB;
  r /= TPB;
  const int i4 = r % G4, m = r / G4;
  const int i = 4 * i4 + ii;
  Wt[idx] = i < T ? W[(long)i * N + t + m * TPB] : 0.0f;
}

bool persistent_shape(int N, int T) { return N >= 256 && N <= 4096 && T >= 4 && T <= MDT_MAX_TAPS; }
int arranged_taps_floats(int N, int T) { return N * 4 * ((T + 3) / 4); }
bool fp8_block_shape(int N) { return N == 512 || N == 1024; }

int arrange_taps(int N, int T, const float *W, float *Wt, hipStream_t st) {
  if (!persistent_shape(N, T)) return fail(MDT_ERR_UNSUPPORTED, "no persistent F-engine for N=%d T=%d", N, T);
  hipLaunchKernelGGL(arrange_taps_kernel, dim3((arranged_taps_floats(N, T) + 255) / 256), dim3(256), 0, st, N, T,
                     W, Wt);
  return check_launch("arrange_taps_kernel");
}

// v as a value the compiler cannot see through: addresses derived from it
// are computed where they are used instead of being hoisted (out of the item
// loop, or to the top of the item) and held in registers across the phases
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// raw buffer resource word 3 for gfx9-family untyped buffer access (32-bit
// data format, no swizzle); loads past num_records return zero
constexpr int kBufferWord3 = 0x00020000;
typedef unsigned v4u __attribute__((ext_vector_type(4)));

// waves per SIMD the register budget targets: 4 (128 VGPRs) for the plain
// outputs; the gain stage, the FP8 power sums at N >= 2048 and 16 points per
// thread (N = 1024: 4 spectra x 16 points are 128 registers) need more room
__device__ __forceinline__ int arranged_taps_floats_dev(int N, int T) { return N * 4 * ((T + 3) / 4); }

template <int N, int SPLIT, bool GAIN>
constexpr int fe_waves_per_simd() {
  return (GAIN || fft_points(N) == 16) ? 2 : ((SPLIT == 2 && N >= 2048) ? 3 : 4);
}

// Frame prefetch by LDS DMA: the next item's NF = SPAN + T - 1 input frames
// (int8 (I, Q), [frame][channel]) are copied into LDS by buffer-to-LDS DMA
// while this item is corner-turned and stored, and the FIR reads its frames
// from LDS.  The FIR's two-byte global loads (88 per thread and item at 4096
// channels) had made it wait on memory latency several times per item (C3
// F-engine 8.33 ms; with the frames prefetched 7.06 ms, same bits).  During
// the transforms the frame area holds the exchange buffers (the frames are
// read by then); the corner-turn staging has its own area, so the DMA never
// targets LDS in use.  At 4096 channels (8 taps, FP32 and FP16 hi/lo
// outputs) each sub-block advances two transforms together: three of the
// four exchange buffers sit in the frame area, the fourth beside the staging,
// and the last pass's twiddles are held in registers (the LDS table keeps
// passes 1 and 2).  At 1024 and 2048 channels the frame area holds every
// sub-block's exchange buffer; the staging rounds per barrier pair (MR) are
// as many as the remaining LDS allows.  Served where it measured faster
// (F-engine alone, same bits; 64 inputs x 8192 spectra at 1024 channels, 40
// and 64 inputs x 4096 spectra at 2048): 1024 channels every output
// (0.85-0.97x the time); 2048 channels the FP16 and FP8 outputs (0.94x at 4
// taps, 0.34x at 8 taps, where the taps do not fit in LDS and the FIR's tap
// loads queued behind its frame loads); not the FP32 output at 2048 channels
// (1.12x: two waves per SIMD spill beside the larger staging rounds) nor 512
// channels (1.07-1.23x: those items are short, the DMA is not hidden).
template <int N, int T, int SPLIT, bool GAIN>
struct FeDma {
  using C = PCfg<N, SPLIT, GAIN>;
  static constexpr int NF = C::SPAN + T - 1;
  static constexpr int FRB = NF * N / 4;  // frame area, float2
  static constexpr bool PAIR = N == 4096;
  static constexpr int L = full_passes(N, C::P);
  static constexpr int TWN = PAIR ? pass_tw_offset<C::P>(L - 1) : C::TWN;
  static constexpr bool TAPS = T * N * 4 <= 32768;
  static constexpr int ROWS = C::TPB * C::ROWW;  // staging float2 per channel round
  static constexpr int XB_MIN = PAIR ? C::FFT_LDS : 0;
  static constexpr int FIXED = TWN * 8 + FRB * 8 + (TAPS ? T * N * 4 : 0) + (SPLIT == 2 ? C::Q * N * 4 : 0) +
                               (SPLIT == 3 ? C::THREADS / 16 : 0) + 64;
  static constexpr int xb(int mr) { return XB_MIN > mr * ROWS ? XB_MIN : mr * ROWS; }
  static constexpr bool fits(int mr) {
    return mr <= C::MR && (mr * C::TPB * C::CPR) % C::THREADS == 0 && xb(mr) * 8 + FIXED <= 163840;
  }
  static constexpr int MR = fits(4) ? 4 : (fits(2) ? 2 : 1);
  static constexpr int XB = xb(MR);
  static constexpr bool ENABLED =
      T > 0 && fits(MR) && (PAIR ? (T == 8 && !GAIN && SPLIT <= 1 && C::Q == 2 && FRB >= 3 * C::FFT_LDS &&
                           C::FFT_LDS >= ROWS && N / ipow_c(C::P, L) == 1)
                        : ((N == 1024 || (N == 2048 && SPLIT != 0)) && FRB >= C::Q * C::FFT_LDS));
};

// s_barrier after this wave's LDS accesses completed, without the release
// fence of __syncthreads(): that fence also waits for every load in flight,
// the frame DMA included (vmcnt(0)); the staging barriers only order LDS
template <bool RAW>
__device__ __forceinline__ void stage_barrier() {
  if constexpr (RAW) {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt, expcnt unchanged
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  } else {
    __syncthreads();
  }
}

// Code-generation choices per output format, each the faster one measured
// (C3 F-engine: FP16 hi/lo and FP32 outputs at N = 4096; C4: the FP8 path's
// output at N = 1024).  At the 128-VGPR budget the register allocation
// decides between them:
//  pin:   each transform's outputs are materialised right after it, so the
//         compiler does not sink the last butterflies into the store rounds
//         (FP16 hi/lo 8.62 against 8.85 ms; slower for the other outputs);
//  lanes: each lane's chunk offsets are computed once per item instead of
//         in every store round (40 % fewer integer instructions; FP16 hi/lo
//         with pin, FP8 12.9-13.1 against 13.1 ms; FP32 10.2 against 9.8 ms);
//  stream: the X input written with non-temporal (streaming) 16-byte stores:
//         it is read back only after the whole launch, far more data than
//         the caches hold (FP16 hi/lo 8.35 against 8.46 ms, FP8 12.14
//         against 13.05 ms; in the FP32 output's per-round stores 10.46
//         against 9.86 ms, and with lanes 10.50 against 9.80 ms).  Taken
//         together with lanes.
template <int SPLIT>
struct FeTuning {  // split 3 (e4m3) follows split 2's choices
  static constexpr bool pin = SPLIT == 1;
  static constexpr bool lanes = SPLIT != 0;
  static constexpr bool stream = SPLIT != 0;
};

// an int8 (I, Q) pair in the low 16 bits of u -> two floats, one SDWA
// conversion of a sign-extended byte each (exact).  Written out because the
// compiler reaches one byte through a bit-field extract or a shift first:
// four VALU instructions per pair instead of two.
__device__ __forceinline__ float2 i8pair_f32(unsigned u) {
  float re, im;
  asm("v_cvt_f32_i32_sdwa %0, sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_0" : "=v"(re) : "v"(u));
  asm("v_cvt_f32_i32_sdwa %0, sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_1" : "=v"(im) : "v"(u));
  return make_float2(re, im);
}

// one 16-byte store of four 32-bit words, streaming or not
template <bool STREAM>
__device__ __forceinline__ void store16(void *dst, unsigned w0, unsigned w1, unsigned w2, unsigned w3) {
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  if constexpr (STREAM)
    __builtin_nontemporal_store(u32x4{w0, w1, w2, w3}, (u32x4 *)dst);
  else
    *(u32x4 *)dst = u32x4{w0, w1, w2, w3};
}

// two complex samples (a, b) -> e4m3 bytes (a.re, a.im, b.re, b.im) of x / scale
// (the scaled conversion divides by its scale operand; |x / scale| <= 448 by
// the block exponent, so nothing saturates), nearest-even
typedef short fe_short2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned e4m3x2(float2 a, float2 b, float scale) {
  fe_short2 r = {0, 0};
  r = __builtin_amdgcn_cvt_scalef32_pk_fp8_f32(r, a.x, a.y, scale, false);
  r = __builtin_amdgcn_cvt_scalef32_pk_fp8_f32(r, b.x, b.y, scale, true);
  return __builtin_bit_cast(unsigned, r);
}

// X input layout with its block and row pitch as constants: XL_ROWS, XL_PAIRS
// (B = 64 bytes of samples, row pitch 2B, block pitch (N/2) cp) or XL_BLOCKS
// (B = 128 bytes of samples, row pitch B, block pitch N cp)
enum { XL_ROWS = 0, XL_PAIRS = 1, XL_BLOCKS = 2 };
template <int N, int T, int SPLIT, bool GAIN, int XL>
__global__ __launch_bounds__((PCfg<N, SPLIT, GAIN>::THREADS), (fe_waves_per_simd<N, SPLIT, GAIN>())) void
fengine_persistent_kernel(FArgs a, long nitems) {
  using C = PCfg<N, SPLIT, GAIN>;
  constexpr int P = C::P, TPB = C::TPB, SB = C::SB, SPAN = C::SPAN, ROWW = C::ROWW;
  constexpr int SPC = C::SPC, CPR = C::CPR;
  constexpr int L = full_passes(N, P);
  constexpr int RL = N / ipow_c(P, L);
  constexpr int NPASS = L + (RL > 1 ? 1 : 0);
  // the arranged taps in LDS where they fit (up to 32 KB: 1024 channels x 8
  // taps, 2048 x 4): read with ds_read_b128, so the FIR's only vector-memory
  // loads are its frames and each wait counts down through them in order;
  // read through L1/L2 otherwise, where a tap load queued behind frame loads
  // waits for all of them
  using D = FeDma<N, T, SPLIT, GAIN>;
  constexpr bool DMA = D::ENABLED;                       // frames prefetched into LDS
  constexpr bool PAIR = DMA && D::PAIR;                  // two transforms per sub-block
  constexpr bool TAPS_LDS =
      DMA ? D::TAPS : (T > 0 && T * N * 4 <= 32768 && (C::TWN + C::BUF) * 8 + T * N * 4 <= 150 * 1024);
  constexpr int NF = SPAN + T - 1;                       // input frames per item
  constexpr int TWN = DMA ? D::TWN : C::TWN;
  constexpr int MR = DMA ? D::MR : C::MR;                // staging rounds per barrier pair
  __shared__ float2 twl[TWN];
  __shared__ float2 buf[DMA ? 1 : C::BUF];
  __shared__ float2 frb[DMA ? D::FRB : 1];  // DMA: the item's frames, [f][channel] int8 (I, Q)
  __shared__ float2 xb[DMA ? D::XB : 1];    // DMA: the staging (PAIR: and the fourth exchange buffer)
  float2 *const sbuf = DMA ? xb : buf;     // corner-turn staging
  __shared__ float wmax[SPLIT == 3 ? C::THREADS / 64 : 1];  // split 3: each wave's largest |re|, |im|
  __shared__ v4u wtl[TAPS_LDS ? T * N / 4 : 1];
  // split 2: the channel powers summed in LDS, one row per sub-block, each
  // entry owned by one thread, items added in order (P more registers do not
  // fit beside the transforms: a 1024-thread workgroup has 128 a thread)
  constexpr bool PW_LDS = SPLIT == 2;
  __shared__ float pwl[PW_LDS ? C::Q * N : 1];
  const int tid = threadIdx.x;
  if constexpr (TAPS_LDS) {
    for (int idx = tid; idx < T * N / 4; idx += C::THREADS) wtl[idx] = ((const v4u *)a.Wt)[idx];
  }
  if constexpr (PW_LDS) {
    for (int idx = tid; idx < C::Q * N; idx += C::THREADS) pwl[idx] = 0.f;
  }
  // the FP16 formats' scale: with gains it covers |g| x the bound (computed on
  // the device before the launch)
  const float split_sc = GAIN ? *a.dscale : a.split_scale;
  // sub-block: uniform within a wave (TPB >= 64); made a scalar explicitly so
  // that frame offsets and buffer bases stay in scalar registers
  const int q = __builtin_amdgcn_readfirstlane(tid / TPB);
  const int t = tid % TPB;

  // per-pass twiddle table: pass p >= 1 (stage NS = 8^p, radix R) holds
  // W_{NS*R}^{jm*r} at [offset(p) + r*NS + jm]
  for (int idx = tid; idx < TWN; idx += C::THREADS) {
    int pass = 1, off = 0, ns = P;
    while (pass < NPASS - 1 && idx >= off + ns * P) {
      off += ns * P;
      ns *= P;
      pass++;
    }
    const int R = (pass < L) ? P : RL;
    const int rel = idx - off, r = rel / ns, jm = rel % ns;
    twl[idx] = a.tw[jm * r * (N / (ns * R))];
  }
  LastRegTw<N, P> tw2{twl, {}};  // PAIR: the last pass's twiddles in registers
  if constexpr (PAIR) {
#pragma unroll
    for (int r = 0; r < P; r++) tw2.w[r] = a.tw[t * r];
  }
  __syncthreads();

  const int nsb = (a.Ns + SPAN - 1) / SPAN;
  const long G = gridDim.x;
  const long b = blockIdx.x;
  const long spec = (long)a.NN * N;
  // the arranged window: T x N floats, or its zero-padded groups of four
  const __amdgpu_buffer_rsrc_t rtap = __builtin_amdgcn_make_buffer_rsrc(
      (void *)a.Wt, (short)0, (T > 0 ? T * N : arranged_taps_floats_dev(N, a.T)) * 4, kBufferWord3);
  float2 *fbuf = (DMA ? frb : buf) + q * C::FFT_LDS;  // this sub-block's exchange buffer
  // XCD-aware: workgroup b runs on XCD b % 8; consecutive logical items (the
  // spectrum blocks of one input: shared FIR frames, adjacent output pieces)
  // go to the same XCD.
  auto logical_of = [&](long base) { return (G % 8 == 0) ? base + (b % 8) * (G / 8) + b / 8 : base + b; };
  // DMA mode: item lg's NF frames (clamped to the input's last frame) into
  // frb, frame f at frb bytes [f N 2, (f + 1) N 2); each wave-instruction
  // moves 1 KB (64 lanes x 16 bytes, written in lane order from the LDS base
  // in M0), 8 per frame
  auto dma_frames = [&](long lg) {
    const int sb_ = (int)(lg % nsb), e_ = (int)(lg / nsb);
    const int s0_ = sb_ * SPAN;
    const int fmax_ = a.Ns + a.T - 2 - s0_;
    const __amdgpu_buffer_rsrc_t rin_ = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(a.in + (long)s0_ * spec + (long)e_ * N), (short)0, 0x7fffffff, kBufferWord3);
    const int tid3 = opaque(tid);
    const int wv = __builtin_amdgcn_readfirstlane(tid3 >> 6), lane = tid3 & 63;
    constexpr int KPF = N * 2 >= 1024 ? N * 2 / 1024 : 1;  // wave-instructions per frame (DMA: N >= 1024)
    for (int c = wv; c < NF * KPF; c += C::THREADS / 64) {
      const int f = c / KPF, k = c % KPF;
      const int fr = f < fmax_ ? f : fmax_;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rin_, (__attribute__((address_space(3))) void *)((char *)frb + (long)c * 1024), 16,
          lane * 16 + k * 1024, (int)(2 * (long)fr * spec), 0, 0);
    }
  };
  if constexpr (DMA) {
    if (logical_of(0) < nitems) dma_frames(logical_of(0));
  }
  // |x|^2 of this thread's channel t + m TPB added to its power sum
  auto add_power = [&](int m, float xr, float xi) {
    float &p = pwl[q * N + t + m * TPB];
    p = __builtin_fmaf(xr, xr, __builtin_fmaf(xi, xi, p));
  };
  for (long base = 0; base < nitems; base += G) {
    const long logical = logical_of(base);
    if (logical >= nitems) break;
    const int sb = (int)(logical % nsb);
    const int e = (int)(logical / nsb);
    const int s0 = sb * SPAN;             // the item's first spectrum
    const int sq = s0 + SB * q;           // this sub-block's first spectrum
    const int scount = min(SB, a.Ns - sq);  // <= 0: a sub-block past Ns (garbage, not stored)

    // FIR, one point at a time: its T taps, then its SB + T - 1 frames; frame
    // f feeds spectrum sl through tap i = f - sl, and each output accumulates
    // its taps in order i = 0..T-1.  Buffer loads: scalar item base and frame
    // offsets, one 32-bit lane offset per point.  Frames past the input
    // (partial last block) are clamped; their spectra are never stored.
    const __amdgpu_buffer_rsrc_t rin = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(a.in + (long)s0 * spec + (long)e * N), (short)0, 0x7fffffff, kBufferWord3);
    const int fmax = a.Ns + a.T - 2 - s0;  // last frame of the input, relative to s0
    if constexpr (DMA) {
      // this item's frames (loaded during the previous item, or before the
      // loop) are in LDS: every wave's DMA done, then visible to all
      asm volatile("" ::: "memory");
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
      __builtin_amdgcn_s_barrier();
      asm volatile("" ::: "memory");
    }
    int twoff = 0;
    asm volatile("" : "+s"(twoff));
    const PassTw<P> tws{twl + twoff};
    int tl = opaque(t);
    float2 acc[SB][P];
    if constexpr (T == 0) {
      // any tap count: the arranged window is zero-padded to a multiple of 4
      // taps and the FIR runs over groups of four, each with its SB + 3
      // frames.  A zero tap adds +0 or -0 to a sum that is never -0, so the
      // sums are those of the T-tap FIR in tap order, bit for bit.
      const int G4 = (a.T + 3) >> 2;
#pragma unroll
      for (int m = 0; m < P; m++) {
        float2 y[SB];
#pragma unroll
        for (int sl = 0; sl < SB; sl++) y[sl] = make_float2(0.f, 0.f);
        for (int g = 0; g < G4; g++) {
          const v4u qv = __builtin_amdgcn_raw_buffer_load_b128(rtap, 16 * tl, 16 * ((m * G4 + g) * TPB), 0);
          const float w[4] = {__uint_as_float(qv[0]), __uint_as_float(qv[1]), __uint_as_float(qv[2]),
                              __uint_as_float(qv[3])};
          float2 x[SB + 3];
#pragma unroll
          for (int f = 0; f < SB + 3; f++) {
            const int ff = SB * q + 4 * g + f;
            const int fr = ff < fmax ? ff : fmax;
            const unsigned short u =
                __builtin_amdgcn_raw_buffer_load_b16(rin, 2 * (tl + m * TPB), (int)(2 * (long)fr * spec), 0);
            x[f] = i8pair_f32(u);
          }
#pragma unroll
          for (int sl = 0; sl < SB; sl++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
              y[sl].x = __builtin_fmaf(w[i], x[sl + i].x, y[sl].x);
              y[sl].y = __builtin_fmaf(w[i], x[sl + i].y, y[sl].y);
            }
        }
#pragma unroll
        for (int sl = 0; sl < SB; sl++) acc[sl][m] = y[sl];
      }
    } else {
#pragma unroll
    for (int m = 0; m < P; m++) {
      float w[T];
#pragma unroll
      for (int i4 = 0; i4 < T / 4; i4++) {
        v4u qv;
        if constexpr (TAPS_LDS)
          qv = wtl[(m * (T / 4) + i4) * TPB + tl];
        else
          qv = __builtin_amdgcn_raw_buffer_load_b128(rtap, 16 * tl, 16 * ((m * (T / 4) + i4) * TPB), 0);
        w[4 * i4] = __uint_as_float(qv[0]);
        w[4 * i4 + 1] = __uint_as_float(qv[1]);
        w[4 * i4 + 2] = __uint_as_float(qv[2]);
        w[4 * i4 + 3] = __uint_as_float(qv[3]);
      }
      float2 x[SB + T - 1];
#pragma unroll
      for (int f = 0; f < SB + T - 1; f++) {
        const int fr = SB * q + f < fmax ? SB * q + f : fmax;
        unsigned short u;
        if constexpr (DMA)  // clamped when copied
          u = ((const unsigned short *)frb)[(SB * q + f) * N + tl + m * TPB];
        else
          u = __builtin_amdgcn_raw_buffer_load_b16(rin, 2 * (tl + m * TPB), (int)(2 * (long)fr * spec), 0);
        x[f] = i8pair_f32(u);  // (I, Q) = sign-extended low and high bytes
      }
#pragma unroll
      for (int sl = 0; sl < SB; sl++) {
        float2 y = make_float2(0.f, 0.f);
#pragma unroll
        for (int i = 0; i < T; i++) {
          y.x = __builtin_fmaf(w[i], x[sl + i].x, y.x);
          y.y = __builtin_fmaf(w[i], x[sl + i].y, y.y);
        }
        acc[sl][m] = y;
      }
    }
    }
    // transforms, one at a time, all SB of them: inside `if (sl < scount)`
    // the FIR sums of the later spectra would be sunk into those branches,
    // keeping every frame's samples live until the last transform
    // DMA: every wave's frame reads done before the frame area becomes the
    // exchange buffers
    if constexpr (DMA) __syncthreads();
    if constexpr (PAIR) {
      // sub-block 0 uses frb slots 0 and 1, sub-block 1 slot 2 and xb
      float2 *l0 = frb + 2 * q * C::FFT_LDS;
      float2 *l1 = q == 0 ? frb + C::FFT_LDS : xb;
#pragma unroll
      for (int pr = 0; pr < SB; pr += 2) {
        float2(&xx)[2][P] = *reinterpret_cast<float2(*)[2][P]>(&acc[pr]);
        fft_block_pair<N, P, false>(xx, opaque(t), l0, l1, tw2);
        if constexpr (FeTuning<SPLIT>::pin) {
#pragma unroll
          for (int v = 0; v < 2; v++)
#pragma unroll
            for (int m = 0; m < P; m++) asm volatile("" : "+v"(xx[v][m].x), "+v"(xx[v][m].y));
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
#pragma unroll
    for (int sl = 0; sl < (PAIR ? 0 : SB); sl++) {
      fft_block_tw<N, P, false>(acc[sl], opaque(t), fbuf, tws);
      if constexpr (FeTuning<SPLIT>::pin) {
#pragma unroll
        for (int m = 0; m < P; m++) asm volatile("" : "+v"(acc[sl][m].x), "+v"(acc[sl][m].y));
      }
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (SPLIT == 2 && !GAIN) {
        // branch-free: spectra past the input weigh zero (finite values
        // times 0, so pw is unchanged bit for bit)
        const float wgt = sl < scount ? split_sc : 0.f;
#pragma unroll
        for (int m = 0; m < P; m++) {
          const float xr = acc[sl][m].x * wgt, xi = acc[sl][m].y * wgt;
          add_power(m, xr, xi);
        }
      }
    }
    if constexpr (DMA) {
      __syncthreads();  // every exchange read done: the frame area is free
      const long nx = logical_of(base + G);
      if (nx < nitems) dma_frames(nx);
    }

    if constexpr (GAIN) {
      tl = opaque(t);
      const float2 g = a.gain[e];
      const float tau = a.delay ? a.delay[e] : 0.0f;
#pragma unroll
      for (int m = 0; m < P; m++) {
        const int k = tl + m * TPB;
        const float2 ph = cmul(g, delay_phasor(signed_bin(k, N), tau, N));
#pragma unroll
        for (int sl = 0; sl < SB; sl++) acc[sl][m] = cmul(acc[sl][m], ph);
      }
      if constexpr (SPLIT == 2) {
#pragma unroll
        for (int m = 0; m < P; m++)
#pragma unroll
          for (int sl = 0; sl < SB; sl++) {
            const float wgt = sl < scount ? split_sc : 0.f;
            const float xr = acc[sl][m].x * wgt, xi = acc[sl][m].y * wgt;
            add_power(m, xr, xi);
          }
      }
    }

    // split 3: the largest |re| or |im| of the spectra in the input (no
    // spectrum past Ns), per wave into LDS; the item's blocks of 16 spectra
    // take the maximum over their sub-blocks' waves after the next barrier
    if constexpr (SPLIT == 3) {
      float mx = 0.f;
#pragma unroll
      for (int sl = 0; sl < SB; sl++) {
        const float wgt = sl < scount ? 1.f : 0.f;
#pragma unroll
        for (int m = 0; m < P; m++)
          mx = fmaxf(mx, fmaxf(fabsf(acc[sl][m].x), fabsf(acc[sl][m].y)) * wgt);
      }
#pragma unroll
      for (int msk = 32; msk >= 1; msk >>= 1) mx = fmaxf(mx, __shfl_xor(mx, msk, 64));
      if ((tid & 63) == 0) wmax[tid >> 6] = mx;
    }

    // corner turn through LDS, MR channel rounds per barrier pair: row
    // (channel) r of a round holds the item's SPAN spectra; each 16-byte
    // chunk of a row goes to one lane, the CPR chunks of a row to consecutive
    // lanes, so each store instruction writes whole SPAN-sample row pieces
    // whole 16-byte chunks: always in the pairs and blocks layouts (cp and B
    // and the chunk offsets are multiples of the 4 / 2 samples of a chunk),
    // in the rows layout when the pitches allow
    const long pitches = a.xg.rp | a.xg.cp;
    const bool vec = XL != XL_ROWS || (SPLIT == 2 ? ((pitches & 3) == 0) : ((pitches & 1) == 0));
    const int scnt_item = min(SPAN, a.Ns - s0);
    // Chunk c = tid + v*THREADS of a round is row c / CPR of the staging
    // buffer, channel k = row % TPB + (m0 + row / TPB)*TPB, spectra
    // s1 = (c % CPR)*SPC onwards.  Its X-input offset
    // xg.chan(k) + xg.at(e, s0 + s1) is, in the layouts' own terms,
    //   rows:   k*cp + e*rp + s0 + s1
    //   pairs:  (blk*(N/2) + (k >> 1))*cp + (2e + (k & 1))*B + in_blk
    //   blocks: (blk*N + k)*cp + e*B + in_blk
    // with blk, in_blk the block of s0 + s1 and the place in it (s0, a
    // multiple of SPAN, and e are uniform).
    // samples per block of the layout: 64 bytes of 8-byte or 4-byte samples in
    // the pairs layout, 128 in the blocks layout; split 3 (blocks only) 64
    // bytes of 2-byte samples, one K chunk of the FP8 X-engine
    constexpr int LGB = SPLIT == 3 ? 5 : (SPLIT == 2 ? 4 : 3) + (XL == XL_BLOCKS ? 1 : 0), B = 1 << LGB;
    if constexpr (FeTuning<SPLIT>::lanes) {
      // = ibase + m0*rstride + v*vstride + xoff, xoff the lane's part,
      // computed once per item: with RV = THREADS / CPR rows per v, row =
      // row0 + v*RV and k = row + m0*TPB; in the pairs layout k >> 1 =
      // (row0 >> 1) + v*RV/2 + m0*TPB/2 and k & 1 = row0 & 1 (RV, TPB even)
      constexpr int NV = (MR * TPB * CPR) / C::THREADS, RV = C::THREADS / CPR;
      static_assert(C::THREADS % CPR == 0 && RV % 8 == 0 && TPB % 2 == 0, "chunk geometry");
      const bool full = vec && scnt_item == SPAN;  // whole chunks only (uniform)
      const int tid2 = opaque(tid);
      const int s1 = (tid2 % CPR) * SPC;  // the same for every v
      const int row0 = tid2 / CPR;
      const float2 *src0 = sbuf + stage_idx<C>(row0, s1);  // the lane's chunk; rows row0 + v*RV share its swizzle
      long xoff, ibase;
      if constexpr (XL == XL_ROWS) {
        xoff = (long)row0 * a.xg.cp + s1;
        ibase = (long)e * a.xg.rp + s0;
      } else if constexpr (XL == XL_PAIRS) {
        xoff = (long)(row0 >> 1) * a.xg.cp + (row0 & 1) * B +
               (SPAN >= B ? (long)(s1 >> LGB) * (N / 2) * a.xg.cp + (s1 & (B - 1)) : (long)s1);
        ibase = (long)(s0 >> LGB) * (N / 2) * a.xg.cp + e * 2 * B + (SPAN >= B ? 0 : (s0 & (B - 1)));
      } else {
        xoff = (long)row0 * a.xg.cp + (SPAN >= B ? (long)(s1 >> LGB) * N * a.xg.cp + (s1 & (B - 1)) : (long)s1);
        ibase = (long)(s0 >> LGB) * N * a.xg.cp + e * B + (SPAN >= B ? 0 : (s0 & (B - 1)));
      }
      const long rstride = (XL == XL_PAIRS ? (long)(TPB / 2) : (long)TPB) * a.xg.cp;  // per channel block of TPB
      const long vstride = (XL == XL_PAIRS ? (long)(RV / 2) : (long)RV) * a.xg.cp;    // per RV rows
      float f8scale = 1.f;  // split 3: 2^eb of the lane's block of 16 spectra
#pragma unroll
      for (int m0 = 0; m0 < P; m0 += MR) {
        stage_barrier<DMA>();
        tl = opaque(t);
#pragma unroll
        for (int mr = 0; mr < MR; mr++)
#pragma unroll
          for (int sl = 0; sl < SB; sl++)
            sbuf[stage_idx<C>(mr * TPB + tl, SB * q + sl)] =
                (SPLIT == 1) ? split_sample(acc[sl][m0 + mr], split_sc) : acc[sl][m0 + mr];
        stage_barrier<DMA>();
        if constexpr (SPLIT == 3) {
          if (m0 == 0) {  // the waves' maxima are visible: each block's exponent
            constexpr int WPS = TPB / 64;  // waves per sub-block
            auto block_max = [&](int b) {
              float bm = 0.f;
              for (int w = 4 * b * WPS; w < min(C::Q, 4 * b + 4) * WPS; w++) bm = fmaxf(bm, wmax[w]);
              return bm;
            };
            const int eb = fp8_block_exponent(block_max(s1 >> 4));
            f8scale = __builtin_ldexpf(1.0f, eb);
            if (tid < SPAN / 16 && (s0 >> 4) + tid < a.nblk16)
              a.scales[(long)((s0 >> 4) + tid) * a.NN + e] =
                  (unsigned char)(fp8_block_exponent(block_max(tid)) + 127);
          }
        }
        const long rbase = ibase + (long)m0 * rstride + xoff;
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const long o = rbase + v * vstride;
          const float2 *src_ = src0 + v * RV * ROWW;
          if constexpr (SPLIT == 3) {  // 2-byte e4m3 samples, 8 per chunk, spectra past Ns zero
            float2 v[8];
#pragma unroll
            for (int i = 0; i < 8; i++) v[i] = s1 + i < scnt_item ? src_[i] : make_float2(0.f, 0.f);
            store16<FeTuning<SPLIT>::stream>((char *)a.out + 2 * o, e4m3x2(v[0], v[1], f8scale),
                                             e4m3x2(v[2], v[3], f8scale), e4m3x2(v[4], v[5], f8scale),
                                             e4m3x2(v[6], v[7], f8scale));
          } else if constexpr (SPLIT == 2) {  // 4-byte FP16 samples, 4 per chunk
            unsigned *d4 = (unsigned *)a.out + o;
            const unsigned h0 = half_sample(src_[0], split_sc), h1 = half_sample(src_[1], split_sc);
            const unsigned h2 = half_sample(src_[2], split_sc), h3 = half_sample(src_[3], split_sc);
            if (full || (vec && s1 + 3 < scnt_item)) {
              store16<FeTuning<SPLIT>::stream>(d4, h0, h1, h2, h3);
            } else {
              if (s1 < scnt_item) d4[0] = h0;
              if (s1 + 1 < scnt_item) d4[1] = h1;
              if (s1 + 2 < scnt_item) d4[2] = h2;
              if (s1 + 3 < scnt_item) d4[3] = h3;
            }
          } else {  // 8-byte samples, 2 per chunk
            float2 *dst = a.out + o;
            const float2 v0 = src_[0], v1 = src_[1];
            if (full || (vec && s1 + 1 < scnt_item)) {
              store16<FeTuning<SPLIT>::stream>(dst, __float_as_uint(v0.x), __float_as_uint(v0.y),
                                               __float_as_uint(v1.x), __float_as_uint(v1.y));
            } else {
              if (s1 < scnt_item) dst[0] = v0;
              if (s1 + 1 < scnt_item) dst[1] = v1;
            }
          }
        }
      }
    } else {  // the offsets recomputed in every round
#pragma unroll
      for (int m0 = 0; m0 < P; m0 += MR) {
        stage_barrier<DMA>();
        tl = opaque(t);
#pragma unroll
        for (int mr = 0; mr < MR; mr++)
#pragma unroll
          for (int sl = 0; sl < SB; sl++)
            sbuf[stage_idx<C>(mr * TPB + tl, SB * q + sl)] =
                (SPLIT == 1) ? split_sample(acc[sl][m0 + mr], split_sc) : acc[sl][m0 + mr];
        stage_barrier<DMA>();
        const int tid2 = opaque(tid);
#pragma unroll
        for (int v = 0; v < (MR * TPB * CPR) / C::THREADS; v++) {
          const int c = tid2 + v * C::THREADS;
          const int row = c / CPR, cc = c % CPR;
          const int k = row % TPB + (m0 + row / TPB) * TPB;
          const int s1 = cc * SPC;  // first spectrum of the chunk, relative to s0
          long o;
          if constexpr (XL == XL_ROWS) {
            o = (long)k * a.xg.cp + (long)e * a.xg.rp + s0 + s1;
          } else {
            // keep the block index uniform where a block holds whole items
            const int blk = SPAN >= B ? (s0 >> LGB) + (s1 >> LGB) : (s0 >> LGB);
            const int in_blk = SPAN >= B ? (s1 & (B - 1)) : (s0 & (B - 1)) + s1;
            if constexpr (XL == XL_PAIRS)
              o = ((long)blk * (N / 2) + (k >> 1)) * a.xg.cp + ((e * 2 + (k & 1)) * B + in_blk);
            else
              o = ((long)blk * N + k) * a.xg.cp + (e * B + in_blk);
          }
          const float2 *src_ = sbuf + stage_idx<C>(row, s1);
          if constexpr (SPLIT == 2) {  // 4-byte FP16 samples, 4 per chunk
            unsigned *d4 = (unsigned *)a.out + o;
            const unsigned h0 = half_sample(src_[0], split_sc), h1 = half_sample(src_[1], split_sc);
            const unsigned h2 = half_sample(src_[2], split_sc), h3 = half_sample(src_[3], split_sc);
            if (vec && s1 + 3 < scnt_item) {
              *(uint4 *)d4 = make_uint4(h0, h1, h2, h3);
            } else {
              if (s1 < scnt_item) d4[0] = h0;
              if (s1 + 1 < scnt_item) d4[1] = h1;
              if (s1 + 2 < scnt_item) d4[2] = h2;
              if (s1 + 3 < scnt_item) d4[3] = h3;
            }
          } else {  // 8-byte samples, 2 per chunk
            float2 *dst = a.out + o;
            const float2 v0 = src_[0], v1 = src_[1];
            if (vec && s1 + 1 < scnt_item) {
              *(float4 *)dst = make_float4(v0.x, v0.y, v1.x, v1.y);
            } else {
              if (s1 < scnt_item) dst[0] = v0;
              if (s1 + 1 < scnt_item) dst[1] = v1;
            }
          }
        }
      }
    }
    stage_barrier<DMA>();  // the staging buffer is the next item's exchange buffers
  }
  if constexpr (SPLIT == 2) {
    // the workgroup's channel powers: its sub-blocks' sums added in the
    // order q = 0..Q-1 from LDS, one row of the partial table per
    // workgroup (its items are fixed by its index), reduced in order by
    // power_reduce_kernel: the same bits on every run
    __syncthreads();
    for (int k = tid; k < N; k += C::THREADS) {
      float sum = 0.f;
      for (int qq = 0; qq < C::Q; qq++) sum += pwl[qq * N + k];
      a.power_part[(long)blockIdx.x * N + k] = sum;
    }
  }
}

__global__ void power_reduce_kernel(int N, int rows, const float *__restrict__ part, float *__restrict__ power) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  float sum = 0.f;
  for (int r = 0; r < rows; r++) sum += part[(long)r * N + k];
  power[k] = sum;
}

template <int N, int T, int SPLIT, bool GAIN, int XL>
static int launch_persistent_k(const FArgs &a, hipStream_t st) {
  using C = PCfg<N, SPLIT, GAIN>;
  static_assert((C::MR * C::TPB * C::CPR) % C::THREADS == 0, "store chunks must divide among the threads");
  const long nitems = (long)a.NN * ((a.Ns + C::SPAN - 1) / C::SPAN);
  // resident workgroups per device (CUs x workgroups per CU), queried once per
  // device for this kernel and kept
  constexpr int kDevs = 64;
  static std::atomic<int> resident[kDevs];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDevs)
    return fail(MDT_ERR_DEVICE, "current device for the F-engine grid");
  int full = resident[dev].load(std::memory_order_relaxed);
  if (full == 0) {
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fengine_persistent_kernel<N, T, SPLIT, GAIN, XL>,
                                                     C::THREADS, 0) != hipSuccess)
      return fail(MDT_ERR_DEVICE, "device attributes for the F-engine grid");
    full = cus * (per_cu > 0 ? per_cu : 1);
    resident[dev].store(full, std::memory_order_relaxed);
  }
  long grid = full;
  if (grid > nitems) grid = nitems;
  if (SPLIT == 2 && grid > a.power_slots) grid = a.power_slots;
  hipLaunchKernelGGL((fengine_persistent_kernel<N, T, SPLIT, GAIN, XL>), dim3((unsigned)grid), dim3(C::THREADS), 0,
                     st, a, nitems);
  MDT_TRY(check_launch("fengine_persistent_kernel"));
  if constexpr (SPLIT == 2) {
    hipLaunchKernelGGL(power_reduce_kernel, dim3((N + 255) / 256), dim3(256), 0, st, N, (int)grid, a.power_part,
                       a.power);
    MDT_TRY(check_launch("power_reduce_kernel"));
  }
  return MDT_SUCCESS;
}

template <int N, int T, int SPLIT>
static int launch_persistent_s(const FArgs &a, hipStream_t st) {
  const bool gain = a.gain != nullptr;
  const XGeom &g = a.xg;
  if constexpr (SPLIT == 3) {  // e4m3: blocks of 32 two-byte samples only
    if (g.lgG == 0 && g.lgB == 5 && g.rp == 32 && g.bp == (long)N * g.cp)
      return gain ? launch_persistent_k<N, T, 3, true, XL_BLOCKS>(a, st)
                  : launch_persistent_k<N, T, 3, false, XL_BLOCKS>(a, st);
    return fail(MDT_ERR_INVALID_VALUE, "the FP8 block-scaled X input is in blocks of 32 spectra");
  }
  constexpr int LGB = SPLIT == 2 ? 4 : 3;  // pairs; blocks one more
  if (g.lgG == 1 && g.lgB == LGB && g.rp == (2L << LGB) && g.bp == (long)(N / 2) * g.cp)
    return gain ? launch_persistent_k<N, T, SPLIT, true, XL_PAIRS>(a, st)
                : launch_persistent_k<N, T, SPLIT, false, XL_PAIRS>(a, st);
  if (g.lgG == 0 && g.lgB == LGB + 1 && g.rp == (2L << LGB) && g.bp == (long)N * g.cp)
    return gain ? launch_persistent_k<N, T, SPLIT, true, XL_BLOCKS>(a, st)
                : launch_persistent_k<N, T, SPLIT, false, XL_BLOCKS>(a, st);
  if (g.lgG == 0 && g.bp == 0)
    return gain ? launch_persistent_k<N, T, SPLIT, true, XL_ROWS>(a, st)
                : launch_persistent_k<N, T, SPLIT, false, XL_ROWS>(a, st);
  return fail(MDT_ERR_INVALID_VALUE, "X input geometry not served by the persistent F-engine");
}

template <int N, int T>
static int launch_persistent(const FArgs &a, hipStream_t st) {
  if (a.Wt == nullptr) return fail(MDT_ERR_INVALID_VALUE, "persistent F-engine needs the arranged taps");
  switch (a.split) {
    case 0: return launch_persistent_s<N, T, 0>(a, st);
    case 1: return launch_persistent_s<N, T, 1>(a, st);
    case 2:
      if constexpr (N >= 2048) return launch_persistent_s<N, T, 2>(a, st);
      break;
    case 3:
      if constexpr (N >= 512 && N <= 1024) return launch_persistent_s<N, T, 3>(a, st);
      break;
    default: break;
  }
  return fail(MDT_ERR_INVALID_VALUE, "F-engine output format %d is not served at Nchannels=%d", a.split, N);
}

template <int N>
static int launch_fengine_n(const FArgs &a, hipStream_t st) {
  using C = FCfg<N>;
  const bool fin = a.fin != nullptr;
  if constexpr (N >= 256 && N <= 4096) {
    if (!fin && a.T == 4) return launch_persistent<N, 4>(a, st);
    if (!fin && a.T == 8) return launch_persistent<N, 8>(a, st);
    if (!fin && persistent_shape(N, a.T)) return launch_persistent<N, 0>(a, st);
  }
  if (a.split >= 2)
    return fail(MDT_ERR_UNSUPPORTED,
                "FP8 X-engine input needs int8 samples and Nchannels in [512, 4096]");
  const long nsb = (a.Ns + C::SB - 1) / C::SB;
  const long neg = (a.NN + C::EB - 1) / C::EB;
  const long total = nsb * neg;
  MDT_TRY(check_grid(total, C::THREADS, "F-engine"));
  dim3 grid((unsigned)total), block(C::THREADS);
  if (fin) {
    switch (a.T) {
      case 4: hipLaunchKernelGGL((fengine_kernel<N, 4, float2>), grid, block, 0, st, a); break;
      case 8: hipLaunchKernelGGL((fengine_kernel<N, 8, float2>), grid, block, 0, st, a); break;
      default: hipLaunchKernelGGL((fengine_kernel<N, 0, float2>), grid, block, 0, st, a); break;
    }
  } else {
    switch (a.T) {
      case 4: hipLaunchKernelGGL((fengine_kernel<N, 4, char2>), grid, block, 0, st, a); break;
      case 8: hipLaunchKernelGGL((fengine_kernel<N, 8, char2>), grid, block, 0, st, a); break;
      default: hipLaunchKernelGGL((fengine_kernel<N, 0, char2>), grid, block, 0, st, a); break;
    }
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
