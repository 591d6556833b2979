// pps.hip -- a2: the polyphase FIR front half of the PFB, drop-in for
// PpS_Batch (MDT_device_functions.cu:113-161, launch
// MDT_OneStation_client.cu:1083).
//
// out[s][e][c] = sum_{i<T} W[i*N + c] * in[s+i][e][c], FP32 fmaf in tap order
// i = 0..T-1 starting from 0 (MDT_device_functions.cu:140-152) -> bit-exact
// with the reference arithmetic for the same window values.
//
// MI355X layout: one thread owns one (input e, channel c) column and walks a
// run of RUN consecutive spectra, keeping the last T input samples in a
// register ring.  Every input frame is read once per run (plus the T-1 halo),
// instead of T times as in a thread-per-output mapping; loads and stores are
// 64-lane contiguous along channels.  Taps live in registers.
#include "common.hpp"

namespace mdt {

constexpr int PPS_THREADS = 256;
constexpr int PPS_RUN = 32;

__device__ __forceinline__ float2 load_sample(const float2 *p) { return *p; }
__device__ __forceinline__ float2 load_sample(const char2 *p) {
  char2 v = *p;
  return make_float2((float)v.x, (float)v.y);
}

template <int T, typename In>
__global__ __launch_bounds__(PPS_THREADS) void pps_ring_kernel(int N, const float *__restrict__ W,
                                                              const In *__restrict__ in,
                                                              float2 *__restrict__ out, int NN,
                                                              int Ns) {
  const int c = blockIdx.x * PPS_THREADS + threadIdx.x;
  if (c >= N) return;
  const int e = blockIdx.y;
  const int s0 = blockIdx.z * PPS_RUN;
  const int s1 = min(Ns, s0 + PPS_RUN);
  const long spec = (long)NN * N;
  const In *p = in + (long)s0 * spec + (long)e * N + c;
  float2 *o = out + (long)s0 * spec + (long)e * N + c;
  float w[T];
#pragma unroll
  for (int i = 0; i < T; i++) w[i] = W[(long)i * N + c];
  float2 xb[T];
#pragma unroll
  for (int i = 0; i < T - 1; i++) xb[i] = load_sample(p + (long)i * spec);
  for (int s = s0; s < s1; s++) {
    xb[T - 1] = load_sample(p + (long)(s - s0 + T - 1) * spec);
    float re = 0.0f, im = 0.0f;
#pragma unroll
    for (int i = 0; i < T; i++) {
      re = __builtin_fmaf(w[i], xb[i].x, re);
      im = __builtin_fmaf(w[i], xb[i].y, im);
    }
    store_stream(o + (long)(s - s0) * spec, make_float2(re, im));
#pragma unroll
    for (int i = 0; i < T - 1; i++) xb[i] = xb[i + 1];
  }
}

// any T in [4, 128]: taps re-read from L1/L2 per output
template <typename In>
__global__ __launch_bounds__(PPS_THREADS) void pps_generic_kernel(int N, int T,
                                                                 const float *__restrict__ W,
                                                                 const In *__restrict__ in,
                                                                 float2 *__restrict__ out, int NN,
                                                                 int Ns) {
  const int c = blockIdx.x * PPS_THREADS + threadIdx.x;
  if (c >= N) return;
  const int e = blockIdx.y;
  const int s0 = blockIdx.z * PPS_RUN;
  const int s1 = min(Ns, s0 + PPS_RUN);
  const long spec = (long)NN * N;
  for (int s = s0; s < s1; s++) {
    const In *p = in + (long)s * spec + (long)e * N + c;
    float re = 0.0f, im = 0.0f;
    for (int i = 0; i < T; i++) {
      float2 x = load_sample(p + (long)i * spec);
      float w = W[(long)i * N + c];
      re = __builtin_fmaf(w, x.x, re);
      im = __builtin_fmaf(w, x.y, im);
    }
    out[(long)s * spec + (long)e * N + c] = make_float2(re, im);
  }
}

template <typename In>
static int launch_pps(int N, int T, const float *W, const In *in, float2 *out, int NN, int Ns,
                      hipStream_t st) {
  dim3 grid((N + PPS_THREADS - 1) / PPS_THREADS, NN, (Ns + PPS_RUN - 1) / PPS_RUN);
  dim3 block(PPS_THREADS);
  switch (T) {
    case 4:
      hipLaunchKernelGGL((pps_ring_kernel<4, In>), grid, block, 0, st, N, W, in, out, NN, Ns);
      break;
    case 8:
      hipLaunchKernelGGL((pps_ring_kernel<8, In>), grid, block, 0, st, N, W, in, out, NN, Ns);
      break;
    case 16:
      hipLaunchKernelGGL((pps_ring_kernel<16, In>), grid, block, 0, st, N, W, in, out, NN, Ns);
      break;
    default:
      hipLaunchKernelGGL((pps_generic_kernel<In>), grid, block, 0, st, N, T, W, in, out, NN, Ns);
  }
  return check_launch("pps kernel");
}

static int check_pps_args(int N, int T, const void *W, const void *in, const void *out, int NN,
                          int Ns) {
  MDT_TRY(check_channels(N));
  MDT_TRY(check_taps(T));
  MDT_TRY(check_inputs(NN));
  MDT_TRY(check_spectra(Ns));
  MDT_REQUIRE_PTR(W);
  MDT_REQUIRE_PTR(in);
  MDT_REQUIRE_PTR(out);
  return MDT_SUCCESS;
}

}  // namespace mdt

using namespace mdt;

extern "C" {

int mdt_pps_batch(int Nchannels, int Ntaps, const float *Window, const mdt_complex *InSignal,
                  mdt_complex *PolyStruct, int NpolsxNelements, int Nspectra, hipStream_t stream) {
  MDT_TRY(check_pps_args(Nchannels, Ntaps, Window, InSignal, PolyStruct, NpolsxNelements, Nspectra));
  if (((uintptr_t)InSignal | (uintptr_t)PolyStruct) & 7)
    return fail(MDT_ERR_MISALIGNED, "complex buffers must be 8-byte aligned");
  return launch_pps(Nchannels, Ntaps, Window, (const float2 *)InSignal, (float2 *)PolyStruct,
                    NpolsxNelements, Nspectra, stream);
}

int mdt_pps_batch_i8(int Nchannels, int Ntaps, const float *Window, const int8_t *InSignal,
                     mdt_complex *PolyStruct, int NpolsxNelements, int Nspectra,
                     hipStream_t stream) {
  MDT_TRY(check_pps_args(Nchannels, Ntaps, Window, InSignal, PolyStruct, NpolsxNelements, Nspectra));
  if (((uintptr_t)InSignal & 1) || ((uintptr_t)PolyStruct & 7))
    return fail(MDT_ERR_MISALIGNED, "int8 I/Q input must be 2-byte aligned, output 8-byte");
  return launch_pps(Nchannels, Ntaps, Window, (const char2 *)InSignal, (float2 *)PolyStruct,
                    NpolsxNelements, Nspectra, stream);
}

}  // extern "C"
