// fft.hip -- a3: batched C2C FFT, drop-in for cufftPlan1d / cufftSetStream /
// cufftExecC2C (MDT_OneStation_client.cu:622, 630, 1105).  Unnormalised,
// natural order, forward sign -1, in place or out of place.
//
// One workgroup holds FPB transforms of N points (N/P threads each, P points
// per thread, see fft_core.hpp); rows are read and written straight from/to
// global memory by the first/last Stockham pass, 64-lane contiguous.
#include <new>

#include "fft_any.hpp"
#include "fft_core.hpp"

namespace mdt {

template <int N, bool INV>
__global__ __launch_bounds__(FftCfg<N>::THREADS) void fft_kernel(const float2 *in, float2 *out,
                                                                 long batch,
                                                                 const float2 *__restrict__ tw) {
  using C = FftCfg<N>;
  __shared__ float2 lds[C::FPB * lds_entries<N>()];
  const int t = threadIdx.x % C::TPB;
  const int f = threadIdx.x / C::TPB;
  const long row = (long)blockIdx.x * C::FPB + f;
  const bool valid = row < batch;
  float2 x[C::P];
  const float2 *src = in + row * N + t;
#pragma unroll
  for (int m = 0; m < C::P; m++) x[m] = valid ? src[m * C::TPB] : make_float2(0.f, 0.f);
  fft_block<N, C::P, INV>(x, t, lds + f * lds_entries<N>(), tw);
  if (valid) {
    float2 *dst = out + row * N + t;
#pragma unroll
    for (int m = 0; m < C::P; m++) store_stream(dst + m * C::TPB, x[m]);
  }
}

template <int N>
static int launch_fft(const float2 *in, float2 *out, long batch, const float2 *tw, bool inv,
                      hipStream_t st) {
  using C = FftCfg<N>;
  long blocks = (batch + C::FPB - 1) / C::FPB;
  MDT_TRY(check_grid(blocks, C::THREADS, "FFT batch"));
  if (inv)
    hipLaunchKernelGGL((fft_kernel<N, true>), dim3((unsigned)blocks), dim3(C::THREADS), 0, st, in,
                       out, batch, tw);
  else
    hipLaunchKernelGGL((fft_kernel<N, false>), dim3((unsigned)blocks), dim3(C::THREADS), 0, st, in,
                       out, batch, tw);
  return check_launch("fft_kernel");
}

int fft_dispatch(int N, const float2 *in, float2 *out, long batch, const float2 *tw, bool inv,
                 hipStream_t st) {
  switch (N) {
    case 8: return launch_fft<8>(in, out, batch, tw, inv, st);
    case 16: return launch_fft<16>(in, out, batch, tw, inv, st);
    case 32: return launch_fft<32>(in, out, batch, tw, inv, st);
    case 64: return launch_fft<64>(in, out, batch, tw, inv, st);
    case 128: return launch_fft<128>(in, out, batch, tw, inv, st);
    case 256: return launch_fft<256>(in, out, batch, tw, inv, st);
    case 512: return launch_fft<512>(in, out, batch, tw, inv, st);
    case 1024: return launch_fft<1024>(in, out, batch, tw, inv, st);
    case 2048: return launch_fft<2048>(in, out, batch, tw, inv, st);
    case 4096: return launch_fft<4096>(in, out, batch, tw, inv, st);
    case 8192: return launch_fft<8192>(in, out, batch, tw, inv, st);
    case 16384: return launch_fft<16384>(in, out, batch, tw, inv, st);
    default: return fail(MDT_ERR_UNSUPPORTED, "FFT length %d is not a power of two in [8, 16384]", N);
  }
}

}  // namespace mdt

struct mdt_fft_plan_s {
  int nx = 0;
  long batch = 0;
  hipStream_t stream = 0;
  float2 *tw = nullptr;  // powers of two: W_nx^k
  mdt::FftAny any;       // other lengths: mixed radix or Bluestein tables
};

using namespace mdt;

extern "C" {

int mdt_fft_plan_1d(mdt_fft_plan *plan, int nx, int type, int batch) {
  MDT_REQUIRE_PTR(plan);
  *plan = nullptr;
  if (type != MDT_FFT_C2C) return fail(MDT_ERR_UNSUPPORTED, "only C2C transforms are provided");
  if (batch < 1) return fail(MDT_ERR_INVALID_VALUE, "batch=%d must be >= 1", batch);
  const bool pow2 = is_pow2(nx) && nx >= 8 && nx <= 16384;
  if (!pow2 && (nx < MDT_MIN_CHANNELS || nx > MDT_MAX_CHANNELS))
    return fail(MDT_ERR_UNSUPPORTED, "FFT length %d: powers of two in [8, 16384], other lengths in [%d, %d]", nx,
                MDT_MIN_CHANNELS, MDT_MAX_CHANNELS);
  mdt_fft_plan p = new (std::nothrow) mdt_fft_plan_s;
  if (!p) return fail(MDT_ERR_ALLOC, "plan");
  p->nx = nx;
  p->batch = batch;
  int st = pow2 ? make_twiddles(nx, &p->tw) : fft_any_create(nx, &p->any);
  if (st != MDT_SUCCESS) {
    fft_any_destroy(&p->any);
    delete p;
    return st;
  }
  *plan = p;
  return MDT_SUCCESS;
}

int mdt_fft_set_stream(mdt_fft_plan plan, hipStream_t stream) {
  MDT_REQUIRE_PTR(plan);
  plan->stream = stream;
  return MDT_SUCCESS;
}

int mdt_fft_exec_c2c(mdt_fft_plan plan, mdt_complex *idata, mdt_complex *odata, int direction) {
  MDT_REQUIRE_PTR(plan);
  MDT_REQUIRE_PTR(idata);
  MDT_REQUIRE_PTR(odata);
  if (direction != MDT_FFT_FORWARD && direction != MDT_FFT_INVERSE)
    return fail(MDT_ERR_INVALID_VALUE, "direction must be -1 (forward) or +1 (inverse)");
  if (((uintptr_t)idata | (uintptr_t)odata) & 7)
    return fail(MDT_ERR_MISALIGNED, "complex buffers must be 8-byte aligned");
  if (plan->any.kind != 0)
    return fft_any_exec(plan->any, (const float2 *)idata, (float2 *)odata, plan->batch, direction == MDT_FFT_INVERSE,
                        plan->stream);
  return fft_dispatch(plan->nx, (const float2 *)idata, (float2 *)odata, plan->batch, plan->tw,
                      direction == MDT_FFT_INVERSE, plan->stream);
}

int mdt_fft_destroy(mdt_fft_plan plan) {
  if (!plan) return MDT_SUCCESS;
  if (plan->tw) (void)hipFree(plan->tw);
  fft_any_destroy(&plan->any);
  delete plan;
  return MDT_SUCCESS;
}

}  // extern "C"
