// fft.hip -- a3: batched C2C FFT, drop-in for cufftPlan1d / cufftSetStream /
// cufftExecC2C (MDT_OneStation_client.cu:622, 630, 1105).  Unnormalised,
// natural order, forward sign -1, in place or out of place.
//
// One workgroup holds FPB transforms of N points (N/P threads each, P points
// per thread, see fft_core.hpp); rows are read and written straight from/to
// global memory by the first/last Stockham pass, 64-lane contiguous.  At
// N = 256 .. 8192 the workgroups are persistent and load their next rows
// while transforming the current ones (fft_stream_kernel).
#include <atomic>
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

// Streaming form: persistent workgroups, each taking the groups of FPB rows
// g = blockIdx.x, + gridDim.x, ...; group g + gridDim.x is loaded into
// registers while group g is transformed, so the HBM reads stay in flight
// through the passes instead of each workgroup's reads, transforms and
// writes following one another.  For that nothing inside the transform may
// wait for those loads: the exchanges use bare barriers (which wait for LDS
// operations only) and the twiddles come from the per-pass LDS table, filled
// once per workgroup from the same W_N^k table.  The butterflies, twiddle
// values and operation order are fft_kernel's, so the bits are too.
// The twiddle source of fft_stream_kernel<N>: the whole per-pass table in
// LDS where it fits beside the exchange buffers (N <= 4096); otherwise
// (8192 points: 100 KB of table beside 68 KB of exchange buffer) the table
// without the remainder pass, whose twiddles each thread keeps in registers
// (RemRegTw), loaded once per workgroup.
template <int N>
constexpr bool stream_rem_tw() {
  using C = FftCfg<N>;
  return (C::FPB * lds_entries<N>() + pass_tw_entries<N, C::P>()) * 8 > 160 * 1024;
}
template <int N, bool REM = stream_rem_tw<N>()>
struct StreamTw {
  using type = PassTw<FftCfg<N>::P>;
  static constexpr int ENTRIES = pass_tw_entries<N, FftCfg<N>::P>();
  __device__ static type make(const float2 *tab, const float2 *, int) { return type{tab}; }
};
template <int N>
struct StreamTw<N, true> {
  using type = RemRegTw<N, FftCfg<N>::P>;
  static constexpr int ENTRIES = type::ENTRIES;
  __device__ static type make(const float2 *tab, const float2 *tw, int t) {
    type r{tab, {}};
    r.load(tw, t);
    return r;
  }
};

template <int N, bool INV>
__global__ __launch_bounds__(FftCfg<N>::THREADS) void fft_stream_kernel(const float2 *in, float2 *out,
                                                                        long batch,
                                                                        const float2 *__restrict__ tw) {
  using C = FftCfg<N>;
  using ST = StreamTw<N>;
  constexpr int P = C::P, TPB = C::TPB;
  __shared__ float2 lds[C::FPB * lds_entries<N>()];
  __shared__ float2 twl[ST::ENTRIES];
  fill_pass_tw<N, P, ST::ENTRIES>(twl, tw, threadIdx.x, C::THREADS);
  // (register twiddles: loaded before the barrier, whose release fence waits
  // for them, so that no wait inside the loop counts them)
  const auto twv = ST::make(twl, tw, threadIdx.x % TPB);
  __syncthreads();
  const int t = threadIdx.x % TPB;
  const int f = threadIdx.x / TPB;
  const long ngroups = (batch + C::FPB - 1) / C::FPB;
  // Every step issues the same P loads and P stores, so that the wait for the
  // prefetched rows counts only the stores behind them: a row past the batch
  // (only in the last group, which also holds row batch - 1) reads row
  // batch - 1 and stores its identical transform there, and a prefetch past
  // the last group re-reads the current one (discarded).  Two named buffers,
  // unrolled by two, so that no register copy of the rows in flight (which
  // would wait for them) lands inside a transform.
  auto load = [&](float2 (&xv)[P], long grp) {
    grp = grp < ngroups ? grp : grp - gridDim.x;
    long row = grp * C::FPB + f;
    row = row < batch ? row : batch - 1;
    const float2 *src = in + row * N + t;
#pragma unroll
    for (int m = 0; m < P; m++) xv[m] = src[m * TPB];
  };
  auto step = [&](float2 (&xv)[P], long grp) {
    fft_block_tw<N, P, INV, true>(xv, t, lds + f * lds_entries<N>(), twv);
    long row = grp * C::FPB + f;
    row = row < batch ? row : batch - 1;
    float2 *dst = out + row * N + t;
#pragma unroll
    for (int m = 0; m < P; m++) store_stream(dst + m * TPB, xv[m]);
  };
  const long G = gridDim.x;
  float2 xa[P], xb[P];
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

// resident workgroups of fft_stream_kernel<N, INV> on the current device
// (CUs x occupancy), queried once per device (threads that race here store
// the same value); 0 if the query fails
template <int N, bool INV>
static long stream_grid() {
  static std::atomic<int> cached[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  int grid = cached[dev].load(std::memory_order_relaxed);
  if (grid == 0) {
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fft_stream_kernel<N, INV>, FftCfg<N>::THREADS, 0) !=
            hipSuccess ||
        cus <= 0 || per <= 0)
      return 0;
    grid = cus * per;
    cached[dev].store(grid, std::memory_order_relaxed);
  }
  return grid;
}

template <int N, bool INV>
static int launch_fft_stream(const float2 *in, float2 *out, long batch, const float2 *tw, long grid,
                             hipStream_t st) {
  using C = FftCfg<N>;
  hipLaunchKernelGGL((fft_stream_kernel<N, INV>), dim3((unsigned)grid), dim3(C::THREADS), 0, st, in, out, batch,
                     tw);
  return check_launch("fft_stream_kernel");
}

template <int N>
static int launch_fft(const float2 *in, float2 *out, long batch, const float2 *tw, bool inv,
                      hipStream_t st) {
  using C = FftCfg<N>;
  if constexpr (N >= 256 && N <= 8192) {
    const long groups = (batch + C::FPB - 1) / C::FPB;
    const long grid = inv ? stream_grid<N, true>() : stream_grid<N, false>();
    if (grid > 0) {
      const long g = grid < groups ? grid : groups;
      return inv ? launch_fft_stream<N, true>(in, out, batch, tw, g, st)
                 : launch_fft_stream<N, false>(in, out, batch, tw, g, st);
    }
  }
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
