// calib.hip -- gain/delay calibration on the visibilities (no counterpart in
// the reference; exercised at 64 inputs, 4 taps, 1024 channels, 512 spectra).
//
// Model: a per-input complex gain g_a and delay tau_a (samples) applied to the
// channelised signal, X'_a[c] = g_a phi_a(c) X_a[c], phi_a(c) =
// exp(-2 pi i k_c tau_a / N), k_c the signed FFT bin (c < N/2 ? c : c - N).
// With G_a(c) = g_a phi_a(c) the visibilities become V'_ab = G_a conj(G_b) V_ab
// and the loss is L = sum_{c, a >= b} |V'_ab - M_ab|^2.
//
// Adjoint (PyTorch convention grad = dL/dRe + i dL/dIm = 2 dL/dconj):
//   grad_g[z]   = 2 sum_c conj(phi_z) [ sum_{b<=z} R_zb G_b conj(V_zb)
//                                      + sum_{a>=z} conj(R_az) G_a V_az ]
//   grad_tau[z] = sum_c 2 f_c [ sum_{b<=z} Im(conj(R_zb) V'_zb)
//                             - sum_{a>=z} Im(conj(R_az) V'_az) ],  f_c = 2 pi k_c / N
// with R = V' - M.  All sums are FP32 with Neumaier compensation and a fixed
// reduction order (deterministic); the phasor argument k*tau is reduced modulo
// N exactly before sincospi.
#include <mutex>
#include <new>

#include "calib.hpp"
#include "common.hpp"

namespace mdt {

struct Kahan {
  float s = 0.f, c = 0.f;
  __device__ void add(float v) {
    const float t = s + v;
    c += (fabsf(s) >= fabsf(v)) ? ((s - t) + v) : ((v - t) + s);
    s = t;
  }
  __device__ float value() const { return s + c; }
};

__device__ __forceinline__ long bl_of(int a, int b) { return (long)a * (a + 1) / 2 + b; }

constexpr int CAL_THREADS = 256, CAL_TILE = 64;      // apply kernel threads, inputs per tile
constexpr int CAL_LG_THREADS = 1024, CAL_WAVES = 16;  // loss/adjoint kernel

// sum over the 64 lanes of a wave by an xor butterfly: every lane ends with
// the same bits (each step adds the same two values in either order)
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// G_a(c) = g_a phi_a(c) of every input a for channel c, into a wave's LDS row
__device__ __forceinline__ void wave_gains(float2 *Gw, int NN, int N, int c, const float2 *g, const float *tau,
                                           int lane) {
  const int kk = signed_bin(c, N);
  for (int a = lane; a < NN; a += 64) Gw[a] = cmul(g[a], delay_phasor(kk, tau ? tau[a] : 0.f, N));
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// V' = G_a conj(G_b) V for every (channel, baseline): one block per channel,
// the channel's gains in LDS, threads over its baselines (coalesced)
__global__ __launch_bounds__(CAL_THREADS) void calib_apply_kernel(int NN, int N, long nb, const float2 *g,
                                                                  const float *tau, const float2 *V, float2 *Vp) {
  extern __shared__ float2 Gs[];
  const int c = blockIdx.x;
  const int kk = signed_bin(c, N);
  for (int a = threadIdx.x; a < NN; a += CAL_THREADS) Gs[a] = cmul(g[a], delay_phasor(kk, tau ? tau[a] : 0.f, N));
  __syncthreads();
  for (long bl = threadIdx.x; bl < nb; bl += CAL_THREADS) {
    int a = (int)((sqrt(8.0 * (double)bl + 1.0) - 1.0) / 2.0);
    while ((long)a * (a + 1) / 2 > bl) a--;
    while ((long)(a + 1) * (a + 2) / 2 <= bl) a++;
    const int b = (int)(bl - (long)a * (a + 1) / 2);
    const long i = (long)c * nb + bl;
    Vp[i] = cmul(cmulc(Gs[a], Gs[b]), V[i]);
  }
}

// Loss and adjoint partials.  Block (channel chunk, input tile of 64):
// CAL_CH channels of the chunk in flight, CAL_CW waves on each (slot cs =
// w / CAL_CW, part pw = w % CAL_CW); lane l is input z = 64 tile + l.  Per
// channel, with the channel's gains G in the slot's LDS row and R = V' - M:
//   rows:    for each input zz of the tile with (zz - z0) % CAL_CW == pw, its
//            row of baselines (zz, p <= zz) (contiguous, lanes over p): |R|^2
//            into the loss, and R G_p conj(V), Im(conj(R) V') summed over the
//            row for input zz;
//   columns: for each input p >= the tile's first with (p - z0) % CAL_CW ==
//            pw, its row's baselines (p, z) with z in the tile (contiguous,
//            lane z): conj(R) G_p V and -Im(conj(R) V') for input z.
// Every baseline enters the loss once (rows partition the triangle) and the
// gradient of each of its two inputs once; the diagonal (z, z) enters both
// sums, as its two terms.  The slot's first wave adds its waves' sums for
// lane z in order, then 2 S conj(phi_z(c)) and 2 f_c T to its channel sums.
// Fixed order throughout: compensated per-lane sums, a butterfly over each
// row, the waves combined in order; partials per (input, chunk) and per
// block, summed in a fixed order by calib_finish_kernel.  Splitting each
// channel over four waves keeps four waves per SIMD in flight at the
// calibration shape (64 inputs, 1024 channels: one wave per channel left one
// wave per SIMD, latency-bound: 0.092 ms).
constexpr int CAL_CW = 4, CAL_CH = CAL_WAVES / CAL_CW;  // waves per channel, channels in flight
template <bool GRAD>
__global__ __launch_bounds__(CAL_LG_THREADS) void calib_lg_kernel(int NN, int N, long nb, int nchunk,
                                                               const float2 *g, const float *tau,
                                                               const float2 *V, const float2 *M,
                                                               float *gpart, float *lpart) {
  extern __shared__ float2 Gs[];  // [CAL_CH][NN]
  __shared__ float part[CAL_WAVES][3][64];
  __shared__ float red[CAL_WAVES][4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int cs = w / CAL_CW, pw = w % CAL_CW;
  const int z0 = blockIdx.y * CAL_TILE, z = z0 + lane;
  const int zend = min(z0 + CAL_TILE, NN);
  const int cpc = (N + nchunk - 1) / nchunk;
  const int c0 = blockIdx.x * cpc, c1 = min(N, c0 + cpc);
  float2 *Gw = Gs + (long)cs * NN;
  Kahan sgr, sgi, stt, sl;
  const int nround = (cpc + CAL_CH - 1) / CAL_CH;  // uniform over the block
  for (int k = 0; k < nround; k++) {
    const int c = c0 + cs + k * CAL_CH;
    const bool cv = c < c1;  // uniform over the slot's waves
    __syncthreads();  // the previous round's reads of the gains are done
    if (cv) {
      const int kk = signed_bin(c, N);
      for (int a = pw * 64 + lane; a < NN; a += CAL_CW * 64) Gw[a] = cmul(g[a], delay_phasor(kk, tau ? tau[a] : 0.f, N));
    }
    __syncthreads();
    Kahan ar, ai, at;  // this wave's part of lane z's sums
    if (cv) {
      const long cb = (long)c * nb;
      for (int zz = z0 + pw; zz < zend; zz += CAL_CW) {
        const float2 Gzz = Gw[zz];
        const long rb = cb + bl_of(zz, 0);
        Kahan rr, ri, rt;
        for (int p = lane; p <= zz; p += 64) {
          const float2 v = V[rb + p], m = M[rb + p];
          const float2 Gp = Gw[p];
          const float2 vp = cmul(cmulc(Gzz, Gp), v);
          const float2 r = make_float2(vp.x - m.x, vp.y - m.y);
          sl.add(__builtin_fmaf(r.x, r.x, r.y * r.y));
          if constexpr (GRAD) {
            const float2 t = cmulc(cmul(r, Gp), v);
            rr.add(t.x);
            ri.add(t.y);
            if (p != zz) rt.add(__builtin_fmaf(r.x, vp.y, -(r.y * vp.x)));
          }
        }
        if constexpr (GRAD) {
          const float sr = wave_sum(rr.value()), si = wave_sum(ri.value()), st = wave_sum(rt.value());
          if (lane == zz - z0) {
            ar.add(sr);
            ai.add(si);
            at.add(st);
          }
        }
      }
      if constexpr (GRAD) {
        const float2 Gz = z < NN ? Gw[z] : make_float2(0.f, 0.f);
        for (int p = z0 + pw; p < NN; p += CAL_CW) {
          if (z <= p) {
            const float2 Gp = Gw[p];
            const long i = cb + bl_of(p, z);
            const float2 v = V[i], m = M[i];
            const float2 vp = cmul(cmulc(Gp, Gz), v);
            const float2 r = make_float2(vp.x - m.x, vp.y - m.y);
            const float2 t = cmul(make_float2(r.x, -r.y), cmul(Gp, v));
            ar.add(t.x);
            ai.add(t.y);
            if (p != z) at.add(-__builtin_fmaf(r.x, vp.y, -(r.y * vp.x)));
          }
        }
      }
    }
    if constexpr (GRAD) {
      part[w][0][lane] = ar.value();
      part[w][1][lane] = ai.value();
      part[w][2][lane] = at.value();
      __syncthreads();
      if (cv && pw == 0 && z < NN) {
        Kahan sa, sb, st;
        for (int q = 0; q < CAL_CW; q++) {
          sa.add(part[w + q][0][lane]);
          sb.add(part[w + q][1][lane]);
          st.add(part[w + q][2][lane]);
        }
        const int kk = signed_bin(c, N);
        const float f = 6.283185307179586f * (float)kk / (float)N;
        const float2 ph = delay_phasor(kk, tau ? tau[z] : 0.f, N);
        const float2 t = cmulc(make_float2(sa.value(), sb.value()), ph);  // S conj(phi)
        sgr.add(2.0f * t.x);
        sgi.add(2.0f * t.y);
        stt.add(2.0f * f * st.value());
      }
    }
  }
  red[w][0][lane] = sgr.value();
  red[w][1][lane] = sgi.value();
  red[w][2][lane] = stt.value();
  red[w][3][lane] = sl.value();
  __syncthreads();
  if (w == 0) {
    Kahan q[4];
    for (int ww = 0; ww < CAL_WAVES; ww++) {
      if (ww % CAL_CW == 0)  // the slots' channel sums
        for (int k = 0; k < 3; k++) q[k].add(red[ww][k][lane]);
      q[3].add(red[ww][3][lane]);  // every wave's loss terms
    }
    if (GRAD && z < NN) {
      float *o = gpart + ((long)z * nchunk + blockIdx.x) * 3;
      o[0] = q[0].value();
      o[1] = q[1].value();
      o[2] = q[2].value();
    }
    const float ls = wave_sum(q[3].value());
    if (lane == 0) lpart[(long)blockIdx.x * gridDim.y + blockIdx.y] = ls;
  }
}

// Final sums in a fixed order (the same bits on every run).  Gradients: 16
// threads per input, thread `sub` summing chunks sub, sub + 16, ... (all its
// loads issued before the first add), then the input's first thread adds the
// 16 partial sums in order.  Loss: every thread sums a strided share of the
// block partials, and the 1024 sums are combined in two ordered levels.
constexpr int CAL_FT = 1024, CAL_FSUB = 16, CAL_FLD = 16;  // threads, threads per input, loads per thread
__global__ __launch_bounds__(CAL_FT) void calib_finish_kernel(int NN, int nchunk, int nlossblk, const float *gpart,
                                                             const float *lpart, float *loss, float2 *grad_g,
                                                             float *grad_tau) {
  __shared__ float sh[3][CAL_FT];
  const int tid = threadIdx.x, sub = tid % CAL_FSUB;
  if (grad_g) {
    for (int z0 = 0; z0 < NN; z0 += CAL_FT / CAL_FSUB) {  // uniform
      const int z = z0 + tid / CAL_FSUB;
      const int zc = z < NN ? z : NN - 1;
      float va[CAL_FLD], vb[CAL_FLD], vt[CAL_FLD];
#pragma unroll
      for (int i = 0; i < CAL_FLD; i++) {
        const int k = sub + CAL_FSUB * i;
        const float *o = gpart + ((long)zc * nchunk + (k < nchunk ? k : nchunk - 1)) * 3;  // in range
        va[i] = o[0];
        vb[i] = o[1];
        vt[i] = o[2];
      }
      Kahan a, b, t;
#pragma unroll
      for (int i = 0; i < CAL_FLD; i++) {
        if (sub + CAL_FSUB * i < nchunk) {
          a.add(va[i]);
          b.add(vb[i]);
          t.add(vt[i]);
        }
      }
      sh[0][tid] = a.value();
      sh[1][tid] = b.value();
      sh[2][tid] = t.value();
      __syncthreads();
      if (sub == 0 && z < NN) {
        Kahan fa, fb, ft;
        for (int j = 0; j < CAL_FSUB; j++) {
          fa.add(sh[0][tid + j]);
          fb.add(sh[1][tid + j]);
          ft.add(sh[2][tid + j]);
        }
        grad_g[z] = make_float2(fa.value(), fb.value());
        if (grad_tau) grad_tau[z] = ft.value();
      }
      __syncthreads();
    }
  }
  if (loss) {
    Kahan l;
    for (int k0 = 0; k0 < nlossblk; k0 += CAL_FT * CAL_FLD) {  // uniform
      float v[CAL_FLD];
#pragma unroll
      for (int i = 0; i < CAL_FLD; i++) {
        const int k = k0 + tid + CAL_FT * i;
        v[i] = lpart[k < nlossblk ? k : nlossblk - 1];
      }
#pragma unroll
      for (int i = 0; i < CAL_FLD; i++)
        if (k0 + tid + CAL_FT * i < nlossblk) l.add(v[i]);
    }
    sh[0][tid] = l.value();
    __syncthreads();
    if (tid < 32) {
      Kahan m;
      for (int j = 0; j < 32; j++) m.add(sh[0][tid * 32 + j]);
      sh[1][tid] = m.value();
    }
    __syncthreads();
    if (tid == 0) {
      Kahan m;
      for (int j = 0; j < 32; j++) m.add(sh[1][j]);
      *loss = m.value();
    }
  }
}

// launch geometry of the loss/adjoint for NN inputs and N channels
struct CalGeom {
  int nchunk, ntile;
  long nb;
  size_t lds;
};
constexpr int CAL_MAX_CHUNKS = 256;
static_assert(CAL_MAX_CHUNKS <= CAL_FSUB * CAL_FLD, "the finish kernel's loads cover every chunk");
__host__ inline CalGeom cal_geom(int NN, int N) {
  CalGeom q;
  q.nchunk = (N + 3) / 4 < CAL_MAX_CHUNKS ? (N + 3) / 4 : CAL_MAX_CHUNKS;
  q.ntile = (NN + CAL_TILE - 1) / CAL_TILE;
  q.nb = (long)NN * (NN + 1) / 2;
  q.lds = sizeof(float2) * (size_t)CAL_CH * NN;
  return q;
}
// workspace in floats: gradient partials [NN][nchunk][3], then loss partials [nchunk][ntile]
__host__ inline size_t cal_scratch_floats(int NN, int N) {
  const CalGeom q = cal_geom(NN, N);
  return 3 * (size_t)NN * q.nchunk + (size_t)q.nchunk * q.ntile;
}

// the calibration loss / adjoint on a given workspace (no allocation)
int calib_loss_grad_run(int NN, int N, const float2 *V, const float2 *M, const float2 *g, const float *tau,
                        float *loss, float2 *grad_g, float *grad_tau, float *work, hipStream_t st) {
  const CalGeom q = cal_geom(NN, N);
  float *gpart = work, *lpart = work + 3 * (size_t)NN * q.nchunk;
  static bool attr_set = false;  // LDS above 64 KiB for the largest arrays
  if (!attr_set) {
    if (hipFuncSetAttribute((const void *)calib_lg_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(float2) * CAL_CH * MDT_MAX_INPUTS)) != hipSuccess ||
        hipFuncSetAttribute((const void *)calib_lg_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(float2) * CAL_CH * MDT_MAX_INPUTS)) != hipSuccess)
      return fail(MDT_ERR_DEVICE, "LDS size of the calibration kernels");
    attr_set = true;
  }
  const dim3 grid(q.nchunk, q.ntile);
  if (grad_g)
    hipLaunchKernelGGL(calib_lg_kernel<true>, grid, dim3(CAL_LG_THREADS), q.lds, st, NN, N, q.nb, q.nchunk, g, tau, V, M,
                       gpart, lpart);
  else
    hipLaunchKernelGGL(calib_lg_kernel<false>, grid, dim3(CAL_LG_THREADS), q.lds, st, NN, N, q.nb, q.nchunk, g, tau, V,
                       M, gpart, lpart);
  MDT_TRY(check_launch("calib_lg_kernel"));
  hipLaunchKernelGGL(calib_finish_kernel, dim3(1), dim3(CAL_FT), 0, st, NN, q.nchunk, q.nchunk * q.ntile, gpart, lpart,
                     loss, grad_g, grad_tau);
  return check_launch("calib_finish_kernel");
}

}  // namespace mdt

// Workspace of the allocation-free entry point (mdt_calib_loss_grad_noalloc):
// sized for the largest shape the limits allow (4096 inputs, 256 channel
// chunks, 64 input tiles), 3 162 112 floats = 12.06 MiB per device, held for
// the life of the process.  Host code never uses the array's name as an
// address: each device's copy is reached through the pointer that
// hipGetSymbolAddress returns with that device current.
constexpr size_t kCalScratchFloats =
    3 * (size_t)MDT_MAX_INPUTS * mdt::CAL_MAX_CHUNKS + (size_t)mdt::CAL_MAX_CHUNKS * (MDT_MAX_INPUTS / mdt::CAL_TILE);
static __device__ __align__(256) float scratch[kCalScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch[kMaxDevices] = {};  // per device: the address of its scratch

int device_scratch(float **ptr) {
  int dev = 0;
  MDT_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return mdt::fail(MDT_ERR_DEVICE, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev] == nullptr) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return mdt::fail(MDT_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch[dev] = (float *)p;
  }
  *ptr = g_scratch[dev];
  return MDT_SUCCESS;
}
}  // namespace

struct mdt_calib_plan_s {
  int NN = 0, N = 0;
  hipStream_t stream = 0;
  float *work = nullptr;  // loss / adjoint partials (cal_scratch_floats)
};

using namespace mdt;

static int check_calib_args(const mdt_complex *V, const mdt_complex *M, const mdt_complex *g, float *loss,
                            mdt_complex *grad_g, float *grad_tau) {
  MDT_REQUIRE_PTR(V);
  MDT_REQUIRE_PTR(M);
  MDT_REQUIRE_PTR(g);
  if (!loss && !grad_g) return fail(MDT_ERR_NULL_POINTER, "nothing to compute (loss and grad_g NULL)");
  if (grad_tau && !grad_g) return fail(MDT_ERR_NULL_POINTER, "grad_tau needs grad_g");
  return MDT_SUCCESS;
}

extern "C" {

int mdt_calib_plan_create(mdt_calib_plan *plan, int NpolsxNelements, int Nchannels) {
  MDT_REQUIRE_PTR(plan);
  *plan = nullptr;
  MDT_TRY(check_inputs(NpolsxNelements));
  MDT_TRY(check_channels(Nchannels));
  mdt_calib_plan p = new (std::nothrow) mdt_calib_plan_s;
  if (!p) return fail(MDT_ERR_ALLOC, "calibration plan");
  p->NN = NpolsxNelements;
  p->N = Nchannels;
  if (hipMalloc(&p->work, sizeof(float) * cal_scratch_floats(p->NN, p->N)) != hipSuccess) {
    delete p;
    return fail(MDT_ERR_ALLOC, "calibration workspace");
  }
  *plan = p;
  return MDT_SUCCESS;
}

int mdt_calib_set_stream(mdt_calib_plan plan, hipStream_t stream) {
  MDT_REQUIRE_PTR(plan);
  plan->stream = stream;
  return MDT_SUCCESS;
}

int mdt_calib_apply(mdt_calib_plan plan, const mdt_complex *V, const mdt_complex *g, const float *tau,
                    mdt_complex *Vp) {
  MDT_REQUIRE_PTR(plan);
  MDT_REQUIRE_PTR(V);
  MDT_REQUIRE_PTR(g);
  MDT_REQUIRE_PTR(Vp);
  const long nb = (long)plan->NN * (plan->NN + 1) / 2;
  hipLaunchKernelGGL(calib_apply_kernel, dim3(plan->N), dim3(CAL_THREADS), sizeof(float2) * plan->NN, plan->stream,
                     plan->NN, plan->N, nb, (const float2 *)g, tau, (const float2 *)V, (float2 *)Vp);
  return check_launch("calib_apply_kernel");
}

int mdt_calib_loss_grad(mdt_calib_plan plan, const mdt_complex *V, const mdt_complex *M,
                        const mdt_complex *g, const float *tau, float *loss, mdt_complex *grad_g,
                        float *grad_tau) {
  MDT_REQUIRE_PTR(plan);
  MDT_TRY(check_calib_args(V, M, g, loss, grad_g, grad_tau));
  return calib_loss_grad_run(plan->NN, plan->N, (const float2 *)V, (const float2 *)M, (const float2 *)g, tau, loss,
                             (float2 *)grad_g, grad_tau, plan->work, plan->stream);
}

int mdt_calib_loss_grad_noalloc(int NpolsxNelements, int Nchannels, const mdt_complex *V, const mdt_complex *M,
                                const mdt_complex *g, const float *tau, float *loss, mdt_complex *grad_g,
                                float *grad_tau, hipStream_t stream) {
  MDT_TRY(check_inputs(NpolsxNelements));
  MDT_TRY(check_channels(Nchannels));
  MDT_TRY(check_calib_args(V, M, g, loss, grad_g, grad_tau));
  const size_t need = cal_scratch_floats(NpolsxNelements, Nchannels);
  if (need > kCalScratchFloats)
    return fail(MDT_ERR_INVALID_VALUE, "%d inputs x %d channels need %zu scratch floats, the maximum is %zu",
                NpolsxNelements, Nchannels, need, kCalScratchFloats);
  float *work = nullptr;
  MDT_TRY(device_scratch(&work));
  return calib_loss_grad_run(NpolsxNelements, Nchannels, (const float2 *)V, (const float2 *)M, (const float2 *)g, tau,
                             loss, (float2 *)grad_g, grad_tau, work, stream);
}

int mdt_calib_plan_destroy(mdt_calib_plan plan) {
  if (!plan) return MDT_SUCCESS;
  if (plan->work) (void)hipFree(plan->work);
  delete plan;
  return MDT_SUCCESS;
}

}  // extern "C"
