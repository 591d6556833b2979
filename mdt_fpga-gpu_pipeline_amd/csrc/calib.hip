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

// G[a][c] = g_a phi_a(c)
__global__ void calib_gains_kernel(int NN, int N, const float2 *g, const float *tau, float2 *G) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)NN * N) return;
  const int a = (int)(i / N), c = (int)(i % N);
  const float2 ph = delay_phasor(signed_bin(c, N), tau ? tau[a] : 0.f, N);
  G[i] = cmul(g[a], ph);
}

__global__ void calib_apply_kernel(int NN, int N, long nb, const float2 *G, const float2 *V,
                                   float2 *Vp) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)N * nb) return;
  const int c = (int)(i / nb);
  const long bl = i % nb;
  int a = (int)((sqrt(8.0 * (double)bl + 1.0) - 1.0) / 2.0);
  while ((long)a * (a + 1) / 2 > bl) a--;
  while ((long)(a + 1) * (a + 2) / 2 <= bl) a++;
  const int b = (int)(bl - (long)a * (a + 1) / 2);
  const float2 ga = G[(long)a * N + c], gb = G[(long)b * N + c];
  Vp[i] = cmul(cmulc(ga, gb), V[i]);
}

// loss partials: one block per chunk of channels, threads over baselines
constexpr int CAL_THREADS = 256;

__device__ float block_sum(float v, float *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = CAL_THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const float r = sh[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(CAL_THREADS) void calib_loss_kernel(int NN, int N, long nb, int cpb,
                                                                 const float2 *G, const float2 *V,
                                                                 const float2 *M, float *partial) {
  __shared__ float sh[CAL_THREADS];
  Kahan acc;
  const int c0 = blockIdx.x * cpb;
  for (int c = c0; c < min(N, c0 + cpb); c++) {
    for (int a = 0; a < NN; a++) {
      const float2 ga = G[(long)a * N + c];
      for (int b = threadIdx.x; b <= a; b += CAL_THREADS) {
        const long i = (long)c * nb + bl_of(a, b);
        const float2 vp = cmul(cmulc(ga, G[(long)b * N + c]), V[i]);
        const float rr = vp.x - M[i].x, ri = vp.y - M[i].y;
        acc.add(__builtin_fmaf(rr, rr, ri * ri));
      }
    }
  }
  const float s = block_sum(acc.value(), sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// gradient partials: block (channel chunk, z); one thread per channel,
// looping over the partner input p; fixed-order tree reduction over the chunk
__global__ __launch_bounds__(CAL_THREADS) void calib_grad_kernel(int NN, int N, long nb, int nchunk,
                                                                 const float2 *G, const float *tau,
                                                                 const float2 *V, const float2 *M,
                                                                 float *partial) {
  __shared__ float sh[CAL_THREADS];
  const int z = blockIdx.y;
  const int c = blockIdx.x * CAL_THREADS + threadIdx.x;
  float cr = 0.f, ci = 0.f, ct = 0.f;
  if (c < N) {
    const float2 Gz = G[(long)z * N + c];
    const int kk = signed_bin(c, N);
    const float f = 6.283185307179586f * (float)kk / (float)N;
    Kahan sr, si, st;
    for (int p = 0; p < NN; p++) {
      const float2 Gp = G[(long)p * N + c];
      if (p <= z) {  // baseline (a = z, b = p): term R G_p conj(V)
        const long i = (long)c * nb + bl_of(z, p);
        const float2 v = V[i], m = M[i];
        const float2 vp = cmul(cmulc(Gz, Gp), v);
        const float2 r = make_float2(vp.x - m.x, vp.y - m.y);
        const float2 t = cmulc(cmul(r, Gp), v);
        sr.add(t.x);
        si.add(t.y);
        if (p != z) st.add(__builtin_fmaf(r.x, vp.y, -(r.y * vp.x)));  // Im(conj(R) V')
      }
      if (p >= z) {  // baseline (a = p, b = z): term conj(R) G_p V
        const long i = (long)c * nb + bl_of(p, z);
        const float2 v = V[i], m = M[i];
        const float2 vp = cmul(cmulc(Gp, Gz), v);
        const float2 r = make_float2(vp.x - m.x, vp.y - m.y);
        const float2 t = cmul(make_float2(r.x, -r.y), cmul(Gp, v));
        sr.add(t.x);
        si.add(t.y);
        if (p != z) st.add(-__builtin_fmaf(r.x, vp.y, -(r.y * vp.x)));
      }
    }
    // grad_g[z] gets 2 conj(phi_z(c)) S(c); tau gradient 2 f (sum)
    const float2 ph = delay_phasor(kk, tau ? tau[z] : 0.f, N);
    const float2 S = make_float2(sr.value(), si.value());
    const float2 t = cmulc(S, ph);  // S conj(phi)
    cr = 2.0f * t.x;
    ci = 2.0f * t.y;
    ct = 2.0f * f * st.value();
  }
  const float r0 = block_sum(cr, sh);
  const float r1 = block_sum(ci, sh);
  const float r2 = block_sum(ct, sh);
  if (threadIdx.x == 0) {
    float *o = partial + ((long)z * nchunk + blockIdx.x) * 3;
    o[0] = r0;
    o[1] = r1;
    o[2] = r2;
  }
}

// final fixed-order sums: thread z over its chunk partials; thread 0 the loss
__global__ void calib_finish_kernel(int NN, int nchunk, int nlossblk, const float *gpart,
                                    const float *lpart, float *loss, float2 *grad_g,
                                    float *grad_tau) {
  for (int z = threadIdx.x; z < NN; z += blockDim.x) {
    if (grad_g) {
      Kahan a, b, t;
      for (int k = 0; k < nchunk; k++) {
        const float *o = gpart + ((long)z * nchunk + k) * 3;
        a.add(o[0]);
        b.add(o[1]);
        t.add(o[2]);
      }
      grad_g[z] = make_float2(a.value(), b.value());
      if (grad_tau) grad_tau[z] = t.value();
    }
  }
  if (threadIdx.x == 0 && loss) {
    Kahan l;
    for (int k = 0; k < nlossblk; k++) l.add(lpart[k]);
    *loss = l.value();
  }
}

}  // namespace mdt

struct mdt_calib_plan_s {
  int NN = 0, N = 0;
  hipStream_t stream = 0;
  float2 *G = nullptr;    // [NN][N] combined gains
  float *gpart = nullptr; // [NN][nchunk][3]
  float *lpart = nullptr; // [nlossblk]
  int nchunk = 0, nlossblk = 0, cpb = 0;
};

using namespace mdt;

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
  p->nchunk = (Nchannels + CAL_THREADS - 1) / CAL_THREADS;
  p->cpb = 4;
  p->nlossblk = (Nchannels + p->cpb - 1) / p->cpb;
  bool ok = hipMalloc(&p->G, sizeof(float2) * (size_t)p->NN * p->N) == hipSuccess &&
            hipMalloc(&p->gpart, sizeof(float) * 3 * (size_t)p->NN * p->nchunk) == hipSuccess &&
            hipMalloc(&p->lpart, sizeof(float) * (size_t)p->nlossblk) == hipSuccess;
  if (!ok) {
    mdt_calib_plan_destroy(p);
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

static int make_G(mdt_calib_plan p, const mdt_complex *g, const float *tau) {
  const long n = (long)p->NN * p->N;
  hipLaunchKernelGGL(calib_gains_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, p->NN,
                     p->N, (const float2 *)g, tau, p->G);
  return check_launch("calib_gains_kernel");
}

int mdt_calib_apply(mdt_calib_plan plan, const mdt_complex *V, const mdt_complex *g, const float *tau,
                    mdt_complex *Vp) {
  MDT_REQUIRE_PTR(plan);
  MDT_REQUIRE_PTR(V);
  MDT_REQUIRE_PTR(g);
  MDT_REQUIRE_PTR(Vp);
  MDT_TRY(make_G(plan, g, tau));
  const long nb = (long)plan->NN * (plan->NN + 1) / 2;
  const long n = nb * plan->N;
  hipLaunchKernelGGL(calib_apply_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, plan->stream,
                     plan->NN, plan->N, nb, plan->G, (const float2 *)V, (float2 *)Vp);
  return check_launch("calib_apply_kernel");
}

int mdt_calib_loss_grad(mdt_calib_plan plan, const mdt_complex *V, const mdt_complex *M,
                        const mdt_complex *g, const float *tau, float *loss, mdt_complex *grad_g,
                        float *grad_tau) {
  MDT_REQUIRE_PTR(plan);
  MDT_REQUIRE_PTR(V);
  MDT_REQUIRE_PTR(M);
  MDT_REQUIRE_PTR(g);
  if (!loss && !grad_g) return fail(MDT_ERR_NULL_POINTER, "nothing to compute (loss and grad_g NULL)");
  if (grad_tau && !grad_g) return fail(MDT_ERR_NULL_POINTER, "grad_tau needs grad_g");
  MDT_TRY(make_G(plan, g, tau));
  const long nb = (long)plan->NN * (plan->NN + 1) / 2;
  if (loss) {
    hipLaunchKernelGGL(calib_loss_kernel, dim3(plan->nlossblk), dim3(CAL_THREADS), 0, plan->stream,
                       plan->NN, plan->N, nb, plan->cpb, plan->G, (const float2 *)V,
                       (const float2 *)M, plan->lpart);
    MDT_TRY(check_launch("calib_loss_kernel"));
  }
  if (grad_g) {
    hipLaunchKernelGGL(calib_grad_kernel, dim3(plan->nchunk, plan->NN), dim3(CAL_THREADS), 0,
                       plan->stream, plan->NN, plan->N, nb, plan->nchunk, plan->G, tau,
                       (const float2 *)V, (const float2 *)M, plan->gpart);
    MDT_TRY(check_launch("calib_grad_kernel"));
  }
  hipLaunchKernelGGL(calib_finish_kernel, dim3(1), dim3(256), 0, plan->stream, plan->NN, plan->nchunk,
                     plan->nlossblk, plan->gpart, plan->lpart, loss, (float2 *)grad_g, grad_tau);
  return check_launch("calib_finish_kernel");
}

int mdt_calib_plan_destroy(mdt_calib_plan plan) {
  if (!plan) return MDT_SUCCESS;
  if (plan->G) (void)hipFree(plan->G);
  if (plan->gpart) (void)hipFree(plan->gpart);
  if (plan->lpart) (void)hipFree(plan->lpart);
  delete plan;
  return MDT_SUCCESS;
}

}  // extern "C"
