/* client_dump.c -- the per-dump calls of the MDT client's rank-2 loop written
 * in C against libmdtfx: set-up as in MDT_OneStation_client.cu:584-705
 * (window, FFT plan, X-engine handle, pointer tables, alpha = 1/Ns), then for
 * each dump the history carry (client.cu:1170) and PpS_Batch ->
 * cufftExecC2C -> ReorderFOutput -> cublasCgemmBatched (client.cu:1083-1128)
 * on the float F input, which the client builds on the host from the wire
 * samples (client.cu:898-907).  The same samples also go through the fused
 * native path; the program exits 0 when the two agree to 1e-5 (largest
 * |V_drop-in - V_native| / sqrt(V_aa V_bb) over channels and baselines).
 *
 * Shape: the deployed station (38 inputs = 19 elements x 2 pols, 8-tap PFB,
 * 1024 channels), 256 spectra per dump, 3 dumps of synthetic int8 samples. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "mdtfx.h"

#define CHECK(x)                                                                                   \
  do {                                                                                             \
    int s_ = (x);                                                                                  \
    if (s_) {                                                                                      \
      fprintf(stderr, "%s:%d: %s -> %d (%s)\n", __FILE__, __LINE__, #x, s_, mdt_last_error_string()); \
      return 1;                                                                                    \
    }                                                                                              \
  } while (0)
#define HIPCHECK(x)                                                                                \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) {                                                                        \
      fprintf(stderr, "%s:%d: %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_));           \
      return 1;                                                                                    \
    }                                                                                              \
  } while (0)

enum { NN = 38, N = 1024, T = 8, NS = 256, DUMPS = 3 };

/* int8 (I, Q) frames -> the client's float F input rows */
static void to_float(const int8_t *x8, mdt_complex *f, size_t n) {
  for (size_t i = 0; i < n; i++) {
    f[i].x = (float)x8[2 * i];
    f[i].y = (float)x8[2 * i + 1];
  }
}

int main(void) {
  const size_t frame = (size_t)NN * N;           /* samples per spectrum (all inputs) */
  const size_t nframes = (size_t)DUMPS * NS + T - 1;
  const size_t nb = (size_t)NN * (NN + 1) / 2;
  hipStream_t st;
  HIPCHECK(hipStreamCreate(&st));

  /* the wire samples of all dumps, int8 (I, Q) [frames][NN][N][2] */
  int8_t *d_x8, *h_x8;
  HIPCHECK(hipMalloc((void **)&d_x8, nframes * frame * 2));
  mdt_synth_params sp;
  memset(&sp, 0, sizeof sp);
  sp.seed = 38;
  sp.sigma = 16.0f;
  sp.rho = 0.3f;
  sp.max_delay = 4;
  CHECK(mdt_synth_i8(&sp, NN, N, (int)nframes, d_x8, st));
  h_x8 = (int8_t *)malloc(nframes * frame * 2);
  if (!h_x8) return 1;
  HIPCHECK(hipMemcpyAsync(h_x8, d_x8, nframes * frame * 2, hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));

  /* client set-up (client.cu:584-705) */
  float *d_W;
  mdt_complex *d_FInput, *d_FOutput, *d_XInput, *d_XOutput, *h_F, *h_X;
  HIPCHECK(hipMalloc((void **)&d_W, sizeof(float) * T * N));
  HIPCHECK(hipMalloc((void **)&d_FInput, sizeof(mdt_complex) * (NS + T - 1) * frame));
  HIPCHECK(hipMalloc((void **)&d_FOutput, sizeof(mdt_complex) * NS * frame));
  HIPCHECK(hipMalloc((void **)&d_XInput, sizeof(mdt_complex) * NS * frame));
  HIPCHECK(hipMalloc((void **)&d_XOutput, sizeof(mdt_complex) * (size_t)N * NN * NN));
  h_F = (mdt_complex *)malloc(sizeof(mdt_complex) * (NS + T - 1) * frame);
  h_X = (mdt_complex *)malloc(sizeof(mdt_complex) * (size_t)N * NN * NN);
  if (!h_F || !h_X) return 1;
  CHECK(mdt_pfb_window(N, T, d_W, MDT_WINDOW_BH4, st));
  mdt_fft_plan fft;
  CHECK(mdt_fft_plan_1d(&fft, N, MDT_FFT_C2C, NS * NN));
  CHECK(mdt_fft_set_stream(fft, st));
  mdt_xengine xe;
  CHECK(mdt_xengine_create(&xe));
  CHECK(mdt_xengine_set_stream(xe, st));
  /* pointer tables, one matrix per channel (client.cu:656-677) */
  const mdt_complex **h_A = (const mdt_complex **)malloc(sizeof(void *) * N);
  mdt_complex **h_C = (mdt_complex **)malloc(sizeof(void *) * N);
  const mdt_complex **d_A;
  mdt_complex **d_C;
  if (!h_A || !h_C) return 1;
  for (int c = 0; c < N; c++) {
    h_A[c] = d_XInput + (size_t)c * NN * NS;
    h_C[c] = d_XOutput + (size_t)c * NN * NN;
  }
  HIPCHECK(hipMalloc((void **)&d_A, sizeof(void *) * N));
  HIPCHECK(hipMalloc((void **)&d_C, sizeof(void *) * N));
  HIPCHECK(hipMemcpy(d_A, h_A, sizeof(void *) * N, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_C, h_C, sizeof(void *) * N, hipMemcpyHostToDevice));
  const mdt_complex alpha = {1.0f / NS, 0.0f}, beta = {0.0f, 0.0f}; /* client.cu:683-686 */

  /* the native path on the same samples */
  mdt_fx_plan p;
  CHECK(mdt_fx_plan_create(&p, NN, N, T, NS, MDT_WINDOW_BH4, MDT_XPREC_FP32));
  CHECK(mdt_fx_set_stream(p, st));
  size_t xbytes;
  CHECK(mdt_fx_xinput_bytes(p, &xbytes));
  void *d_xin;
  mdt_complex *d_vis, *h_vis;
  HIPCHECK(hipMalloc(&d_xin, xbytes));
  HIPCHECK(hipMalloc((void **)&d_vis, sizeof(mdt_complex) * N * nb));
  h_vis = (mdt_complex *)malloc(sizeof(mdt_complex) * N * nb);
  if (!h_vis) return 1;

  double worst = 0.0;
  for (int d = 0; d < DUMPS; d++) {
    if (d == 0) { /* first block: history and stack from the wire */
      to_float(h_x8, h_F, (NS + T - 1) * frame);
      HIPCHECK(hipMemcpyAsync(d_FInput, h_F, sizeof(mdt_complex) * (NS + T - 1) * frame,
                              hipMemcpyHostToDevice, st));
    } else { /* history carry, then the new stack behind it */
      CHECK(mdt_history_shift(d_FInput, NN, N, T, NS, (int)sizeof(mdt_complex), st));
      to_float(h_x8 + ((size_t)d * NS + T - 1) * frame * 2, h_F, NS * frame);
      HIPCHECK(hipMemcpyAsync(d_FInput + (size_t)(T - 1) * frame, h_F, sizeof(mdt_complex) * NS * frame,
                              hipMemcpyHostToDevice, st));
    }
    CHECK(mdt_pps_batch(N, T, d_W, d_FInput, d_FOutput, NN, NS, st));
    CHECK(mdt_fft_exec_c2c(fft, d_FOutput, d_FOutput, MDT_FFT_FORWARD));
    CHECK(mdt_reorder_foutput(NN, N, d_FOutput, d_XInput, NS, st));
    CHECK(mdt_cgemm_batched(xe, MDT_OP_N, MDT_OP_C, NN, NN, NS, &alpha, d_A, NN, d_A, NN, &beta, d_C, NN, N));
    CHECK(mdt_fx_dump(p, d_x8 + (size_t)d * NS * frame * 2, (mdt_complex *)d_xin, d_vis, NULL));
    HIPCHECK(hipMemcpyAsync(h_X, d_XOutput, sizeof(mdt_complex) * (size_t)N * NN * NN, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(h_vis, d_vis, sizeof(mdt_complex) * N * nb, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    double err = 0.0;
    for (int c = 0; c < N; c++) {
      const mdt_complex *X = h_X + (size_t)c * NN * NN, *v = h_vis + (size_t)c * nb;
      for (int a = 0; a < NN; a++)
        for (int b = 0; b <= a; b++) {
          const mdt_complex x = X[a + b * NN], y = v[a * (a + 1) / 2 + b];
          const double norm = sqrt((double)X[a + a * NN].x * (double)X[b + b * NN].x);
          const double e = hypot((double)x.x - y.x, (double)x.y - y.y) / norm;
          if (!(e <= err)) err = e; /* a NaN counts as the worst */
        }
    }
    printf("dump %d: drop-in vs native max normalised difference %.3e\n", d, err);
    if (!(err <= worst)) worst = err;
  }

  CHECK(mdt_fx_plan_destroy(p));
  CHECK(mdt_fft_destroy(fft));
  CHECK(mdt_xengine_destroy(xe));
  HIPCHECK(hipFree(d_xin));
  HIPCHECK(hipFree(d_vis));
  HIPCHECK(hipFree(d_A));
  HIPCHECK(hipFree(d_C));
  HIPCHECK(hipFree(d_W));
  HIPCHECK(hipFree(d_FInput));
  HIPCHECK(hipFree(d_FOutput));
  HIPCHECK(hipFree(d_XInput));
  HIPCHECK(hipFree(d_XOutput));
  HIPCHECK(hipFree(d_x8));
  HIPCHECK(hipStreamDestroy(st));
  free(h_x8);
  free(h_F);
  free(h_X);
  free(h_vis);
  free(h_A);
  free(h_C);
  const int ok = worst < 1e-5;
  printf("%s: %d dumps, max normalised difference %.3e\n", ok ? "ok" : "FAILED", DUMPS, worst);
  return ok ? 0 : 2;
}
