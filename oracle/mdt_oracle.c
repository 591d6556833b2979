/*
 * mdt_oracle.c -- CPU restatement of the MDT FX correlator chain.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libmdtfx, never the product.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Built by oracle/Makefile with gcc -O3 -ffp-contract=off so that
 * every fused multiply-add is an explicit fmaf()/fma() call.
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   orc_window_*      closed-form special values (centre tap = 1, sinc zeros
 *                     at integer arguments, symmetry), numpy golden vectors
 *   orc_pps_*         impulse -> window column, constant -> tap sum,
 *                     linearity, FP64 vs FP32 agreement
 *   orc_fft_* / dft   brute-force DFT at small N, numpy.fft (pocketfft)
 *                     golden vectors (powers of two and 1000, 1536, 3000,
 *                     4099), Parseval, tones, shift theorem
 *   orc_reorder       index formula, inverse permutation, raw bit patterns
 *   orc_xengine_*     Hermitian symmetry, Cauchy-Schwarz, identical inputs,
 *                     numpy einsum golden vectors; the complex128 branch
 *                     (orc_xengine_full_f64) against the same golden vectors
 *                     and against the complex64 branch on identical values
 *   orc_unpack_u32    exhaustive 2^32 comparison of the literal reference
 *                     expression against the two's-complement closed form
 *   orc_depacketise   hand-built packets with known destinations
 *   orc_calib_*       FP64 central finite differences, closed-form cases
 *   orc_fx_chain_*    numpy FP64 FIR -> numpy.fft -> einsum golden chain
 * The test-side e4m3 quantiser (tests/e4m3.py) is pinned against torch's
 * float8_e4m3fn conversion.
 */
#include "mdt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------- sinpi/cospi
 * Exact argument reduction: sinpi(n) == 0 and cospi(n + 1/2) == 0 exactly, so
 * the sinc zeros of the window (devfn.cu:70-75) come out as exact zeros. */
double orc_sinpi(double x) {
  double n = nearbyint(x);
  double r = x - n; /* exact, |r| <= 0.5 */
  double s;
  if (r == 0.0) return 0.0;
  if (fabs(r) <= 0.25)
    s = sin(M_PI * r);
  else
    s = copysign(cos(M_PI * (0.5 - fabs(r))), r);
  return (fmod(fabs(n), 2.0) == 1.0) ? -s : s;
}

double orc_cospi(double x) {
  double n = nearbyint(x);
  double r = fabs(x - n);
  double c;
  if (r <= 0.25)
    c = cos(M_PI * r);
  else
    c = sin(M_PI * (0.5 - r));
  return (fmod(fabs(n), 2.0) == 1.0) ? -c : c;
}

/* ---------------------------------------------------------------- window
 * devfn.cu:52-88 (PFB_Window_min4term_BH): W[i] = sinc((i - TN/2)/N) * w(i),
 * w = 0.35875 - 0.48829 cos(2 pi i/TN) + 0.14128 cos(4 pi i/TN)
 *            - 0.01168 cos(6 pi i/TN); sinc(0) := 1 (devfn.cu:72-75).
 * The Hamming variant replaces w by 0.54 - 0.46 cos(2 pi i/TN). */
void orc_window_f64(int N, int T, int kind, double *W) {
  long TN = (long)T * N;
  for (long i = 0; i < TN; i++) {
    double x = ((double)i - 0.5 * (double)TN) / (double)N; /* devfn.cu:70 */
    double sinc = (x == 0.0) ? 1.0 : orc_sinpi(x) / (M_PI * x);
    double u = 2.0 * (double)i / (double)TN; /* devfn.cu:81 */
    double w;
    if (kind == ORC_WINDOW_HAMMING)
      w = 0.54 - 0.46 * orc_cospi(u);
    else
      w = 0.35875 - 0.48829 * orc_cospi(u) + 0.14128 * orc_cospi(2.0 * u) -
          0.01168 * orc_cospi(3.0 * u); /* devfn.cu:82 */
    W[i] = sinc * w;
  }
}

void orc_window_f32(int N, int T, int kind, float *W) {
  long TN = (long)T * N;
  double *tmp = (double *)malloc(sizeof(double) * TN);
  orc_window_f64(N, T, kind, tmp);
  for (long i = 0; i < TN; i++) W[i] = (float)tmp[i];
  free(tmp);
}

/* ---------------------------------------------------------------- PpS_Batch
 * devfn.cu:113-161: out[s][e][c] = sum_{i=0}^{T-1} W[i*N + c] * in[s+i][e][c],
 * real and imaginary parts separately, fmaf in tap order i = 0..T-1
 * (devfn.cu:143-152).  Input row stride between spectra is NN*N
 * (devfn.cu:129: stride_spec = gridDim.z * Nchannels). */
void orc_pps_f32(int N, int T, const float *W, const float *in, float *out, int NN, int Ns) {
  long spec = (long)NN * N;
#pragma omp parallel for collapse(2) schedule(static)
  for (int s = 0; s < Ns; s++)
    for (int e = 0; e < NN; e++)
      for (int c = 0; c < N; c++) {
        float re = 0.0f, im = 0.0f;
        for (int i = 0; i < T; i++) {
          long idx = (long)(s + i) * spec + (long)e * N + c;
          float w = W[(long)i * N + c];
          re = fmaf(w, in[2 * idx], re);
          im = fmaf(w, in[2 * idx + 1], im);
        }
        long o = (long)s * spec + (long)e * N + c;
        out[2 * o] = re;
        out[2 * o + 1] = im;
      }
}

/* same FIR with int8 I/Q input; int8 -> float is exact */
void orc_pps_i8_f32(int N, int T, const float *W, const int8_t *in, float *out, int NN, int Ns) {
  long spec = (long)NN * N;
#pragma omp parallel for collapse(2) schedule(static)
  for (int s = 0; s < Ns; s++)
    for (int e = 0; e < NN; e++)
      for (int c = 0; c < N; c++) {
        float re = 0.0f, im = 0.0f;
        for (int i = 0; i < T; i++) {
          long idx = (long)(s + i) * spec + (long)e * N + c;
          float w = W[(long)i * N + c];
          re = fmaf(w, (float)in[2 * idx], re);
          im = fmaf(w, (float)in[2 * idx + 1], im);
        }
        long o = (long)s * spec + (long)e * N + c;
        out[2 * o] = re;
        out[2 * o + 1] = im;
      }
}

void orc_pps_f64(int N, int T, const double *W, const double *in, double *out, int NN, int Ns) {
  long spec = (long)NN * N;
#pragma omp parallel for collapse(2) schedule(static)
  for (int s = 0; s < Ns; s++)
    for (int e = 0; e < NN; e++)
      for (int c = 0; c < N; c++) {
        double re = 0.0, im = 0.0;
        for (int i = 0; i < T; i++) {
          long idx = (long)(s + i) * spec + (long)e * N + c;
          double w = W[(long)i * N + c];
          re = fma(w, in[2 * idx], re);
          im = fma(w, in[2 * idx + 1], im);
        }
        long o = (long)s * spec + (long)e * N + c;
        out[2 * o] = re;
        out[2 * o + 1] = im;
      }
}

/* ---------------------------------------------------------------- FFT
 * cufftPlan1d(N, CUFFT_C2C, batch) + cufftExecC2C (client.cu:622, 1105):
 * X[k] = sum_n x[n] exp(sign * 2 pi i n k / N), unnormalised, natural order,
 * sign = -1 forward.  Batch rows are contiguous N-point vectors. */
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

void orc_dft_f64(int N, long batch, const double *in, double *out, int sign) {
  double *cs = (double *)malloc(sizeof(double) * 2 * N);
  for (int m = 0; m < N; m++) {
    cs[2 * m] = orc_cospi(2.0 * m / N);
    cs[2 * m + 1] = (double)sign * orc_sinpi(2.0 * m / N);
  }
#pragma omp parallel for schedule(static)
  for (long b = 0; b < batch; b++) {
    const double *x = in + 2 * b * N;
    double *X = out + 2 * b * N;
    for (int k = 0; k < N; k++) {
      double re = 0.0, im = 0.0;
      for (int n = 0; n < N; n++) {
        long m = ((long)n * k) % N; /* exact angle reduction */
        double c = cs[2 * m], s = cs[2 * m + 1];
        re += x[2 * n] * c - x[2 * n + 1] * s;
        im += x[2 * n] * s + x[2 * n + 1] * c;
      }
      X[2 * k] = re;
      X[2 * k + 1] = im;
    }
  }
  free(cs);
}

static void bitrev_perm(int N, int *rev) {
  int lg = 0;
  while ((1 << lg) < N) lg++;
  for (int i = 0; i < N; i++) {
    int r = 0;
    for (int b = 0; b < lg; b++)
      if (i & (1 << b)) r |= 1 << (lg - 1 - b);
    rev[i] = r;
  }
}

/* iterative radix-2 decimation-in-time on one row (double) */
static void fft_row_f64(int N, double *a, const double *tw, const int *rev) {
  for (int i = 0; i < N; i++) {
    int j = rev[i];
    if (j > i) {
      double t0 = a[2 * i], t1 = a[2 * i + 1];
      a[2 * i] = a[2 * j];
      a[2 * i + 1] = a[2 * j + 1];
      a[2 * j] = t0;
      a[2 * j + 1] = t1;
    }
  }
  for (int len = 2; len <= N; len <<= 1) {
    int half = len >> 1, step = N / len;
    for (int i = 0; i < N; i += len)
      for (int k = 0; k < half; k++) {
        double wr = tw[2 * k * step], wi = tw[2 * k * step + 1];
        double *u = a + 2 * (i + k), *v = a + 2 * (i + k + half);
        double vr = v[0] * wr - v[1] * wi, vi = v[0] * wi + v[1] * wr;
        v[0] = u[0] - vr;
        v[1] = u[1] - vi;
        u[0] += vr;
        u[1] += vi;
      }
  }
}

static void fft_row_f32(int N, float *a, const float *tw, const int *rev) {
  for (int i = 0; i < N; i++) {
    int j = rev[i];
    if (j > i) {
      float t0 = a[2 * i], t1 = a[2 * i + 1];
      a[2 * i] = a[2 * j];
      a[2 * i + 1] = a[2 * j + 1];
      a[2 * j] = t0;
      a[2 * j + 1] = t1;
    }
  }
  for (int len = 2; len <= N; len <<= 1) {
    int half = len >> 1, step = N / len;
    for (int i = 0; i < N; i += len)
      for (int k = 0; k < half; k++) {
        float wr = tw[2 * k * step], wi = tw[2 * k * step + 1];
        float *u = a + 2 * (i + k), *v = a + 2 * (i + k + half);
        float vr = v[0] * wr - v[1] * wi, vi = v[0] * wi + v[1] * wr;
        v[0] = u[0] - vr;
        v[1] = u[1] - vi;
        u[0] += vr;
        u[1] += vi;
      }
  }
}

static double *make_tw_f64(int N, int sign) {
  double *tw = (double *)malloc(sizeof(double) * N); /* N/2 complex */
  for (int k = 0; k < N / 2; k++) {
    tw[2 * k] = orc_cospi(2.0 * k / N);
    tw[2 * k + 1] = (double)sign * orc_sinpi(2.0 * k / N);
  }
  return tw;
}

void orc_fft_f64(int N, long batch, const double *in, double *out, int sign) {
  if (!is_pow2(N)) {
    orc_dft_f64(N, batch, in, out, sign);
    return;
  }
  double *tw = make_tw_f64(N, sign);
  int *rev = (int *)malloc(sizeof(int) * N);
  bitrev_perm(N, rev);
  if (out != in) memcpy(out, in, sizeof(double) * 2 * N * batch);
#pragma omp parallel for schedule(static)
  for (long b = 0; b < batch; b++) fft_row_f64(N, out + 2 * b * N, tw, rev);
  free(tw);
  free(rev);
}

void orc_fft_f32(int N, long batch, const float *in, float *out, int sign) {
  double *twd = make_tw_f64(N, sign);
  float *tw = (float *)malloc(sizeof(float) * N);
  for (int k = 0; k < N; k++) tw[k] = (float)twd[k]; /* FP64 twiddles rounded once */
  int *rev = (int *)malloc(sizeof(int) * N);
  bitrev_perm(N, rev);
  if (out != in) memcpy(out, in, sizeof(float) * 2 * N * batch);
#pragma omp parallel for schedule(static)
  for (long b = 0; b < batch; b++) fft_row_f32(N, out + 2 * b * N, tw, rev);
  free(twd);
  free(tw);
  free(rev);
}

/* ---------------------------------------------------------------- corner turn
 * devfn.cu:183-208: XIn[(c*Ns + s)*NN + e] = FOut[(s*NN + e)*N + c]
 * (read index devfn.cu:197, write index devfn.cu:207 with gridDim.z = Ns). */
void orc_reorder(int NN, int N, int Ns, int elem_bytes, const void *FOut, void *XIn) {
  const char *src = (const char *)FOut;
  char *dst = (char *)XIn;
#pragma omp parallel for schedule(static)
  for (int s = 0; s < Ns; s++)
    for (int e = 0; e < NN; e++)
      for (int c = 0; c < N; c++) {
        long from = ((long)s * NN + e) * N + c;
        long to = ((long)c * Ns + s) * NN + e;
        memcpy(dst + to * elem_bytes, src + from * elem_bytes, (size_t)elem_bytes);
      }
}

/* ---------------------------------------------------------------- X-engine
 * cublasCgemmBatched(h, OP_N, OP_C, NN, NN, Ns, alpha, A[], NN, A[], NN, beta=0,
 * C[], NN, N) with A[c] = XIn + c*NN*Ns, C[c] = XOut + c*NN*NN
 * (client.cu:656-660, 683-686, 1128):
 *   C_c[a + b*NN] = alpha * sum_s A_c[a + s*NN] * conj(A_c[b + s*NN]).
 * Plain double accumulation: |error| <= Ns * 2^-53 * sum|terms|, below
 * 1e-11 relative for Ns <= 16384. */
void orc_xengine_full_f64(int NN, int Ns, int N, double alpha, const double *XIn, double *XOut) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < N; c++) {
    const double *A = XIn + 2L * c * NN * Ns;
    double *C = XOut + 2L * c * NN * NN;
    for (int b = 0; b < NN; b++)
      for (int a = 0; a < NN; a++) {
        double re = 0.0, im = 0.0;
        for (int s = 0; s < Ns; s++) {
          const double *xa = A + 2L * ((long)s * NN + a);
          const double *xb = A + 2L * ((long)s * NN + b);
          re += xa[0] * xb[0] + xa[1] * xb[1];
          im += xa[1] * xb[0] - xa[0] * xb[1];
        }
        C[2L * (a + (long)b * NN)] = alpha * re;
        C[2L * (a + (long)b * NN) + 1] = alpha * im;
      }
  }
}

void orc_xengine_full_from_f32(int NN, int Ns, int N, double alpha, const float *XIn, double *XOut) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < N; c++) {
    const float *A = XIn + 2L * c * NN * Ns;
    double *C = XOut + 2L * c * NN * NN;
    for (int b = 0; b < NN; b++)
      for (int a = 0; a < NN; a++) {
        double re = 0.0, im = 0.0;
        for (int s = 0; s < Ns; s++) {
          const float *xa = A + 2L * ((long)s * NN + a);
          const float *xb = A + 2L * ((long)s * NN + b);
          re += (double)xa[0] * xb[0] + (double)xa[1] * xb[1];
          im += (double)xa[1] * xb[0] - (double)xa[0] * xb[1];
        }
        C[2L * (a + (long)b * NN)] = alpha * re;
        C[2L * (a + (long)b * NN) + 1] = alpha * im;
      }
  }
}

/* one batch entry of cublasCgemm(OP_N, OP_C): C = alpha A B^H + beta C0,
 * column-major, A m x k (lda), B n x k (ldb), C m x n (ldc); complex alpha/beta */
void orc_cgemm_nc_f64(int m, int n, int k, const double *alpha, const float *A, int lda,
                      const float *B, int ldb, const double *beta, const float *C0, int ldc,
                      double *C) {
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) {
      double re = 0.0, im = 0.0;
      for (int p = 0; p < k; p++) {
        const float *a = A + 2L * (i + (long)p * lda);
        const float *b = B + 2L * (j + (long)p * ldb);
        re += (double)a[0] * b[0] + (double)a[1] * b[1];
        im += (double)a[1] * b[0] - (double)a[0] * b[1];
      }
      long o = 2L * (i + (long)j * ldc);
      double cr = C0 ? C0[o] : 0.0, ci = C0 ? C0[o + 1] : 0.0;
      C[o] = alpha[0] * re - alpha[1] * im + beta[0] * cr - beta[1] * ci;
      C[o + 1] = alpha[0] * im + alpha[1] * re + beta[0] * ci + beta[1] * cr;
    }
}

/* ---------------------------------------------------------------- unpack
 * client.cu:898-907, transcribed literally: the constant 4294967295 is a
 * (64-bit) long, so 4294967295 - w + 1 is 2^32 - w, and the (int) conversion
 * of 2^31 wraps to INT_MIN; the word 0x80000000 therefore decodes to +4096.0.
 * The product is formed in double and stored to float. */
float orc_unpack_u32(uint32_t w) {
  if (w >= 2147483648u) {
    long t = 4294967295L - (long)w + 1L;
    int s = (int)(uint32_t)t; /* GCC: modulo 2^32 conversion */
    return (float)(-1.9073486328125e-06 * s);
  }
  return (float)(1.9073486328125e-06 * (int)w);
}

void orc_unpack_u32_array(const uint32_t *w, long n, float *out) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) out[i] = orc_unpack_u32(w[i]);
}

/* Closed form, independent of the literal expression above: the word is the
 * two's-complement int32 v scaled by 2^-19, except that v = INT_MIN maps to
 * +2^31 * 2^-19.  Returns the number of words (of all 2^32) on which the two
 * disagree bit-for-bit. */
int orc_count_unpack_mismatch_all(const float *unused, long *first_bad) {
  (void)unused;
  long bad = 0, first = -1;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (long hi = 0; hi < 65536; hi++) {
    for (long lo = 0; lo < 65536; lo++) {
      uint32_t w = (uint32_t)((hi << 16) | lo);
      int32_t v = (int32_t)w;
      double mag = (v == INT32_MIN) ? 2147483648.0 : (double)v;
      float ref = (float)ldexp(mag, -19);
      float got = orc_unpack_u32(w);
      uint32_t a, b;
      memcpy(&a, &ref, 4);
      memcpy(&b, &got, 4);
      if (a != b) {
        bad++;
#pragma omp critical
        {
          if (first < 0 || (long)w < first) first = (long)w;
        }
      }
    }
  }
  if (first_bad) *first_bad = first;
  return (int)(bad > 2147483647L ? 2147483647L : bad);
}

/* ---------------------------------------------------------------- depacketiser
 * client.cu:846-911 (first block: history samples go to stack 1) and
 * client.cu:945-1005 (steady state: stack = (sample / (Ns*N)) % 2).
 * Packet: word 0 station index, word 1 SNAP index, word 2 packet counter,
 * word 3 unused, then SperElement x InputsPerSNAP[snap] (re, im) word pairs
 * (client.cu:888-891).  Element offset = station*38 + sum of inputs of lower
 * SNAP indices + j (client.cu:877-885; client.h:64-68).  stacks holds
 * 2 * Ns * NN * N complex floats; nsamples[0..1] counts samples per stack.
 * Where the reference would write outside its stacks (SNAP >= 4, elements
 * beyond NN, a payload longer than the packet, a history longer than a stack,
 * a steady-state index before the first sample) the sample is skipped and
 * counted in nsamples[2], the rule the device kernel implements. */
static const int kInputsPerSnap[4] = {10, 10, 10, 8};

void orc_depacketise(const uint32_t *pkts, int npkts, int words_per_pkt, int SperElement,
                     int NN, int N, int Ns, int T, int first_block, float *stacks,
                     long *nsamples) {
  long begins = (long)(T - 1) * N;
  long per_stack = (long)Ns * N;
  for (int p = 0; p < npkts; p++) {
    const uint32_t *buf = pkts + (long)p * words_per_pkt;
    uint32_t station = buf[0], snap = buf[1], cnt = buf[2];
    int hdr = snap < 4 && station < 65536u;
    long elem0 = 0;
    int nin = 0;
    if (hdr) {
      for (uint32_t k = 0; k < snap; k++) elem0 += kInputsPerSnap[k];
      elem0 += (long)station * 38;
      nin = kInputsPerSnap[snap];
      hdr = elem0 + nin <= NN && 4 + 2L * SperElement * nin <= words_per_pkt;
    }
    for (int i = 0; i < SperElement; i++) {
      long sample, stack, spec;
      int ok = hdr;
      if (first_block) {
        sample = (long)cnt * SperElement + i;
        if (sample < begins) {
          stack = 1;
        } else {
          sample -= begins;
          stack = 0;
        }
        spec = sample / N;
        ok = ok && spec < Ns;
      } else {
        sample = (long)cnt * SperElement - begins + i;
        ok = ok && sample >= 0;
        stack = ok ? (sample / per_stack) % 2 : 0;
        spec = ok ? (sample / N) % Ns : 0;
      }
      if (!ok) {
        nsamples[2]++;
        continue;
      }
      long stride = (stack * Ns + spec) * (long)NN * N + elem0 * N + sample % N;
      for (int j = 0; j < nin; j++) {
        long sb = 4 + 2L * ((long)i * nin + j);
        long dst = stride + (long)j * N;
        stacks[2 * dst] = orc_unpack_u32(buf[sb]);
        stacks[2 * dst + 1] = orc_unpack_u32(buf[sb + 1]);
      }
      nsamples[stack]++;
    }
  }
}

/* ---------------------------------------------------------------- chains */
static inline long bl_index(int a, int b) { return (long)a * (a + 1) / 2 + b; }

/* FP64 chain from int8 input [Ns+T-1][NN][N][2] to triangle visibilities for
 * the listed channels: vis[ci][bl] complex, alpha = 1/Ns (client.cu:683). */
void orc_fx_chain_f64(int NN, int N, int T, int Ns, const double *W, const int8_t *in,
                      const int *chans, int nchan, double *vis) {
  long spec = (long)NN * N;
  double *x = (double *)calloc((size_t)2 * nchan * Ns * NN, sizeof(double)); /* [ci][s][e] */
  double *tw = make_tw_f64(N, -1);
  int *rev = (int *)malloc(sizeof(int) * N);
  if (is_pow2(N)) bitrev_perm(N, rev);
#pragma omp parallel
  {
    double *y = (double *)malloc(sizeof(double) * 2 * N);
    double *Y = (double *)malloc(sizeof(double) * 2 * N);
#pragma omp for collapse(2) schedule(static)
    for (int s = 0; s < Ns; s++)
      for (int e = 0; e < NN; e++) {
        for (int c = 0; c < N; c++) {
          double re = 0.0, im = 0.0;
          for (int i = 0; i < T; i++) {
            long idx = (long)(s + i) * spec + (long)e * N + c;
            re = fma(W[(long)i * N + c], (double)in[2 * idx], re);
            im = fma(W[(long)i * N + c], (double)in[2 * idx + 1], im);
          }
          y[2 * c] = re;
          y[2 * c + 1] = im;
        }
        if (is_pow2(N)) {
          fft_row_f64(N, y, tw, rev);
          memcpy(Y, y, sizeof(double) * 2 * N);
        } else {
          orc_dft_f64(N, 1, y, Y, -1);
        }
        for (int ci = 0; ci < nchan; ci++) {
          long o = ((long)ci * Ns + s) * NN + e;
          x[2 * o] = Y[2 * chans[ci]];
          x[2 * o + 1] = Y[2 * chans[ci] + 1];
        }
      }
    free(y);
    free(Y);
  }
  long nb = (long)NN * (NN + 1) / 2;
  double alpha = 1.0 / Ns;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int ci = 0; ci < nchan; ci++)
    for (int a = 0; a < NN; a++) {
      const double *A = x + 2L * ci * Ns * NN;
      for (int b = 0; b <= a; b++) {
        double re = 0.0, im = 0.0;
        for (int s = 0; s < Ns; s++) {
          const double *xa = A + 2L * ((long)s * NN + a);
          const double *xb = A + 2L * ((long)s * NN + b);
          re += xa[0] * xb[0] + xa[1] * xb[1];
          im += xa[1] * xb[0] - xa[0] * xb[1];
        }
        long o = (long)ci * nb + bl_index(a, b);
        vis[2 * o] = alpha * re;
        vis[2 * o + 1] = alpha * im;
      }
    }
  free(x);
  free(tw);
  free(rev);
}

/* FP32 restatement of the whole chain (the CPU baseline "port"): fmaf FIR in
 * tap order, radix-2 FP32 FFT with FP64-rounded twiddles, corner turn,
 * FP32-accumulated triangle correlation scaled by 1/Ns.  vis: [N][nb] complex. */
void orc_fx_chain_f32(int NN, int N, int T, int Ns, const float *W, const int8_t *in,
                      float *vis, int nthreads) {
#ifdef _OPENMP
  int prev = omp_get_max_threads();
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  long spec = (long)NN * N;
  float *F = (float *)malloc(sizeof(float) * 2 * Ns * spec);
  float *X = (float *)malloc(sizeof(float) * 2 * Ns * spec);
  orc_pps_i8_f32(N, T, W, in, F, NN, Ns);
  orc_fft_f32(N, (long)Ns * NN, F, F, -1);
  orc_reorder(NN, N, Ns, 8, F, X);
  long nb = (long)NN * (NN + 1) / 2;
  float alpha = 1.0f / (float)Ns;
#pragma omp parallel
  {
    float *acc = (float *)malloc(sizeof(float) * 2 * nb);
#pragma omp for schedule(dynamic, 1)
    for (int c = 0; c < N; c++) {
      memset(acc, 0, sizeof(float) * 2 * nb);
      const float *A = X + 2L * c * Ns * NN;
      for (int s = 0; s < Ns; s++) {
        const float *xs = A + 2L * s * NN;
        for (int a = 0; a < NN; a++) {
          float ar = xs[2 * a], ai = xs[2 * a + 1];
          float *row = acc + 2 * bl_index(a, 0);
          for (int b = 0; b <= a; b++) {
            float br = xs[2 * b], bi = xs[2 * b + 1];
            row[2 * b] += ar * br + ai * bi;
            row[2 * b + 1] += ai * br - ar * bi;
          }
        }
      }
      float *o = vis + 2L * c * nb;
      for (long q = 0; q < 2 * nb; q++) o[q] = alpha * acc[q];
    }
    free(acc);
  }
  free(F);
  free(X);
#ifdef _OPENMP
  omp_set_num_threads(prev);
#endif
}

/* ---------------------------------------------------------------- calibration
 * No reference counterpart.  Per-input complex gain g_a and delay tau_a (in
 * samples) applied to the channelised signal, X'_a[c] = g_a e^{-2 pi i k_c
 * tau_a / N} X_a[c], with k_c the signed FFT bin (c < N/2 ? c : c - N).  Then
 * V'_ab = g_a conj(g_b) e^{-2 pi i k_c (tau_a - tau_b)/N} V_ab and
 * L = sum_{c, a>=b} |V'_ab - M_ab|^2.  grad_g = dL/dRe g + i dL/dIm g. */
void orc_calib_apply_f64(int NN, int N, const double *V, const double *g, const double *tau,
                         double *Vp) {
  long nb = (long)NN * (NN + 1) / 2;
  for (int c = 0; c < N; c++) {
    double k = (c < N / 2) ? c : c - N;
    for (int a = 0; a < NN; a++)
      for (int b = 0; b <= a; b++) {
        long o = (long)c * nb + bl_index(a, b);
        double ph = -2.0 * k * (tau[a] - tau[b]) / N;
        double pr = orc_cospi(ph), pi = orc_sinpi(ph);
        double gr = g[2 * a] * g[2 * b] + g[2 * a + 1] * g[2 * b + 1];
        double gi = g[2 * a + 1] * g[2 * b] - g[2 * a] * g[2 * b + 1];
        double fr = gr * pr - gi * pi, fi = gr * pi + gi * pr;
        Vp[2 * o] = fr * V[2 * o] - fi * V[2 * o + 1];
        Vp[2 * o + 1] = fr * V[2 * o + 1] + fi * V[2 * o];
      }
  }
}

double orc_calib_loss_grad_f64(int NN, int N, const double *V, const double *M, const double *g,
                               const double *tau, double *grad_g, double *grad_tau) {
  long nb = (long)NN * (NN + 1) / 2;
  double L = 0.0;
  if (grad_g) memset(grad_g, 0, sizeof(double) * 2 * NN);
  if (grad_tau) memset(grad_tau, 0, sizeof(double) * NN);
  for (int c = 0; c < N; c++) {
    double k = (c < N / 2) ? c : c - N;
    for (int a = 0; a < NN; a++)
      for (int b = 0; b <= a; b++) {
        long o = (long)c * nb + bl_index(a, b);
        double vr = V[2 * o], vi = V[2 * o + 1];
        double ph = -2.0 * k * (tau[a] - tau[b]) / N;
        double pr = orc_cospi(ph), pi = orc_sinpi(ph);
        /* P = phi * V */
        double Pr = pr * vr - pi * vi, Pi = pr * vi + pi * vr;
        double gar = g[2 * a], gai = g[2 * a + 1], gbr = g[2 * b], gbi = g[2 * b + 1];
        /* G = g_a conj(g_b) */
        double Gr = gar * gbr + gai * gbi, Gi = gai * gbr - gar * gbi;
        double Vpr = Gr * Pr - Gi * Pi, Vpi = Gr * Pi + Gi * Pr;
        double Rr = Vpr - M[2 * o], Ri = Vpi - M[2 * o + 1];
        L += Rr * Rr + Ri * Ri;
        if (grad_g) {
          /* to g_a: 2 R g_b conj(P) */
          double tr = Rr * gbr - Ri * gbi, ti = Rr * gbi + Ri * gbr; /* R g_b */
          grad_g[2 * a] += 2.0 * (tr * Pr + ti * Pi);
          grad_g[2 * a + 1] += 2.0 * (ti * Pr - tr * Pi);
          /* to g_b: 2 conj(R) g_a P */
          double ur = Rr * gar + Ri * gai, ui = Rr * gai - Ri * gar; /* conj(R) g_a */
          grad_g[2 * b] += 2.0 * (ur * Pr - ui * Pi);
          grad_g[2 * b + 1] += 2.0 * (ur * Pi + ui * Pr);
        }
        if (grad_tau) {
          /* dV'/dtau_a = V' * (-2 pi i k / N); dL/dtau_a = 2 Re(conj(R) dV') */
          double f = 2.0 * M_PI * k / N;
          /* conj(R) * V' * (-i f): Re = f * Im(conj(R) V') */
          double cre_im = Rr * Vpi - Ri * Vpr; /* Im(conj(R) V') */
          grad_tau[a] += 2.0 * f * cre_im;
          grad_tau[b] -= 2.0 * f * cre_im;
        }
      }
  }
  return L;
}
