/*
 * mdt_oracle.h -- CPU restatement of the MDT FX correlator chain.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product library (libmdtfx).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links or calls
 * it.
 *
 * Every function cites the reference file:line whose semantics it restates
 * (paths relative to the aragorn2101/MDT_FPGA-GPU_pipeline checkout):
 *   devfn.cu  = MDT_OneStation_client/MDT_device_functions.cu
 *   client.cu = MDT_OneStation_client/MDT_OneStation_client.cu
 *
 * Layouts (all complex data interleaved re,im):
 *   F input   [Ns+T-1][NN][N]      (devfn.cu:124-135)
 *   F output  [Ns][NN][N]          (devfn.cu:155-159)
 *   X input   [N][Ns][NN]          (devfn.cu:207)
 *   X output  [N][NN][NN] column-major, V[a + b*NN] = alpha*sum_s x_a conj(x_b)
 *                                  (client.cu:656-660, 1128)
 *   triangle  [N][NN*(NN+1)/2], baseline index bl(a,b) = a*(a+1)/2 + b, a >= b
 */
#ifndef MDT_ORACLE_H
#define MDT_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* window kinds */
#define ORC_WINDOW_BH4 0     /* sinc x minimum 4-term Blackman-Harris, devfn.cu:52-88 */
#define ORC_WINDOW_HAMMING 1 /* sinc x periodic Hamming (the 40-input Hamming case) */

double orc_sinpi(double x);
double orc_cospi(double x);
void orc_window_f64(int N, int T, int kind, double *W);
void orc_window_f32(int N, int T, int kind, float *W);

/* PpS_Batch, devfn.cu:113-161 */
void orc_pps_f32(int N, int T, const float *W, const float *in, float *out, int NN, int Ns);
void orc_pps_i8_f32(int N, int T, const float *W, const int8_t *in, float *out, int NN, int Ns);
void orc_pps_f64(int N, int T, const double *W, const double *in, double *out, int NN, int Ns);

/* batched C2C FFT, client.cu:622 + 1105 (cufftPlan1d / cufftExecC2C) */
void orc_dft_f64(int N, long batch, const double *in, double *out, int sign);
void orc_fft_f64(int N, long batch, const double *in, double *out, int sign);
void orc_fft_f32(int N, long batch, const float *in, float *out, int sign);

/* ReorderFOutput, devfn.cu:183-208 (element size in bytes: byte-exact copy) */
void orc_reorder(int NN, int N, int Ns, int elem_bytes, const void *FOut, void *XIn);

/* X-engine, client.cu:683-686 + 1128 */
void orc_xengine_full_f64(int NN, int Ns, int N, double alpha, const double *XIn, double *XOut);
void orc_xengine_full_from_f32(int NN, int Ns, int N, double alpha, const float *XIn, double *XOut);
void orc_cgemm_nc_f64(int m, int n, int k, const double *alpha, const float *A, int lda,
                      const float *B, int ldb, const double *beta, const float *C0, int ldc,
                      double *C);

/* unpack, client.cu:898-907 */
float orc_unpack_u32(uint32_t w);
void orc_unpack_u32_array(const uint32_t *w, long n, float *out);
int orc_count_unpack_mismatch_all(const float *table_lo16, long *first_bad);

/* depacketiser, client.cu:846-911 (first block) and 945-1005 (steady state);
 * nsamples[3]: samples to stack 0, to stack 1, and skipped (they would land
 * outside the stacks) */
void orc_depacketise(const uint32_t *pkts, int npkts, int words_per_pkt, int SperElement,
                     int Nelements_total, int N, int Ns, int T, int first_block,
                     float *stacks, long *nsamples);

/* full chains.  vis: triangle layout, complex, only channels listed in chans[] */
void orc_fx_chain_f64(int NN, int N, int T, int Ns, const double *W, const int8_t *in,
                      const int *chans, int nchan, double *vis);
void orc_fx_chain_f32(int NN, int N, int T, int Ns, const float *W, const int8_t *in,
                      float *vis, int nthreads);

/* calibration (no reference counterpart): V' = g_a conj(g_b) exp(-2 pi i k_c (tau_a-tau_b)/N) V */
double orc_calib_loss_grad_f64(int NN, int N, const double *V, const double *M, const double *g,
                               const double *tau, double *grad_g, double *grad_tau);
void orc_calib_apply_f64(int NN, int N, const double *V, const double *g, const double *tau,
                         double *Vp);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
