/*
 * mdtfx.h -- C ABI of the MI355X-native MDT FX correlator (libmdtfx.so).
 *
 * Drop-in replacement for the four GPU calls of the MDT_OneStation_client
 * rank-2 loop (MDT_OneStation_client.cu:1083, 1105, 1114, 1128) plus the
 * native fused path (int8 samples -> PFB+FFT+corner turn -> triangle X-engine).
 *
 * Conventions (all entry points):
 *   - every function returns an int status (MDT_SUCCESS == 0) and never aborts;
 *     mdt_last_error_string() describes the last failure of the calling thread.
 *     This replaces the reference's cudaErrCheck()/MPI_Abort convention
 *     (MDT_host_functions.cu:57-71, MDT_OneStation_client.cu:1087-1092).
 *   - device buffers are caller-owned (as in MDT_OneStation_client.cu:560-576);
 *     kernels never allocate.  Plans/handles own only their constant tables
 *     (window, twiddles) and are released with the matching _destroy call.
 *   - work is stream-ordered and asynchronous on the given hipStream_t
 *     (0 = legacy default stream); the library never synchronises the device.
 *   - mdt_complex is layout-identical to float2 / cuComplex / cufftComplex.
 *   - argument limits follow MDT_OneStation_client.h:54-61:
 *     3 <= NpolsxNelements <= 4096, 32 <= Nchannels <= 16384,
 *     4 <= Ntaps <= 128, 2 <= Nspectra <= 16384.
 */
#ifndef MDTFX_H
#define MDTFX_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  float x, y;
} mdt_complex;

enum mdt_status {
  MDT_SUCCESS = 0,
  MDT_ERR_INVALID_VALUE = 1, /* shape or limit outside the contract */
  MDT_ERR_NULL_POINTER = 2,
  MDT_ERR_UNSUPPORTED = 3,   /* valid for the reference, not implemented here */
  MDT_ERR_LAUNCH = 4,        /* hipGetLastError() after a launch */
  MDT_ERR_ALLOC = 5,         /* plan/table allocation failed */
  MDT_ERR_DEVICE = 6,        /* other HIP runtime failure */
  MDT_ERR_MISALIGNED = 7
};

/* limits, MDT_OneStation_client.h:54-61 (inputs = Npols x Nelements) */
#define MDT_MIN_INPUTS 3
#define MDT_MAX_INPUTS 4096
#define MDT_MIN_CHANNELS 32
#define MDT_MAX_CHANNELS 16384
#define MDT_MIN_TAPS 4
#define MDT_MAX_TAPS 128
#define MDT_MIN_SPECTRA 2
#define MDT_MAX_SPECTRA 16384

/* window kinds */
#define MDT_WINDOW_BH4 0     /* sinc x min 4-term Blackman-Harris (MDT_device_functions.cu:52-88) */
#define MDT_WINDOW_HAMMING 1 /* sinc x periodic Hamming */

/* FFT direction and type, values as cufft.h (CUFFT_FORWARD, CUFFT_INVERSE, CUFFT_C2C) */
#define MDT_FFT_FORWARD (-1)
#define MDT_FFT_INVERSE (1)
#define MDT_FFT_C2C 0x29

/* X-engine operations, values as cublas_v2.h (CUBLAS_OP_N / _T / _C) */
#define MDT_OP_N 0
#define MDT_OP_T 1
#define MDT_OP_C 2

/* X-engine precisions of the native path */
#define MDT_XPREC_FP32 0       /* v_mfma_f32_32x32x2f32 */
#define MDT_XPREC_FP16_SPLIT 1 /* hi+lo FP16, three v_mfma_f32_32x32x16_f16 products */
#define MDT_XPREC_FP8 2        /* e4m3 on v_mfma_scale 32x32x64: block exponents (512, 1024 channels) or
                                  per-channel power-of-two RMS scaling (2048, 4096 channels) */

const char *mdt_last_error_string(void);
int mdt_version(int *version);

/* ---------------------------------------------------------------- a1 window
 * Replaces PFB_Window_min4term_BH<<<...>>>(Nchannels, Ntaps, Window)
 * (MDT_device_functions.cu:52-88, MDT_device_functions.h:44; called as
 * PFB_Window_ExBlackman at MDT_OneStation_client.cu:589).  Window is a device
 * array of Ntaps*Nchannels floats, tap-major.  Values are evaluated in FP64
 * and rounded once to FP32. */
int mdt_pfb_window(int Nchannels, int Ntaps, float *Window, int kind, hipStream_t stream);
/* The name the client calls, PFB_Window_ExBlackman<<<...>>>(Nchannels, Ntaps,
 * d_Window) (MDT_OneStation_client.cu:589), which the reference never defines;
 * the window it provides is the minimum 4-term Blackman-Harris one, so this is
 * mdt_pfb_window(..., MDT_WINDOW_BH4, stream). */
int mdt_pfb_window_exblackman(int Nchannels, int Ntaps, float *Window, hipStream_t stream);

/* ---------------------------------------------------------------- a2 PFB FIR
 * Replaces PpS_Batch<<<(ceil(N/1024), Nspectra, NN), 1024>>>(Nchannels, Ntaps,
 * Window, InSignal, PolyStruct) (MDT_device_functions.cu:113-161,
 * MDT_device_functions.h:45, launch MDT_OneStation_client.cu:1083).  The grid
 * extents the reference passes implicitly (gridDim.y = Nspectra, gridDim.z =
 * NpolsxNelements) are explicit arguments.  InSignal: (Nspectra+Ntaps-1) x NN
 * x Nchannels; PolyStruct: Nspectra x NN x Nchannels.  Bit-exact: FP32 fmaf in
 * tap order. */
int mdt_pps_batch(int Nchannels, int Ntaps, const float *Window, const mdt_complex *InSignal,
                  mdt_complex *PolyStruct, int NpolsxNelements, int Nspectra, hipStream_t stream);
/* Same FIR on int8 (I, Q) input samples, same layout with 2 bytes per sample. */
int mdt_pps_batch_i8(int Nchannels, int Ntaps, const float *Window, const int8_t *InSignal,
                     mdt_complex *PolyStruct, int NpolsxNelements, int Nspectra,
                     hipStream_t stream);

/* ---------------------------------------------------------------- a3 FFT
 * Replaces cufftPlan1d(&plan, Nchannels, CUFFT_C2C, Nspectra*NN),
 * cufftSetStream, cufftExecC2C(plan, in, out, CUFFT_FORWARD)
 * (MDT_OneStation_client.cu:622, 630, 1105).  Batched, unnormalised, natural
 * bin order; in place or out of place.  nx: every length the client accepts,
 * [32, 16384] (MDT_OneStation_client.h:56-57), and powers of two from 8:
 * powers of two in radix-16/8/4 Stockham passes, other 7-smooth lengths
 * (2^a 3^b 5^c 7^d) in mixed-radix passes, the rest by Bluestein's chirp-z
 * convolution (MDT_ERR_UNSUPPORTED outside these). */
typedef struct mdt_fft_plan_s *mdt_fft_plan;
int mdt_fft_plan_1d(mdt_fft_plan *plan, int nx, int type, int batch);
int mdt_fft_set_stream(mdt_fft_plan plan, hipStream_t stream);
int mdt_fft_exec_c2c(mdt_fft_plan plan, mdt_complex *idata, mdt_complex *odata, int direction);
int mdt_fft_destroy(mdt_fft_plan plan);

/* ---------------------------------------------------------------- a4 corner turn
 * Replaces ReorderFOutput<<<...>>>(NpolsxNelements, Nchannels, FOutput,
 * XInput) (MDT_device_functions.cu:183-208, MDT_device_functions.h:46, launch
 * MDT_OneStation_client.cu:1114).  XInput[(c*Ns + s)*NN + e] =
 * FOutput[(s*NN + e)*N + c]; a bit-exact 8-byte permutation. */
int mdt_reorder_foutput(int NpolsxNelements, int Nchannels, const mdt_complex *FOutput,
                        mdt_complex *XInput, int Nspectra, hipStream_t stream);

/* ---------------------------------------------------------------- a5 X-engine
 * Replaces cublasCreate / cublasSetStream / cublasCgemmBatched
 * (MDT_OneStation_client.cu:690, 699, 1128).  mdt_cgemm_batched takes the
 * reference's arguments unchanged (device pointer tables, column-major):
 *   C[i] = alpha * op(A[i]) * op(B[i]) + beta * C[i]
 * Supported: transa = MDT_OP_N, transb = MDT_OP_C (the reference call).  When
 * Aarray == Barray the product is Hermitian and only its lower triangle is
 * computed (the upper one is written as the conjugate mirror). */
typedef struct mdt_xengine_s *mdt_xengine;
int mdt_xengine_create(mdt_xengine *handle);
int mdt_xengine_set_stream(mdt_xengine handle, hipStream_t stream);
int mdt_xengine_destroy(mdt_xengine handle);
int mdt_cgemm_batched(mdt_xengine handle, int transa, int transb, int m, int n, int k,
                      const mdt_complex *alpha, const mdt_complex *const Aarray[], int lda,
                      const mdt_complex *const Barray[], int ldb, const mdt_complex *beta,
                      mdt_complex *const Carray[], int ldc, int batchCount);
/* Strided form of the same Hermitian product without pointer tables:
 * XOutput[c*NN*NN + a + b*NN] = alpha * sum_s x[c][s][a] conj(x[c][s][b]),
 * x[c][s][e] = XInput[(c*Nspectra + s)*NN + e]. */
int mdt_xengine_herk(mdt_xengine handle, int NpolsxNelements, int Nspectra, float alpha,
                     const mdt_complex *XInput, mdt_complex *XOutput, int Nchannels);

/* ---------------------------------------------------------------- a6 PFB history
 * The (Ntaps-1)-spectrum carry between dumps: copies the last Ntaps-1 spectra
 * of the F input to its front (MDT_OneStation_client.cu:1170).  elem_bytes = 8
 * for the float2 input of mdt_pps_batch, 2 for int8 (I, Q). */
int mdt_history_shift(void *FInput, int NpolsxNelements, int Nchannels, int Ntaps, int Nspectra,
                      int elem_bytes, hipStream_t stream);

/* ---------------------------------------------------------------- a7 unpack
 * Decodes the SNAP wire format: each uint32 word is a two's-complement fixed
 * point value scaled by 2^-19, computed as in MDT_OneStation_client.cu:898-907
 * (including its mapping of 0x80000000 to +4096.0).  n words -> n floats. */
int mdt_unpack_u32(const uint32_t *words, float *out, size_t n, hipStream_t stream);

/* ---------------------------------------------------------------- depacketiser
 * Replaces the client's rank-1 loop (MDT_OneStation_client.cu:846-911 for the
 * first block, first_block = 1; 945-1005 in steady state, first_block = 0) on
 * the device.  npkts packets of words_per_pkt uint32 words each: station,
 * SNAP, packet counter, unused, then SperElement x InputsPerSNAP[snap]
 * (re, im) Fix_32_19 word pairs.  Each sample goes to
 * Stacks[stack][spectrum][element][channel] (2 x Nspectra x NpolsxNelements x
 * Nchannels complex64, the reference's h_Stacks), decoded as mdt_unpack_u32.
 * counts (device, 3 x uint64) accumulates the samples written to stacks 0 and
 * 1 (the reference's Nsamples[]) and the samples dropped because their header
 * or index would land outside the stacks (the reference writes those out of
 * bounds).  Two packets that address the same sample race, as in any
 * parallel order. */
int mdt_depacketise(const uint32_t *pkts, int npkts, int words_per_pkt, int SperElement,
                    int NpolsxNelements, int Nchannels, int Nspectra, int Ntaps, int first_block,
                    mdt_complex *Stacks, unsigned long long *counts, hipStream_t stream);

/* ---------------------------------------------------------------- egress (host)
 * The client's CSV writer (MDT_OneStation_client.cu:1202-1272) and a binary
 * dump, from visibilities in host memory.  Layouts: the reference's
 * column-major cuBLAS output [Nchannels][NN][NN], or the native packed lower
 * triangle [Nchannels][NN(NN+1)/2] (NN = Nelements * Npols).  CSV lines
 * "j,p,k,q,z,re,im" for element j <= k, pol p <= q, channel z in fftshift
 * order, V = alpha sum x_(j,p) conj(x_(k,q)), "%.6f".  all_pol_pairs = 0
 * reproduces the reference's enumeration, which skips (j, 1, k, 0) for
 * k > j; 1 writes those rows too.  mdt_csv_filename formats
 * "MDT_%Y%m%d_%H%M%S.csv" in local time (client.cu:1209).  The binary dump
 * is a 64-byte header (magic "MDTVIS01", int32 layout, dtype (0 complex64,
 * 1 complex128), Nelements, Npols, Nchannels, Nspectra, int64 unix time) and
 * the visibilities as stored. */
#define MDT_VIS_FULL_COLMAJOR 0
#define MDT_VIS_TRIANGLE 1
int mdt_csv_filename(char *name, size_t size, long long unix_time);
int mdt_write_csv(const char *path, const mdt_complex *vis, int layout, int Nelements, int Npols,
                  int Nchannels, int all_pol_pairs);
int mdt_write_binary(const char *path, const void *vis, int layout, int dtype, int Nelements, int Npols,
                     int Nchannels, int Nspectra, long long unix_time);

/* ---------------------------------------------------------------- native fused path
 * One dump: int8 (I,Q) samples [Ns+T-1][NN][N][2] -> fused PFB FIR + FFT +
 * corner turn -> channel-major X input (layouts below) -> MFMA
 * X-engine over the lower triangle -> visibilities [N][NN(NN+1)/2] complex
 * FP32, baseline bl(a,b) = a(a+1)/2 + b for a >= b, V = (1/Ns) sum_s
 * x_a conj(x_b) (the alpha of MDT_OneStation_client.cu:683).  Optionally adds
 * V into an FP64 complex accumulator of the same layout (cross-dump
 * integration).  The plan owns the window and twiddle tables only.
 * Nchannels must be a power of two here (MDT_ERR_UNSUPPORTED otherwise: the
 * fused F-engine's transforms are radix-8 Stockham); other channel counts the
 * client allows run through the drop-in calls (mdt_pps_batch, the any-length
 * mdt_fft_exec_c2c, mdt_reorder_foutput, mdt_cgemm_batched). */
typedef struct mdt_fx_plan_s *mdt_fx_plan;
int mdt_fx_plan_create(mdt_fx_plan *plan, int NpolsxNelements, int Nchannels, int Ntaps,
                       int Nspectra, int window_kind, int xprec);
int mdt_fx_set_stream(mdt_fx_plan plan, hipStream_t stream);
/* The X input's sample size, and under MDT_XIN_AUTO its layout, follow the
 * precision: query mdt_fx_xinput_bytes again after changing it. */
int mdt_fx_set_precision(mdt_fx_plan plan, int xprec);
/* bytes of the X input buffer the caller must provide (layout and precision
 * as set when it is filled) */
int mdt_fx_xinput_bytes(mdt_fx_plan plan, size_t *bytes);
/* X input layouts.  In all, sample (channel k, input e, spectrum s) sits at
 * element (s / block) * block_pitch + (k / chan_group) * chan_pitch
 *         + e * row_pitch + (k % chan_group) * block + s % block
 * (mdt_fx_xinput_strides returns the five values).
 * MDT_XIN_PAIRS: chan_group = 2, block = the spectra in 64 bytes (8 samples
 * of 8 bytes, 16 of the FP8 path's 4 bytes), row_pitch = 2*block, chan_pitch
 * = NN*row_pitch, block_pitch = ceil(N/2)*chan_pitch: a block of spectra of
 * channels 2j and 2j+1 of one input is one 128-byte piece.
 * MDT_XIN_BLOCKS: chan_group = 1, block = the spectra in 128 bytes (16 or
 * 32 samples; the block-scaled FP8 format: 32 samples of 2 bytes), row_pitch
 * = block, chan_pitch = NN*block, block_pitch = N*chan_pitch.
 * MDT_XIN_ROWS: chan_group = 1, block_pitch = 0 (block = 2^30 > s): element
 * k*chan_pitch + e*row_pitch + s, row_pitch = Ns rounded up to even +
 * row_pad, chan_pitch = NN*row_pitch + chan_pad (FP8 path: both rounded up
 * to a multiple of 4).  The padding (even, in samples) keeps rows and
 * channels off power-of-two strides; defaults 8 and 64.
 * MDT_XIN_AUTO (default): the layout whose F-engine writes and X-engine
 * reads were fastest: pairs for the FP16 split up to 256 inputs, blocks for
 * the FP16 split above 256 inputs and for FP8, rows for FP32.  The
 * block-scaled FP8 format (512 and 1024 channels) is always in blocks; the
 * layout set here then applies when the precision changes. */
#define MDT_XIN_AUTO (-1)
#define MDT_XIN_PAIRS 0
#define MDT_XIN_ROWS 1
#define MDT_XIN_BLOCKS 2
int mdt_fx_set_xinput_layout(mdt_fx_plan plan, int layout);
int mdt_fx_xinput_strides(mdt_fx_plan plan, size_t *block, size_t *block_pitch, size_t *chan_group,
                          size_t *chan_pitch, size_t *row_pitch);
/* rows layout only: padding, and the two pitches (MDT_ERR_INVALID_VALUE in
 * the other layouts) */
int mdt_fx_set_xinput_padding(mdt_fx_plan plan, int row_pad, int chan_pad);
int mdt_fx_xinput_pitch(mdt_fx_plan plan, size_t *row_pitch, size_t *chan_pitch);
/* copies the plan's FP32 window [Ntaps][Nchannels] to a device buffer */
int mdt_fx_get_window(mdt_fx_plan plan, float *Window);
/* The X input's sample format follows the plan's precision: FP32 (re, im) for
 * MDT_XPREC_FP32; for MDT_XPREC_FP16_SPLIT each 8-byte sample holds FP16
 * (hi_re, hi_im, lo_re, lo_im) of x*s, s the plan's power-of-two scale,
 * hi = fp16(x*s), lo = fp16(x*s - hi).  MDT_XPREC_FP8 (Nchannels in
 * [512, 4096]) has two formats by channel count:
 *   512 and 1024 channels: each sample is 2 bytes, e4m3 (re, im) of x / 2^eb
 *     (nearest-even), eb one exponent per (input, block of 16 spectra) shared
 *     by all channels, the smallest with max |re|, |im| over the block <=
 *     448 2^eb; the blocks layout of 32 spectra (64-byte runs) always; the
 *     exponents as E8M0 bytes eb + 127, [ceil(Ns/16)][NN], follow the samples
 *     in the same buffer (mdt_fx_xinput_scales) and are the MFMA's scale
 *     operands.  Every sample keeps e4m3's 3-bit precision whatever its
 *     block's power, bursts included; results are the same bits on every run.
 *   2048 and 4096 channels (an F-engine item spans fewer than 16 spectra):
 *     each sample is 4 bytes, FP16 (re, im) of x*s, and the F-engine also
 *     sums each channel's power (in a fixed order); the X-engine scales each
 *     channel by a power of two to RMS ~ 32 and quantises on its way in.
 * Run the F-engine at the precision the X-engine will use. */
int mdt_fx_fengine(mdt_fx_plan plan, const int8_t *InSignal, mdt_complex *XInputNative);
/* The same F-engine on the reference's float F input, d_FInput [Ns+T-1][NN][N]
 * complex64 as decoded from the wire (MDT_OneStation_client.cu:898-907, 1022,
 * 1048): replaces the client's PpS_Batch + cufftExecC2C + ReorderFOutput
 * (client.cu:1083-1114) in one pass, bit-identical to those three drop-in
 * calls.  FP32 and FP16-split precisions (MDT_ERR_UNSUPPORTED for FP8). */
int mdt_fx_fengine_f32(mdt_fx_plan plan, const mdt_complex *InSignal, mdt_complex *XInputNative);
/* Per-input complex gain g_e and delay tau_e (samples) applied by the F-engine
 * to each channel k: X'_e[k] = g_e exp(-2 pi i k_s tau_e / N) X_e[k], k_s the
 * signed bin.  Device arrays of NN entries, read at every F-engine launch
 * (they may be updated in place between launches); gains = NULL switches the
 * stage off.  The FP16 / FP8 sample formats' power-of-two scale must then
 * cover max|g| as well: each F-engine launch with gains derives it on the
 * device, in stream order, and the X-engine divides by the scale the last
 * F-engine launch used.  No host synchronisation.  No counterpart in the
 * reference (calibration extension). */
int mdt_fx_set_gains(mdt_fx_plan plan, const mdt_complex *gains, const float *delays);
int mdt_fx_xengine(mdt_fx_plan plan, const mdt_complex *XInputNative, mdt_complex *Vis,
                   double *Acc64);
int mdt_fx_dump(mdt_fx_plan plan, const int8_t *InSignal, mdt_complex *XInputNative,
                mdt_complex *Vis, double *Acc64);
/* FP8 plans at 2048 and 4096 channels: the per-channel exponents e_c the
 * X-engine applies (device int array of Nchannels): samples are quantised as
 * e4m3(x*s*2^e_c), RMS ~ 32. */
int mdt_fx_get_channel_exponents(mdt_fx_plan plan, int *exponents);
/* FP8 plans at 512 and 1024 channels: where the block exponent bytes sit in
 * the X input (byte offset) and how many blocks of 16 spectra there are (the
 * bytes are [nblocks][NpolsxNelements]). */
int mdt_fx_xinput_scales(mdt_fx_plan plan, size_t *offset, size_t *nblocks);
/* the power-of-two scale s of the FP16 sample formats: the plan's, or with
 * gains the one the last F-engine launch used (then this synchronises the
 * plan's stream) */
int mdt_fx_get_split_scale(mdt_fx_plan plan, float *scale);
/* the plan's shape and precision (NULL outputs are skipped) */
int mdt_fx_plan_shape(mdt_fx_plan plan, int *NpolsxNelements, int *Nchannels, int *Ntaps, int *Nspectra,
                      int *xprec);
int mdt_fx_plan_destroy(mdt_fx_plan plan);

/* ---------------------------------------------------------------- streaming correlator
 * Consecutive dumps through one plan, replacing the client's four-rank MPI
 * loop (MDT_OneStation_client.cu:916-1276): the rank-1 ping-pong pinned
 * stacks (client.cu:526, 1009-1057), the rank-2 F- and X-engines
 * (client.cu:1083-1128), the rank-3 history shift (client.cu:1170) and the
 * visibility read-back (client.cu:1199).  One process, three HIP streams:
 * host-to-device copies, F+X compute and device-to-host copies overlap, ordered
 * by events.
 *
 * Two host slots, each a pinned input stack of Nspectra new spectra (int8
 * (I, Q) [Nspectra][NN][Nchannels][2], the F-engine's sample layout) and a
 * pinned visibility buffer [Nchannels][NN(NN+1)/2] complex64 (the plan's
 * triangle).  Each slot has its own device input buffer of Ntaps-1 + Nspectra
 * spectra: a dump first copies the previous dump's last Ntaps-1 spectra to the
 * front (the PFB history carry), then the slot's new stack behind them.  Before
 * the first dump the history is what mdt_stream_prime loaded (zeros at
 * creation).  Use: fill a slot's stack, mdt_stream_submit (returns at once),
 * later mdt_stream_wait on that slot, read its visibilities, refill.
 * Submitting a slot that has not been waited for since its last submit is an
 * error.  The stream takes over the plan's stream; while the stream exists,
 * use the plan only through it.  Needs Nspectra >= Ntaps-1. */
typedef struct mdt_stream_s *mdt_stream;
typedef struct {
  float h2d_ms;     /* history carry + stack copy */
  float fengine_ms; /* F-engine */
  float xengine_ms; /* X-engine (+ FP64 accumulation) */
  float d2h_ms;     /* visibility copy */
  float total_ms;   /* start of the input copies to the end of the visibility copy */
} mdt_dump_times;
int mdt_stream_create(mdt_stream *stream, mdt_fx_plan plan);
int mdt_stream_host_stack(mdt_stream stream, int slot, int8_t **stack);
int mdt_stream_host_vis(mdt_stream stream, int slot, mdt_complex **vis);
/* history = Ntaps-1 spectra (host memory, same layout as a stack), NULL for
 * zeros; the next dump starts a new stream of spectra.  No dump may be in
 * flight (the client's first block, client.cu:859-916). */
int mdt_stream_prime(mdt_stream stream, const int8_t *history);
/* Ordering with the caller's work.  The stream's three HIP streams are
 * non-blocking; each submit first makes the dump wait for everything queued
 * so far on the caller stream (the legacy default stream 0 unless set here),
 * so buffers the caller prepared there -- a zeroed Acc64, gains -- are ready
 * when the dump reads them. */
int mdt_stream_set_caller_stream(mdt_stream stream, hipStream_t caller);
/* Acc64: optional device FP64 accumulator [Nchannels][NN(NN+1)/2] complex,
 * V added in (as mdt_fx_xengine).  Every dump in flight adds into it: the
 * caller may read or write Acc64 only while no slot is in flight (every
 * submitted slot waited for); zeroing it on the caller stream right before
 * the first submit is ordered by the rule above. */
int mdt_stream_submit(mdt_stream stream, int slot, double *Acc64);
/* blocks until the slot's dump is done: its stack may be refilled and its
 * visibilities read; times (optional) gets the dump's phase times */
int mdt_stream_wait(mdt_stream stream, int slot, mdt_dump_times *times);
int mdt_stream_destroy(mdt_stream stream);

/* ---------------------------------------------------------------- packet streaming correlator
 * The client's whole per-dump loop on its own wire data
 * (MDT_OneStation_client.cu:846-1276): SNAP packets -> the rank-1 decode and
 * scatter into float stacks (client.cu:846-1005, 898-907) -> d_FInput =
 * PFB history + stack (client.cu:1022, 1048) -> PpS_Batch / cufftExecC2C /
 * ReorderFOutput / cublasCgemmBatched (client.cu:1083-1128) -> history shift
 * (client.cu:1170) -> visibilities to the host (client.cu:1199), in one
 * process.  Two host slots, each a pinned packet buffer of max_pkts packets
 * of words_per_pkt uint32 words (the mdt_depacketise format) and a pinned
 * visibility buffer [Nchannels][NN(NN+1)/2] complex64.  A submit takes the
 * packets of one dump: they are copied to the device, decoded there into the
 * slot's float F input (mdt_fx_fengine_f32) and correlated, on three HIP
 * streams ordered by events (as mdt_stream_*).
 * Dumps are numbered from 0 by submit order.  Sample i of a packet with
 * counter cnt has index u = cnt*SperElement + i - (Ntaps-1)*Nchannels (the
 * client's steady-state index, client.cu:953); dump d takes u in
 * [d*Ns*N, (d+1)*Ns*N), and dump 0, the first block (client.cu:855), also
 * u in [-(Ntaps-1)*N, 0), its PFB history; later dumps carry the previous
 * dump's last Ntaps-1 spectra on the device.  A sample no packet carries is 0
 * (the client would keep the stack's stale value).  Samples of another dump,
 * bad headers and short payloads are dropped.  counts (mdt_pstream_wait,
 * nullable, 3 entries): samples written to the dump's spectra, to its history
 * spectra (dump 0 only), dropped -- per (packet, sample), as the client's
 * Nsamples[].  The plan must use the FP32 or FP16-split precision; the stream
 * binds the plan's stream while it exists.  Needs Nspectra >= Ntaps-1. */
typedef struct mdt_pstream_s *mdt_pstream;
int mdt_pstream_create(mdt_pstream *stream, mdt_fx_plan plan, int words_per_pkt, int SperElement, int max_pkts);
int mdt_pstream_host_packets(mdt_pstream stream, int slot, uint32_t **pkts);
int mdt_pstream_host_vis(mdt_pstream stream, int slot, mdt_complex **vis);
/* as mdt_stream_set_caller_stream */
int mdt_pstream_set_caller_stream(mdt_pstream stream, hipStream_t caller);
/* the next submit is dump 0 again (a new first block); no dump may be in flight */
int mdt_pstream_restart(mdt_pstream stream);
/* Acc64 as in mdt_stream_submit (same ordering rules) */
int mdt_pstream_submit(mdt_pstream stream, int slot, int npkts, double *Acc64);
int mdt_pstream_wait(mdt_pstream stream, int slot, mdt_dump_times *times, unsigned long long *counts);
int mdt_pstream_destroy(mdt_pstream stream);

/* ---------------------------------------------------------------- calibration (a9)
 * No counterpart in the reference.  For raw triangle visibilities V (as from
 * mdt_fx_xengine) and per-input gains g, delays tau: G_a(c) = g_a
 * exp(-2 pi i k_c tau_a / N), V'_ab = G_a conj(G_b) V_ab (what the F-engine
 * gain stage produces), L = sum_{c, a>=b} |V' - M|^2.  mdt_calib_loss_grad
 * writes L and the adjoint grad_g = dL/dRe g + i dL/dIm g, grad_tau = dL/dtau
 * (each output nullable, grad_tau needs grad_g).  FP32 with compensated,
 * fixed-order (deterministic) sums; the plan owns its workspace. */
typedef struct mdt_calib_plan_s *mdt_calib_plan;
int mdt_calib_plan_create(mdt_calib_plan *plan, int NpolsxNelements, int Nchannels);
int mdt_calib_set_stream(mdt_calib_plan plan, hipStream_t stream);
int mdt_calib_apply(mdt_calib_plan plan, const mdt_complex *V, const mdt_complex *g, const float *tau,
                    mdt_complex *Vp);
int mdt_calib_loss_grad(mdt_calib_plan plan, const mdt_complex *V, const mdt_complex *M,
                        const mdt_complex *g, const float *tau, float *loss, mdt_complex *grad_g,
                        float *grad_tau);
int mdt_calib_plan_destroy(mdt_calib_plan plan);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mdt_calib_loss_grad without a plan, computing the same
 * loss and adjoint with the same kernels (bit-identical results) on a
 * workspace the library holds in a static device array of 12.06 MiB per GPU
 * (3 162 112 floats: the partial sums of any shape within the limits,
 * 3 <= NpolsxNelements <= 4096, 32 <= Nchannels <= 16384; larger needs are
 * refused with MDT_ERR_INVALID_VALUE).  It allocates nothing, ever.  Calls on
 * one GPU share its workspace: calls that overlap on one GPU (on different
 * streams, or from several threads) race; the workspace needs no clearing
 * (every partial is written before it is read). */
int mdt_calib_loss_grad_noalloc(int NpolsxNelements, int Nchannels, const mdt_complex *V, const mdt_complex *M,
                                const mdt_complex *g, const float *tau, float *loss, mdt_complex *grad_g,
                                float *grad_tau, hipStream_t stream);

/* ---------------------------------------------------------------- synthetic input
 * Deterministic counter-based generator of int8 (I, Q) samples for the
 * operating points: x_e[t] = clip(round(sigma*(sqrt(1-rho) n_e[t] +
 * sqrt(rho) c[t - delay_e]) + tone)), n and c unit complex Gaussians from
 * splitmix64 + Box-Muller.  Layout [nframes][NN][N][2]; frame f holds samples
 * t = f*N .. f*N+N-1 of every input (MDT_OneStation_client.cu:855-856). */
typedef struct {
  uint64_t seed;
  float sigma;
  float rho;        /* common-component power fraction, 0 = independent inputs */
  int max_delay;    /* per-input integer delay of the common component, 0..max_delay */
  int ntones;       /* up to 4 tones, added to every input */
  float tone_freq[4]; /* cycles per sample */
  float tone_amp[4];
  long long t0;     /* first sample index (for streaming dumps) */
} mdt_synth_params;
int mdt_synth_i8(const mdt_synth_params *p, int NpolsxNelements, int Nchannels, int nframes,
                 int8_t *out, hipStream_t stream);

/* ---------------------------------------------------------------- HBM probe
 * A measurement aid with no counterpart in the reference: one streaming pass
 * over `bytes` (a multiple of 16, buffers 16-byte aligned) with 16-byte lanes
 * and a grid-stride loop over `grid` workgroups of 256 threads, to measure
 * what this device's HBM delivers for each kind of traffic the kernels move:
 * MDT_PROBE_READ reads src (dst unused), MDT_PROBE_WRITE writes dst with
 * streaming stores (src unused), MDT_PROBE_COPY copies src to dst. */
#define MDT_PROBE_READ 0
#define MDT_PROBE_WRITE 1
#define MDT_PROBE_COPY 2
int mdt_hbm_probe(int kind, const void *src, void *dst, size_t bytes, int grid, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* MDTFX_H */
