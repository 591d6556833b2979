// fxplan.hip -- the native one-dump plan: window + twiddle tables and the
// F-engine -> X-engine sequence of the MDT_OneStation_client rank-2 loop
// (MDT_OneStation_client.cu:1083-1128) as two kernels on one stream.
#include <cmath>
#include <new>
#include <vector>

#include "fengine.hpp"
#include "fft_core.hpp"
#include "xengine.hpp"

namespace mdt {
int launch_window(int N, int T, float *W, int kind, hipStream_t stream);
}

struct mdt_fx_plan_s {
  int NN = 0, N = 0, T = 0, Ns = 0, kind = 0, xprec = 0;
  hipStream_t stream = 0;
  float *W = nullptr;
  float2 *tw = nullptr;
  const float2 *gain = nullptr;
  const float *delay = nullptr;
  float f16_scale = 1.0f;  // exact power of two keeping |X| * scale <= 2^15 (no gains)
  double wbound = 0.0;     // rigorous bound of |X| without gains
  float *dscale = nullptr; // device: the scale of the last F-engine launch (f16_scale, or with gains)
  bool dscale_base = true; // dscale holds f16_scale
  float *power = nullptr;  // FP8 path, per-channel format: per-channel power from the F-engine
  float *power_part = nullptr;  // its per-workgroup partial sums [power_slots][N]
  int power_slots = 0;
  float *Wt = nullptr;     // window arranged for the persistent F-engine (when it serves the shape)
  int layout = MDT_XIN_AUTO;
  int row_pad = 8, chan_pad = 64;  // rows layout: padding (samples) after each row / channel
  // rows-layout pitches in samples; 16-byte aligned rows: even for 8-byte
  // samples, a multiple of 4 for the 4-byte FP8-path samples
  long rp() const {
    const long r = (long)Ns + (Ns & 1) + row_pad;
    return xprec == MDT_XPREC_FP8 ? (r + 3) & ~3L : r;
  }
  long cp() const {
    const long c = (long)NN * rp() + chan_pad;
    return xprec == MDT_XPREC_FP8 ? (c + 3) & ~3L : c;
  }
  // FP8 sample formats: e4m3 with block exponents (512 and 1024 channels),
  // FP16 with per-channel power (2048 and 4096)
  bool f8block() const { return xprec == MDT_XPREC_FP8 && mdt::fp8_block_shape(N); }
  // the layout in use: MDT_XIN_AUTO picks, per X-engine, the one it reads
  // best with the F-engine's writes (measured): pairs only where each
  // channel's samples are read once (FP16 split, NN <= 256: one tile or the
  // whole-channel kernel); the tiled X-engines that re-read a channel do
  // not reuse the partner channel's half of each 128-byte line.  The
  // block-scaled FP8 format has one layout, blocks of 32 spectra.
  int xin_layout() const {
    if (f8block()) return MDT_XIN_BLOCKS;
    if (layout != MDT_XIN_AUTO) return layout;
    if (xprec == MDT_XPREC_FP32) return MDT_XIN_ROWS;
    if (xprec == MDT_XPREC_FP16_SPLIT && NN <= 256) return MDT_XIN_PAIRS;
    return MDT_XIN_BLOCKS;
  }
  // the X input geometry of the current layout and precision (XGeom)
  mdt::XGeom geom() const {
    mdt::XGeom g;
    const int lay = xin_layout();
    const int lg_esize = xprec == MDT_XPREC_FP8 ? 2 : 3;
    if (f8block()) {  // 32 spectra of 2-byte e4m3 samples: 64-byte rows, one X-engine K chunk
      g.lgB = 5;
      g.rp = 32;
      g.cp = (long)NN * g.rp;
      g.bp = (long)N * g.cp;
      return g;
    }
    if (lay == MDT_XIN_ROWS) {
      g.rp = rp();
      g.cp = cp();
    } else if (lay == MDT_XIN_PAIRS) {  // two channels x 64 bytes of spectra per piece
      g.lgB = 6 - lg_esize;
      g.lgG = 1;
      g.rp = 2L << g.lgB;
      g.cp = (long)NN * g.rp;
      g.bp = (long)((N + 1) / 2) * g.cp;
    } else {  // MDT_XIN_BLOCKS: one channel x 128 bytes of spectra per piece
      g.lgB = 7 - lg_esize;
      g.rp = 1L << g.lgB;
      g.cp = (long)NN * g.rp;
      g.bp = (long)N * g.cp;
    }
    return g;
  }
  // block-scaled FP8: the exponent bytes [nblk16][NN] after the samples
  int nblk16() const { return (Ns + 15) / 16; }
  size_t scales_offset() const { return ((size_t)xin_samples() * 2 + 255) & ~(size_t)255; }
  // X input extent in samples
  long xin_samples() const {
    const mdt::XGeom g = geom();
    if (xin_layout() == MDT_XIN_ROWS) return (long)N * g.cp;
    const long B = 1L << g.lgB;
    return (((long)Ns + B - 1) / B) * g.bp;
  }
};

using namespace mdt;

// Rigorous bound of the channelised samples for int8 input: |X[k]| <=
// sum_n |y[n]| <= |x|max * sum_{i,n} |W[i][n]|, |x|max = |(-128, -128)|.  The
// FP16 pre-scale is the power of two that maps this bound to <= 2^15, below
// the FP16 maximum 65504 (f16_prescale).  Computed once, at plan creation.
static int choose_f16_scale(mdt_fx_plan p) {
  std::vector<float> w((size_t)p->T * p->N);
  if (hipMemcpy(w.data(), p->W, sizeof(float) * w.size(), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(MDT_ERR_DEVICE, "window read-back");
  double sum = 0.0;
  for (float v : w) sum += std::fabs((double)v);
  p->wbound = 181.02 * sum;
  p->f16_scale = f16_prescale(p->wbound);
  if (hipMemcpy(p->dscale, &p->f16_scale, sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return fail(MDT_ERR_DEVICE, "scale upload");
  return MDT_SUCCESS;
}

// With gains the bound grows by max(max_e |g_e|, 1): the F-engine's scale is
// computed on the device, in stream order before the F-engine reads it, and
// stays there for the X-engine that divides by it.  gains = NULL writes the
// plan's scale back.  One workgroup, fixed-order maximum.
__global__ void gain_scale_kernel(const float2 *gains, int NN, double wbound, float *dscale) {
  __shared__ double red[256];
  double m = 0.0;
  if (gains != nullptr)
    for (int e = threadIdx.x; e < NN; e += 256) {
      const float2 g = gains[e];
      m = fmax(m, sqrt((double)g.x * g.x + (double)g.y * g.y));
    }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *dscale = f16_prescale(wbound * fmax(red[0], 1.0));
}

__global__ void zero_kernel(float *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

static size_t elem_bytes(const mdt_fx_plan_s *p) { return p->f8block() ? 2 : (p->xprec == MDT_XPREC_FP8 ? 4 : 8); }

static int check_xprec(int xprec, int N, int T) {
  if (xprec != MDT_XPREC_FP32 && xprec != MDT_XPREC_FP16_SPLIT && xprec != MDT_XPREC_FP8)
    return fail(MDT_ERR_INVALID_VALUE, "unknown X-engine precision %d", xprec);
  // the FP8 sample format and channel powers come from the persistent F-engine
  if (xprec == MDT_XPREC_FP8 && (N < 512 || !persistent_shape(N, T)))
    return fail(MDT_ERR_UNSUPPORTED, "FP8 path needs Nchannels in [512, 4096]");
  return MDT_SUCCESS;
}

extern "C" {

int mdt_fx_plan_create(mdt_fx_plan *plan, int NpolsxNelements, int Nchannels, int Ntaps,
                       int Nspectra, int window_kind, int xprec) {
  MDT_REQUIRE_PTR(plan);
  *plan = nullptr;
  MDT_TRY(check_inputs(NpolsxNelements));
  MDT_TRY(check_channels(Nchannels));
  MDT_TRY(check_taps(Ntaps));
  MDT_TRY(check_spectra(Nspectra));
  if (!is_pow2(Nchannels))
    return fail(MDT_ERR_UNSUPPORTED, "Nchannels=%d: the FFT needs a power of two", Nchannels);
  if (window_kind != MDT_WINDOW_BH4 && window_kind != MDT_WINDOW_HAMMING)
    return fail(MDT_ERR_INVALID_VALUE, "unknown window kind %d", window_kind);
  MDT_TRY(check_xprec(xprec, Nchannels, Ntaps));
  mdt_fx_plan p = new (std::nothrow) mdt_fx_plan_s;
  if (!p) return fail(MDT_ERR_ALLOC, "plan");
  p->NN = NpolsxNelements;
  p->N = Nchannels;
  p->T = Ntaps;
  p->Ns = Nspectra;
  p->kind = window_kind;
  p->xprec = xprec;
  if (hipMalloc(&p->W, sizeof(float) * (size_t)Ntaps * Nchannels) != hipSuccess) {
    delete p;
    return fail(MDT_ERR_ALLOC, "window table");
  }
  if (hipMalloc(&p->power, sizeof(float) * (size_t)Nchannels) != hipSuccess) {
    (void)hipFree(p->W);
    delete p;
    return fail(MDT_ERR_ALLOC, "channel power");
  }
  if (hipMalloc(&p->dscale, sizeof(float)) != hipSuccess) {
    (void)hipFree(p->W);
    (void)hipFree(p->power);
    delete p;
    return fail(MDT_ERR_ALLOC, "scale");
  }
  int st = launch_window(Nchannels, Ntaps, p->W, window_kind, 0);
  if (st == MDT_SUCCESS) {
    hipLaunchKernelGGL(zero_kernel, dim3((Nchannels + 255) / 256), dim3(256), 0, 0, p->power, Nchannels);
    st = check_launch("zero_kernel");
  }
  if (st == MDT_SUCCESS && persistent_shape(Nchannels, Ntaps)) {
    if (hipMalloc(&p->Wt, sizeof(float) * (size_t)arranged_taps_floats(Nchannels, Ntaps)) != hipSuccess)
      st = fail(MDT_ERR_ALLOC, "arranged window");
    else
      st = arrange_taps(Nchannels, Ntaps, p->W, p->Wt, 0);
  }
  if (st == MDT_SUCCESS && persistent_shape(Nchannels, Ntaps) && !fp8_block_shape(Nchannels)) {
    // the per-channel FP8 format's partial power table: one row per
    // resident workgroup of the persistent F-engine (at most 4 per CU)
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      st = fail(MDT_ERR_DEVICE, "device attributes");
    p->power_slots = 4 * (cus > 0 ? cus : 256);
    if (st == MDT_SUCCESS &&
        hipMalloc(&p->power_part, sizeof(float) * (size_t)p->power_slots * Nchannels) != hipSuccess)
      st = fail(MDT_ERR_ALLOC, "partial channel powers");
  }
  if (st == MDT_SUCCESS) st = make_twiddles(Nchannels, &p->tw);
  if (st == MDT_SUCCESS && hipStreamSynchronize(0) != hipSuccess)
    st = fail(MDT_ERR_DEVICE, "plan table build");
  if (st == MDT_SUCCESS) st = choose_f16_scale(p);
  if (st != MDT_SUCCESS) {
    (void)hipFree(p->W);
    (void)hipFree(p->power);
    (void)hipFree(p->dscale);
    if (p->power_part) (void)hipFree(p->power_part);
    if (p->Wt) (void)hipFree(p->Wt);
    if (p->tw) (void)hipFree(p->tw);
    delete p;
    return st;
  }
  *plan = p;
  return MDT_SUCCESS;
}

int mdt_fx_set_stream(mdt_fx_plan plan, hipStream_t stream) {
  MDT_REQUIRE_PTR(plan);
  plan->stream = stream;
  return MDT_SUCCESS;
}

int mdt_fx_set_precision(mdt_fx_plan plan, int xprec) {
  MDT_REQUIRE_PTR(plan);
  MDT_TRY(check_xprec(xprec, plan->N, plan->T));
  plan->xprec = xprec;
  return MDT_SUCCESS;
}

int mdt_fx_set_gains(mdt_fx_plan plan, const mdt_complex *gains, const float *delays) {
  MDT_REQUIRE_PTR(plan);
  if (!gains && delays) return fail(MDT_ERR_NULL_POINTER, "delays need gains");
  plan->gain = (const float2 *)gains;  // read (and the scale derived) at each F-engine launch
  plan->delay = delays;
  return MDT_SUCCESS;
}

int mdt_fx_xinput_bytes(mdt_fx_plan plan, size_t *bytes) {
  MDT_REQUIRE_PTR(plan);
  MDT_REQUIRE_PTR(bytes);
  *bytes = plan->f8block() ? plan->scales_offset() + (size_t)plan->nblk16() * plan->NN
                           : elem_bytes(plan) * (size_t)plan->xin_samples();
  return MDT_SUCCESS;
}

int mdt_fx_set_xinput_layout(mdt_fx_plan plan, int layout) {
  MDT_REQUIRE_PTR(plan);
  if (layout != MDT_XIN_AUTO && layout != MDT_XIN_PAIRS && layout != MDT_XIN_ROWS && layout != MDT_XIN_BLOCKS)
    return fail(MDT_ERR_INVALID_VALUE, "unknown X input layout %d", layout);
  plan->layout = layout;
  return MDT_SUCCESS;
}

int mdt_fx_xinput_strides(mdt_fx_plan plan, size_t *block, size_t *block_pitch, size_t *chan_group,
                          size_t *chan_pitch, size_t *row_pitch) {
  MDT_REQUIRE_PTR(plan);
  MDT_REQUIRE_PTR(block);
  MDT_REQUIRE_PTR(block_pitch);
  MDT_REQUIRE_PTR(chan_group);
  MDT_REQUIRE_PTR(chan_pitch);
  MDT_REQUIRE_PTR(row_pitch);
  const XGeom g = plan->geom();
  *block = (size_t)1 << g.lgB;
  *block_pitch = (size_t)g.bp;
  *chan_group = (size_t)1 << g.lgG;
  *chan_pitch = (size_t)g.cp;
  *row_pitch = (size_t)g.rp;
  return MDT_SUCCESS;
}

int mdt_fx_set_xinput_padding(mdt_fx_plan plan, int row_pad, int chan_pad) {
  MDT_REQUIRE_PTR(plan);
  if (row_pad < 0 || chan_pad < 0 || (row_pad & 1) || (chan_pad & 1) || row_pad > 4096 ||
      chan_pad > (1 << 20))
    return fail(MDT_ERR_INVALID_VALUE, "padding must be even and in [0, 4096] / [0, 2^20]");
  plan->row_pad = row_pad;
  plan->chan_pad = chan_pad;
  return MDT_SUCCESS;
}

int mdt_fx_xinput_pitch(mdt_fx_plan plan, size_t *row_pitch, size_t *chan_pitch) {
  MDT_REQUIRE_PTR(plan);
  MDT_REQUIRE_PTR(row_pitch);
  MDT_REQUIRE_PTR(chan_pitch);
  if (plan->xin_layout() != MDT_XIN_ROWS)
    return fail(MDT_ERR_INVALID_VALUE, "the X input is not in the rows layout (mdt_fx_xinput_strides)");
  *row_pitch = (size_t)plan->rp();
  *chan_pitch = (size_t)plan->cp();
  return MDT_SUCCESS;
}

int mdt_fx_get_window(mdt_fx_plan plan, float *Window) {
  MDT_REQUIRE_PTR(plan);
  MDT_REQUIRE_PTR(Window);
  MDT_HIP_TRY(hipMemcpyAsync(Window, plan->W, sizeof(float) * (size_t)plan->T * plan->N,
                             hipMemcpyDeviceToDevice, plan->stream));
  return MDT_SUCCESS;
}

}  // extern "C"

// the F-engine on int8 (in8) or float2 (fin) input
static int fengine(mdt_fx_plan plan, const int8_t *in8, const mdt_complex *fin, mdt_complex *XInputNative) {
  MDT_REQUIRE_PTR(plan);
  MDT_REQUIRE_PTR(XInputNative);
  if (((uintptr_t)in8 & 1) || ((uintptr_t)fin & 7) || ((uintptr_t)XInputNative & 15))
    return fail(MDT_ERR_MISALIGNED, "input must be 2-byte (int8) or 8-byte (float2), X input 16-byte aligned");
  FArgs a{};
  a.in = (const char2 *)in8;
  a.fin = (const float2 *)fin;
  a.W = plan->W;
  a.Wt = plan->Wt;
  a.tw = plan->tw;
  a.out = (float2 *)XInputNative;
  a.NN = plan->NN;
  a.Ns = plan->Ns;
  a.T = plan->T;
  a.xg = plan->geom();
  a.gain = plan->gain;
  a.delay = plan->delay;
  a.split = (plan->xprec == MDT_XPREC_FP16_SPLIT) ? 1 : (plan->xprec == MDT_XPREC_FP8 ? 2 : 0);
  a.split_scale = plan->f16_scale;
  a.dscale = plan->dscale;
  a.power = plan->power;
  a.power_part = plan->power_part;
  a.power_slots = plan->power_slots;
  if (plan->f8block()) {
    a.split = 3;
    a.scales = (unsigned char *)XInputNative + plan->scales_offset();
    a.nblk16 = plan->nblk16();
  }
  if ((a.split == 1 || a.split == 2) && (plan->gain != nullptr || !plan->dscale_base)) {
    hipLaunchKernelGGL(gain_scale_kernel, dim3(1), dim3(256), 0, plan->stream, (const float2 *)plan->gain,
                       plan->NN, plan->wbound, plan->dscale);
    MDT_TRY(check_launch("gain_scale_kernel"));
    plan->dscale_base = plan->gain == nullptr;
  }
  return launch_fengine(a, plan->N, plan->stream);
}

extern "C" {

int mdt_fx_fengine(mdt_fx_plan plan, const int8_t *InSignal, mdt_complex *XInputNative) {
  MDT_REQUIRE_PTR(InSignal);
  return fengine(plan, InSignal, nullptr, XInputNative);
}

int mdt_fx_fengine_f32(mdt_fx_plan plan, const mdt_complex *InSignal, mdt_complex *XInputNative) {
  MDT_REQUIRE_PTR(InSignal);
  if (plan && plan->xprec == MDT_XPREC_FP8)
    return fail(MDT_ERR_UNSUPPORTED, "the FP8 path takes int8 samples");
  return fengine(plan, nullptr, InSignal, XInputNative);
}

int mdt_fx_xengine(mdt_fx_plan plan, const mdt_complex *XInputNative, mdt_complex *Vis,
                   double *Acc64) {
  MDT_REQUIRE_PTR(plan);
  MDT_REQUIRE_PTR(XInputNative);
  MDT_REQUIRE_PTR(Vis);
  XArgs g{};
  const int NN = plan->NN;
  g.m = g.n = NN;
  g.k = plan->Ns;
  g.batch = plan->N;
  g.alpha = make_float2(1.0f / (float)plan->Ns, 0.f);  // MDT_OneStation_client.cu:683
  g.beta = make_float2(0.f, 0.f);
  const XGeom xg = plan->geom();
  g.Abase = g.Bbase = (const float2 *)XInputNative;
  g.Abatch = g.Bbatch = xg.cp;
  g.ldI = g.ldJ = xg.rp;
  g.xlgB = xg.lgB;
  g.xlgG = xg.lgG;
  g.xbp = xg.bp;
  g.Cbase = (float2 *)Vis;
  g.Cbatch = (long)NN * (NN + 1) / 2;
  g.herm = 1;
  g.acc64 = Acc64;
  g.prec = plan->xprec;
  g.in_scale = plan->f16_scale;
  g.scale_dev = plan->dscale;  // as the last F-engine launch used it
  if (plan->xprec == MDT_XPREC_FP16_SPLIT) return xeng_launch_f16split(g, plan->stream);
  if (plan->f8block()) {
    g.f8scales = (const unsigned char *)XInputNative + plan->scales_offset();
    g.nblk16 = plan->nblk16();
    return xeng_launch_f8_block(g, plan->stream);
  }
  if (plan->xprec == MDT_XPREC_FP8) return xeng_launch_f8(g, plan->power, plan->stream);
  return xeng_launch(g, XL_MK, XO_TRI, plan->stream);
}

int mdt_fx_dump(mdt_fx_plan plan, const int8_t *InSignal, mdt_complex *XInputNative,
                mdt_complex *Vis, double *Acc64) {
  MDT_TRY(mdt_fx_fengine(plan, InSignal, XInputNative));
  return mdt_fx_xengine(plan, XInputNative, Vis, Acc64);
}

int mdt_fx_get_channel_exponents(mdt_fx_plan plan, int *exponents) {
  MDT_REQUIRE_PTR(plan);
  MDT_REQUIRE_PTR(exponents);
  if (plan->xprec != MDT_XPREC_FP8) return fail(MDT_ERR_INVALID_VALUE, "plan precision is not FP8");
  if (plan->f8block())
    return fail(MDT_ERR_INVALID_VALUE, "FP8 at %d channels scales per block of spectra (mdt_fx_xinput_scales)",
                plan->N);
  return launch_chan_exponents(plan->N, plan->power, (float)plan->NN * (float)plan->Ns, exponents,
                               plan->stream);
}

int mdt_fx_xinput_scales(mdt_fx_plan plan, size_t *offset, size_t *nblocks) {
  MDT_REQUIRE_PTR(plan);
  MDT_REQUIRE_PTR(offset);
  MDT_REQUIRE_PTR(nblocks);
  if (!plan->f8block())
    return fail(MDT_ERR_INVALID_VALUE, "the X input has no block exponents (FP8 at 512 or 1024 channels only)");
  *offset = plan->scales_offset();
  *nblocks = (size_t)plan->nblk16();
  return MDT_SUCCESS;
}

int mdt_fx_plan_shape(mdt_fx_plan plan, int *NpolsxNelements, int *Nchannels, int *Ntaps, int *Nspectra,
                      int *xprec) {
  MDT_REQUIRE_PTR(plan);
  if (NpolsxNelements) *NpolsxNelements = plan->NN;
  if (Nchannels) *Nchannels = plan->N;
  if (Ntaps) *Ntaps = plan->T;
  if (Nspectra) *Nspectra = plan->Ns;
  if (xprec) *xprec = plan->xprec;
  return MDT_SUCCESS;
}

int mdt_fx_get_split_scale(mdt_fx_plan plan, float *scale) {
  MDT_REQUIRE_PTR(plan);
  MDT_REQUIRE_PTR(scale);
  if (plan->dscale_base) {
    *scale = plan->f16_scale;
    return MDT_SUCCESS;
  }
  MDT_HIP_TRY(hipMemcpyAsync(scale, plan->dscale, sizeof(float), hipMemcpyDeviceToHost, plan->stream));
  MDT_HIP_TRY(hipStreamSynchronize(plan->stream));
  return MDT_SUCCESS;
}

int mdt_fx_plan_destroy(mdt_fx_plan plan) {
  if (!plan) return MDT_SUCCESS;
  if (plan->W) (void)hipFree(plan->W);
  if (plan->power) (void)hipFree(plan->power);
  if (plan->dscale) (void)hipFree(plan->dscale);
  if (plan->power_part) (void)hipFree(plan->power_part);
  if (plan->Wt) (void)hipFree(plan->Wt);
  if (plan->tw) (void)hipFree(plan->tw);
  delete plan;
  return MDT_SUCCESS;
}

}  // extern "C"
