// pstream.hip -- the packet streaming correlator: the client's whole per-dump
// loop on its own wire data (MDT_OneStation_client.cu:846-1276) in one
// process.  The client receives SNAP packets on rank 0, decodes and scatters
// them into ping-pong float stacks on rank 1 (client.cu:846-1005), copies the
// stack behind the PFB history into d_FInput on rank 2 (client.cu:1022, 1048),
// runs PpS_Batch / cufftExecC2C / ReorderFOutput / cublasCgemmBatched
// (client.cu:1083-1128), shifts the history (client.cu:1170) and reads the
// visibilities back (client.cu:1199).  Here the raw packets of one dump go to
// the device and everything else happens there:
//
//   copy-in:  wait in_free[b]; packets host -> device; history: the previous
//             dump's last T-1 spectra (device to device; the first dump takes
//             its history from its own packets); the dump's spectra zeroed
//             (a sample no packet carries stays 0); the depacketiser decodes
//             each Fix_32_19 word pair (client.cu:898-907) into the slot's F
//             input; record h2d_done[b]
//   compute:  wait h2d_done[b], vis_free[b]; float-input F-engine (FIR + FFT
//             + corner turn, bit-identical to the three drop-in calls) -> X
//             input; record in_free[b]; X-engine -> vis[b] (+ FP64 acc);
//             record x_done[b]
//   copy-out: wait x_done[b]; visibilities and sample counts to the host;
//             record vis_free[b]
//
// Sample indices.  Sample i of a packet with counter cnt is element-sample
// u = cnt*S + i - (T-1)*N (the client's steady-state index, client.cu:953;
// its first block, client.cu:855, is u + (T-1)*N).  Dump d takes u in
// [d*Ns*N, (d+1)*Ns*N), at spectrum T-1 + (u - d*Ns*N) / N of its F input,
// channel u % N; the first dump (d = 0) also takes u in [-(T-1)*N, 0), its
// history spectra.  Samples of another dump (late or early packets),
// malformed headers and payloads are dropped and counted.
#include <new>

#include "common.hpp"
#include "wire.hpp"

namespace mdt {

__global__ void fill_zero_kernel(unsigned long long *p, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0ull;
}

// one thread per (packet, sample, SNAP input); counts: samples written to the
// dump's spectra, to the history spectra, dropped
__global__ void depacket_dump_kernel(const uint32_t *__restrict__ pkts, int npkts, int words_per_pkt, int S,
                                     int NN, int N, int Ns, int T, long dump, float2 *__restrict__ fin,
                                     unsigned long long *__restrict__ counts) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per_pkt = (long)S * 10;  // 10 = most inputs of a SNAP
  const long p = gid / per_pkt;
  if (p >= npkts) return;
  const int i = (int)((gid % per_pkt) / 10), j = (int)(gid % 10);
  const uint32_t *buf = pkts + p * (long)words_per_pkt;
  const uint32_t station = buf[0], snap = buf[1], cnt = buf[2];
  int elem0 = 0, nin = 0;
  const bool hdr = header_ok(station, snap, S, NN, words_per_pkt, &elem0, &nin);
  const long begins = (long)(T - 1) * N, per_dump = (long)Ns * N;
  const long lo = dump * per_dump - (dump == 0 ? begins : 0), hi = (dump + 1) * per_dump;
  if (i == 0 && j == 0) {
    unsigned long long cs = 0, ch = 0, bad = 0;
    for (int ii = 0; ii < S; ii++) {
      const long u = (long)cnt * S + ii - begins;
      if (!hdr || u < lo || u >= hi) bad++;
      else if (u >= dump * per_dump) cs++;
      else ch++;
    }
    if (cs) atomicAdd(&counts[0], cs);
    if (ch) atomicAdd(&counts[1], ch);
    if (bad) atomicAdd(&counts[2], bad);
  }
  if (!hdr || j >= nin) return;
  const long u = (long)cnt * S + i - begins;
  if (u < lo || u >= hi) return;
  const long v = u - dump * per_dump + begins;  // >= 0: row v / N of the F input
  const long dst = ((v / N) * NN + elem0 + j) * (long)N + v % N;
  const long sb = 4 + 2L * ((long)i * nin + j);
  fin[dst] = make_float2(unpack_word(buf[sb]), unpack_word(buf[sb + 1]));
}

}  // namespace mdt

using mdt::fail;

namespace {

enum { kSlots = 2 };

struct PEv {
  hipEvent_t t0 = nullptr, h2d_done = nullptr, f0 = nullptr, in_free = nullptr, x_done = nullptr,
             vis_free = nullptr;
};

}  // namespace

struct mdt_pstream_s {
  mdt_fx_plan plan = nullptr;
  int NN = 0, N = 0, T = 0, Ns = 0;
  int words = 0, S = 0, max_pkts = 0;
  size_t spectrum = 0;  // bytes of one float2 spectrum: NN * N * 8
  size_t vis_bytes = 0, pkt_bytes = 0;
  hipStream_t cin = nullptr, comp = nullptr, cout = nullptr;
  uint32_t *host_pkts[kSlots] = {};
  mdt_complex *host_vis[kSlots] = {};
  unsigned long long *host_counts[kSlots] = {};
  uint32_t *pkts[kSlots] = {};
  float2 *in[kSlots] = {};  // (T-1 + Ns) float2 spectra each
  mdt_complex *vis[kSlots] = {};
  unsigned long long *counts[kSlots] = {};
  void *xin = nullptr;
  size_t xin_bytes = 0;
  PEv ev[kSlots];
  hipStream_t caller = 0;
  hipEvent_t caller_ev = nullptr;
  bool busy[kSlots] = {};
  long dump = 0;  // index of the next submitted dump (0: first block)
  int prev = -1;
};

static void release(mdt_pstream s) {
  if (s->cin) (void)hipStreamSynchronize(s->cin);
  if (s->comp) (void)hipStreamSynchronize(s->comp);
  if (s->cout) (void)hipStreamSynchronize(s->cout);
  for (int b = 0; b < kSlots; b++) {
    if (s->host_pkts[b]) (void)hipHostFree(s->host_pkts[b]);
    if (s->host_vis[b]) (void)hipHostFree(s->host_vis[b]);
    if (s->host_counts[b]) (void)hipHostFree(s->host_counts[b]);
    if (s->pkts[b]) (void)hipFree(s->pkts[b]);
    if (s->in[b]) (void)hipFree(s->in[b]);
    if (s->vis[b]) (void)hipFree(s->vis[b]);
    if (s->counts[b]) (void)hipFree(s->counts[b]);
    hipEvent_t *e[] = {&s->ev[b].t0, &s->ev[b].h2d_done, &s->ev[b].f0,
                       &s->ev[b].in_free, &s->ev[b].x_done, &s->ev[b].vis_free};
    for (hipEvent_t *p : e)
      if (*p) (void)hipEventDestroy(*p);
  }
  if (s->caller_ev) (void)hipEventDestroy(s->caller_ev);
  if (s->xin) (void)hipFree(s->xin);
  if (s->cin) (void)hipStreamDestroy(s->cin);
  if (s->comp) (void)hipStreamDestroy(s->comp);
  if (s->cout) (void)hipStreamDestroy(s->cout);
  delete s;
}

static int check_slot(mdt_pstream s, int slot) {
  if (!s) return fail(MDT_ERR_NULL_POINTER, "packet stream is NULL");
  if (slot < 0 || slot >= kSlots) return fail(MDT_ERR_INVALID_VALUE, "slot %d outside [0, %d)", slot, kSlots);
  return MDT_SUCCESS;
}

static float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  return hipEventElapsedTime(&ms, a, b) == hipSuccess ? ms : -1.f;
}

static int zero(void *p, size_t bytes, hipStream_t st) {  // a kernel, not a memset node
  const long n = (long)(bytes / 8);
  if (n == 0) return MDT_SUCCESS;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(mdt::fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (unsigned long long *)p, n);
  return mdt::check_launch("fill_zero_kernel");
}

extern "C" {

int mdt_pstream_create(mdt_pstream *stream, mdt_fx_plan plan, int words_per_pkt, int SperElement, int max_pkts) {
  MDT_REQUIRE_PTR(stream);
  MDT_REQUIRE_PTR(plan);
  *stream = nullptr;
  if (words_per_pkt < 4 || SperElement < 1 || max_pkts < 1 || (long)SperElement * 10 > (1L << 24))
    return fail(MDT_ERR_INVALID_VALUE, "packet shape (words=%d, SperElement=%d, max_pkts=%d)", words_per_pkt,
                SperElement, max_pkts);
  int xprec = 0;
  mdt_pstream s = new (std::nothrow) mdt_pstream_s;
  if (!s) return fail(MDT_ERR_ALLOC, "packet stream");
  s->plan = plan;
  int st = mdt_fx_plan_shape(plan, &s->NN, &s->N, &s->T, &s->Ns, &xprec);
  if (st == MDT_SUCCESS && s->Ns < s->T - 1)
    st = fail(MDT_ERR_UNSUPPORTED, "streaming needs Nspectra >= Ntaps-1 (%d < %d)", s->Ns, s->T - 1);
  if (st == MDT_SUCCESS && xprec == MDT_XPREC_FP8)
    st = fail(MDT_ERR_UNSUPPORTED, "the packet stream's float samples need the FP32 or FP16-split precision");
  size_t xbytes = 0;
  if (st == MDT_SUCCESS) st = mdt_fx_xinput_bytes(plan, &xbytes);
  s->words = words_per_pkt;
  s->S = SperElement;
  s->max_pkts = max_pkts;
  s->spectrum = sizeof(float2) * (size_t)s->NN * s->N;
  s->vis_bytes = sizeof(mdt_complex) * (size_t)s->N * ((size_t)s->NN * (s->NN + 1) / 2);
  s->pkt_bytes = sizeof(uint32_t) * (size_t)words_per_pkt * max_pkts;
  const size_t fin = s->spectrum * (size_t)(s->T - 1 + s->Ns);
  auto ok = [&](hipError_t e, const char *what) {
    if (st == MDT_SUCCESS && e != hipSuccess)
      st = fail(e == hipErrorOutOfMemory ? MDT_ERR_ALLOC : MDT_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
  };
  if (st == MDT_SUCCESS) {
    ok(hipStreamCreateWithFlags(&s->cin, hipStreamNonBlocking), "copy-in stream");
    ok(hipStreamCreateWithFlags(&s->comp, hipStreamNonBlocking), "compute stream");
    ok(hipStreamCreateWithFlags(&s->cout, hipStreamNonBlocking), "copy-out stream");
    ok(hipEventCreateWithFlags(&s->caller_ev, hipEventDisableTiming), "caller event");
    ok(hipMalloc(&s->xin, xbytes), "X input");
    s->xin_bytes = xbytes;
    for (int b = 0; b < kSlots; b++) {
      ok(hipHostMalloc((void **)&s->host_pkts[b], s->pkt_bytes, hipHostMallocDefault), "pinned packets");
      ok(hipHostMalloc((void **)&s->host_vis[b], s->vis_bytes, hipHostMallocDefault), "pinned visibilities");
      ok(hipHostMalloc((void **)&s->host_counts[b], 3 * sizeof(unsigned long long), hipHostMallocDefault),
         "pinned counts");
      ok(hipMalloc(&s->pkts[b], s->pkt_bytes), "device packets");
      ok(hipMalloc(&s->in[b], fin), "device F input");
      ok(hipMalloc(&s->vis[b], s->vis_bytes), "device visibilities");
      ok(hipMalloc(&s->counts[b], 3 * sizeof(unsigned long long)), "device counts");
      hipEvent_t *e[] = {&s->ev[b].t0, &s->ev[b].h2d_done, &s->ev[b].f0,
                         &s->ev[b].in_free, &s->ev[b].x_done, &s->ev[b].vis_free};
      for (hipEvent_t *p : e) ok(hipEventCreate(p), "event");
    }
  }
  if (st == MDT_SUCCESS && hipDeviceSynchronize() != hipSuccess) st = fail(MDT_ERR_DEVICE, "packet stream set-up");
  if (st != MDT_SUCCESS) {
    release(s);
    return st;
  }
  *stream = s;
  return MDT_SUCCESS;
}

int mdt_pstream_host_packets(mdt_pstream stream, int slot, uint32_t **pkts) {
  MDT_TRY(check_slot(stream, slot));
  MDT_REQUIRE_PTR(pkts);
  *pkts = stream->host_pkts[slot];
  return MDT_SUCCESS;
}

int mdt_pstream_host_vis(mdt_pstream stream, int slot, mdt_complex **vis) {
  MDT_TRY(check_slot(stream, slot));
  MDT_REQUIRE_PTR(vis);
  *vis = stream->host_vis[slot];
  return MDT_SUCCESS;
}

int mdt_pstream_set_caller_stream(mdt_pstream stream, hipStream_t caller) {
  MDT_REQUIRE_PTR(stream);
  stream->caller = caller;
  return MDT_SUCCESS;
}

int mdt_pstream_restart(mdt_pstream stream) {
  MDT_REQUIRE_PTR(stream);
  for (int b = 0; b < kSlots; b++)
    if (stream->busy[b]) return fail(MDT_ERR_INVALID_VALUE, "slot %d has a dump in flight", b);
  stream->dump = 0;
  stream->prev = -1;
  return MDT_SUCCESS;
}

static int enqueue(mdt_pstream s, int slot, int npkts, double *Acc64) {
  size_t xbytes = 0;
  MDT_TRY(mdt_fx_xinput_bytes(s->plan, &xbytes));
  if (xbytes > s->xin_bytes)
    return fail(MDT_ERR_INVALID_VALUE, "the plan's X input (%zu bytes) outgrew the stream's (%zu bytes)", xbytes,
                s->xin_bytes);
  PEv &e = s->ev[slot];
  const size_t hist = s->spectrum * (s->T - 1), stack = s->spectrum * s->Ns;
  char *in = (char *)s->in[slot];
  // copy-in: after the caller's earlier work and the last F-engine that read this slot
  MDT_HIP_TRY(hipEventRecord(s->caller_ev, s->caller));
  MDT_HIP_TRY(hipStreamWaitEvent(s->cin, s->caller_ev, 0));
  MDT_HIP_TRY(hipStreamWaitEvent(s->comp, s->caller_ev, 0));
  MDT_HIP_TRY(hipStreamWaitEvent(s->cin, e.in_free, 0));
  MDT_HIP_TRY(hipEventRecord(e.t0, s->cin));
  if (npkts > 0)
    MDT_HIP_TRY(hipMemcpyAsync(s->pkts[slot], s->host_pkts[slot], sizeof(uint32_t) * (size_t)s->words * npkts,
                               hipMemcpyHostToDevice, s->cin));
  if (s->dump == 0 || s->prev < 0) {
    MDT_TRY(zero(in, hist + stack, s->cin));  // history from this dump's own packets
  } else {
    // the previous dump's last T-1 spectra (read before this slot's spectra
    // are cleared: with prev == slot they lie in them, past the head written)
    if (hist > 0)
      MDT_HIP_TRY(hipMemcpyAsync(in, (const char *)s->in[s->prev] + stack, hist, hipMemcpyDeviceToDevice, s->cin));
    MDT_TRY(zero(in + hist, stack, s->cin));
  }
  MDT_TRY(zero(s->counts[slot], 3 * sizeof(unsigned long long), s->cin));
  if (npkts > 0) {
    const long total = (long)npkts * s->S * 10;
    const long blocks = (total + 255) / 256;
    MDT_TRY(mdt::check_grid(blocks, 256, "depacketiser"));
    hipLaunchKernelGGL(mdt::depacket_dump_kernel, dim3((unsigned)blocks), dim3(256), 0, s->cin, s->pkts[slot], npkts,
                       s->words, s->S, s->NN, s->N, s->Ns, s->T, s->dump, s->in[slot], s->counts[slot]);
    MDT_TRY(mdt::check_launch("depacket_dump_kernel"));
  }
  MDT_HIP_TRY(hipEventRecord(e.h2d_done, s->cin));
  // compute
  MDT_HIP_TRY(hipStreamWaitEvent(s->comp, e.h2d_done, 0));
  MDT_HIP_TRY(hipStreamWaitEvent(s->comp, e.vis_free, 0));
  MDT_HIP_TRY(hipEventRecord(e.f0, s->comp));
  MDT_TRY(mdt_fx_set_stream(s->plan, s->comp));
  MDT_TRY(mdt_fx_fengine_f32(s->plan, (const mdt_complex *)s->in[slot], (mdt_complex *)s->xin));
  MDT_HIP_TRY(hipEventRecord(e.in_free, s->comp));
  MDT_TRY(mdt_fx_xengine(s->plan, (const mdt_complex *)s->xin, s->vis[slot], Acc64));
  MDT_HIP_TRY(hipEventRecord(e.x_done, s->comp));
  // copy-out
  MDT_HIP_TRY(hipStreamWaitEvent(s->cout, e.x_done, 0));
  MDT_HIP_TRY(hipMemcpyAsync(s->host_vis[slot], s->vis[slot], s->vis_bytes, hipMemcpyDeviceToHost, s->cout));
  MDT_HIP_TRY(hipMemcpyAsync(s->host_counts[slot], s->counts[slot], 3 * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, s->cout));
  MDT_HIP_TRY(hipEventRecord(e.vis_free, s->cout));
  return MDT_SUCCESS;
}

int mdt_pstream_submit(mdt_pstream stream, int slot, int npkts, double *Acc64) {
  MDT_TRY(check_slot(stream, slot));
  mdt_pstream s = stream;
  if (s->busy[slot]) return fail(MDT_ERR_INVALID_VALUE, "slot %d submitted again before mdt_pstream_wait", slot);
  if (npkts < 0 || npkts > s->max_pkts)
    return fail(MDT_ERR_INVALID_VALUE, "npkts=%d outside [0, %d]", npkts, s->max_pkts);
  const int st = enqueue(s, slot, npkts, Acc64);
  if (st != MDT_SUCCESS) {
    (void)hipStreamSynchronize(s->cin);
    (void)hipStreamSynchronize(s->comp);
    (void)hipStreamSynchronize(s->cout);
    return st;
  }
  s->busy[slot] = true;
  s->prev = slot;
  s->dump++;
  return MDT_SUCCESS;
}

int mdt_pstream_wait(mdt_pstream stream, int slot, mdt_dump_times *times, unsigned long long *counts) {
  MDT_TRY(check_slot(stream, slot));
  mdt_pstream s = stream;
  if (!s->busy[slot]) return fail(MDT_ERR_INVALID_VALUE, "slot %d has no dump in flight", slot);
  MDT_HIP_TRY(hipEventSynchronize(s->ev[slot].vis_free));
  s->busy[slot] = false;
  if (times) {
    const PEv &e = s->ev[slot];
    times->h2d_ms = elapsed(e.t0, e.h2d_done);
    times->fengine_ms = elapsed(e.f0, e.in_free);
    times->xengine_ms = elapsed(e.in_free, e.x_done);
    times->d2h_ms = elapsed(e.x_done, e.vis_free);
    times->total_ms = elapsed(e.t0, e.vis_free);
  }
  if (counts)
    for (int k = 0; k < 3; k++) counts[k] = s->host_counts[slot][k];
  return MDT_SUCCESS;
}

int mdt_pstream_destroy(mdt_pstream stream) {
  if (!stream) return MDT_SUCCESS;
  (void)mdt_fx_set_stream(stream->plan, 0);
  release(stream);
  return MDT_SUCCESS;
}

}  // extern "C"
