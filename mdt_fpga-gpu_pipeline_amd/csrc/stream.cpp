// stream.cpp -- the streaming correlator: consecutive dumps through one plan,
// with the input copies, the F+X compute and the visibility copies on three
// HIP streams ordered by events.  It replaces the client's four MPI ranks
// (MDT_OneStation_client.cu:916-1276), which pass d_FInput between processes
// and run every copy and kernel on one stream per process (client.cu:333-334).
//
// Per dump from slot b (prev = the slot of the dump before, or the prime
// buffer):
//   copy-in:  wait in_free[b] (the last F-engine that read in[b]);
//             in[b][0, T-1) <- prev's last T-1 spectra (device to device);
//             in[b][T-1, T-1+Ns) <- host stack b;  record h2d_done[b]
//   compute:  wait h2d_done[b], vis_free[b]; F-engine in[b] -> X input;
//             record in_free[b]; X-engine -> vis[b] (+ FP64 accumulator);
//             record x_done[b]
//   copy-out: wait x_done[b]; host vis b <- vis[b]; record vis_free[b]
// The history copy runs before the stack copy, so with prev == b (one slot
// used twice in a row) the tail it reads is not yet overwritten: the tail
// starts at spectrum Ns >= T-1, past the head it writes.  The one X input is
// shared by all dumps: its writer (F) and reader (X) are on the compute
// stream, which orders them.  The three streams are non-blocking, so before
// anything else a dump's copy-in and compute streams wait for an event
// recorded on the caller's stream at submit: what the caller queued there
// first (zeroing the FP64 accumulator, writing gains) is complete before the
// dump reads it.
#include <new>

#include "common.hpp"

using mdt::fail;

namespace {

enum { kSlots = 2 };

struct Ev {
  hipEvent_t t0 = nullptr, h2d_done = nullptr, f0 = nullptr, in_free = nullptr, x_done = nullptr,
             vis_free = nullptr;
};

}  // namespace

struct mdt_stream_s {
  mdt_fx_plan plan = nullptr;
  int NN = 0, N = 0, T = 0, Ns = 0;
  size_t spectrum = 0;  // bytes of one input spectrum: NN * N * 2
  size_t vis_bytes = 0;
  hipStream_t cin = nullptr, comp = nullptr, cout = nullptr;
  int8_t *host_stack[kSlots] = {};
  mdt_complex *host_vis[kSlots] = {};
  int8_t *in[kSlots] = {};  // (T-1 + Ns) spectra each
  int8_t *prime = nullptr;  // T-1 spectra: history of the first dump
  void *xin = nullptr;
  size_t xin_bytes = 0;  // the X input allocated for the plan as it was at creation
  mdt_complex *vis[kSlots] = {};
  Ev ev[kSlots];
  hipStream_t caller = 0;        // submits order after the work queued here
  hipEvent_t caller_ev = nullptr;
  bool busy[kSlots] = {};
  int prev = -1;  // slot of the last submitted dump, -1: the prime buffer
};

static void release(mdt_stream s) {
  if (s->cin) (void)hipStreamSynchronize(s->cin);
  if (s->comp) (void)hipStreamSynchronize(s->comp);
  if (s->cout) (void)hipStreamSynchronize(s->cout);
  for (int b = 0; b < kSlots; b++) {
    if (s->host_stack[b]) (void)hipHostFree(s->host_stack[b]);
    if (s->host_vis[b]) (void)hipHostFree(s->host_vis[b]);
    if (s->in[b]) (void)hipFree(s->in[b]);
    if (s->vis[b]) (void)hipFree(s->vis[b]);
    hipEvent_t *e[] = {&s->ev[b].t0, &s->ev[b].h2d_done, &s->ev[b].f0,
                       &s->ev[b].in_free, &s->ev[b].x_done, &s->ev[b].vis_free};
    for (hipEvent_t *p : e)
      if (*p) (void)hipEventDestroy(*p);
  }
  if (s->caller_ev) (void)hipEventDestroy(s->caller_ev);
  if (s->prime) (void)hipFree(s->prime);
  if (s->xin) (void)hipFree(s->xin);
  if (s->cin) (void)hipStreamDestroy(s->cin);
  if (s->comp) (void)hipStreamDestroy(s->comp);
  if (s->cout) (void)hipStreamDestroy(s->cout);
  delete s;
}

static int check_slot(mdt_stream s, int slot) {
  if (!s) return fail(MDT_ERR_NULL_POINTER, "stream is NULL");
  if (slot < 0 || slot >= kSlots) return fail(MDT_ERR_INVALID_VALUE, "slot %d outside [0, %d)", slot, kSlots);
  return MDT_SUCCESS;
}

static float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  return hipEventElapsedTime(&ms, a, b) == hipSuccess ? ms : -1.f;
}

extern "C" {

int mdt_stream_create(mdt_stream *stream, mdt_fx_plan plan) {
  MDT_REQUIRE_PTR(stream);
  MDT_REQUIRE_PTR(plan);
  *stream = nullptr;
  mdt_stream s = new (std::nothrow) mdt_stream_s;
  if (!s) return fail(MDT_ERR_ALLOC, "stream");
  s->plan = plan;
  int st = mdt_fx_plan_shape(plan, &s->NN, &s->N, &s->T, &s->Ns, nullptr);
  if (st == MDT_SUCCESS && s->Ns < s->T - 1)
    st = fail(MDT_ERR_UNSUPPORTED, "streaming needs Nspectra >= Ntaps-1 (%d < %d)", s->Ns, s->T - 1);
  size_t xbytes = 0;
  if (st == MDT_SUCCESS) st = mdt_fx_xinput_bytes(plan, &xbytes);
  s->spectrum = (size_t)s->NN * s->N * 2;
  s->vis_bytes = sizeof(mdt_complex) * (size_t)s->N * ((size_t)s->NN * (s->NN + 1) / 2);
  const size_t stack = s->spectrum * s->Ns, hist = s->spectrum * (s->T - 1);
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
    ok(hipMalloc(&s->prime, hist > 0 ? hist : 16), "history buffer");
    if (st == MDT_SUCCESS && hist > 0) ok(hipMemset(s->prime, 0, hist), "history zeroing");
    for (int b = 0; b < kSlots; b++) {
      ok(hipHostMalloc((void **)&s->host_stack[b], stack, hipHostMallocDefault), "pinned input stack");
      ok(hipHostMalloc((void **)&s->host_vis[b], s->vis_bytes, hipHostMallocDefault), "pinned visibilities");
      ok(hipMalloc(&s->in[b], stack + hist), "device input");
      ok(hipMalloc(&s->vis[b], s->vis_bytes), "device visibilities");
      hipEvent_t *e[] = {&s->ev[b].t0, &s->ev[b].h2d_done, &s->ev[b].f0,
                         &s->ev[b].in_free, &s->ev[b].x_done, &s->ev[b].vis_free};
      for (hipEvent_t *p : e) ok(hipEventCreate(p), "event");
    }
  }
  if (st == MDT_SUCCESS) st = mdt_fx_set_stream(plan, s->comp);
  if (st == MDT_SUCCESS && hipDeviceSynchronize() != hipSuccess) st = fail(MDT_ERR_DEVICE, "stream set-up");
  if (st != MDT_SUCCESS) {
    release(s);
    return st;
  }
  *stream = s;
  return MDT_SUCCESS;
}

int mdt_stream_host_stack(mdt_stream stream, int slot, int8_t **stack) {
  MDT_TRY(check_slot(stream, slot));
  MDT_REQUIRE_PTR(stack);
  *stack = stream->host_stack[slot];
  return MDT_SUCCESS;
}

int mdt_stream_host_vis(mdt_stream stream, int slot, mdt_complex **vis) {
  MDT_TRY(check_slot(stream, slot));
  MDT_REQUIRE_PTR(vis);
  *vis = stream->host_vis[slot];
  return MDT_SUCCESS;
}

int mdt_stream_prime(mdt_stream stream, const int8_t *history) {
  MDT_REQUIRE_PTR(stream);
  mdt_stream s = stream;
  for (int b = 0; b < kSlots; b++)
    if (s->busy[b]) return fail(MDT_ERR_INVALID_VALUE, "slot %d has a dump in flight", b);
  const size_t hist = s->spectrum * (s->T - 1);
  if (hist > 0) {
    if (history)
      MDT_HIP_TRY(hipMemcpy(s->prime, history, hist, hipMemcpyHostToDevice));
    else
      MDT_HIP_TRY(hipMemset(s->prime, 0, hist));
  }
  s->prev = -1;
  return MDT_SUCCESS;
}

// the enqueue sequence of one dump (see the top of this file)
static int enqueue(mdt_stream s, int slot, double *Acc64) {
  size_t xbytes = 0;  // the plan's layout or precision may have changed since creation
  MDT_TRY(mdt_fx_xinput_bytes(s->plan, &xbytes));
  if (xbytes > s->xin_bytes)
    return fail(MDT_ERR_INVALID_VALUE, "the plan's X input (%zu bytes) outgrew the stream's (%zu bytes)", xbytes,
                s->xin_bytes);
  Ev &e = s->ev[slot];
  const size_t hist = s->spectrum * (s->T - 1), stack = s->spectrum * s->Ns;
  const int8_t *src = s->prev < 0 ? s->prime : s->in[s->prev] + stack;  // previous dump's last T-1 spectra
  // the caller's earlier work (accumulator zeroing, gains) comes first
  MDT_HIP_TRY(hipEventRecord(s->caller_ev, s->caller));
  MDT_HIP_TRY(hipStreamWaitEvent(s->cin, s->caller_ev, 0));
  MDT_HIP_TRY(hipStreamWaitEvent(s->comp, s->caller_ev, 0));
  // copy-in
  MDT_HIP_TRY(hipStreamWaitEvent(s->cin, e.in_free, 0));
  MDT_HIP_TRY(hipEventRecord(e.t0, s->cin));
  if (hist > 0) MDT_HIP_TRY(hipMemcpyAsync(s->in[slot], src, hist, hipMemcpyDeviceToDevice, s->cin));
  MDT_HIP_TRY(hipMemcpyAsync(s->in[slot] + hist, s->host_stack[slot], stack, hipMemcpyHostToDevice, s->cin));
  MDT_HIP_TRY(hipEventRecord(e.h2d_done, s->cin));
  // compute
  MDT_HIP_TRY(hipStreamWaitEvent(s->comp, e.h2d_done, 0));
  MDT_HIP_TRY(hipStreamWaitEvent(s->comp, e.vis_free, 0));
  MDT_HIP_TRY(hipEventRecord(e.f0, s->comp));
  MDT_TRY(mdt_fx_set_stream(s->plan, s->comp));  // whatever the plan was last bound to
  MDT_TRY(mdt_fx_fengine(s->plan, s->in[slot], (mdt_complex *)s->xin));
  MDT_HIP_TRY(hipEventRecord(e.in_free, s->comp));
  MDT_TRY(mdt_fx_xengine(s->plan, (const mdt_complex *)s->xin, s->vis[slot], Acc64));
  MDT_HIP_TRY(hipEventRecord(e.x_done, s->comp));
  // copy-out
  MDT_HIP_TRY(hipStreamWaitEvent(s->cout, e.x_done, 0));
  MDT_HIP_TRY(hipMemcpyAsync(s->host_vis[slot], s->vis[slot], s->vis_bytes, hipMemcpyDeviceToHost, s->cout));
  MDT_HIP_TRY(hipEventRecord(e.vis_free, s->cout));
  return MDT_SUCCESS;
}

int mdt_stream_set_caller_stream(mdt_stream stream, hipStream_t caller) {
  MDT_REQUIRE_PTR(stream);
  stream->caller = caller;
  return MDT_SUCCESS;
}

int mdt_stream_submit(mdt_stream stream, int slot, double *Acc64) {
  MDT_TRY(check_slot(stream, slot));
  mdt_stream s = stream;
  if (s->busy[slot]) return fail(MDT_ERR_INVALID_VALUE, "slot %d submitted again before mdt_stream_wait", slot);
  const int st = enqueue(s, slot, Acc64);
  if (st != MDT_SUCCESS) {
    // part of the dump may be queued: let it drain, so that the slot's
    // buffers are free again when this returns (the error text is kept)
    (void)hipStreamSynchronize(s->cin);
    (void)hipStreamSynchronize(s->comp);
    (void)hipStreamSynchronize(s->cout);
    return st;
  }
  s->busy[slot] = true;
  s->prev = slot;
  return MDT_SUCCESS;
}

int mdt_stream_wait(mdt_stream stream, int slot, mdt_dump_times *times) {
  MDT_TRY(check_slot(stream, slot));
  mdt_stream s = stream;
  if (!s->busy[slot]) return fail(MDT_ERR_INVALID_VALUE, "slot %d has no dump in flight", slot);
  MDT_HIP_TRY(hipEventSynchronize(s->ev[slot].vis_free));
  s->busy[slot] = false;
  if (times) {
    const Ev &e = s->ev[slot];
    times->h2d_ms = elapsed(e.t0, e.h2d_done);
    times->fengine_ms = elapsed(e.f0, e.in_free);
    times->xengine_ms = elapsed(e.in_free, e.x_done);
    times->d2h_ms = elapsed(e.x_done, e.vis_free);
    times->total_ms = elapsed(e.t0, e.vis_free);
  }
  return MDT_SUCCESS;
}

int mdt_stream_destroy(mdt_stream stream) {
  if (!stream) return MDT_SUCCESS;
  (void)mdt_fx_set_stream(stream->plan, 0);
  release(stream);
  return MDT_SUCCESS;
}

}  // extern "C"
