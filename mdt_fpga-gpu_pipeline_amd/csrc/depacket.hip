// depacket.hip -- GPU depacketiser: SNAP packets -> the F-input stacks, the
// work of the client's rank-1 loop (MDT_OneStation_client.cu:846-911 for the
// first block, 945-1005 in steady state) done on the device.
//
// Packet p (words_per_pkt uint32 words): word 0 station, word 1 SNAP, word 2
// packet counter, word 3 unused, then SperElement x InputsPerSNAP[snap]
// (re, im) word pairs, sample-major (client.cu:888-891).  Sample i of the
// packet has the element-sample index
//   first block:  cnt*SperElement + i; below (T-1)*N it is history (stack 1),
//                 else minus (T-1)*N in stack 0;
//   steady state: cnt*SperElement - (T-1)*N + i, stack (index / (Ns*N)) % 2,
//                 spectrum (index / N) % Ns;
// channel index % N; element station*38 + sum of lower SNAPs' inputs + j.
// Destination Stacks[stack][spectrum][element][channel] (complex64).
//
// One thread per (packet, sample, SNAP input).  counts[0] and counts[1] add
// the samples that went to stacks 0 and 1 (the reference's Nsamples[]),
// counts[2] the samples dropped: headers come from the network, and unlike the
// reference a sample whose station, SNAP, payload size or index would land
// outside the stacks is dropped, never written.
#include "common.hpp"
#include "wire.hpp"

namespace mdt {

// where sample i of a packet goes: stack and spectrum, or false when it would
// land outside the stacks
struct SampleDest {
  long sample, stack, spec;
  bool ok;
};
__device__ __forceinline__ SampleDest sample_dest(uint32_t cnt, int i, int S, int N, int Ns, int T,
                                                  int first_block) {
  const long begins = (long)(T - 1) * N, per_stack = (long)Ns * N;
  SampleDest d;
  if (first_block) {
    d.sample = (long)cnt * S + i;
    if (d.sample < begins) {
      d.stack = 1;
    } else {
      d.sample -= begins;
      d.stack = 0;
    }
    d.spec = d.sample / N;
    d.ok = d.spec < Ns;
  } else {
    d.sample = (long)cnt * S - begins + i;
    d.ok = d.sample >= 0;
    d.stack = d.ok ? (d.sample / per_stack) % 2 : 0;
    d.spec = d.ok ? (d.sample / N) % Ns : 0;
  }
  return d;
}

__global__ void depacket_kernel(const uint32_t *__restrict__ pkts, int npkts, int words_per_pkt, int S,
                                int NN, int N, int Ns, int T, int first_block, float2 *__restrict__ stacks,
                                unsigned long long *__restrict__ counts) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per_pkt = (long)S * 10;  // (sample, input) slots per packet; 10 = most inputs of a SNAP
  const long p = gid / per_pkt;
  if (p >= npkts) return;
  const int i = (int)((gid % per_pkt) / 10), j = (int)(gid % 10);
  const uint32_t *buf = pkts + p * (long)words_per_pkt;
  const uint32_t station = buf[0], snap = buf[1], cnt = buf[2];
  int elem0 = 0, nin = 0;
  const bool hdr = header_ok(station, snap, S, NN, words_per_pkt, &elem0, &nin);
  if (i == 0 && j == 0) {  // the packet's sample counts: at most three atomics
    unsigned long long c0 = 0, c1 = 0, bad = 0;
    for (int ii = 0; ii < S; ii++) {
      const SampleDest d = sample_dest(cnt, ii, S, N, Ns, T, first_block);
      if (!hdr || !d.ok) bad++;
      else if (d.stack == 0) c0++;
      else c1++;
    }
    if (c0) atomicAdd(&counts[0], c0);
    if (c1) atomicAdd(&counts[1], c1);
    if (bad) atomicAdd(&counts[2], bad);
  }
  if (!hdr || j >= nin) return;
  const SampleDest d = sample_dest(cnt, i, S, N, Ns, T, first_block);
  if (!d.ok) return;
  const long dst = (d.stack * Ns + d.spec) * (long)NN * N + (long)(elem0 + j) * N + d.sample % N;
  const long sb = 4 + 2L * ((long)i * nin + j);
  stacks[dst] = make_float2(unpack_word(buf[sb]), unpack_word(buf[sb + 1]));
}

}  // namespace mdt

using namespace mdt;

extern "C" {

int mdt_depacketise(const uint32_t *pkts, int npkts, int words_per_pkt, int SperElement, int NpolsxNelements,
                    int Nchannels, int Nspectra, int Ntaps, int first_block, mdt_complex *Stacks,
                    unsigned long long *counts, hipStream_t stream) {
  MDT_REQUIRE_PTR(pkts);
  MDT_REQUIRE_PTR(Stacks);
  MDT_REQUIRE_PTR(counts);
  MDT_TRY(check_inputs(NpolsxNelements));
  MDT_TRY(check_channels(Nchannels));
  MDT_TRY(check_spectra(Nspectra));
  MDT_TRY(check_taps(Ntaps));
  if (npkts < 0 || SperElement < 1 || words_per_pkt < 4)
    return fail(MDT_ERR_INVALID_VALUE, "packet batch shape (npkts=%d, SperElement=%d, words=%d)", npkts,
                SperElement, words_per_pkt);
  if ((long)SperElement * 10 > (1L << 24)) return fail(MDT_ERR_INVALID_VALUE, "SperElement too large");
  if (npkts == 0) return MDT_SUCCESS;
  const long total = (long)npkts * SperElement * 10;
  const long blocks = (total + 255) / 256;
  MDT_TRY(check_grid(blocks, 256, "depacketiser"));
  hipLaunchKernelGGL(depacket_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, pkts, npkts, words_per_pkt,
                     SperElement, NpolsxNelements, Nchannels, Nspectra, Ntaps, first_block, (float2 *)Stacks,
                     counts);
  return check_launch("depacket_kernel");
}

}  // extern "C"
