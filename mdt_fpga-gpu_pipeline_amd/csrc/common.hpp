// common.hpp -- shared helpers of libmdtfx: status/error reporting, argument
// validation against the reference limits, complex arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdint>

#include "mdtfx.h"

namespace mdt {

// Thread-local error text, returned by mdt_last_error_string().
void set_error_text(const char *fmt, ...);

// Record an error and return its status code.
int fail(int code, const char *fmt, ...);

// Status of the last launch on this thread (hipGetLastError), as MDT_ERR_LAUNCH.
int check_launch(const char *what);

inline bool is_pow2(long v) { return v > 0 && (v & (v - 1)) == 0; }

// The dispatch packet carries each grid dimension in work-items as a 32-bit
// value: a 1-D launch of `blocks` x `threads` must stay below 2^32 work-items
// (beyond that the grid silently wraps).  Launches whose size follows the data
// check this first.
int check_grid(long blocks, long threads, const char *what);

inline int ilog2(long v) {
  int l = 0;
  while ((1L << l) < v) l++;
  return l;
}

// Reference limits, MDT_OneStation_client.h:54-61.
int check_inputs(int NN);
int check_channels(int N);
int check_taps(int T);
int check_spectra(int Ns);

// Non-temporal (streaming) store of one 8-byte element, for the drop-in
// calls' outputs: each is written once, read back only by a later launch,
// and at the station's shape (2.7 GB) far larger than the caches (PpS_Batch
// 1.04 -> 0.96 ms, C2C FFT 1.375 -> 1.35 ms, corner turn 1.235 -> 1.11 ms).
__device__ __forceinline__ void store_stream(float2 *p, float2 v) {
  typedef float f32x2_t __attribute__((ext_vector_type(2)));
  __builtin_nontemporal_store(f32x2_t{v.x, v.y}, (f32x2_t *)p);
}
__device__ __forceinline__ void store_stream(unsigned long long *p, unsigned long long v) {
  __builtin_nontemporal_store(v, p);
}

// -------------------------------------------------- native X input geometry
// Sample (channel k, input e, spectrum s) of the fused path's X input sits at
// element chan(k) + at(e, s), in samples:
//   chan(k)  = (k >> lgG) * cp + (k mod 2^lgG) * 2^lgB
//   at(e, s) = (s >> lgB) * bp + e * rp + (s mod 2^lgB)
// Rows layout (lgG = 0, lgB = 30, bp = 0): k*cp + e*rp + s.  Pairs layout
// (lgG = 1): the two channels of a pair side by side per input, spectra in
// blocks of B = 2^lgB (64 bytes of samples), rp = 2B, cp = NN*2B, bp =
// ceil(N/2)*cp, so that B consecutive spectra of two adjacent channels of one
// input are one 128-byte piece.
struct XGeom {
  long rp = 0, cp = 0, bp = 0;
  int lgB = 30, lgG = 0;
  __host__ __device__ long chan(int k) const {
    return (long)(k >> lgG) * cp + ((long)(k & ((1 << lgG) - 1)) << lgB);
  }
  __host__ __device__ long at(int e, int s) const {
    return (long)(s >> lgB) * bp + (long)e * rp + (s & ((1 << lgB) - 1));
  }
  // spectrum part of at() (for loads whose row part is added separately)
  __host__ __device__ long spec(int s) const { return (long)(s >> lgB) * bp + (s & ((1 << lgB) - 1)); }
};

// ----------------------------------------------------------- complex helpers
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// The library is compiled with -ffp-contract=off: every fused multiply-add is
// written out, so the same expression rounds identically in every kernel that
// inlines it (fused and drop-in paths stay bit-identical).
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// Exact-argument sin(pi x), cos(pi x) in FP64: exact zeros at the integers
// (sin) and half-integers (cos), so windowed-sinc zeros are exact.
__device__ __forceinline__ double dsinpi(double x) {
  double n = rint(x);
  double r = x - n;
  if (r == 0.0) return 0.0;
  double s = (fabs(r) <= 0.25) ? sin(M_PI * r) : copysign(cos(M_PI * (0.5 - fabs(r))), r);
  return (fmod(fabs(n), 2.0) == 1.0) ? -s : s;
}
__device__ __forceinline__ double dcospi(double x) {
  double n = rint(x);
  double r = fabs(x - n);
  double c = (r <= 0.25) ? cos(M_PI * r) : sin(M_PI * (0.5 - r));
  return (fmod(fabs(n), 2.0) == 1.0) ? -c : c;
}

}  // namespace mdt

#define MDT_HIP_TRY(call)                                                              \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return ::mdt::fail(MDT_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define MDT_REQUIRE_PTR(p)                                                  \
  do {                                                                      \
    if ((p) == nullptr) return ::mdt::fail(MDT_ERR_NULL_POINTER, "%s is NULL", #p); \
  } while (0)

#define MDT_TRY(expr)            \
  do {                           \
    int s_ = (expr);             \
    if (s_ != MDT_SUCCESS) return s_; \
  } while (0)
