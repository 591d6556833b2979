// wire.hpp -- the SNAP wire format shared by the unpack and depacketiser
// kernels.
#pragma once

#include <cstdint>

#include "common.hpp"

namespace mdt {

// One Fix_32_19 word -> float, MDT_OneStation_client.cu:898-907:
// w >= 2^31 -> -2^-19 * (int)(2^32 - w), else 2^-19 * (int)w, the product in
// double, stored as float.  (int)(2^31) wraps to INT_MIN, so 0x80000000 maps
// to +4096.0 exactly as in the reference.
__device__ __forceinline__ float unpack_word(uint32_t w) {
  if (w >= 2147483648u) {
    int v = (int)(uint32_t)(0u - w);  // 2^32 - w, modulo 2^32
    return (float)(-1.9073486328125e-06 * (double)v);
  }
  return (float)(1.9073486328125e-06 * (double)(int)w);
}

// station / SNAP geometry, MDT_OneStation_client.h:64-68
constexpr int kInputsPerStation = 38;
constexpr int kSnapsPerStation = 4;
__device__ constexpr int kInputsPerSnap[kSnapsPerStation] = {10, 10, 10, 8};

// Packet header check (uniform per packet): the station and SNAP indices
// give the first element station*38 + sum of the lower SNAPs' inputs
// (MDT_OneStation_client.cu:877-885) and the input count; the packet must
// hold S samples of that many inputs and the elements must exist.
__device__ __forceinline__ bool header_ok(uint32_t station, uint32_t snap, int S, int NN, int words_per_pkt,
                                          int *elem0, int *nin) {
  if (snap >= (uint32_t)kSnapsPerStation || station >= 65536u) return false;
  int e = 0;
  for (uint32_t k = 0; k < snap; k++) e += kInputsPerSnap[k];
  const long e0 = (long)station * kInputsPerStation + e;
  *nin = kInputsPerSnap[snap];
  *elem0 = (int)e0;
  return e0 + *nin <= NN && 4 + 2L * S * *nin <= words_per_pkt;
}

}  // namespace mdt
