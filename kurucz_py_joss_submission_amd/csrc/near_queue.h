// The record of one queued (list entry, pixel) pair of the Voigt-zone kernels
// (k_near): 64 bits, packed and unpacked here so that a CPU build can check
// the layout at its bounds.
//
//   bits  0-31  the entry's index within its depth's lists (< 2^32)
//   bits 32-51  |k|, the pixel's offset from the line centre, 1 <= |k| <= the
//               zone's half-width <= MAXPROF = 10^6 < 2^20
//   bits 52-57  the pixel's lane in the wave (0..63)
//   bits 58-63  a tag (0..63): the pixel's slot in k_near
//
// The header compiles as HIP device code and as plain C++ (for CPU checks).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define SYNTHE_NQ_HD __host__ __device__ __forceinline__
#else
#define SYNTHE_NQ_HD inline
#endif

namespace synthe {
namespace nq {

constexpr int kOffsetBits = 20;
constexpr long kMaxZone = 1000000;  // MAXPROF: the set-up clamps every Voigt zone to it
static_assert(kMaxZone < (1L << kOffsetBits), "a zone half-width must fit the offset field");

struct Record {
  long j_rel;  // entry index within the depth's lists
  long ka;     // |k|
  int lane, tag;
};

SYNTHE_NQ_HD uint64_t pack(long j_rel, long ka, int lane, int tag) {
  return (uint64_t)(uint32_t)j_rel | ((uint64_t)ka << 32) | ((uint64_t)lane << 52) | ((uint64_t)tag << 58);
}

SYNTHE_NQ_HD Record unpack(uint64_t r) {
  Record q;
  q.j_rel = (long)(r & 0xffffffffull);
  q.ka = (long)((r >> 32) & ((1ull << kOffsetBits) - 1));
  q.lane = (int)((r >> 52) & 63);
  q.tag = (int)(r >> 58);
  return q;
}

}  // namespace nq
}  // namespace synthe
