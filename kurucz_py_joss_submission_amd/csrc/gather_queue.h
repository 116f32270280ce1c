// The record of one queued list entry of the wing gather (k_gather's LDS
// queue of the entries it evaluates by rows): 16 bytes, packed and unpacked
// here so that a CPU build can check the layout at its bounds.
//
//   xs               the far wing's X as the list head stores it (a double)
//   pk bits  0-20    dc = c - t0, the line centre relative to the tile's first
//                    pixel, 21-bit two's complement
//   pk bits 21-40    r, the Voigt zone's half-width (<= MAXSTEP)
//   pk bits 41-60    MAXSTEP (<= MAXPROF = 10^6 < 2^20)
//
// An entry that reaches a tile of T pixels has -MAXSTEP <= dc <= MAXSTEP + T - 1,
// so |dc| < 2^20 for tiles up to 2^12 pixels.  The header compiles as HIP
// device code and as plain C++ (for CPU checks).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define SYNTHE_GQ_HD __host__ __device__ __forceinline__
#else
#define SYNTHE_GQ_HD inline
#endif

namespace synthe {
namespace gq {

constexpr int kFieldBits = 20;
constexpr long kMaxZone = 1000000;  // MAXPROF: the set-up clamps every MAXSTEP to it
constexpr long kMaxTile = 4096;     // pixels per tile the centre field allows
static_assert(kMaxZone + kMaxTile < (1L << kFieldBits), "a centre offset must fit 21 signed bits");

struct Fields {
  int dc, r, m;
};

SYNTHE_GQ_HD uint64_t pack(int dc, int r, int m) {
  return (uint64_t)((uint32_t)dc & 0x1fffffu) | ((uint64_t)(uint32_t)r << 21) | ((uint64_t)(uint32_t)m << 41);
}

SYNTHE_GQ_HD Fields unpack(uint64_t pk) {
  Fields f;
  f.dc = ((int)(uint32_t)(pk << 11)) >> 11;  // bits 0-20, sign-extended
  f.r = (int)((pk >> 21) & 0xfffffu);
  f.m = (int)((pk >> 41) & 0xfffffu);
  return f;
}

}  // namespace gq
}  // namespace synthe
