// CPU build of the wing gather's queue record (kurucz_py_joss_submission_amd/csrc/gather_queue.h)
// for checking its bit layout at the bounds without a GPU.  Test infrastructure.
#include "gather_queue.h"

extern "C" {
int gq_roundtrip(int dc, int r, int m, long* out) {
  const synthe::gq::Fields f = synthe::gq::unpack(synthe::gq::pack(dc, r, m));
  out[0] = f.dc;
  out[1] = f.r;
  out[2] = f.m;
  return 0;
}
long gq_max_zone(void) { return synthe::gq::kMaxZone; }
long gq_max_tile(void) { return synthe::gq::kMaxTile; }
}
