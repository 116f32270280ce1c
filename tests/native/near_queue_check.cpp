// CPU build of the Voigt-zone queue record (kurucz_py_joss_submission_amd/csrc/near_queue.h)
// for checking its bit layout at the bounds without a GPU.  Test infrastructure.
#include "near_queue.h"

extern "C" {
int nq_roundtrip(long j_rel, long ka, int lane, int tag, long* out) {
  const synthe::nq::Record r = synthe::nq::unpack(synthe::nq::pack(j_rel, ka, lane, tag));
  out[0] = r.j_rel;
  out[1] = r.ka;
  out[2] = r.lane;
  out[3] = r.tag;
  return 0;
}
long nq_max_zone(void) { return synthe::nq::kMaxZone; }
}
