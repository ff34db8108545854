// Default schedules of the hand-written kernels, and the bf16 MFMA SpMM
// entry used when the MFMA translation unit is not part of the build.

#include "common.h"

extern "C" void taco_hip_default_schedule(int32_t op, int32_t m, int64_t nnz,
                                          int32_t k, taco_hip_schedule* out) {
  taco_hip_schedule s = {};
  (void)m;
  (void)nnz;
  if (op == 0) {
    // SpMV: merge path, 8 merge items per thread (2048 per workgroup).
    s.algo = TACO_HIP_MERGE_PATH;
    s.block_size = 256;
    s.items_per_thread = 8;
  } else {
    // SpMM: merge path per lane team, lanes along k, 256 merge items per
    // team (long teams keep the carries of long rows few).
    s.algo = TACO_HIP_MERGE_PATH;
    s.block_size = 256;
    s.items_per_thread = 256;
    int g = 8;
    while (g < k && g < 64) g <<= 1;
    s.group_size = g;
  }
  *out = s;
}
