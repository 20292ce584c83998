// Fast trace forms over the 4-wide SoA BVH (placeholder until implemented).
#include "rtn_internal.hpp"

namespace rtn {
int trace_wide_scene(const rtn_scene*, const float* const[6], int64_t, const rtn_trace_params*, const rtn_hits*,
                     unsigned long long*, hipStream_t) {
  set_error("rtn_trace: wide/persistent forms not built yet");
  return RTN_EUNSUPPORTED;
}
}  // namespace rtn
