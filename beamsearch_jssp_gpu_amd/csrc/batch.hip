// batch.hip -- batched solves (placeholder: sequential loop over jssp_solve).
#include <hip/hip_runtime.h>
#include "engine.hpp"

namespace jssp {
int solve_any(jssp_engine* e, const jssp_instance* in, const jssp_params* p, jssp_result* res,
              int32_t* trail_out, int32_t* pcount_out);
}

extern "C" int jssp_solve_batch(jssp_engine* e, int32_t count, const jssp_instance* insts,
                                const jssp_params* params, jssp_result* results) {
  if (!e || count < 0 || (count > 0 && (!insts || !params || !results))) return JSSP_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  for (int i = 0; i < count; ++i) {
    int rc = jssp::solve_any(e, &insts[i], &params[i], &results[i], nullptr, nullptr);
    if (rc) return rc;
  }
  return JSSP_OK;
}
