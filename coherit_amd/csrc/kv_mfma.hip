// int8-limb MFMA operator path (filled in after the FP32 path is validated).
#include "kv_plan.hpp"

namespace coh {

int plan_mfma(const coh_panel*, int precision, int64_t, int32_t, int32_t, KvPlan*) {
  return fail(COH_ERR_VALUE, "precision %d is not available in this build", precision);
}
int run_phase1_mfma(const coh_panel*, const KvPlan&, const float*, int64_t, int64_t,
                    const CompScale&, int, uint8_t*, hipStream_t) {
  return fail(COH_ERR_VALUE, "MFMA path not available");
}
int run_phase2_mfma(const coh_panel*, const KvPlan&, int64_t, int, float*, int64_t, int64_t, float,
                    uint8_t*, hipStream_t) {
  return fail(COH_ERR_VALUE, "MFMA path not available");
}

}  // namespace coh
