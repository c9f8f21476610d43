// gauge_force.hip -- Symanzik one-loop gauge force (placeholder until the kernel lands).
#include "handles.h"

extern "C" {
void QPHIX_D3_symanzik_1loop_gauge_force(QPHIX_info_t *info, QPHIX_D3_GaugeField *, QPHIX_D3_Force *,
                                         QPHIX_gauge_coeffs_t *, double) {
  qphix::fail(__FILE__, __LINE__, "gauge force not built yet");
  if (info) info->status = QPHIX_FAIL;
}
void QPHIX_F3_symanzik_1loop_gauge_force(QPHIX_info_t *info, QPHIX_F3_GaugeField *, QPHIX_F3_Force *,
                                         QPHIX_gauge_coeffs_t *, float) {
  qphix::fail(__FILE__, __LINE__, "gauge force not built yet");
  if (info) info->status = QPHIX_FAIL;
}
}
