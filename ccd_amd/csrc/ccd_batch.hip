// ccd_batch.hip -- device-resident batched engine (placeholder until implemented).
#include <hip/hip_runtime.h>

#include "ccd_host.h"

using namespace ccdamd;

extern "C" int ccdh_batch_fit(ccdh_data_t, int, const ccdh_fit_spec*, double*, ccdh_fit_result*,
                              const ccdh_batch_options*) {
  return setError(CCDH_E_INVALID, "batched engine not built");
}
