// Multipole path (spherical.rs): host dispatch.
#include "capi_internal.hpp"

namespace csx {

csx_status multipole_dispatch(csx_structure* s, const double* q, size_t nq, int l_max,
                              bool implicit_h, const csx_solvent_params* p, int precision,
                              double* out) {
  (void)s; (void)q; (void)nq; (void)l_max; (void)implicit_h; (void)p; (void)precision; (void)out;
  return fail(CSX_INVALID_INPUT, "multipole path is not built yet");
}

}  // namespace csx
