// sop_shim.h -- device-context helpers shared by the shim and the step loops.
#ifndef SOPG_SOP_SHIM_H
#define SOPG_SOP_SHIM_H

#include "../../include/sopg.h"

// The context holding the current topology as chain 0 (created / rebuilt on
// demand; chain parameters refreshed from the globals on every call).
sopg_ctx* sopg_shim_context();
// Host globals pos / unc_pos / vel / force -> device, and back.
void sopg_shim_put_state();
void sopg_shim_get_state();
// 3 * nbead * nforce_evals replayed deviates drawn from `generator`, or NULL
// when the device generator is selected (`set noise device;`).
const double* sopg_shim_noise(long nforce_evals);
void sopg_shim_check(int rc, const char* what);
void sopg_shim_die(const char* what);

#endif
