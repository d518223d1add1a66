// Library-internal entry between the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/soapfilm.h"

namespace sfapi {

// sf_step_noalloc on the device region [scratch, scratch + bytes): the
// context's set-up carves every work array from it, runs `nsteps` steps of
// one mixed-precision bubble on `stream` and returns after the device work
// has finished. Errors are reported like every entry point (status code,
// sf_last_error(NULL)).
int step_in_scratch(void* scratch, size_t bytes, const sf_options* opt,
                    const sf_bubble_desc* bubble, double t_seconds, int nsteps,
                    hipStream_t stream, const float* const in[4], const float* uair0,
                    const float* vair0, float* const out[4], sf_bubble_stats* stats);

// The argument checks of sf_step_noalloc, made before any device call.
int check_noalloc(const sf_options* opt, const sf_bubble_desc* bubble, int nsteps,
                  const float* const in[4], const float* uair0, const float* vair0,
                  float* const out[4]);

// Records a failure for sf_last_error(NULL) and returns `code`.
int record_error(int code, const char* msg);

}  // namespace sfapi
