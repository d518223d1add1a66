// sf_step_noalloc: the soap-film step for callers that cannot allocate
// device memory. Every work array of the step (the advected fields, the
// assembled operator, the solver vectors, tables, descriptors and sums) is
// carved from `scratch`, a fixed array in the library's device image; each
// GPU the library runs on holds its own copy until the process ends.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/soapfilm.h"
#include "sf_internal.h"

namespace {

// 48 MiB: the work arrays of a mixed-precision bubble of 2^18 cells (for
// example 256 x 512) with room to spare.
constexpr size_t kScratchBytes = (size_t)48 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
// One lock for every call: the array of a device is shared by all calls on
// it, so calls are serialised (each returns after its device work ends).
std::mutex g_lock;
// The array's device address on each GPU, from hipGetSymbolAddress with that
// GPU current at the first call there. In host code `scratch` names a host
// stand-in, never the device memory.
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

extern "C" int sf_step_noalloc(const sf_options* opt, const sf_bubble_desc* bubble,
                               double t_seconds, int32_t nsteps, void* stream,
                               const float* u_theta, const float* u_phi, const float* eta,
                               const float* gamma, const float* uair0, const float* vair0,
                               float* u_theta_out, float* u_phi_out, float* eta_out,
                               float* gamma_out, sf_bubble_stats* stats) {
    if (!opt || !bubble) return sfapi::record_error(SF_ERR_INVALID, "null argument");
    if (opt->device < 0 || opt->device >= kMaxDevices)
        return sfapi::record_error(SF_ERR_INVALID, "device ordinal out of range");
    const float* const in[4] = {u_theta, u_phi, eta, gamma};
    float* const out[4] = {u_theta_out, u_phi_out, eta_out, gamma_out};
    // sizes and options are checked before the device is touched
    const int rc = sfapi::check_noalloc(opt, bubble, nsteps, in, uair0, vair0, out);
    if (rc != SF_OK) return rc;
    std::lock_guard<std::mutex> lock(g_lock);
    hipError_t e = hipSetDevice(opt->device);
    if (e != hipSuccess) return sfapi::record_error(SF_ERR_HIP, hipGetErrorString(e));
    void*& ptr = g_scratch_ptr[opt->device];
    if (!ptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return sfapi::record_error(SF_ERR_HIP, e != hipSuccess ? hipGetErrorString(e)
                                                                   : "no device address of the work array");
        ptr = p;
    }
    return sfapi::step_in_scratch(ptr, kScratchBytes, opt, bubble, t_seconds, nsteps,
                                  static_cast<hipStream_t>(stream), in, uair0, vair0, out, stats);
}
