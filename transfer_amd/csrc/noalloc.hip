// noalloc.hip -- an inner product that allocates nothing, for callers that
// cannot call hipMalloc (code under HIP graph capture, tight serving loops).
//
// tr_dot (CudaContext::dot) sums per-workgroup partials in the context's
// reduction buffer and copies the result to the host.  tr_dot_noalloc runs
// the same two kernels with the partials in `scratch`, a static device
// array (one per GPU, held until the process ends), and leaves the result
// in a device double of the caller's: no allocation, no host copy, no
// synchronisation.  Every launch writes every partial it later reads, so
// nothing needs clearing between calls; calls that overlap on one GPU
// (other streams or threads) share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "ctx.h"
#include "internal.h"
#include "kernels.h"

namespace tr {

static constexpr int kNoAllocPartials = 1024;  // the partials launch_dot writes at most
static __device__ __align__(256) double scratch[kNoAllocPartials];

static constexpr int kMaxDevices = 64;
static double *g_scratch[kMaxDevices] = {};  // device address of `scratch` on each GPU
static std::mutex g_scratch_mu;

// Resolves this GPU's address of `scratch` (the device current on entry is
// the context's).  Host code never passes `scratch` itself to a kernel: its
// name in host code is a stand-in, not a device address.
int noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(TR_ERR_ARG, "noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return TR_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(TR_ERR_HIP, std::string("noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<double *>(p);
    return TR_OK;
}

}  // namespace tr

using namespace tr;

extern "C" int tr_dot_noalloc(tr_ctx *c, const double *x, const double *y, double *result) {
    if (!c || !x || !y || !result) return fail(TR_ERR_ARG, "tr_dot_noalloc: null argument");
    if (red_partials() > kNoAllocPartials)
        return fail(TR_ERR_ARG, "tr_dot_noalloc: the reduction needs more partials than the scratch holds");
    if (c->device < 0 || c->device >= kMaxDevices) return fail(TR_ERR_ARG, "tr_dot_noalloc: device out of range");
    double *part;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        part = g_scratch[c->device];
    }
    if (!part) return fail(TR_ERR_STATE, "tr_dot_noalloc: scratch not set up for this device");
    // vectors of any length: the partial count is capped at red_partials()
    HIPCHECK(launch_dot(x, y, c->nint(), part, result, c->stream));
    return TR_OK;
}
