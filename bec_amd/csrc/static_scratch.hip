// A fixed device scratch array for the entry points that must not allocate
// (callers under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc), and bec_projections_static, the one entry point that uses it.
// Each GPU the library runs on holds its own copy of the array until the
// process ends.
//
// Host code never passes `scratch` itself to a kernel or a copy: in host code
// the name is a stand-in, not a device address.  scratch_on_current_device()
// resolves the address once per device, with that device current, and every
// user takes the pointer it returns.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "host_common.hpp"

namespace bec {

int ctx_projections(bec_ctx *ctx, int group, double alpha, double *buf, size_t cap, double *xy_a, double *xy_b,
                    double *zy_a, double *zy_b, double *axial_a, double *axial_b);

// 4.25 MiB: the largest projection workspace of a supported lattice
// (256 x 32 x 32: 2 x 262,144 cell sums + 18,944 projection values) with room
static constexpr size_t SCRATCH_DOUBLES = 557056;
static constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

static std::mutex g_mutex;
static double *g_ptr[MAX_DEVICES] = {};   // the array's device address, per device

static hipError_t scratch_on_current_device(double **ptr)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_ptr[dev] = (double *)p;
    }
    *ptr = g_ptr[dev];
    return hipSuccess;
}

} // namespace bec

// bec_projections with its workspace in the fixed scratch array of the current
// device instead of device memory the library allocates: no call allocates.
// Calls that overlap on one GPU share the array and race.
extern "C" int bec_projections_static(bec_ctx *ctx, int group, double alpha, double *xy_a, double *xy_b,
                                      double *zy_a, double *zy_b, double *axial_a, double *axial_b)
{
    if (!ctx) return BEC_ERR_INVALID;
    double *scr = nullptr;
    hipError_t e = bec::scratch_on_current_device(&scr);
    if (e != hipSuccess) {
        bec::set_last_error(std::string("bec_projections_static: scratch address: ") + hipGetErrorString(e));
        return bec::hip_status(e);
    }
    return bec::ctx_projections(ctx, group, alpha, scr, bec::SCRATCH_DOUBLES, xy_a, xy_b, zy_a, zy_b, axial_a,
                                axial_b);
}
