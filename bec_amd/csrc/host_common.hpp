// Host-side structures shared by the C ABI implementation files.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/bec_api.h"
#include "backend.hpp"

namespace bec {

struct Group {
    bec_calc_params calc{};
    DevGroup dev{};
    long cells = 0;
    TrajParams hp{};
    double tables_dt = -1.0;     // dt the real-time kinetic tables were built for
    double tables_dts = -1.0;    // dt_steady of the imaginary-time table
    double *d_partial = nullptr; // per-workgroup partial sums
    long partial_slots = 0;
    double *d_sums = nullptr;    // [ntraj][4]
    double *d_proj = nullptr;    // projection workspace (projection_doubles), allocated on first use
    void *d_avg = nullptr;       // calculateAverages workspace: x-space a, b and [cells][6], on first use
    bool initialized = false;
};

struct DeviceBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

hipError_t dev_alloc(void **p, size_t bytes);
void dev_free(void *&p);

// Build the group's device-side parameters and twiddles (no state allocation).
int group_setup(Group &g, const bec_calc_params &c, int prec, int ntraj, bool alloc_state);
void group_release(Group &g);
int group_build_tables(Group &g, double dt, hipStream_t s);
TrajParams make_traj_params(const bec_calc_params &c, double dt);

// host copies with precision conversion (host side is always complex128)
int upload_complex(const double *h, void *d, long n, int prec, hipStream_t s);
int download_complex(const void *d, double *h, long n, int prec, hipStream_t s);

double k_of(int i, double dk, int n);

int hip_status(hipError_t e);
void set_last_error(const std::string &msg);

} // namespace bec

#define BEC_HIP(call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            bec::set_last_error(std::string(#call) + ": " + hipGetErrorString(e_));                   \
            return bec::hip_status(e_);                                                                \
        }                                                                                              \
    } while (0)
