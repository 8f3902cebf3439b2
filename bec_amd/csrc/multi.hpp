// One launch per hot pass over every group of a context (the uneven grids of
// one run): the launch arguments carry a table of the groups' tile ranges.
// Internal interface between bec_api.cpp and multi.hip.
#pragma once

#include "backend.hpp"

namespace bec {

constexpr int MULTI_MAX_GROUPS = 8;

// z pass of one group: OpKinetic, one component per tile
struct MultiZGroup {
    void *a, *b;
    const void *kxy, *kz;        // kinetic tables [KT_COUNT][records][..]
    long kxy_stride, kz_stride;  // complex values per table variant
    const void *tw;
    long tile0;                  // first tile of the group in the launch
    int nxy, nz, rec_shift;
};

struct MultiZArgs {
    MultiZGroup g[MULTI_MAX_GROUPS];
    int n_groups;
    long total_tiles;   // every group's tiles: 2 components * ntraj * nxy / PW
};

// xy pass of one group: the real-time nonlinear step (OpNonlinRT)
struct MultiXyGroup {
    void *a, *b;
    const TrajParams *params;
    unsigned *step;   // loss-noise step counter, or null
    long tile0;
    uint32_t k0, k1, group;
    int n, nz, ntraj, rec_shift;
};

struct MultiXyArgs {
    MultiXyGroup g[MULTI_MAX_GROUPS];
    int n_groups;
    long total_tiles;   // every group's tiles: ntraj * nz / (64 / n)
};

// The two launches of a set of groups and their resident grids.
struct MultiPlan {
    MultiZArgs z;
    MultiXyArgs xy;
    int n_groups, nmax, itmax1, noise;
    int z_grid, xy_grid;
};

// Fills plan for these groups; false when one launch cannot step them
// together: more than MULTI_MAX_GROUPS groups, an FP64 state, an unsupported
// lattice, or groups that differ in itmax == 1 or in the loss noise.
bool multi_plan(const DevGroup *const *groups, int n, MultiPlan *plan);
// the kinetic z pass (table variant KT_*) of every group in one launch
hipError_t run_z_multi(const MultiPlan &plan, int variant, hipStream_t s);
// the nonlinear xy pass of every group in one launch (then the noise counters)
hipError_t run_xy_multi(const MultiPlan &plan, hipStream_t s);

} // namespace bec
