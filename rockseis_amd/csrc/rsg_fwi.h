// Internal declarations of the FWI driver.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/rsgpu.h"

namespace rsg {

// Binomial checkpointing schedule (revolve.cpp:8-305).
struct Revolve {
    enum Action { ADVANCE = 0, TAKESHOT, RESTORE, FIRSTURN, YOUTURN, TERMINATE, ERROR };
    static constexpr int kCheckUp = 512;  // revolve.h:17
    static constexpr int kRepsUp = 64;    // revolve.h:18
    int steps, snaps, capo, fine, check, turn, oldfine;
    std::vector<int> ch;
    Revolve(int steps, int snaps);
    Action next();
    static int maxrange(int ss, int tt);
    static int adjust(int steps);
    static int numforw(int steps, int snaps);
};

// Misfit (added to *misfit in double) and residuals of nshot shots of ntr
// traces (shot k's traces follow shot k-1's) for type 0 DIFFERENCE or
// 1 CORRELATION, optional data weights; map holds dim indices per trace
// (misfit.hip).
int acoustic_misfit_residuals(int type, const float *mod, const float *rec, const float *wei, float *res,
                              const int *map, int dim, int ntr, int nt, int nshot, int scaling, double *misfit,
                              hipStream_t st);

}  // namespace rsg
