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

// DIFFERENCE misfit (double, into *misfit) and residuals of ntr traces;
// map holds dim indices per trace (fwi.hip).
__global__ void misfit_residual_kernel(const float *__restrict__ mod, const float *__restrict__ rec,
                                       float *__restrict__ res, const int *__restrict__ map, int dim,
                                       int ntr, int nt, int scaling, double *misfit);

}  // namespace rsg
