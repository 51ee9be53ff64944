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

// One forward acoustic 2D step of a batch out of place: the previous state
// is read from `in` (left unchanged), the new one written to `out` (fd2d.hip).
int a2d_forward_step_from(const rsg_grid *g, const rsg_state *in, const rsg_state *out, const rsg_batch *b,
                          hipStream_t stream);

// Misfit (added to *misfit in double) and residuals of nshot shots of ntr
// traces (shot k's traces follow shot k-1's) for type 0 DIFFERENCE,
// 1 CORRELATION, 2 ADAPTIVE_GAUSS, 3 ADAPTIVE_LINEAR (4 ADAPTIVE_SINGLE is
// the reference's default branch, DIFFERENCE), optional data weights; map
// holds dim indices per trace; dim 2 clips types 1-3 at the 97th percentile
// as FwiAcoustic2D does, dim 3 does not, as FwiAcoustic3D (misfit.hip).
int acoustic_misfit_residuals(int type, const float *mod, const float *rec, const float *wei, float *res,
                              const int *map, int dim, int ntr, int nt, int nshot, int scaling, double *misfit,
                              hipStream_t st);

// The P (0), Ux (1) and Uz (2) records of one elastic shot, [nrec][nt] each
// on the device; wei[c] may be null; map: the Ux receiver map (a trace
// counts in the misfit when both its indices are >= 0).
struct E2DRecs {
    int nrec;
    const int *map;
    const float *mod[3], *obs[3], *wei[3];
    float *res[3];
    float *norms;  // scratch of CORRELATION (9 nrec floats), set internally
};

// Elastic misfit (added to *misfit in double) and residuals of every shot
// (FwiElastic2D::computeMisfit / computeResiduals, fwi.cpp:2852-3773), type
// 0..4 as rs_fwimisfit; the percentile clips are per shot.  Types 1-4
// allocate scratch and synchronize the stream; DIFFERENCE only enqueues.
int elastic_misfit_residuals(int type, const std::vector<E2DRecs> &shots, int nt, double *misfit,
                             hipStream_t st);

// backwardstepVelocity of the elastic 3D state with a scratch of six padded
// grids (fd3d_elastic.hip): the moduli-weighted stresses formed once per
// cell, then the forward velocity kernel on them.
int e3d_backward_velocity_premul(const rsg_grid *g, const rsg_state_e3d *s, float *scratch, hipStream_t st);

// The Ux (0), Uy (1) and Uz (2) records of one elastic 3D shot, [nrec][nt]
// each on the device; wei[c] may be null.
struct E3DRecs {
    int nrec;
    const float *mod[3], *obs[3], *wei[3];
    float *res[3];
};

// Elastic 3D misfit (added to *misfit in double) and residuals
// (FwiElastic3D::computeMisfit / computeResiduals, fwi.cpp:5632-6196), type
// 0..4 as rs_fwimisfit (3 and 4: the default branch).  Types 1 and 2
// allocate scratch and synchronize the stream; the others only enqueue.
int elastic3d_misfit_residuals(int type, const E3DRecs &r, int nt, double *misfit, hipStream_t st);

}  // namespace rsg
