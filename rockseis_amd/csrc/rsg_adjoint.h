// Internal interface of the exact discrete adjoint of the acoustic 2D FWI
// gradient (adjoint_a2d.hip), driven by the Revolve loop of fwi.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/rsgpu.h"

namespace rsg {

// Device work arrays, each nshot padded grids (shot stride = field stride)
// except gvp / grho (one padded grid, FP64).
struct A2DExactBufs {
    float *W1, *W2;         // adjoints of the half step's x / z derivatives
    float *gL, *gRx, *gRz;  // time-summed gradients of the staggered model
    double *gvp, *grho;     // padded-grid (vp, rho) gradients
};

// One reverse step at time t: fw holds the forward state before step t and
// is advanced by the velocity half step; bw holds the adjoint state after
// step t and is carried back to before it.  res: residual traces (injected
// with weight 1), src: the source wavelet (PRESSURE); wavgrad may be null.
int a2d_exact_step(const rsg_grid *g, const rsg_state *fw, const rsg_state *bw, const rsg_batch *b,
                   const A2DExactBufs &e, const rsg_traces *res, const rsg_traces *src, float *wavgrad, int t,
                   hipStream_t st);

// (gL, gRx, gRz) summed over the shots -> (vp, rho) gradients of the nx*nz
// model, added to vpgrad / rhograd.
int a2d_exact_model_grad(const rsg_grid *g, int nx, int nz, const float *Vp, const float *Rho,
                         const A2DExactBufs &e, int nshot, size_t fstride, float *vpgrad, float *rhograd,
                         hipStream_t st);

}  // namespace rsg
