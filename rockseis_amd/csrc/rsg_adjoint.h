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

// ---- acoustic 3D (adjoint_a3d.hip), one shot

// Work arrays on the padded grid: the adjoints of the half step's three
// derivatives, the time-summed gradients of the staggered model, and the
// padded-grid (vp, rho) gradients in FP64.
struct A3DExactBufs {
    float *W[3];
    float *gL, *gR[3];
    double *gvp, *grho;
};

// The source terms and the reverse stress and velocity half steps of
// reverse step t (the forward velocity half step on fw and the residual
// injection into bw come first); pad: the map offset (lpml).
int a3d_exact_reverse(const rsg_grid *g, const rsg_state *fw, const rsg_state *bw, const A3DExactBufs &e,
                      const rsg_traces *src, float *wavgrad, int t, int pad, hipStream_t st);
// (gL, gRx, gRy, gRz) -> (vp, rho) gradients of the nz*ny*nx model, added to
// vpgrad / rhograd.
int a3d_exact_model_grad(const rsg_grid *g, int nx, int ny, int nz, const float *Vp, const float *Rho,
                         const A3DExactBufs &e, float *vpgrad, float *rhograd, hipStream_t st);

// ---- elastic 2D (adjoint_e2d.hip), ragged batches

// Work arrays of one shot, each a padded grid of that shot: the adjoints of
// the half step's four derivatives and the time-summed gradients of the
// staggered model.
struct E2DExactWork {
    float *W[4];
    float *gL, *gL2M, *gM, *gRx, *gRz;
};

// Device descriptors of a batch for the reverse-step kernels.
struct E2DExact {
    int nshot = 0, order = 0, total = 0;
    void *d_desc = nullptr;
    int *d_tile0 = nullptr;
};

// fw / bw: each shot's forward and adjoint states (fields, strips, model);
// work: each shot's work arrays.
int e2d_exact_create(int nshot, const rsg_grid *grids, const rsg_state *fw, const rsg_state *bw,
                     const E2DExactWork *work, E2DExact *out, hipStream_t st);
void e2d_exact_destroy(E2DExact *p);
// Reverse stress and velocity half steps of every shot (steps 4 and 5 of the
// reverse step: the forward velocity half step, the residual injection and
// the source terms come first).
int e2d_exact_reverse(const E2DExact &p, hipStream_t st);
// One window's gradients (scale = 1: vp, vs, rho; 0: lambda, mu, rho) from
// its work arrays, added to o0..o2 (nz*nx); pad3: 3 padded grids of doubles.
int e2d_exact_model_grad(const E2DExactWork &w, int nx, int nz, int lpml, int fs, int scale, const float *Vp,
                         const float *Vs, const float *Rho, double *pad3, float *o0, float *o1, float *o2,
                         hipStream_t st);

}  // namespace rsg
