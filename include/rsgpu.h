/*
 * rsgpu.h -- C-ABI of librsgpu.so, the MI355X (gfx950) implementation of
 * RockSeis's staggered-grid finite-difference time step, source injection,
 * receiver gathers and FWI gradient kernels.
 *
 * Plain C: opaque stream pointers (hipStream_t passed as void*), device
 * pointers and sizes; no HIP or torch types.  Every function returns 0 on
 * success and a negative code on failure; rsg_last_error() then holds a
 * message (thread-local).  Host callers map a non-zero return to the
 * reference's fatal path (rs_error: print and exit, utils.cpp:51-62).
 *
 * Layout is the reference's (waves.cpp:4-13): padded grid, x fastest,
 * I2D(i,j)=j*nx+i, I3D(i,j,k)=k*nx*ny+j*nx+i; PML strips I2D_lr=j*lpml+i,
 * I2D_tb=j*nx+i, I3D_lr=k*lpml*ny+j*lpml+i, I3D_fb=k*nx*lpml+j*nx+i,
 * I3D_tb=k*nx*ny+j*nx+i.  Traces are trace-major, nt fastest
 * (Idat(it,i)=i*nt+it).
 */
#ifndef RSGPU_H
#define RSGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSG_OK 0
#define RSG_EINVAL (-1)
#define RSG_EHIP (-2)
#define RSG_ENOMEM (-3)

/* Field kinds, numbered like rs_field (utils.h:22). */
#define RSG_PRESSURE 0
#define RSG_VX 1
#define RSG_VY 2
#define RSG_VZ 3

/* Propagator kinds */
#define RSG_KIND_A2D 1 /* WavesAcoustic2D */
#define RSG_KIND_E2D 2 /* WavesElastic2D  */
#define RSG_KIND_A3D 3 /* WavesAcoustic3D */

/* Grid of one propagator: padded sizes as Der(nx_pml,ny_pml,nz_pml,...)
 * receives them (modelling.cpp:213), ny=1 in 2D. */
typedef struct {
    int nx, ny, nz;
    int lpml;   /* PML width (all six sides applied)                        */
    int order;  /* stencil half-width 1..8 (Der clamps, der.cpp:94-95)      */
    int fs;     /* free surface: Szz row z=lpml zeroed (waves.cpp:1666-1674) */
    float dx, dy, dz, dt;
    int domdec; /* domain-decomposition mode of a one-domain model built by
                   Model*::getDomainmodel (model.cpp:1511-1620): nx, ny, nz
                   already hold the PML padding, source/receiver maps index
                   the grid with no lpml offset (waves.cpp:411-418) and
                   FWI images and snapshots cover the whole grid.  The PML
                   strips are still applied on every side (lpml wide).    */
} rsg_grid;

/* PML memory strip slots of rsg_state.psi[] */
enum {
    /* WavesAcoustic2D / PmlAcoustic2D (pml.h:103-111) */
    RSG_A2D_P_LEFT = 0, RSG_A2D_P_RIGHT, RSG_A2D_P_TOP, RSG_A2D_P_BOTTOM,
    RSG_A2D_VXX_LEFT, RSG_A2D_VXX_RIGHT, RSG_A2D_VZZ_TOP, RSG_A2D_VZZ_BOTTOM,
    RSG_A2D_NSTRIP,
    /* WavesElastic2D / PmlElastic2D (pml.h:159-178) */
    RSG_E2D_SXX_LEFT = 0, RSG_E2D_SXX_RIGHT, RSG_E2D_SXZX_LEFT, RSG_E2D_SXZX_RIGHT,
    RSG_E2D_SZZ_TOP, RSG_E2D_SZZ_BOTTOM, RSG_E2D_SXZZ_TOP, RSG_E2D_SXZZ_BOTTOM,
    RSG_E2D_VXX_LEFT, RSG_E2D_VXX_RIGHT, RSG_E2D_VZX_LEFT, RSG_E2D_VZX_RIGHT,
    RSG_E2D_VZZ_TOP, RSG_E2D_VZZ_BOTTOM, RSG_E2D_VXZ_TOP, RSG_E2D_VXZ_BOTTOM,
    RSG_E2D_NSTRIP,
    /* WavesAcoustic3D / PmlAcoustic3D (pml.h:132-149) */
    RSG_A3D_P_LEFT = 0, RSG_A3D_P_RIGHT, RSG_A3D_P_FRONT, RSG_A3D_P_BACK,
    RSG_A3D_P_TOP, RSG_A3D_P_BOTTOM, RSG_A3D_VXX_LEFT, RSG_A3D_VXX_RIGHT,
    RSG_A3D_VYY_FRONT, RSG_A3D_VYY_BACK, RSG_A3D_VZZ_TOP, RSG_A3D_VZZ_BOTTOM,
    RSG_A3D_NSTRIP
};
#define RSG_MAX_STRIPS 16

/* PML memory strip slots of rsg_state_e3d.psi[] (WavesElastic3D /
 * PmlElastic3D, pml.h:197-234): x faces (lpml*ny*nz each, [z][y][i]), y
 * faces (nx*lpml*nz, [z][i][x]), z faces (nx*ny*lpml, [i][y][x]). */
enum {
    RSG_E3D_SXX_LEFT = 0, RSG_E3D_SXX_RIGHT, RSG_E3D_SXZX_LEFT, RSG_E3D_SXZX_RIGHT,
    RSG_E3D_SXYX_LEFT, RSG_E3D_SXYX_RIGHT, RSG_E3D_VXX_LEFT, RSG_E3D_VXX_RIGHT,
    RSG_E3D_VZX_LEFT, RSG_E3D_VZX_RIGHT, RSG_E3D_VYX_LEFT, RSG_E3D_VYX_RIGHT,
    RSG_E3D_SYY_FRONT, RSG_E3D_SYY_BACK, RSG_E3D_SXYY_FRONT, RSG_E3D_SXYY_BACK,
    RSG_E3D_SYZY_FRONT, RSG_E3D_SYZY_BACK, RSG_E3D_VYY_FRONT, RSG_E3D_VYY_BACK,
    RSG_E3D_VXY_FRONT, RSG_E3D_VXY_BACK, RSG_E3D_VZY_FRONT, RSG_E3D_VZY_BACK,
    RSG_E3D_SZZ_TOP, RSG_E3D_SZZ_BOTTOM, RSG_E3D_SXZZ_TOP, RSG_E3D_SXZZ_BOTTOM,
    RSG_E3D_SYZZ_TOP, RSG_E3D_SYZZ_BOTTOM, RSG_E3D_VZZ_TOP, RSG_E3D_VZZ_BOTTOM,
    RSG_E3D_VXZ_TOP, RSG_E3D_VXZ_BOTTOM, RSG_E3D_VYZ_TOP, RSG_E3D_VYZ_BOTTOM,
    RSG_E3D_NSTRIP
};

/* Device pointers of one propagator's state (a Waves<T> object plus the
 * staggered model arrays it reads).  Members a kind does not use are NULL. */
typedef struct {
    float *P, *Vx, *Vy, *Vz, *Sxx, *Szz, *Sxz;   /* fields                        */
    const float *L, *L2M, *M, *Rx, *Ry, *Rz;    /* staggered model (padded grid) */
    const float *tab;  /* 12*lpml PML tables in pml.h:60-74 order: A,B,C ltf,
                          A,B,C ltf_stag, A,B,C rbb, A,B,C rbb_stag           */
    float *psi[RSG_MAX_STRIPS]; /* memory strips, slots above               */
    float *wrk;        /* optional scratch nx*ny*nz (unused by the fused kernels) */
} rsg_state;

/* A set of traces on the grid: a source (data = wavelet) or receivers
 * (data = record), as Data2D/Data3D plus the Geometry maps of makeMap. */
typedef struct {
    int field;           /* RSG_PRESSURE / RSG_VX / RSG_VY / RSG_VZ            */
    int ntrace;
    int nt;              /* samples per trace                                  */
    const int *map;      /* device, ntrace*dim grid indices (unpadded), -1 off  */
    const float *shift;  /* device, ntrace*dim fractional offsets              */
    float *data;         /* device, ntrace*nt                                  */
    const float *factor; /* elastic pressure records: per-trace divisor or NULL */
} rsg_traces;

/* ---- runtime ---------------------------------------------------------- */
const char *rsg_last_error(void);
int rsg_device_count(int *n);
int rsg_set_device(int dev);
int rsg_stream_sync(void *stream);
int rsg_malloc(void **ptr, size_t bytes);
int rsg_free(void *ptr);
int rsg_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int rsg_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int rsg_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
/* Zero with a fill kernel (safe inside hipGraph capture). */
int rsg_zero(void *ptr, size_t bytes, void *stream);

/* Launch-shape knobs: "rz2d" (rows per thread of the 2D tiles: 1, 2, 4, 8),
 * "rz2d_ragged" (the same for ragged batches, default 2), "pml_cost" (weight
 * in percent of a PML / edge cell or tile against an interior one when the
 * tiles are split into eight XCD ranges of equal cost, 2D and 3D, default
 * 200), "zc3d" (z planes per 3D block, 0 = automatic), "xcd3d" (3D tile
 * order: 1 = each XCD marches its own contiguous group of (x, y) columns
 * through z, 0 = grid order; default 1), "fuse_io" (modelling runs: sources
 * and velocity gathers in the stress kernel, 1 = on).  Results do not depend
 * on them; rockseis_amd.autotune picks rz2d, pml_cost and zc3d per
 * workload. */
int rsg_set_tuning(const char *key, int value);
int rsg_get_tuning(const char *key, int *value);

/* ---- host-side set-up, bit-exact with the reference --------------------- */
/* Pml<T>::computeABC(sign) with Amax=0, Kmax=2 and the given Smax
 * (pml.cpp:79-126; SMAX=1200 forced by modelling.cpp:216-218). */
int rsg_pml_tables(int lpml, float dt, float smax, int sign, float *tab_host);
/* Der<T> coefficient table, 36 entries, order k at k(k-1)/2 (der.cpp:3-53). */
int rsg_der_coeffs(float *out36);
/* Der<T>::dd{x,y,z}_{fw,bw} (der.cpp:155-540) on a device array: df = the
 * staggered derivative of f along axis 0 (x), 1 (y) or 2 (z) of the
 * nx x ny x nz grid with step d, forward (backward = 0) or backward, with
 * the reference's reduced orders at the ends of the axis and its
 * operation order (equal to the reference's floats).  order is clamped to
 * 1..8 as Der does; the axis must hold at least 2*order+1 cells. */
int rsg_der_apply(int axis, int backward, const float *f, float *df, int nx, int ny, int nz, float d,
                  int order, void *stream);
/* Model*::staggerModels (model.cpp:1350-1411, 2720-2801, 1965-2040):
 * unpadded [nz][nx] (or [nz][ny][nx]) host inputs -> padded host outputs. */
int rsg_stagger_a2d(const float *vp, const float *rho, int nx, int nz, int lpml, int fs,
                    float *L, float *Rx, float *Rz);
int rsg_stagger_e2d(const float *vp, const float *vs, const float *rho, int nx, int nz, int lpml,
                    int fs, float *L, float *L2M, float *M, float *Rx, float *Rz);
int rsg_stagger_a3d(const float *vp, const float *rho, int nx, int ny, int nz, int lpml, int fs,
                    float *L, float *Rx, float *Ry, float *Rz);
/* The same staggering (model.cpp:1350-1411, 1965-2040) on the device, equal
 * to rsg_stagger_a2d / rsg_stagger_e2d bit for bit: device inputs of nx x nz
 * cells with row stride ld >= nx (a window of a wider model), device outputs
 * of (nx+2*lpml) x (nz+2*lpml).  Queued on `stream`; returns after the
 * stream has drained (the zero-density check reads one flag back). */
int rsg_stagger_a2d_dev(const float *vp, const float *rho, int ld, int nx, int nz, int lpml, int fs,
                        float *L, float *Rx, float *Rz, void *stream);
int rsg_stagger_e2d_dev(const float *vp, const float *vs, const float *rho, int ld, int nx, int nz, int lpml,
                        int fs, float *L, float *L2M, float *M, float *Rx, float *Rz, void *stream);
/* Geometry2D::makeMap for one coordinate set (geometry.cpp:137-235), with
 * the field shift of Data2D::makeMap (-0.5 along x for VX, z for VZ). */
int rsg_makemap2d(const float *coords, int ntrace, int nx, int nz, float dx, float dz, float ox,
                  float oz, int field, int *map, float *shift);
int rsg_makemap3d(const float *coords, int ntrace, int nx, int ny, int nz, float dx, float dy,
                  float dz, float ox, float oy, float oz, int field, int *map, float *shift);

/* ---- one time step (the Waves* member functions) ------------------------ */
/* WavesAcoustic2D::forwardstepVelocity / forwardstepStress (waves.cpp:217-404) */
int rsg_a2d_forward_velocity(const rsg_grid *g, const rsg_state *s, void *stream);
int rsg_a2d_forward_stress(const rsg_grid *g, const rsg_state *s, void *stream);
/* WavesElastic2D::forwardstep* / backwardstep* (waves.cpp:1416-2080) */
int rsg_e2d_forward_velocity(const rsg_grid *g, const rsg_state *s, void *stream);
int rsg_e2d_forward_stress(const rsg_grid *g, const rsg_state *s, void *stream);
int rsg_e2d_backward_velocity(const rsg_grid *g, const rsg_state *s, void *stream);
int rsg_e2d_backward_stress(const rsg_grid *g, const rsg_state *s, void *stream);
/* WavesAcoustic3D::forwardstepVelocity / forwardstepStress (waves.cpp:710-996) */
int rsg_a3d_forward_velocity(const rsg_grid *g, const rsg_state *s, void *stream);
int rsg_a3d_forward_stress(const rsg_grid *g, const rsg_state *s, void *stream);

/* Launch one half step `n` times back to back (which: 0 forward velocity,
 * 1 forward stress, 2 adjoint-state velocity, 3 adjoint-state stress), for
 * timing a kernel with events on `stream` without per-launch host overhead. */
int rsg_repeat_step(int kind, int which, const rsg_grid *g, const rsg_state *s, int n,
                    void *stream);

/* Many shots on one grid in one launch: every field and strip of shot k sits
 * at k * stride from shot 0's pointer in rsg_state; model arrays and PML tables
 * are shared.  No counterpart in the reference, which runs one shot per MPI
 * rank (parallel.cpp:690-732); results per shot equal the single-shot calls. */
typedef struct {
    int nshot;
    size_t field_stride;  /* elements between shots of P/Vx/Vz/Sxx/Szz/Sxz   */
    size_t xstrip_stride; /* between shots of an x-side strip (lpml*nz)      */
    size_t zstrip_stride; /* between shots of a z-side strip (nx*lpml)       */
} rsg_batch;
int rsg_a2d_forward_velocity_batch(const rsg_grid *g, const rsg_state *s, const rsg_batch *b,
                                   void *stream);
int rsg_a2d_forward_stress_batch(const rsg_grid *g, const rsg_state *s, const rsg_batch *b,
                                 void *stream);
/* adjoint = 0: forwardstep*, 1: backwardstep* */
int rsg_e2d_velocity_batch(const rsg_grid *g, const rsg_state *s, const rsg_batch *b, int adjoint,
                           void *stream);
int rsg_e2d_stress_batch(const rsg_grid *g, const rsg_state *s, const rsg_batch *b, int adjoint,
                         void *stream);

/* Ragged batch: nshot shots, each on its own grid (an aperture window of a
 * larger model: its own nx, nz, model arrays, fields and strips), one launch
 * per half step for all of them.  grids[k] / states[k] describe shot k as the
 * single-shot calls take them; every shot uses one stencil order.  The plan
 * copies the pointers: recreate it when they change.  No counterpart in the
 * reference, which runs one window per MPI worker (inversion.cpp:332-600);
 * results per shot equal the single-shot calls. */
typedef struct rsg_ragged rsg_ragged;
int rsg_ragged2d_create(int kind, int nshot, const rsg_grid *grids, const rsg_state *states,
                        rsg_ragged **plan);
/* which: 0 forward velocity, 1 forward stress, 2 adjoint-state velocity,
 * 3 adjoint-state stress (the last two elastic only) */
int rsg_ragged2d_step(const rsg_ragged *plan, int which, void *stream);
int rsg_ragged2d_destroy(rsg_ragged *plan);

/* insertSource (waves.cpp:407-497, 998-1137, 2085-2171): kind is
 * 1 acoustic 2D, 2 elastic 2D, 3 acoustic 3D; the field and its modulus are
 * chosen from src->field as the reference does. */
int rsg_insert_source(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src,
                      int it, void *stream);
/* recordData (waves.cpp:500-595, 1140-1275, 2174-2288) into rec->data. */
int rsg_record_data(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *rec,
                    int it, void *stream);

/* Interp<T>::interp (interp.cpp:110-186) on device records: from[ntr][n1]
 * sampled at o1 + i*d1 -> to[ntr][n2] at o2 + j*d2.  mode numbered like
 * rs_interpmode (utils.h:25): 0 LINEAR, 1 BSPLINE, 2 SINC. */
int rsg_interp(int mode, const float *from, int ntr, int n1, float d1, float o1, float *to, int n2, float d2,
               float o2, void *stream);

/* ---- whole modelling run (Modelling*::run time loop, modelling.cpp:242-286)
 * Steps it0..it0+nsteps-1: velocity, stress, insertSource(it), recordData(it)
 * for each receiver set.  graph_chunk > 0 captures that many steps in one
 * hipGraph and replays it (time index read from device memory). */
int rsg_modelling_run(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src,
                      int nrec, const rsg_traces *recs, int it0, int nsteps, int graph_chunk,
                      void *stream);
/* The same loop with snapshots (Modelling*::setSnapP/Vx/Vz/Sxx/Szz/Sxz,
 * modelling.h:79-81, 155-160, written at modelling.cpp:272-282 through
 * Snapshot2D/3D::writeSnap, snap.cpp:638-669): after the gathers of step
 * it with it % snapinc == 0, out + (it / snapinc) * interior receives
 * f[0] (+ f[1]) (+ f[2]) over the interior of the grid (the cells at
 * lpml .. n-lpml-1 on every padded axis; the whole grid with g->domdec),
 * x fastest.  out holds (nt-1)/snapinc + 1 slots. */
typedef struct {
    const float *f[3]; /* device fields summed into the snapshot (NULL: unused) */
    float *out;        /* device, [snapnt][interior]                             */
    int snapinc;
} rsg_snapset;
int rsg_modelling_run_snap(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src,
                           int nrec, const rsg_traces *recs, int nsnap, const rsg_snapset *snaps, int it0,
                           int nsteps, int graph_chunk, void *stream);

/* ---- modelling plan: the same time loop (Modelling*::run, modelling.cpp:
 * 242-286, with the snapshots of rsg_modelling_run_snap) captured once as
 * hipGraphs of 1, 2, 4, ..., max_chunk steps (max_chunk rounded down to a
 * power of two, 1..4096) and replayed for any range of steps: run() replays
 * the largest graph while it fits and then the binary digits of the rest,
 * so every step runs from a graph and no capture happens after create.
 * The graphs hold the pointers and launch shapes in force at create: the
 * fields, model, traces and snapshot buffers must outlive the plan, and a
 * retuned launch shape needs a new plan.  One run at a time per plan;
 * runs are queued on `stream` and ordered with the caller's other work
 * there.  destroy() waits for the device before it releases the graphs. */
typedef struct rsg_modelling_plan rsg_modelling_plan;
int rsg_modelling_plan_create(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src,
                              int nrec, const rsg_traces *recs, int nsnap, const rsg_snapset *snaps,
                              int max_chunk, void *stream, rsg_modelling_plan **out);
int rsg_modelling_plan_run(rsg_modelling_plan *p, int it0, int nsteps, void *stream);
int rsg_modelling_plan_destroy(rsg_modelling_plan *p);

/* ---- WavesElastic3D (waves.cpp:2310-4403): the nine fields, the
 * staggered model of ModelElastic3D::staggerModels (model.cpp:3371-3473)
 * and the 36 strips; padded grids [z][y][x]. */
typedef struct {
    float *Vx, *Vy, *Vz, *Sxx, *Syy, *Szz, *Syz, *Sxz, *Sxy;
    const float *L, *L2M, *Mxz, *Myz, *Mxy, *Rx, *Ry, *Rz;
    const float *tab;                 /* PML tables as in rsg_state              */
    float *psi[RSG_E3D_NSTRIP];
} rsg_state_e3d;
/* ModelElastic3D::staggerModels on the host, bit-exact: L, L2M, M_xz, M_yz,
 * M_xy, Rx, Ry, Rz of the (nx+2lpml)(ny+2lpml)(nz+2lpml) padded grid. */
int rsg_stagger_e3d(const float *vp, const float *vs, const float *rho, int nx, int ny, int nz, int lpml, int fs,
                    float *L, float *L2M, float *Mxz, float *Myz, float *Mxy, float *Rx, float *Ry, float *Rz);
/* forwardstepVelocity / forwardstepStress (waves.cpp:2465-3258); g: padded
 * sizes (2 lpml < n, nz <= 65535), fs zeroes Szz on plane lpml. */
int rsg_e3d_forward_velocity(const rsg_grid *g, const rsg_state_e3d *s, void *stream);
int rsg_e3d_forward_stress(const rsg_grid *g, const rsg_state_e3d *s, void *stream);
/* insertSource (waves.cpp:4104-4246): PRESSURE explosive into Sxx, Syy,
 * Szz (-dt wav w), VX / VY / VZ with dt R; traces add with atomics. */
int rsg_e3d_insert_source(const rsg_grid *g, const rsg_state_e3d *s, const rsg_traces *src, int it, void *stream);
/* recordData (waves.cpp:4247-4403): VX / VY / VZ gathers, or PRESSURE
 * (Sxx + Syy + Szz) / 3 divided by (3 rho vp^2 - 4 rho vs^2) / 3 at the
 * trace's cell of the unpadded model Vp, Vs, Rho (device, [nz][ny][nx]). */
int rsg_e3d_record_data(const rsg_grid *g, const rsg_state_e3d *s, const rsg_traces *rec, const float *Vp,
                        const float *Vs, const float *Rho, int it, void *stream);
/* backwardstepVelocity / backwardstepStress (waves.cpp:3260-4103), the
 * adjoint-state steps of FwiElastic3D: velocity from the derivatives of the
 * moduli-weighted stresses (L2M Sxx + L Syy + L Szz, M_xy Sxy, M_xz Sxz, ...
 * with dt R), stress from bare velocity derivatives (dt); the forward
 * steps' strips and free surface. */
int rsg_e3d_backward_velocity(const rsg_grid *g, const rsg_state_e3d *s, void *stream);
int rsg_e3d_backward_stress(const rsg_grid *g, const rsg_state_e3d *s, void *stream);
/* backwardstepVelocity with a caller workspace of 6 nx*ny*nz floats (device;
 * its contents are overwritten): the six weighted stresses are formed once
 * per cell, as the reference's wrk arrays, and the forward velocity kernel
 * runs on them (about 2.5x faster than the entry above, which forms them at
 * every stencil tap).  No reference counterpart beyond backwardstepVelocity
 * itself: the workspace replaces the reference's internal wrk array. */
int rsg_e3d_backward_velocity_ws(const rsg_grid *g, const rsg_state_e3d *s, float *workspace, void *stream);

/* ---- storage of FWI snapshots and Revolve checkpoints (the fp16 member of
 * the FWI settings below; the reference keeps them as T = float, revolve.cpp
 * 648-1271): stencil arithmetic stays FP32 in every case.
 *   RSG_STORE_FP32  plain copies;
 *   RSG_STORE_FP16  FP16 with one power-of-two scale per stored array (per
 *                   field or PML strip of a shot): 11 significant bits;
 *   RSG_STORE_S16   16-bit block floating point: blocks of 256 consecutive
 *                   values share a power-of-two exponent, each value is a
 *                   signed 16-bit integer (error <= 2^-16 of the block's
 *                   largest magnitude), 2 bytes a value plus 2 a block.
 * Both 16-bit kinds halve the memory of FP32; S16 keeps the gradient error
 * against an FP64 computation about five times lower than FP16. */
#define RSG_STORE_FP32 0
#define RSG_STORE_FP16 1
#define RSG_STORE_S16 2

/* ---- FWI gradient of one shot (FwiAcoustic2D::run_optimal, fwi.cpp:1159-1332,
 * and FwiAcoustic2D::run, fwi.cpp:1030-1156) */
#define RSG_FWI_REVOLVE 0 /* run_optimal: Revolve checkpoints in HBM (ncheck)        */
#define RSG_FWI_STORE 1   /* run_optimal's gradient, forward pressure of every step
                             kept in HBM (forward sweep once)                       */
#define RSG_FWI_FULL 2    /* run: pressure every snapinc-th step kept in HBM, in
                             place of the snapshot file                             */

/* Adjoint of the acoustic 2D gradient.  REFERENCE: the reference's
 * run_optimal / run (adjoint advanced with the forward operator, crossCorr
 * imaging).  EXACT (no reference counterpart): the transposed discrete step
 * (transposed staggered stencils and edge stencils, time-reversed PML,
 * residuals injected at the receivers with weight 1) and the transposed model
 * staggering, so vpgrad/rhograd are the gradient of the DIFFERENCE misfit of
 * the discrete forward map and wavgrad its gradient in the source wavelet.
 * EXACT takes RSG_FWI_REVOLVE and DIFFERENCE only; the image gradients are
 * summed over time in FP32 per shot and mapped to (Vp, rho) in FP64. */
#define RSG_ADJOINT_REFERENCE 0
#define RSG_ADJOINT_EXACT 1

typedef struct {
    const float *L, *Rx, *Rz, *tab; /* staggered model and PML tables (device)    */
    const float *Vp, *Rho;          /* unpadded model nz*nx (device), for crossCorr;
                                       with g->domdec the model on the grid       */
    int nx, nz;                     /* image sizes: the unpadded sizes, grid
                                       (nx+2lpml)(nz+2lpml); with g->domdec the
                                       grid's own sizes (crossCorr over the
                                       domain with no pad, fwi.cpp:535-544)       */
    rsg_traces src;                 /* PRESSURE source, wavelet in data             */
    rsg_traces obs;                 /* observed PRESSURE data at the receivers      */
    float *dmod, *dres;             /* out: modelled data, residuals (ntrace*nt)   */
    float *wavgrad;                 /* out: adjoint recorded at the source, or NULL */
    float *vpgrad, *rhograd;        /* in/out: accumulated gradients (nz*nx)        */
    double *misfit;                 /* out (host): DIFFERENCE misfit 0.5*sum r^2    */
    int ncheck;                     /* Revolve snapshots (RSG_FWI_REVOLVE)          */
    int mode;                       /* RSG_FWI_REVOLVE or RSG_FWI_STORE             */
    int *mode_used;                 /* out (host, optional): STORE falls back to
                                       REVOLVE when the history does not fit        */
    int snapinc;                    /* RSG_FWI_FULL: snapshot increment (Fwi snapinc) */
    int fp16;                       /* storage of snapshots / checkpoints
                                       (RSG_STORE_*): FP32, FP16 with a power-of-two
                                       scale per stored array, or S16              */
    int misfit_type;                /* rs_fwimisfit (utils.h:26): 0 DIFFERENCE,
                                       1 CORRELATION, 2 ADAPTIVE_GAUSS,
                                       3 ADAPTIVE_LINEAR, 4 ADAPTIVE_SINGLE (the
                                       acoustic classes' default branch:
                                       DIFFERENCE)                                   */
    const float *weight;            /* data weights like obs (setDataweight) or NULL */
    int adjoint;                    /* RSG_ADJOINT_REFERENCE or RSG_ADJOINT_EXACT    */
    float *srcilum;                 /* in/out (optional, like vpgrad): source
                                       illumination -= (-2/Vp) P_f^2 at every
                                       imaging step (crossCorr with srcilumset,
                                       fwi.cpp:567-569); REFERENCE adjoint only */
} rsg_fwi_a2d;

int rsg_fwi_a2d_gradient(const rsg_grid *g, const rsg_fwi_a2d *f, void *stream);
/* ---- misfit and residuals alone (FwiAcoustic2D/3D::computeMisfit and
 * computeResiduals, fwi.cpp:574-1027, 1468-1900): type as misfit_type above,
 * mod/rec/wei/res device arrays [nshot][ntr][nt], map [nshot][ntr][dim]
 * receiver map (traces with a negative entry count in no misfit), scaling
 * 1 (P) or -1 (VX/VZ data); *misfit (host, optional) = the misfit summed in
 * double over the shots.  dim 2 follows FwiAcoustic2D, whose types 1-3 end
 * with a 97th-percentile clip per shot; dim 3 follows FwiAcoustic3D, which
 * has none. */
int rsg_misfit_residuals(int type, const float *mod, const float *rec, const float *wei, float *res, const int *map,
                         int dim, int ntr, int nt, int nshot, int scaling, double *misfit, void *stream);
/* No counterpart in the reference.  rsg_misfit_residuals for callers that
 * cannot allocate device memory (code under hipGraph capture, serving loops,
 * harnesses that forbid hipMalloc): same arguments and results, no
 * hipMalloc / hipFree and no host allocation.  Its device scratch is one
 * fixed array of RSG_NOALLOC_SCRATCH_BYTES per GPU, held from the library's
 * load to the end of the process.  Limits: types 0 (DIFFERENCE),
 * 1 (CORRELATION) and 4 (its DIFFERENCE branch) -- the ADAPTIVE types need
 * 7*nt floats per trace and take rsg_misfit_residuals; at most 64 shots; the
 * scratch must hold 3*ntr*nshot floats of CORRELATION norms plus about
 * 1 KiB per shot (larger inputs return RSG_EINVAL).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams or
 * threads) race and must be serialized by the caller. */
#define RSG_NOALLOC_SCRATCH_BYTES (4u << 20)
int rsg_misfit_residuals_noalloc(int type, const float *mod, const float *rec, const float *wei, float *res,
                                 const int *map, int dim, int ntr, int nt, int nshot, int scaling, double *misfit,
                                 void *stream);
/* nshot shots on one grid in one set of launches per step (the reference
 * runs one shot per MPI worker, inversion.cpp:332-600, and stacks the shot
 * gradients, image.cpp:533-600): traces of shot k follow shot k-1 in
 * src/obs (map, shift, data), dmod/dres/wavgrad are [nshot][ntrace][nt],
 * vpgrad/rhograd and the misfit are summed over the shots. */
int rsg_fwi_a2d_gradient_batch(const rsg_grid *g, const rsg_fwi_a2d *f, int nshot, void *stream);
/* FwiAcoustic2D::crossCorr (fwi.cpp:511-571) on device arrays, one imaging
 * step: over ix, iz in [1, n-1) of the nx x nz image, vpgrad -= (-2/Vp) wsp
 * (Dx wrx + Dz wrz), rhograd -= (-1/rho) wsp (...) and += the velocity terms,
 * srcilum (optional) -= (-2/Vp) wsp^2.  wsp is padded by pads cells on each
 * side, wrx / wrz / Rx / Rz by padr; domdec = 1 takes pads = padr = 0 and
 * images on the padded domain, as the reference does. */
int rsg_crosscorr_a2d(const float *wsp, int pads, const float *wrx, const float *wrz, int padr,
                      const float *Vp, const float *Rho, const float *Rx, const float *Rz, int nx, int nz,
                      float dx, float dz, int domdec, float *vpgrad, float *rhograd, float *srcilum,
                      void *stream);
/* Snapshot2D/3D::writeSnap (snap.cpp:638-669, 1138-1172): out[k][j][i] =
 * f0 + f1 + f2 (those given, added in that order) at cell (i+pad, j+pad,
 * k+pad) of the g->nx x g->ny x g->nz grid, for the (nx-2pad) x (ny-2pad)
 * x (nz-2pad) interior (2D: ny = 1, no y padding). */
int rsg_snap_interior(const float *f0, const float *f1, const float *f2, const rsg_grid *g, int pad,
                      float *out, void *stream);
/* ---- Acoustic 3D FWI gradient of one shot (FwiAcoustic3D::run_optimal,
 * fwi.cpp:1991-2171): Revolve checkpoints in HBM (FP32, or FP16 with one
 * power-of-two scale per array), adjoint with the forward operator,
 * crossCorr (fwi.cpp:1403-1466) with the reference's 3D loop bounds. */
typedef struct {
    const float *L, *Rx, *Ry, *Rz, *tab; /* staggered model and PML tables (device)  */
    const float *Vp, *Rho;               /* unpadded model nz*ny*nx (device)         */
    int nx, ny, nz;                      /* unpadded sizes                           */
    rsg_traces src;                      /* PRESSURE source, wavelet in data         */
    rsg_traces obs;                      /* observed PRESSURE data                   */
    float *dmod, *dres;                  /* out: modelled data, residuals            */
    float *wavgrad;                      /* out (optional): wavelet gradient         */
    float *vpgrad, *rhograd;             /* in/out: accumulated gradients            */
    double *misfit;                      /* out (host)                               */
    int ncheck;                          /* Revolve snapshots                        */
    int fp16;                            /* checkpoint storage, RSG_STORE_*          */
    int misfit_type;                     /* 0 DIFFERENCE, 1 CORRELATION, 2 ADAPTIVE_GAUSS,
                                            3 ADAPTIVE_LINEAR, 4 (= DIFFERENCE); the
                                            residuals are not percentile-clipped
                                            (FwiAcoustic3D::computeResiduals,
                                            fwi.cpp:1651-1853)                        */
    const float *weight;                 /* data weights like obs, or NULL           */
    int adjoint;                         /* RSG_ADJOINT_REFERENCE (0) or RSG_ADJOINT_EXACT
                                            (no reference counterpart: the exact discrete
                                            adjoint of the forward step and of
                                            ModelAcoustic3D::staggerModels, the gradient
                                            of the DIFFERENCE misfit; misfit_type 0 or 4,
                                            models padded by the library)              */
} rsg_fwi_a3d;

int rsg_fwi_a3d_gradient(const rsg_grid *g, const rsg_fwi_a3d *f, void *stream);

/* ---- Elastic 3D FWI gradient of one shot (FwiElastic3D::run_optimal,
 * fwi.cpp:6770-6991): Revolve checkpoints of the nine fields and 36 strips
 * in HBM (RSG_STORE_*), Ux / Uy / Uz records, misfit and residuals
 * (computeMisfit / computeResiduals, fwi.cpp:5632-6196: DIFFERENCE,
 * CORRELATION, ADAPTIVE_GAUSS, other types the reference's default branch),
 * residuals injected cell by cell in trace order into the adjoint state of
 * backwardstepVelocity / Stress, crossCorr (fwi.cpp:5046-5291, the
 * reference's stencils and indices kept) into lambda / mu / rho kernels and
 * scaleGrad (fwi.cpp:5553-5630) at the end; the scaled Vp, Vs, rho gradients
 * are ADDED to vpgrad / vsgrad / rhograd (zero them for one shot's
 * gradient).  Residual sample 0 of CORRELATION and of the default branch is
 * left as the caller's array holds it, as the reference does. */
typedef struct {
    const float *L, *L2M, *Mxz, *Myz, *Mxy; /* staggered moduli (padded, device)    */
    const float *Rx, *Ry, *Rz, *tab;     /* staggered buoyancies, PML tables        */
    const float *Vp, *Vs, *Rho;          /* unpadded model nz*ny*nx (device)        */
    int nx, ny, nz;                      /* unpadded sizes                          */
    rsg_traces src;                      /* PRESSURE, VX, VY or VZ source            */
    rsg_traces obs[3];                   /* observed Ux, Uy, Uz (fields VX, VY, VZ,
                                            maps of makeMap for each; equal ntrace, nt) */
    float *dmod[3], *dres[3];            /* out: modelled data, residuals [ntrace][nt] */
    const float *weight[3];              /* data weights like obs, or NULL          */
    float *wavgrad;                      /* out (optional): wavelet gradient, src layout,
                                            of the source's field                   */
    float *vpgrad, *vsgrad, *rhograd;    /* in/out: gradients added, nz*ny*nx      */
    double *misfit;                      /* out (host, optional)                    */
    int ncheck;                          /* Revolve snapshots                       */
    int fp16;                            /* checkpoint storage, RSG_STORE_*         */
    int misfit_type;                     /* rs_fwimisfit: 0 DIFFERENCE, 1 CORRELATION,
                                            2 ADAPTIVE_GAUSS, 3 and 4 the default branch */
} rsg_fwi_e3d;

int rsg_fwi_e3d_gradient(const rsg_grid *g, const rsg_fwi_e3d *f, void *stream);

/* ---- Acoustic 2D reverse-time migration of one shot
 * (RtmAcoustic2D::run_optimal, rtm.cpp:423-544): forward steps with the
 * source on the Revolve schedule (FP32 checkpoints in HBM), the receiver
 * wavefield advanced with the forward operator from the recorded data, and
 * the imaging condition with subsurface offsets (RtmAcoustic2D::crossCorr,
 * rtm.cpp:191-219) at every reverse step.  The image is zeroed first, as
 * run_optimal's allocateImage does; the call returns when the image is
 * complete. */
typedef struct {
    const float *L, *Rx, *Rz, *tab;  /* staggered model and PML tables (device)      */
    int nx, nz;                      /* unpadded sizes                               */
    rsg_traces src;                  /* PRESSURE source, wavelet in data             */
    rsg_traces data;                 /* recorded PRESSURE data (GMAP)                */
    float *image;                    /* out: [nhz][nhx][nz][nx] (Image2D ki2D order)  */
    int nhx, nhz;                    /* subsurface offsets, hx = ihx - (nhx-1)/2     */
    int ncheck;                      /* Revolve snapshots                            */
} rsg_rtm_a2d;

int rsg_rtm_a2d_image(const rsg_grid *g, const rsg_rtm_a2d *r, void *stream);
/* nshot shots on one grid in one set of launches per step (the reference
 * migrates one shot per MPI worker and stacks the images, image.cpp): traces
 * of shot k follow shot k-1 in src/data (map, shift, data).  The image is the
 * sum over steps of the per-step sum over shots (shot order), whereas the
 * reference's stack is the sum over shots of whole per-shot images: the
 * two differ in rounding only (relative L2 5.7e-7 for 3 shots in the GPU
 * test), not bit for bit.  Every shot uses the same full grid: the per-shot
 * window offsets of stackImage (image.cpp:577-578) are not supported; shots
 * on their own windows take rsg_rtm_a2d_image one at a time. */
int rsg_rtm_a2d_batch(const rsg_grid *g, const rsg_rtm_a2d *r, int nshot, void *stream);

/* ---- Acoustic 3D reverse-time migration of one shot
 * (RtmAcoustic3D::run_optimal and crossCorr, rtm.cpp): as above on the 3D
 * propagator, with nhx x nhy x nhz subsurface offsets; image
 * [nhz][nhy][nhx][nz][ny][nx] (Image3D ki3D order, rtm.h:32). */
typedef struct {
    const float *L, *Rx, *Ry, *Rz, *tab; /* staggered model and PML tables (device)  */
    int nx, ny, nz;                      /* unpadded sizes (nz <= 65535)             */
    rsg_traces src;                      /* PRESSURE source, wavelet in data         */
    rsg_traces data;                     /* recorded PRESSURE data (GMAP)            */
    float *image;                        /* out, zeroed first                        */
    int nhx, nhy, nhz;                   /* subsurface offsets                       */
    int ncheck;                          /* Revolve snapshots                        */
} rsg_rtm_a3d;

int rsg_rtm_a3d_image(const rsg_grid *g, const rsg_rtm_a3d *r, void *stream);

/* ---- Elastic 2D FWI gradient (FwiElastic2D::run_optimal, fwi.cpp:3937-4143)
 * for a ragged batch of shots, each on its own aperture window.  Per shot:
 * forward elastic steps with the source and P / Ux / Uz gathers, the misfit
 * and residuals of the three records (FwiElastic2D::computeMisfit /
 * computeResiduals, fwi.cpp:2852-3773: DIFFERENCE, CORRELATION,
 * ADAPTIVE_GAUSS, ADAPTIVE_LINEAR, ADAPTIVE_SINGLE, optional data weights),
 * adjoint-state steps driven by the residuals, crossCorr
 * (fwi.cpp:2591-2775) at every reverse step with
 * Revolve checkpoints of all shots in HBM.  The windows' kernels are mapped
 * by scaleGrad (fwi.cpp:2778-2849) when scale = 1, then added into the
 * stacked images at x offset ox (the reference stacks shot images in its
 * inversion driver, image.cpp:533-600). */
typedef struct {
    rsg_grid g;                          /* window grid, padded (nx, nz incl. 2*lpml)  */
    const float *L, *L2M, *M, *Rx, *Rz;  /* staggered window model (device, padded)    */
    const float *Vp, *Vs, *Rho;          /* unpadded window model nz*nx (device)       */
    int ox;                              /* first column of the window in the images   */
    rsg_traces src;                      /* source: PRESSURE (explosive), VX or VZ     */
    rsg_traces obs[3];                   /* observed P, Ux (VX), Uz (VZ) records of one
                                            receiver set (maps and shifts per field)  */
    float *dmod[3], *dres[3];            /* out: modelled records and residuals        */
    float *wavgrad;                      /* out (optional): wavelet gradient           */
    const float *weight[3];              /* data weights of P, Ux, Uz like obs, or NULL
                                            (setDataweightp/x/z)                       */
} rsg_fwi_e2d_shot;

typedef struct {
    const float *tab;  /* PML tables of the shared lpml and dt (device)              */
    int ncheck;        /* Revolve snapshots                                            */
    int fp16;          /* checkpoint storage, RSG_STORE_*                              */
    int scale;         /* 1: Vp / Vs / rho gradients; 0: lambda / mu / rho kernels     */
    int nx, nz;        /* stacked image size                                           */
    float *grad[3];    /* in/out (device): gradients added here, nz*nx each            */
    double *misfit;    /* out (host): misfit summed over the shots                     */
    int misfit_type;   /* 0 DIFFERENCE, 1 CORRELATION, 2 ADAPTIVE_GAUSS, 3 ADAPTIVE_LINEAR,
                          4 ADAPTIVE_SINGLE (rs_fwimisfit, utils.h:26); the 97th-percentile
                          clips of types 1-3 are per shot, as the reference runs one
                          shot at a time; types 2-4 leave the P residual at zero      */
    int adjoint;       /* RSG_ADJOINT_REFERENCE (0): the reference's adjoint-state steps
                          and crossCorr.  RSG_ADJOINT_EXACT (no reference counterpart):
                          the exact discrete adjoint of the forward step and of
                          ModelElastic2D::staggerModels, the gradient of the DIFFERENCE
                          misfit of the discrete forward map (misfit_type 0 and windows
                          padded by the library, no domain models); scale 1 gives the
                          Vp / Vs / rho gradients, 0 those of lambda / mu / rho as
                          independent parameters; wavgrad = dJ/d(wavelet)            */
} rsg_fwi_e2d;

int rsg_fwi_e2d_gradient_batch(const rsg_fwi_e2d *f, const rsg_fwi_e2d_shot *shots, int nshot,
                               void *stream);

/* ---- the same gradient as a persistent plan (no reference counterpart: the
 * reference runs one shot per MPI worker, inversion.cpp:332-600).  create
 * validates the batch, allocates every state and checkpoint buffer and
 * captures the Revolve sweep as two hipGraphs on a stream of its own: up to
 * the forward step of the first turn, and the reverse sweep with the stacking.
 * run replays them, with the misfit and residuals computed between the two
 * (eagerly: the percentile clips read back to the host), and returns the
 * misfit; it waits for work already on `stream` before it starts and returns
 * with the plan's work finished.  The graphs hold the caller's pointers
 * (models, records, residual and gradient arrays): between runs the caller
 * may change their contents (a new model restaggered in place, new
 * observed data), not the arrays themselves.  Gradients are added into
 * f->grad on every run, so zero them in between for a fresh gradient.
 * f->misfit is ignored (run returns it).  One run at a time per plan (its
 * states and graphs are shared by all runs); separate plans are independent. */
typedef struct rsg_fwi_e2d_plan rsg_fwi_e2d_plan;
int rsg_fwi_e2d_plan_create(const rsg_fwi_e2d *f, const rsg_fwi_e2d_shot *shots, int nshot,
                            rsg_fwi_e2d_plan **plan);
int rsg_fwi_e2d_plan_run(rsg_fwi_e2d_plan *plan, double *misfit, void *stream);
int rsg_fwi_e2d_plan_destroy(rsg_fwi_e2d_plan *plan);

/* ---- elastic misfit and residuals alone (FwiElastic2D::computeMisfit and
 * computeResiduals, fwi.cpp:2852-3773), shot by shot as the reference runs
 * them: type as rsg_fwi_e2d.misfit_type; per shot the P, Ux, Uz records
 * [nrec][nt] (device), optional weights, the Ux receiver map [nrec][2]
 * (traces with a negative entry count in no misfit); *misfit (host,
 * optional) = the misfit summed over the shots in double. */
typedef struct {
    int nrec;
    const int *map;                      /* Ux receiver map (device)                   */
    const float *mod[3], *obs[3];        /* modelled and observed P, Ux, Uz (device)    */
    const float *weight[3];              /* data weights or NULL                        */
    float *res[3];                       /* out: residuals (device)                     */
} rsg_e2d_records;

int rsg_misfit_residuals_e2d(int type, const rsg_e2d_records *shots, int nshot, int nt, double *misfit,
                             void *stream);

/* ---- RSS files (rockseis::File, file.cpp:184-260; data.cpp:463-634) into
 * and out of device memory: one read / write of the whole file through
 * pinned host memory and one strided copy, where the reference's
 * Data2D::read / write go trace by trace.  The layout is the reference's:
 * "R0CKS", int data_format, header_format, type, size_t Nheader, Ndims,
 * 9 size_t n, 9 double d, 9 double o (start of data at byte 249), then the
 * data; DATA2D / DATA3D traces carry 4 (sx, sz, gx, gz) / 6 (sx, sy, sz,
 * gx, gy, gz) header floats before their nt samples.  Errors: RSG_EINVAL
 * for a missing, truncated or mismatching file. */
#define RSG_RSS_GENERIC 0
#define RSG_RSS_REGULAR 1
#define RSG_RSS_DATA2D 2
#define RSG_RSS_DATA3D 3
typedef struct {
    int data_format, header_format, type;  /* bytes per value / header value, rs_datatype */
    size_t nheader, ndims;
    size_t n[9];
    double d[9], o[9];
    size_t start;                          /* byte offset of the data                     */
} rsg_rss_header;
int rsg_rss_read_header(const char *path, rsg_rss_header *h);
/* coords (host, [ntrace][4] or [ntrace][6], may be NULL), data (device,
 * [ntrace][nt]); ntrace and nt must equal the file's n2 and n1. */
int rsg_rss_read_data2d(const char *path, float *coords, float *data, int ntrace, int nt, void *stream);
int rsg_rss_read_data3d(const char *path, float *coords, float *data, int ntrace, int nt, void *stream);
/* Data2D<float>::write / Data3D<float>::write: dt and ot are the float
 * members the reference widens to the header's doubles (the 3D file records
 * o1 = 0, as the reference's does). */
int rsg_rss_write_data2d(const char *path, const float *coords, const float *data, int ntrace, int nt, float dt,
                         float ot, void *stream);
int rsg_rss_write_data3d(const char *path, const float *coords, const float *data, int ntrace, int nt, float dt,
                         void *stream);
/* A REGULAR float array (models, images, Nheader 0) of n values (the product
 * of the file's non-zero sizes) into device memory, and back out with the
 * header h (data_format 4, nheader 0). */
int rsg_rss_read_regular(const char *path, float *data, size_t n, void *stream);
int rsg_rss_write_regular(const char *path, const rsg_rss_header *h, const float *data, void *stream);

/* ---- in-memory image stacking: Image2D::stackImage_parallel
 * (image.cpp:533-600) with both images in device memory.  Each local cell
 * (ix, iz) with padlx <= ix < nx-padhx, padlz <= iz < nz-padhz, for every
 * subsurface offset, is added to the global cell (ix + ix0, iz + iz0) when
 * that lies inside the global image; ix0 = rsg_image_offset(l->ox, l->dx,
 * g->ox, g->dx), iz0 likewise.  Images are [nhz][nhx][nz][nx]; they must
 * share nhx, nhz, dx and dz (RSG_EINVAL otherwise).  Stacking the shots one
 * call after another adds in the reference's order. */
typedef struct {
    int nx, nz, nhx, nhz;
    float dx, dz, ox, oz;
} rsg_image2d;
int rsg_stack_image_2d(float *gimg, const rsg_image2d *g, const float *limg, const rsg_image2d *l, int padlx,
                       int padhx, int padlz, int padhz, void *stream);

/* Model<T>::getLocalsize2d (model.cpp:271-340): the aperture window of a
 * gather.  coords [ntrace][4] hold sx, sz, gx, gz per trace; map 0 (SMAP)
 * centres the window on the sources, 1 (GMAP) on the receivers. */
int rsg_local_window_2d(const float *coords, int ntrace, float dx, float ox, float aperture, int map,
                        long long *start, long long *size);
/* ModelAcoustic2D::getDomainmodel (model.cpp:1511-1620) for one domain:
 * the window [start, start+size) of the global [nz][nx] model, padded by
 * lpml cells with edge values, as the arrays of the domain model
 * ((size+2lpml) x (nz+2lpml)): Vp, R, L = R Vp^2, Rx = 2/(R + R_x+1)
 * (1/R in the last global column), Rz = 2/(R + R_z+1), free surface
 * Rx *= 2, L = 0 on row lpml; *ox_out / *oz_out its origin. */
int rsg_domain_model_a2d(const float *vp, const float *rho, int nx, int nz, int lpml, int fs, float dx,
                         float dz, float ox, float oz, long long start, long long size, float *Vp, float *R,
                         float *L, float *Rx, float *Rz, float *ox_out, float *oz_out);
/* ModelElastic2D::getDomainmodel for one domain: as above, plus Vs, L2M,
 * M (moduli and buoyancies as model.cpp writes them for the domain model). */
int rsg_domain_model_e2d(const float *vp, const float *vs, const float *rho, int nx, int nz, int lpml, int fs,
                         float dx, float dz, float ox, float oz, long long start, long long size, float *Vp,
                         float *Vs, float *R, float *L, float *L2M, float *M, float *Rx, float *Rz,
                         float *ox_out, float *oz_out);
/* Image2D::stackImage_parallel's placement (image.cpp:577-578): the first
 * column (row) of a local image of origin ol and step dl in a global image
 * of origin og and step dg, (int)(ol/dl - og/dg) in float arithmetic. */
int rsg_image_offset(float ol, float dl, float og, float dg);

/* Revolve::revolve() action sequence (revolve.cpp:150-305): advance=0,
 * takeshot=1, restore=2, firsturn=3, youturn=4, terminate=5, error=6. */
int rsg_revolve_schedule(int nt, int snaps, int *actions, int *capos, int *checks, int maxn,
                         int *n);
/* Revolve::numforw (revolve.cpp:31-62) */
int rsg_revolve_numforw(int nt, int snaps);
/* Revolve::adjust (revolve.cpp:114-148): the snapshot count the
 * constructor takes when the requested one is invalid. */
int rsg_revolve_adjust(int nt);
/* The Revolve state machine one action at a time, for callers that drive
 * the schedule themselves (the Revolve class API: constructor,
 * revolve() and getCapo/getCheck/getFine, revolve.h:25-124).  init takes
 * the constructor's arguments (snaps outside [2, nt] replaced by
 * rsg_revolve_adjust(nt), *adjusted set to 1); next returns the action
 * numbered as rsg_revolve_schedule and updates the state. */
typedef struct {
    int steps, snaps, capo, fine, check, turn, oldfine;
    int ch[512];
} rsg_revolve_state;
int rsg_revolve_init(rsg_revolve_state *st, int nt, int snaps, int *adjusted);
int rsg_revolve_next(rsg_revolve_state *st);
/* Snapshot2D/3D cadence (snap.cpp:330-335): snapshots are taken at it % snapinc
 * == 0 and read back at reverse times t with (t - enddiff) % snapinc == 0. */
int rsg_snap_cadence(int nt, float dt, int snapinc, int *snapnt, int *enddiff);

#ifdef __cplusplus
}
#endif
#endif /* RSGPU_H */
