// Elastic 2D FWI gradient (FwiElastic2D::run_optimal, fwi.cpp:3937-4143) on
// the device, for a ragged batch of shots: every shot has its own aperture
// window (grid size, staggered model, fields, strips) and all shots advance
// together, one launch per half step (rsg_ragged2d plans).
//
// Per shot, as in the reference: forward elastic steps with the source and
// the P / Ux / Uz gathers; at the first turn the misfit of the three records
// and their residuals (any of the five misfit types, data weights; the
// percentile clips per shot, fwi.cpp:2852-3773, misfit.hip);
// the adjoint wavefield advanced with backwardstepVelocity/Stress
// (waves.cpp:1738-2080) and driven by the three residuals; at every reverse
// step crossCorr (fwi.cpp:2591-2775) into the lambda / mu / rho kernels of
// the window, plus the wavelet gradient.  Revolve checkpoints (fields and
// strips of every shot) live in HBM, FP32 or FP16 with one scale per array.
// At the end each window's kernels are mapped to Vp / Vs / rho gradients
// (scaleGrad, fwi.cpp:2778-2849) or kept as lambda / mu / rho, and all
// windows are stacked into the caller's images at their x offsets, in
// double, shot by shot.
#include <algorithm>
#include <memory>
#include <vector>

#include "rsg_adjoint.h"
#include "rsg_common.h"
#include "rsg_fwi.h"
#include "rsg_ragged.h"
#include "rsg_srcrec_dev.h"
#include "rsg_store16.h"

// This file is compiled with -ffp-contract=off (Makefile): every multiply and
// add is rounded separately, as in the reference's host arithmetic.

namespace rsg {

namespace {

// Everything a kernel needs of one shot.
struct ShotDev {
    int nxp, nzp, nx, nz, lpml, ox;
    float *fVx, *fVz, *fSxx, *fSzz, *fSxz;
    float *bVx, *bVz, *bSxx, *bSzz, *bSxz;
    const float *Rx, *Rz, *Vp, *Vs, *Rho;
    float *gRx, *gRz;  // exact adjoint: buoyancy gradients (force sources add to them)
    int sfield, nsrc, nrec;
    const int *smap;
    const float *sshift, *wav;
    float *wavgrad;
    const int *gmap[3];
    const float *gshift[3], *dobs[3];
    float *dmod[3], *dres[3];
    float *lam, *mu, *rho;
    // residual injection per touched cell: cells rcell[c][j], their
    // contributions rent[c][rstart[c][j] .. rstart[c][j+1]) = trace*4 + corner
    // (i1*2 + i2), in trace order
    int rcn[3];
    const int *rcell[3], *rstart[3], *rent[3];
};

constexpr int kIoThreads = 64;

// insertSource of the forward wavefield (waves.cpp:2085-2171): PRESSURE is
// explosive (Sxx and Szz -= dt*wav*w), VX / VZ scale by Rx / Rz.
__global__ void e2d_source_kernel(const ShotDev *__restrict__ sd, int it, int nt, float dt) {
    const ShotDev &s = sd[blockIdx.y];
    const int i = blockIdx.x * kIoThreads + threadIdx.x;
    if (i >= s.nsrc) return;
    if (s.sfield == RSG_PRESSURE) {
        inject2d_trace(i, s.fSxx, nullptr, -dt, s.nxp, s.lpml, s.smap, s.sshift, s.wav, nt, it,
                       (size_t)s.nxp * s.nzp);
        inject2d_trace(i, s.fSzz, nullptr, -dt, s.nxp, s.lpml, s.smap, s.sshift, s.wav, nt, it,
                       (size_t)s.nxp * s.nzp);
    } else if (s.sfield == RSG_VX) {
        inject2d_trace(i, s.fVx, s.Rx, dt, s.nxp, s.lpml, s.smap, s.sshift, s.wav, nt, it,
                       (size_t)s.nxp * s.nzp);
    } else {
        inject2d_trace(i, s.fVz, s.Rz, dt, s.nxp, s.lpml, s.smap, s.sshift, s.wav, nt, it,
                       (size_t)s.nxp * s.nzp);
    }
}

// recordData of P (0.5*(Sxx+Szz), no reciprocity scaling), Ux and Uz
// (waves.cpp:2174-2288).
__global__ void e2d_record_kernel(const ShotDev *__restrict__ sd, int it, int nt) {
    const ShotDev &s = sd[blockIdx.y];
    const int i = blockIdx.x * kIoThreads + threadIdx.x;
    if (i >= s.nrec) return;
    record2d_trace(i, s.fSxx, s.fSzz, nullptr, s.nxp, s.lpml, s.gmap[0], s.gshift[0], s.dmod[0], nt, it,
                   (size_t)s.nxp * s.nzp);
    record2d_trace(i, s.fVx, nullptr, nullptr, s.nxp, s.lpml, s.gmap[1], s.gshift[1], s.dmod[1], nt, it,
                   (size_t)s.nxp * s.nzp);
    record2d_trace(i, s.fVz, nullptr, nullptr, s.nxp, s.lpml, s.gmap[2], s.gshift[2], s.dmod[2], nt, it,
                   (size_t)s.nxp * s.nzp);
}

// The three residuals into the adjoint wavefield (insertSource of dataresP,
// dataresUx, dataresUz, fwi.cpp:4069-4071): P into Sxx and Szz with -dt, Ux
// into Vx with dt*Rx, Uz into Vz with dt*Rz, bilinear over each trace's 2x2
// cells.  Neighbouring receivers share corner cells (the staggered Ux / Uz
// positions fall between grid nodes), so one thread per touched cell
// (blockIdx.z = component) adds the cell's contributions in trace order, the
// reference's order, and the result does not depend on scheduling.  EXACT:
// the transpose of the gathers instead (the exact adjoint), P with 1/2 into
// Sxx and Szz, Ux and Uz with weight 1.
template <bool EXACT>
__global__ void e2d_residual_kernel(const ShotDev *__restrict__ sd, int it, int nt, float dt) {
    const ShotDev &s = sd[blockIdx.y];
    const int c = blockIdx.z;
    const int j = blockIdx.x * kIoThreads + threadIdx.x;
    if (j >= s.rcn[c]) return;
    const size_t p = (size_t)s.rcell[c][j];
    const float *shift = s.gshift[c], *res = s.dres[c];
    float *F = c == 0 ? s.bSxx : (c == 1 ? s.bVx : s.bVz);
    const float *Mod = c == 0 ? nullptr : (c == 1 ? s.Rx : s.Rz);
    const float scale = c == 0 ? -dt : dt;
    const float m = EXACT ? (c == 0 ? 0.5f : 1.f) : (Mod ? scale * Mod[p] : scale);
    float a = F[p], b = c == 0 ? s.bSzz[p] : 0.f;
    for (int q = s.rstart[c][j]; q < s.rstart[c][j + 1]; q++) {
        const int e = s.rent[c][q], i = e >> 2, i1 = (e >> 1) & 1, i2 = e & 1;
        const float sx = shift[2 * i], sz = shift[2 * i + 1];
        const float xt = i2 ? sx : 1 - sx, yt = i1 ? sz : 1 - sz;
        const float v = m * res[(size_t)i * nt + it] * xt * yt;
        a += v;
        b += v;
    }
    F[p] = a;
    if (c == 0) s.bSzz[p] = b;
}

// Exact adjoint of insertSource at step it: the wavelet gradient (PRESSURE:
// -dt sum w (lSxx + lSzz); VX / VZ: dt sum w R lV) and, for a force source,
// dt wav w lV into the buoyancy gradient (atomic: traces may share cells).
__global__ void e2d_exact_source_kernel(const ShotDev *__restrict__ sd, int it, int nt, float dt) {
    const ShotDev &s = sd[blockIdx.y];
    const int i = blockIdx.x * kIoThreads + threadIdx.x;
    if (i >= s.nsrc) return;
    const int mx = s.smap[2 * i], mz = s.smap[2 * i + 1];
    if (mx < 0 || mz < 0) return;
    const float sx = s.sshift[2 * i], sz = s.sshift[2 * i + 1];
    const float xtri[2] = {1 - sx, sx}, ytri[2] = {1 - sz, sz};
    const float w = s.wav[(size_t)i * nt + it];
    float acc = 0.f;
    for (int i1 = 0; i1 < 2; i1++)
        for (int i2 = 0; i2 < 2; i2++) {
            const size_t q = (size_t)(s.lpml + mz + i1) * s.nxp + (s.lpml + mx + i2);
            if (q >= (size_t)s.nxp * s.nzp) continue;  // as the forward injection skips it
            const float wt = xtri[i2] * ytri[i1];
            if (s.sfield == RSG_PRESSURE) {
                acc -= dt * wt * (s.bSxx[q] + s.bSzz[q]);
            } else {
                const bool vx = s.sfield == RSG_VX;
                const float l = vx ? s.bVx[q] : s.bVz[q];
                acc += dt * wt * (vx ? s.Rx[q] : s.Rz[q]) * l;
                unsafeAtomicAdd((vx ? s.gRx : s.gRz) + q, dt * w * wt * l);
            }
        }
    if (s.wavgrad) s.wavgrad[(size_t)i * nt + it] += acc;
}

// Wavelet gradient of crossCorr: the adjoint field of the source type
// gathered at source traces whose map falls on an interior cell
// ix, iz in [1, n-1) (fwi.cpp:2696-2758).
__global__ void e2d_wavgrad_kernel(const ShotDev *__restrict__ sd, int it, int nt) {
    const ShotDev &s = sd[blockIdx.y];
    const int i = blockIdx.x * kIoThreads + threadIdx.x;
    if (!s.wavgrad || i >= s.nsrc) return;
    const int mx = s.smap[2 * i], mz = s.smap[2 * i + 1];
    if (mx < 1 || mx >= s.nx - 1 || mz < 1 || mz >= s.nz - 1) return;
    const float sx = s.sshift[2 * i], sz = s.sshift[2 * i + 1];
    const float xtri[2] = {1 - sx, sx}, ytri[2] = {1 - sz, sz};
    float acc = s.wavgrad[(size_t)i * nt + it];
    for (int i1 = 0; i1 < 2; i1++)
        for (int i2 = 0; i2 < 2; i2++) {
            const size_t q = (size_t)(mz + s.lpml + i1) * s.nxp + (mx + s.lpml + i2);
            if (q >= (size_t)s.nxp * s.nzp) continue;
            const float v = s.sfield == RSG_PRESSURE ? s.bSxx[q] + s.bSzz[q]
                                                     : (s.sfield == RSG_VX ? s.bVx[q] : s.bVz[q]);
            acc += xtri[i2] * ytri[i1] * v;
        }
    s.wavgrad[(size_t)i * nt + it] = acc;
}

// crossCorr of one reverse step (fwi.cpp:2676-2772), interior ix, iz in
// [1, n-1) of every window (blockIdx.z = shot), with the reference's
// float / double promotions.
__global__ __launch_bounds__(256) void e2d_crosscorr_kernel(const ShotDev *__restrict__ sd, float dx,
                                                             float dz) {
    const ShotDev &s = sd[blockIdx.z];
    const int ix = blockIdx.x * 64 + threadIdx.x + 1;
    const int iz = blockIdx.y * 4 + threadIdx.y + 1;
    if (ix >= s.nx - 1 || iz >= s.nz - 1) return;
    const int n = s.nxp;
    const size_t q = (size_t)(iz + s.lpml) * n + ix + s.lpml;
    const float *wsx = s.fVx, *wsz = s.fVz, *wsxx = s.fSxx, *wszz = s.fSzz, *wsxz = s.fSxz;
    const float *wrx = s.bVx, *wrz = s.bVz, *rsxx = s.bSxx, *rszz = s.bSzz, *rsxz = s.bSxz;
    const float msxx = (wsx[q] - wsx[q - 1]) / dx;
    const float mszz = (wsz[q] - wsz[q - n]) / dz;
    const size_t m = (size_t)iz * s.nx + ix;
    s.lam[m] -= (msxx + mszz) * (rsxx[q] + rszz[q]);
    const float msxz1 = (wsx[q - 1] - wsx[q - 1 - n]) / dz + (wsz[q - n] - wsz[q - 1 - n]) / dx;
    const float msxz2 = (wsx[q] - wsx[q - n]) / dz + (wsz[q + 1 - n] - wsz[q - n]) / dx;
    const float msxz3 = (wsx[q - 1 + n] - wsx[q - 1]) / dz + (wsz[q] - wsz[q - 1]) / dx;
    const float msxz4 = (wsx[q + n] - wsx[q]) / dz + (wsz[q + 1] - wsz[q]) / dx;
    const float f1 = msxx * rszz[q] + mszz * rsxx[q];
    const float f2 = msxz1 * rsxz[q - 1 - n] + msxz2 * rsxz[q - n] + msxz3 * rsxz[q - 1] + msxz4 * rsxz[q];
    s.mu[m] = (float)((double)s.mu[m] + (2.0 * (double)f1 - 0.25 * (double)f2));
    const double rx0 = s.Rx[q], rx1 = s.Rx[q - 1], rz0 = s.Rz[q], rz1 = s.Rz[q - n];
    float ux = (float)(0.5 * (double)wrx[q] * rx0 * (double)(wsxx[q + 1] - wsxx[q]) / (double)dx);
    ux = (float)((double)ux + 0.5 * (double)wrx[q - 1] * rx1 * (double)(wsxx[q] - wsxx[q - 1]) / (double)dx);
    ux = (float)((double)ux + 0.5 * (double)wrx[q] * rx0 * (double)(wsxz[q] - wsxz[q - n]) / (double)dz);
    ux = (float)((double)ux + 0.5 * (double)wrx[q - 1] * rx1 * (double)(wsxz[q - 1] - wsxz[q - 1 - n]) / (double)dz);
    float uz = (float)(0.5 * (double)wrz[q] * rz0 * (double)(wsxz[q] - wsxz[q - 1]) / (double)dx);
    uz = (float)((double)uz + 0.5 * (double)wrz[q - n] * rz1 * (double)(wsxz[q - n] - wsxz[q - 1 - n]) / (double)dx);
    uz = (float)((double)uz + 0.5 * (double)wrz[q] * rz0 * (double)(wszz[q + n] - wszz[q]) / (double)dz);
    uz = (float)((double)uz + 0.5 * (double)wrz[q - n] * rz1 * (double)(wszz[q] - wszz[q - n]) / (double)dz);
    s.rho[m] -= (ux + uz);
}

// Stack every window's kernels into the caller's images (x offset ox, z
// from 0), shots in order, in double; with `scale` each window value is
// first mapped by scaleGrad with that window's Vp, Vs, rho.
__global__ __launch_bounds__(256) void e2d_stack_kernel(const ShotDev *__restrict__ sd, int nshot, int nx,
                                                         int nz, float *g0, float *g1, float *g2, int scale) {
    const int x = blockIdx.x * 256 + threadIdx.x, z = blockIdx.y;
    if (x >= nx || z >= nz) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int k = 0; k < nshot; k++) {
        const ShotDev &s = sd[k];
        const int wx = x - s.ox;
        if (wx < 0 || wx >= s.nx || z >= s.nz) continue;
        const size_t m = (size_t)z * s.nx + wx;
        float lam = s.lam[m], mu = s.mu[m], rho = s.rho[m];
        if (scale) {
            const float vp = s.Vp[m], vs = s.Vs[m], r = s.Rho[m];
            const float vpscale = (float)(2.0 * (double)r * (double)vp);
            const float vsscale = (float)(2.0 * (double)r * (double)vs);
            const float rhoscale1 = (float)((double)(vp * vp) - 2.0 * (double)vs * (double)vs);
            const float rhoscale2 = vs * vs;
            const float l = lam, u = mu;
            lam = vpscale * l;
            mu = (float)((double)(vsscale * u) - 2.0 * (double)vsscale * (double)l);
            rho = rhoscale1 * l + rhoscale2 * u + rho;
        }
        a0 += lam;
        a1 += mu;
        a2 += rho;
    }
    const size_t p = (size_t)z * nx + x;
    g0[p] = (float)((double)g0[p] + a0);
    g1[p] = (float)((double)g1[p] + a1);
    g2[p] = (float)((double)g2[p] + a2);
}

struct DevMem {
    std::vector<void *> ptrs;
    ~DevMem() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **out, size_t n) {
        void *p = nullptr;
        if (n == 0) n = 1;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return fail(RSG_ENOMEM, "fwi_e2d: device allocation failed");
        ptrs.push_back(p);
        *out = (T *)p;
        return RSG_OK;
    }
};

struct PlanGuard {
    Ragged2D p;
    ~PlanGuard() { ragged2d_destroy(&p); }
};

struct ExactGuard {
    E2DExact p;
    ~ExactGuard() { e2d_exact_destroy(&p); }
};

// Is strip slot i of the elastic state an x-side strip (lpml x nz)?
inline bool e2d_xstrip(int i) { return (i / 4) % 2 == 0; }

}  // namespace

// One batched elastic FWI gradient: the validated inputs, every device buffer
// and the two halves of the Revolve sweep.  Phase A runs from the zeroed
// state through the forward step of the first turn; the misfit and residuals
// follow (eager: the percentile clips need host round trips); phase B runs
// the reverse sweep with its recomputations and stacks the windows' kernels.
// Both phases only enqueue kernels and device-to-device copies on the
// stream, so a plan captures each as a hipGraph once and replays it.
// Map and image offset of a window: lpml, or 0 on a domain model whose
// arrays already hold the padding (FwiElastic2D::crossCorr, fwi.cpp:2666-2675).
inline int shot_pad(const rsg_grid &g) { return g.domdec ? 0 : g.lpml; }

struct E2DFwi {
    const rsg_fwi_e2d *f = nullptr;
    const rsg_fwi_e2d_shot *sh = nullptr;
    int ns = 0, nt = 0, lpml = 0, order = 0;
    float dt = 0.f;
    size_t maxsrc = 0, maxrec = 0, maxrcell = 0;
    int maxnx = 0, maxnz = 0;
    std::vector<size_t> off, segs, img_off;
    size_t total = 0, cstride = 0;
    int nseg = 0, nchunk = 0;
    bool f16 = false;  // FP16 or S16 checkpoints (2 bytes a value)
    int store = kStoreF32;
    // S16: blocks restart at every field and strip of every shot (segs); a
    // checkpoint is sblk blocks of kS16Block values and sblk exponents
    short *cks = nullptr, *ckexp = nullptr;
    size_t *dsblk = nullptr, sblk = 0;
    int *dsseg = nullptr;
    std::unique_ptr<Revolve> rev;
    bool reverse = false;
    DevMem dm;
    float *fwb = nullptr, *bwb = nullptr, *imgs = nullptr, *ck = nullptr;
    half_t *ck16 = nullptr;
    unsigned *ckamax = nullptr, *ckpart = nullptr;
    size_t *dsegs = nullptr;
    Chunk16 *dchunks = nullptr;
    ShotDev *dsd = nullptr;
    double *dmis = nullptr;
    std::vector<E2DRecs> recs;
    PlanGuard pf, pb;
    dim3 iob_s, iob_r, iob_c, xcg;
    // exact adjoint: per-shot work arrays (in wbase), reverse-step kernels,
    // the state before the first turn's forward step, FP64 model-gradient
    // scratch of the largest padded window
    bool exact = false;
    std::vector<E2DExactWork> work;
    float *wbase = nullptr, *turn_state = nullptr;
    size_t wtotal = 0;
    double *pad3 = nullptr;
    ExactGuard ex;

    int init(const rsg_fwi_e2d *f_, const rsg_fwi_e2d_shot *sh_, int ns_, hipStream_t st) {
        f = f_;
        sh = sh_;
        ns = ns_;
        const rsg_grid &g0 = sh[0].g;
        nt = sh[0].src.nt;
        lpml = g0.lpml;
        order = g0.order;
        dt = g0.dt;
        // ---- validation
        for (int k = 0; k < ns; k++) {
            const rsg_fwi_e2d_shot &s = sh[k];
            const int pad = shot_pad(s.g), nx = s.g.nx - 2 * pad, nz = s.g.nz - 2 * pad;
            if (s.g.lpml != lpml || s.g.order != order || s.g.dt != dt || s.g.dx != g0.dx || s.g.dz != g0.dz ||
                s.g.ny != 1)
                return fail(RSG_EINVAL, "fwi_e2d: shots must share lpml, order, dx, dz, dt (2D grids)");
            if (!s.L || !s.L2M || !s.M || !s.Rx || !s.Rz || !s.Vp || !s.Vs || !s.Rho)
                return fail(RSG_EINVAL, "fwi_e2d: null model array");
            if (s.src.nt != nt || s.src.ntrace < 0 ||
                (s.src.ntrace > 0 && (!s.src.map || !s.src.shift || !s.src.data)))
                return fail(RSG_EINVAL, "fwi_e2d: every source needs nt samples and its map/shift/data");
            if (s.src.field != RSG_PRESSURE && s.src.field != RSG_VX && s.src.field != RSG_VZ)
                return fail(RSG_EINVAL, "fwi_e2d: source field must be P, VX or VZ");
            const int want[3] = {RSG_PRESSURE, RSG_VX, RSG_VZ};
            for (int c = 0; c < 3; c++) {
                const rsg_traces &o = s.obs[c];
                if (o.field != want[c] || o.nt != nt || o.ntrace != s.obs[0].ntrace || !o.map || !o.shift ||
                    !o.data || !s.dmod[c] || !s.dres[c])
                    return fail(RSG_EINVAL, "fwi_e2d: obs[0..2] must be P, VX, VZ records of one receiver set");
            }
            if (nx < 3 || nz < 3 || s.ox < 0 || s.ox + nx > f->nx || nz > f->nz)
                return fail(RSG_EINVAL, "fwi_e2d: window outside the stacked image");
            maxsrc = std::max(maxsrc, (size_t)s.src.ntrace);
            maxrec = std::max(maxrec, (size_t)s.obs[0].ntrace);
            maxnx = std::max(maxnx, nx);
            maxnz = std::max(maxnz, nz);
        }
        // ---- state layout: per shot [Vx Vz Sxx Szz Sxz | 16 strips], shots back to back
        off.assign(ns + 1, 0);
        segs.reserve((size_t)ns * 21 + 1);
        for (int k = 0; k < ns; k++) {
            const size_t n = (size_t)sh[k].g.nx * sh[k].g.nz, sx = (size_t)lpml * sh[k].g.nz,
                         sz = (size_t)sh[k].g.nx * lpml;
            size_t o = off[k];
            for (int j = 0; j < 5; j++, o += n) segs.push_back(o);
            for (int i = 0; i < RSG_E2D_NSTRIP; i++) {
                segs.push_back(o);
                o += e2d_xstrip(i) ? sx : sz;
            }
            off[k + 1] = o;
        }
        segs.push_back(off[ns]);
        total = off[ns];
        nseg = (int)segs.size() - 1;
        exact = f->adjoint == RSG_ADJOINT_EXACT;
        if (f->adjoint != RSG_ADJOINT_REFERENCE && !exact)
            return fail(RSG_EINVAL, "fwi_e2d: adjoint must be RSG_ADJOINT_REFERENCE or RSG_ADJOINT_EXACT");
        if (exact) {
            if (f->misfit_type != 0) return fail(RSG_EINVAL, "fwi_e2d: the exact adjoint takes the DIFFERENCE misfit");
            for (int k = 0; k < ns; k++)
                if (sh[k].g.domdec)
                    return fail(RSG_EINVAL, "fwi_e2d: the exact adjoint takes windows padded by the library "
                                            "(no domain models)");
        }
        rev.reset(new Revolve(nt, f->ncheck));
        store = f->fp16;
        if (store != kStoreF32 && store != kStoreF16 && store != kStoreS16)
            return fail(RSG_EINVAL, "fwi_e2d: checkpoint storage must be 0 (FP32), 1 (FP16) or 2 (S16)");
        f16 = store != kStoreF32;
        if (store == kStoreF16 && nseg > 65535)
            return fail(RSG_EINVAL, "fwi_e2d: FP16 checkpoints take at most 3120 shots");
        {
            size_t free_b = 0, total_b = 0;
            RSG_HIP(hipMemGetInfo(&free_b, &total_b));
            size_t img = 0;
            for (int k = 0; k < ns; k++) {
                const int pad = shot_pad(sh[k].g);
                img += 3 * (size_t)(sh[k].g.nx - 2 * pad) * (sh[k].g.nz - 2 * pad);
            }
            size_t need = total * 4 * 2 + img * 4 + total * (f16 ? 2 : 4) * (size_t)rev->snaps;
            if (exact) {
                size_t nf = 0;
                for (int k = 0; k < ns; k++) nf += (size_t)sh[k].g.nx * sh[k].g.nz;
                need += nf * 9 * 4 + total * 4;
            }
            if (need + ((size_t)1 << 30) > free_b)
                return fail(RSG_ENOMEM, "fwi_e2d: states and checkpoints do not fit in device memory");
        }
        if (int rc = dm.alloc(&fwb, total)) return rc;
        if (int rc = dm.alloc(&bwb, total)) return rc;
        img_off.assign(ns + 1, 0);
        for (int k = 0; k < ns; k++) {
            const int pad = shot_pad(sh[k].g);
            img_off[k + 1] = img_off[k] + 3 * (size_t)(sh[k].g.nx - 2 * pad) * (sh[k].g.nz - 2 * pad);
        }
        if (int rc = dm.alloc(&imgs, img_off[ns])) return rc;
        if (exact) {
            size_t maxp = 0;
            for (int k = 0; k < ns; k++) {
                const size_t n = (size_t)sh[k].g.nx * sh[k].g.nz;
                wtotal += 9 * n;
                maxp = std::max(maxp, n);
            }
            if (int rc = dm.alloc(&wbase, wtotal)) return rc;
            if (int rc = dm.alloc(&turn_state, total)) return rc;
            if (int rc = dm.alloc(&pad3, 3 * maxp)) return rc;
            work.resize(ns);
            float *q = wbase;
            for (int k = 0; k < ns; k++) {
                const size_t n = (size_t)sh[k].g.nx * sh[k].g.nz;
                E2DExactWork &w = work[k];
                for (int i = 0; i < 4; i++, q += n) w.W[i] = q;
                w.gL = q;
                w.gL2M = q + n;
                w.gM = q + 2 * n;
                w.gRx = q + 3 * n;
                w.gRz = q + 4 * n;
                q += 5 * n;
            }
        }
        // ---- per-shot states, ragged plans, kernel descriptors
        std::vector<rsg_grid> grids(ns);
        std::vector<rsg_state> sf(ns), sb(ns);
        std::vector<ShotDev> hd(ns);
        for (int k = 0; k < ns; k++) {
            const rsg_fwi_e2d_shot &s = sh[k];
            grids[k] = s.g;
            const size_t n = (size_t)s.g.nx * s.g.nz, sx = (size_t)lpml * s.g.nz, sz = (size_t)s.g.nx * lpml;
            for (int w = 0; w < 2; w++) {
                float *b = (w ? bwb : fwb) + off[k];
                rsg_state &r = w ? sb[k] : sf[k];
                r = rsg_state{};
                r.Vx = b;
                r.Vz = b + n;
                r.Sxx = b + 2 * n;
                r.Szz = b + 3 * n;
                r.Sxz = b + 4 * n;
                float *q = b + 5 * n;
                for (int i = 0; i < RSG_E2D_NSTRIP; i++) {
                    r.psi[i] = q;
                    q += e2d_xstrip(i) ? sx : sz;
                }
                r.L = s.L;
                r.L2M = s.L2M;
                r.M = s.M;
                r.Rx = s.Rx;
                r.Rz = s.Rz;
                r.tab = f->tab;
            }
            ShotDev &d = hd[k];
            d.nxp = s.g.nx;
            d.nzp = s.g.nz;
            // image window and map offset: the unpadded window, or the whole
            // padded domain of a domain model (fwi.cpp:2666-2675)
            d.nx = s.g.nx - 2 * shot_pad(s.g);
            d.nz = s.g.nz - 2 * shot_pad(s.g);
            d.lpml = shot_pad(s.g);
            d.ox = s.ox;
            d.fVx = sf[k].Vx;
            d.fVz = sf[k].Vz;
            d.fSxx = sf[k].Sxx;
            d.fSzz = sf[k].Szz;
            d.fSxz = sf[k].Sxz;
            d.bVx = sb[k].Vx;
            d.bVz = sb[k].Vz;
            d.bSxx = sb[k].Sxx;
            d.bSzz = sb[k].Szz;
            d.bSxz = sb[k].Sxz;
            d.Rx = s.Rx;
            d.Rz = s.Rz;
            d.gRx = exact ? work[k].gRx : nullptr;
            d.gRz = exact ? work[k].gRz : nullptr;
            d.Vp = s.Vp;
            d.Vs = s.Vs;
            d.Rho = s.Rho;
            d.sfield = s.src.field;
            d.nsrc = s.src.ntrace;
            d.nrec = s.obs[0].ntrace;
            d.smap = s.src.map;
            d.sshift = s.src.shift;
            d.wav = s.src.data;
            d.wavgrad = s.wavgrad;
            for (int c = 0; c < 3; c++) {
                d.gmap[c] = s.obs[c].map;
                d.gshift[c] = s.obs[c].shift;
                d.dobs[c] = s.obs[c].data;
                d.dmod[c] = s.dmod[c];
                d.dres[c] = s.dres[c];
            }
            const size_t ni = (size_t)d.nx * d.nz;
            d.lam = imgs + img_off[k];
            d.mu = d.lam + ni;
            d.rho = d.mu + ni;
        }
        // residual injection tables: per shot and component the touched cells,
        // each with its (trace, corner) contributions in trace order
        {
            std::vector<int> tab;
            std::vector<size_t> base(ns * 3 * 3);  // offsets of cells, starts, entries
            std::vector<int> mp;
            for (int k = 0; k < ns; k++) {
                const int nrec = sh[k].obs[0].ntrace, nxp = sh[k].g.nx, pad = shot_pad(sh[k].g);
                const long ncell = (long)sh[k].g.nx * sh[k].g.nz;
                for (int c = 0; c < 3; c++) {
                    mp.resize(2 * (size_t)nrec);
                    if (nrec) {  // the caller may have written the maps on st
                        RSG_HIP(hipMemcpyAsync(mp.data(), sh[k].obs[c].map, sizeof(int) * 2 * nrec,
                                               hipMemcpyDeviceToHost, st));
                        RSG_HIP(hipStreamSynchronize(st));
                    }
                    std::vector<std::pair<int, int>> ent;  // (cell, trace*4 + corner)
                    ent.reserve(4 * (size_t)nrec);
                    for (int i = 0; i < nrec; i++) {
                        const int mx = mp[2 * i], mz = mp[2 * i + 1];
                        if (mx < 0 || mz < 0) continue;
                        if ((long)(pad + mz + 1) * nxp + pad + mx + 1 >= ncell)
                            return fail(RSG_EINVAL, "fwi_e2d: receiver map outside the padded grid");
                        for (int i1 = 0; i1 < 2; i1++)
                            for (int i2 = 0; i2 < 2; i2++)
                                ent.emplace_back((pad + mz + i1) * nxp + (pad + mx + i2), i * 4 + i1 * 2 + i2);
                    }
                    std::stable_sort(ent.begin(), ent.end(),
                                     [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
                                         return x.first < y.first;
                                     });
                    std::vector<int> cells, starts, es;
                    for (size_t q = 0; q < ent.size(); q++) {
                        if (q == 0 || ent[q].first != ent[q - 1].first) {
                            cells.push_back(ent[q].first);
                            starts.push_back((int)q);
                        }
                        es.push_back(ent[q].second);
                    }
                    starts.push_back((int)ent.size());
                    hd[k].rcn[c] = (int)cells.size();
                    maxrcell = std::max(maxrcell, cells.size());
                    const size_t b = 9 * (size_t)k + 3 * c;
                    base[b] = tab.size();
                    tab.insert(tab.end(), cells.begin(), cells.end());
                    base[b + 1] = tab.size();
                    tab.insert(tab.end(), starts.begin(), starts.end());
                    base[b + 2] = tab.size();
                    tab.insert(tab.end(), es.begin(), es.end());
                }
            }
            int *dtab = nullptr;
            if (int rc = dm.alloc(&dtab, tab.size())) return rc;
            RSG_HIP(hipMemcpyAsync(dtab, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice, st));
            RSG_HIP(hipStreamSynchronize(st));  // the host table goes out of scope
            for (int k = 0; k < ns; k++)
                for (int c = 0; c < 3; c++) {
                    const size_t b = 9 * (size_t)k + 3 * c;
                    hd[k].rcell[c] = dtab + base[b];
                    hd[k].rstart[c] = dtab + base[b + 1];
                    hd[k].rent[c] = dtab + base[b + 2];
                }
        }
        // the three records of every shot for the misfit and residuals
        recs.assign(ns, E2DRecs{});
        for (int k = 0; k < ns; k++) {
            E2DRecs &r = recs[k];
            r.nrec = sh[k].obs[0].ntrace;
            r.map = sh[k].obs[1].map;  // traces count in the misfit by their Ux map (fwi.cpp:2911)
            for (int c = 0; c < 3; c++) {
                r.mod[c] = sh[k].dmod[c];
                r.obs[c] = sh[k].obs[c].data;
                r.wei[c] = sh[k].weight[c];
                r.res[c] = sh[k].dres[c];
            }
        }
        if (int rc = ragged2d_create(RSG_KIND_E2D, ns, grids.data(), sf.data(), &pf.p)) return rc;
        if (int rc = ragged2d_create(RSG_KIND_E2D, ns, grids.data(), sb.data(), &pb.p)) return rc;
        if (exact)
            if (int rc = e2d_exact_create(ns, grids.data(), sf.data(), sb.data(), work.data(), &ex.p, st)) return rc;
        if (int rc = dm.alloc(&dsd, ns)) return rc;
        if (int rc = dm.alloc(&dmis, 1)) return rc;
        RSG_HIP(hipMemcpyAsync(dsd, hd.data(), sizeof(ShotDev) * ns, hipMemcpyHostToDevice, st));
        // FP16 checkpoint slots start at multiples of 4 values (8-byte aligned)
        cstride = (total + 3) & ~(size_t)3;
        const std::vector<Chunk16> chunks = store == kStoreF16 ? store16_chunks(segs) : std::vector<Chunk16>();
        nchunk = (int)chunks.size();
        const std::vector<size_t> sblkoff = store == kStoreS16 ? s16_block_offsets(segs) : std::vector<size_t>();
        if (store == kStoreS16) {
            sblk = sblkoff.back();
            if (int rc = dm.alloc(&cks, sblk * kS16Block * (size_t)rev->snaps)) return rc;
            if (int rc = dm.alloc(&ckexp, sblk * rev->snaps)) return rc;
            if (int rc = dm.alloc(&dsegs, segs.size())) return rc;
            if (int rc = dm.alloc(&dsblk, sblkoff.size())) return rc;
            RSG_HIP(hipMemcpyAsync(dsegs, segs.data(), sizeof(size_t) * segs.size(), hipMemcpyHostToDevice, st));
            RSG_HIP(hipMemcpyAsync(dsblk, sblkoff.data(), sizeof(size_t) * sblkoff.size(), hipMemcpyHostToDevice, st));
            const std::vector<int> sseg = s16_block_segments(sblkoff);
            if (int rc = dm.alloc(&dsseg, sseg.size())) return rc;
            RSG_HIP(hipMemcpyAsync(dsseg, sseg.data(), sizeof(int) * sseg.size(), hipMemcpyHostToDevice, st));
            RSG_HIP(hipStreamSynchronize(st));  // sseg goes out of scope
        } else if (f16) {
            if (int rc = dm.alloc(&dsegs, segs.size())) return rc;
            if (int rc = dm.alloc(&dchunks, chunks.size())) return rc;
            if (int rc = dm.alloc(&ckpart, chunks.size())) return rc;
            if (int rc = dm.alloc(&ck16, cstride * (size_t)rev->snaps)) return rc;
            if (int rc = dm.alloc(&ckamax, (size_t)nseg * rev->snaps)) return rc;
            RSG_HIP(hipMemcpyAsync(dsegs, segs.data(), sizeof(size_t) * segs.size(), hipMemcpyHostToDevice, st));
            RSG_HIP(
                hipMemcpyAsync(dchunks, chunks.data(), sizeof(Chunk16) * chunks.size(), hipMemcpyHostToDevice, st));
        } else {
            if (int rc = dm.alloc(&ck, total * (size_t)rev->snaps)) return rc;
        }
        RSG_HIP(hipStreamSynchronize(st));  // the host tables above go out of scope
        iob_s = dim3((unsigned)((maxsrc + kIoThreads - 1) / kIoThreads), ns);
        iob_r = dim3((unsigned)((maxrec + kIoThreads - 1) / kIoThreads), ns);
        iob_c = dim3((unsigned)std::max<size_t>((maxrcell + kIoThreads - 1) / kIoThreads, 1), ns, 3);
        xcg = dim3((maxnx + 63) / 64, (maxnz + 3) / 4, ns);
        return RSG_OK;
    }

    int fstep(int it, bool rec, hipStream_t st) {
        if (int rc = ragged2d_step(&pf.p, 0, st)) return rc;
        if (int rc = ragged2d_step(&pf.p, 1, st)) return rc;
        if (maxsrc) {
            e2d_source_kernel<<<iob_s, kIoThreads, 0, st>>>(dsd, it, nt, dt);
            RSG_LAUNCH_CHECK("e2d_source_kernel");
        }
        if (rec && maxrec) {
            e2d_record_kernel<<<iob_r, kIoThreads, 0, st>>>(dsd, it, nt);
            RSG_LAUNCH_CHECK("e2d_record_kernel");
        }
        return RSG_OK;
    }

    // One exact reverse step at time it: fw holds the forward state before
    // step it (advanced here by its velocity half step), bw the adjoint
    // state after it.
    int exact_step(int it, hipStream_t st) {
        if (int rc = ragged2d_step(&pf.p, 0, st)) return rc;
        if (maxrec) {
            e2d_residual_kernel<true><<<iob_c, kIoThreads, 0, st>>>(dsd, it, nt, dt);
            RSG_LAUNCH_CHECK("e2d_residual_kernel");
        }
        if (maxsrc) {
            e2d_exact_source_kernel<<<iob_s, kIoThreads, 0, st>>>(dsd, it, nt, dt);
            RSG_LAUNCH_CHECK("e2d_exact_source_kernel");
        }
        return e2d_exact_reverse(ex.p, st);
    }

    int reverse_tail(int it, hipStream_t st) {
        if (maxrec) {
            e2d_residual_kernel<false><<<iob_c, kIoThreads, 0, st>>>(dsd, it, nt, dt);
            RSG_LAUNCH_CHECK("e2d_residual_kernel");
        }
        e2d_crosscorr_kernel<<<xcg, dim3(64, 4), 0, st>>>(dsd, sh[0].g.dx, sh[0].g.dz);
        RSG_LAUNCH_CHECK("e2d_crosscorr_kernel");
        if (maxsrc) {
            e2d_wavgrad_kernel<<<iob_s, kIoThreads, 0, st>>>(dsd, it, nt);
            RSG_LAUNCH_CHECK("e2d_wavgrad_kernel");
        }
        return RSG_OK;
    }

    // Phase A: zero every state and output, then the Revolve loop
    // (fwi.cpp:3986-4124) up to and including the forward step of the
    // first turn.  Returns that step's index in *turn.
    int phase_a(hipStream_t st, int *turn) {
        if (int rc = launch_zero(fwb, total * sizeof(float), st)) return rc;
        if (int rc = launch_zero(bwb, total * sizeof(float), st)) return rc;
        if (int rc = launch_zero(imgs, img_off[ns] * sizeof(float), st)) return rc;
        if (int rc = launch_zero(dmis, sizeof(double), st)) return rc;
        if (exact)
            if (int rc = launch_zero(wbase, wtotal * sizeof(float), st)) return rc;
        for (int k = 0; k < ns; k++) {
            for (int c = 0; c < 3; c++)
                if (int rc = launch_zero(sh[k].dmod[c], sizeof(float) * sh[k].obs[0].ntrace * nt, st)) return rc;
            if (sh[k].wavgrad)
                if (int rc = launch_zero(sh[k].wavgrad, sizeof(float) * sh[k].src.ntrace * nt, st)) return rc;
        }
        rev.reset(new Revolve(nt, f->ncheck));
        reverse = false;
        for (;;) {
            const int oldcapo = rev->capo;
            const Revolve::Action a = rev->next();
            const int capo = rev->capo;
            if (a == Revolve::ADVANCE) {
                for (int it = oldcapo; it < capo; it++)
                    if (int rc = fstep(it, true, st)) return rc;
            } else if (a == Revolve::FIRSTURN) {
                if (exact)  // the exact reverse step starts from the state before this step
                    RSG_HIP(hipMemcpyAsync(turn_state, fwb, total * sizeof(float), hipMemcpyDeviceToDevice, st));
                if (int rc = fstep(capo, true, st)) return rc;
                *turn = capo;
                return RSG_OK;
            } else if (a == Revolve::TAKESHOT) {
                if (int rc = takeshot(st)) return rc;
            } else {
                return fail(RSG_EINVAL, "fwi_e2d: unexpected Revolve action before the first turn");
            }
        }
    }

    int takeshot(hipStream_t st) {
        if (store == kStoreS16)
            return pack_s16(fwb, cks + sblk * kS16Block * rev->check, ckexp + sblk * rev->check, dsegs, dsblk, dsseg,
                            sblk, st);
        if (f16)
            return pack16_table(fwb, ck16 + cstride * rev->check, ckamax + (size_t)nseg * rev->check, ckpart, dsegs,
                                dchunks, nchunk, nseg, total, st);
        RSG_HIP(hipMemcpyAsync(ck + total * rev->check, fwb, total * sizeof(float), hipMemcpyDeviceToDevice, st));
        return RSG_OK;
    }

    int misfit(hipStream_t st) {
        if (!maxrec) return RSG_OK;
        return elastic_misfit_residuals(f->misfit_type, recs, nt, dmis, st);
    }

    // Phase B: the first turn's adjoint terms, the rest of the Revolve loop,
    // then every window stacked into the caller's images.
    int phase_b(int turn, hipStream_t st) {
        if (exact) {
            RSG_HIP(hipMemcpyAsync(fwb, turn_state, total * sizeof(float), hipMemcpyDeviceToDevice, st));
            if (int rc = exact_step(turn, st)) return rc;
        } else if (int rc = reverse_tail(turn, st)) {
            return rc;
        }
        reverse = true;
        for (;;) {
            const int oldcapo = rev->capo;
            const Revolve::Action a = rev->next();
            const int capo = rev->capo;
            if (a == Revolve::ADVANCE) {
                for (int it = oldcapo; it < capo; it++)
                    if (int rc = fstep(it, !reverse, st)) return rc;
            } else if (a == Revolve::YOUTURN && exact) {
                // fw is left half stepped: Revolve restores before the next step
                if (int rc = exact_step(capo, st)) return rc;
            } else if (a == Revolve::YOUTURN) {
                if (int rc = ragged2d_step(&pb.p, 2, st)) return rc;
                if (int rc = ragged2d_step(&pb.p, 3, st)) return rc;
                if (int rc = reverse_tail(capo, st)) return rc;
            } else if (a == Revolve::TAKESHOT) {
                if (int rc = takeshot(st)) return rc;
            } else if (a == Revolve::RESTORE) {
                if (store == kStoreS16) {
                    if (int rc = unpack_s16(cks + sblk * kS16Block * rev->check, ckexp + sblk * rev->check, fwb,
                                            dsegs, dsblk, dsseg, sblk, st))
                        return rc;
                } else if (f16) {
                    if (int rc = unpack16_table(ck16 + cstride * rev->check, ckamax + (size_t)nseg * rev->check, fwb,
                                                dsegs, dchunks, nchunk, total, st))
                        return rc;
                } else {
                    RSG_HIP(hipMemcpyAsync(fwb, ck + total * rev->check, total * sizeof(float),
                                           hipMemcpyDeviceToDevice, st));
                }
            } else if (a == Revolve::TERMINATE) {
                break;
            } else {
                return fail(RSG_EINVAL, "fwi_e2d: Revolve returned an error");
            }
        }
        if (exact)  // each window's gradients into its images, already in the requested parameters
            for (int k = 0; k < ns; k++) {
                const int nx = sh[k].g.nx - 2 * lpml, nz = sh[k].g.nz - 2 * lpml;
                float *o = imgs + img_off[k];
                const size_t ni = (size_t)nx * nz;
                if (int rc = e2d_exact_model_grad(work[k], nx, nz, lpml, sh[k].g.fs, f->scale, sh[k].Vp, sh[k].Vs,
                                                  sh[k].Rho, pad3, o, o + ni, o + 2 * ni, st))
                    return rc;
            }
        e2d_stack_kernel<<<dim3((f->nx + 255) / 256, f->nz), 256, 0, st>>>(dsd, ns, f->nx, f->nz, f->grad[0],
                                                                             f->grad[1], f->grad[2],
                                                                             exact ? 0 : f->scale);
        RSG_LAUNCH_CHECK("e2d_stack_kernel");
        return RSG_OK;
    }

    int read_misfit(double *out, hipStream_t st) {
        double m = 0.0;
        RSG_HIP(hipMemcpyAsync(&m, dmis, sizeof(double), hipMemcpyDeviceToHost, st));
        RSG_HIP(hipStreamSynchronize(st));
        if (out) *out = m;
        return RSG_OK;
    }
};

static int fwi_e2d_impl(const rsg_fwi_e2d *f, const rsg_fwi_e2d_shot *sh, int ns, hipStream_t st) {
    E2DFwi r;
    int turn = 0;
    if (int rc = r.init(f, sh, ns, st)) return rc;
    if (int rc = r.phase_a(st, &turn)) return rc;
    if (int rc = r.misfit(st)) return rc;
    if (int rc = r.phase_b(turn, st)) return rc;
    return r.read_misfit(f->misfit, st);
}

// Capture of one phase as a graph: the stream must not be the legacy
// default stream (the plan makes its own).
template <class F>
static int capture(hipStream_t st, hipGraphExec_t *exec, F &&body) {
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(RSG_EHIP, "fwi_e2d plan: hipStreamBeginCapture failed");
    int rc = body();
    const hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc == RSG_OK && e != hipSuccess) rc = check_hip(e, "fwi_e2d plan: hipStreamEndCapture");
    if (rc == RSG_OK) {
        const hipError_t e2 = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
        if (e2 != hipSuccess) rc = check_hip(e2, "fwi_e2d plan: hipGraphInstantiate");
    }
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
}

}  // namespace rsg

extern "C" int rsg_fwi_e2d_gradient_batch(const rsg_fwi_e2d *f, const rsg_fwi_e2d_shot *shots, int nshot,
                                          void *stream) {
    if (!f || !shots || nshot < 1 || nshot > 65535) return rsg::fail(RSG_EINVAL, "rsg_fwi_e2d_gradient_batch: bad arguments");
    if (!f->tab || !f->grad[0] || !f->grad[1] || !f->grad[2] || f->nx < 1 || f->nz < 1)
        return rsg::fail(RSG_EINVAL, "rsg_fwi_e2d_gradient_batch: null tables or images");
    if (f->misfit_type < 0 || f->misfit_type > 4)
        return rsg::fail(RSG_EINVAL,
                         "rsg_fwi_e2d_gradient_batch: misfit type must be DIFFERENCE (0), CORRELATION (1), "
                         "ADAPTIVE_GAUSS (2), ADAPTIVE_LINEAR (3) or ADAPTIVE_SINGLE (4)");
    return rsg::fwi_e2d_impl(f, shots, nshot, (hipStream_t)stream);
}

struct rsg_fwi_e2d_plan {
    rsg_fwi_e2d f;                         // copy of the settings (pointers stay the caller's)
    std::vector<rsg_fwi_e2d_shot> shots;   // copy of the shot table
    rsg::E2DFwi run;
    hipStream_t st = nullptr;
    hipGraphExec_t a = nullptr, b = nullptr;
    int turn = 0;
    ~rsg_fwi_e2d_plan() {
        if (a) (void)hipGraphExecDestroy(a);
        if (b) (void)hipGraphExecDestroy(b);
        if (st) (void)hipStreamDestroy(st);
    }
};

extern "C" int rsg_fwi_e2d_plan_create(const rsg_fwi_e2d *f, const rsg_fwi_e2d_shot *shots, int nshot,
                                       rsg_fwi_e2d_plan **plan) {
    if (!plan) return rsg::fail(RSG_EINVAL, "rsg_fwi_e2d_plan_create: null plan");
    *plan = nullptr;
    if (!f || !shots || nshot < 1 || nshot > 65535) return rsg::fail(RSG_EINVAL, "rsg_fwi_e2d_plan_create: bad arguments");
    if (!f->tab || !f->grad[0] || !f->grad[1] || !f->grad[2] || f->nx < 1 || f->nz < 1)
        return rsg::fail(RSG_EINVAL, "rsg_fwi_e2d_plan_create: null tables or images");
    if (f->misfit_type < 0 || f->misfit_type > 4)
        return rsg::fail(RSG_EINVAL, "rsg_fwi_e2d_plan_create: misfit type must be 0..4");
    std::unique_ptr<rsg_fwi_e2d_plan> p(new rsg_fwi_e2d_plan);
    p->f = *f;
    p->f.misfit = nullptr;
    p->shots.assign(shots, shots + nshot);
    if (hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking) != hipSuccess)
        return rsg::fail(RSG_EHIP, "rsg_fwi_e2d_plan_create: stream creation failed");
    if (int rc = p->run.init(&p->f, p->shots.data(), nshot, p->st)) return rc;
    rsg_fwi_e2d_plan *q = p.get();
    if (int rc = rsg::capture(q->st, &q->a, [q]() { return q->run.phase_a(q->st, &q->turn); })) return rc;
    if (int rc = rsg::capture(q->st, &q->b, [q]() { return q->run.phase_b(q->turn, q->st); })) return rc;
    *plan = p.release();
    return RSG_OK;
}

extern "C" int rsg_fwi_e2d_plan_run(rsg_fwi_e2d_plan *plan, double *misfit, void *stream) {
    if (!plan || !plan->a || !plan->b) return rsg::fail(RSG_EINVAL, "rsg_fwi_e2d_plan_run: no plan");
    hipStream_t st = plan->st;
    // the caller's stream work (new models, zeroed images) comes first
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
        return rsg::fail(RSG_EHIP, "rsg_fwi_e2d_plan_run: event creation failed");
    hipError_t e = hipEventRecord(ev, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, ev, 0);
    (void)hipEventDestroy(ev);
    if (e != hipSuccess) return rsg::check_hip(e, "rsg_fwi_e2d_plan_run: stream ordering");
    RSG_HIP(hipGraphLaunch(plan->a, st));
    if (int rc = plan->run.misfit(st)) return rc;
    RSG_HIP(hipGraphLaunch(plan->b, st));
    return plan->run.read_misfit(misfit, st);  // synchronizes the plan's stream
}

extern "C" int rsg_fwi_e2d_plan_destroy(rsg_fwi_e2d_plan *plan) {
    if (!plan) return RSG_OK;
    (void)hipStreamSynchronize(plan->st);
    delete plan;
    return RSG_OK;
}
