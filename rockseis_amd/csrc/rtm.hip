// Acoustic 2D and 3D reverse-time migration of one shot
// (RtmAcoustic2D::run_optimal, rtm.cpp:423-544, and RtmAcoustic3D::run_optimal)
// on the device.
//
// The source wavefield is stepped forward on the Revolve schedule with its
// checkpoints (pressure, both velocities and the 8 PML strips) in HBM; the
// receiver wavefield is advanced with the forward operator from the recorded
// data injected at the receivers, and every reverse step adds the imaging
// condition with subsurface offsets (RtmAcoustic2D::crossCorr,
// rtm.cpp:191-219) into the image.
#include <vector>

#include "rsg_common.h"
#include "rsg_fwi.h"
#include "rsg_inject_ordered.h"
#include "rsg_srcrec.h"

// This file is compiled with -ffp-contract=off (Makefile): the product and
// the sum of the imaging condition are rounded separately, as in the
// reference's host arithmetic.

namespace rsg {

// image[ihz][ihx][iz][ix] += ws(ix-hx, iz-hz) * wr(ix+hx, iz+hz) with
// hx = ihx - (nhx-1)/2, hz = ihz - (nhz-1)/2, where both points lie in the
// unpadded grid (ki2D / ks2D / kr2D of rtm.h:27-29); ws and wr on the padded
// grid.  blockIdx.z = ihz*nhx + ihx.
__global__ __launch_bounds__(256) void rtm_image_a2d_kernel(const float *__restrict__ ws, const float *__restrict__ wr,
                                                            float *__restrict__ img, int nx, int nz, int lpml,
                                                            int nhx, int nhz) {
    const int ix = blockIdx.x * 64 + threadIdx.x, iz = blockIdx.y * 4 + threadIdx.y;
    if (ix >= nx || iz >= nz) return;
    const int ihx = (int)blockIdx.z % nhx, ihz = (int)blockIdx.z / nhx;
    const int hx = -(nhx - 1) / 2 + ihx, hz = -(nhz - 1) / 2 + ihz;
    if (ix - hx < 0 || ix - hx >= nx || ix + hx < 0 || ix + hx >= nx) return;
    if (iz - hz < 0 || iz - hz >= nz || iz + hz < 0 || iz + hz >= nz) return;
    const size_t nxp = (size_t)nx + 2 * lpml;
    const size_t m = (size_t)blockIdx.z * nx * nz + (size_t)iz * nx + ix;
    img[m] += ws[(size_t)(iz - hz + lpml) * nxp + (ix - hx + lpml)] * wr[(size_t)(iz + hz + lpml) * nxp + (ix + hx + lpml)];
}

// The same in 3D (RtmAcoustic3D::crossCorr, rtm.cpp ki3D / ks3D / kr3D of
// rtm.h:32-35): image[((ihz*nhy + ihy)*nhx + ihx)][iz][iy][ix]; one thread
// per cell, looping over the offsets.
__global__ __launch_bounds__(256) void rtm_image_a3d_kernel(const float *__restrict__ ws, const float *__restrict__ wr,
                                                            float *__restrict__ img, int nx, int ny, int nz,
                                                            int lpml, int nhx, int nhy, int nhz) {
    const int ix = blockIdx.x * 64 + threadIdx.x, iy = blockIdx.y * 4 + threadIdx.y, iz = blockIdx.z;
    if (ix >= nx || iy >= ny || iz >= nz) return;
    const size_t nxp = (size_t)nx + 2 * lpml, nxyp = nxp * ((size_t)ny + 2 * lpml);
    const size_t ncell = (size_t)nx * ny * nz, m = ((size_t)iz * ny + iy) * nx + ix;
    for (int ihz = 0; ihz < nhz; ihz++) {
        const int hz = -(nhz - 1) / 2 + ihz;
        if (iz - hz < 0 || iz - hz >= nz || iz + hz < 0 || iz + hz >= nz) continue;
        for (int ihy = 0; ihy < nhy; ihy++) {
            const int hy = -(nhy - 1) / 2 + ihy;
            if (iy - hy < 0 || iy - hy >= ny || iy + hy < 0 || iy + hy >= ny) continue;
            for (int ihx = 0; ihx < nhx; ihx++) {
                const int hx = -(nhx - 1) / 2 + ihx;
                if (ix - hx < 0 || ix - hx >= nx || ix + hx < 0 || ix + hx >= nx) continue;
                const size_t o = ((size_t)ihz * nhy + ihy) * nhx + ihx;
                const size_t ps = (size_t)(iz - hz + lpml) * nxyp + (size_t)(iy - hy + lpml) * nxp + (ix - hx + lpml);
                const size_t pr = (size_t)(iz + hz + lpml) * nxyp + (size_t)(iy + hy + lpml) * nxp + (ix + hx + lpml);
                img[o * ncell + m] += ws[ps] * wr[pr];
            }
        }
    }
}

// A batch of shots (blockIdx.z = offset): image += ws_k * wr_k for k in shot
// order, shot k's fields at k*fstride; the stacked image of the batch.
__global__ __launch_bounds__(256) void rtm_image_a2d_batch_kernel(const float *__restrict__ ws,
                                                                  const float *__restrict__ wr, float *__restrict__ img,
                                                                  int nx, int nz, int lpml, int nhx, int nhz, int ns,
                                                                  size_t fstride) {
    const int ix = blockIdx.x * 64 + threadIdx.x, iz = blockIdx.y * 4 + threadIdx.y;
    if (ix >= nx || iz >= nz) return;
    const int ihx = (int)blockIdx.z % nhx, ihz = (int)blockIdx.z / nhx;
    const int hx = -(nhx - 1) / 2 + ihx, hz = -(nhz - 1) / 2 + ihz;
    if (ix - hx < 0 || ix - hx >= nx || ix + hx < 0 || ix + hx >= nx) return;
    if (iz - hz < 0 || iz - hz >= nz || iz + hz < 0 || iz + hz >= nz) return;
    const size_t nxp = (size_t)nx + 2 * lpml;
    const size_t m = (size_t)blockIdx.z * nx * nz + (size_t)iz * nx + ix;
    const size_t ps = (size_t)(iz - hz + lpml) * nxp + (ix - hx + lpml), pr = (size_t)(iz + hz + lpml) * nxp + (ix + hx + lpml);
    float a = img[m];
    for (int k = 0; k < ns; k++) a += ws[k * fstride + ps] * wr[k * fstride + pr];
    img[m] = a;
}

namespace {

struct RtmMem {
    std::vector<void *> ptrs;
    ~RtmMem() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **out, size_t n) {
        void *p = nullptr;
        if (hipMalloc(&p, (n ? n : 1) * sizeof(T)) != hipSuccess)
            return fail(RSG_ENOMEM, "rtm: device allocation failed");
        ptrs.push_back(p);
        *out = (T *)p;
        return RSG_OK;
    }
};

// One WavesAcoustic2D state in one allocation: P, Vx, Vz, then the x strips
// (lpml*nz) and the z strips (nx*lpml).
int a2d_state(RtmMem &dm, const rsg_grid *g, const rsg_rtm_a2d *r, rsg_state *s, float **base, size_t *total,
              hipStream_t st) {
    const size_t n = (size_t)g->nx * g->nz, sx = (size_t)g->lpml * g->nz, sz = (size_t)g->nx * g->lpml;
    *total = 3 * n + 4 * sx + 4 * sz;
    if (int rc = dm.alloc(base, *total)) return rc;
    if (int rc = launch_zero(*base, *total * sizeof(float), st)) return rc;
    *s = rsg_state{};
    s->P = *base;
    s->Vx = *base + n;
    s->Vz = *base + 2 * n;
    float *p = *base + 3 * n;
    const int xs[4] = {RSG_A2D_P_LEFT, RSG_A2D_VXX_LEFT, RSG_A2D_P_RIGHT, RSG_A2D_VXX_RIGHT};
    const int zs[4] = {RSG_A2D_P_TOP, RSG_A2D_VZZ_TOP, RSG_A2D_P_BOTTOM, RSG_A2D_VZZ_BOTTOM};
    for (int i = 0; i < 4; i++, p += sx) s->psi[xs[i]] = p;
    for (int i = 0; i < 4; i++, p += sz) s->psi[zs[i]] = p;
    s->L = r->L;
    s->Rx = r->Rx;
    s->Rz = r->Rz;
    s->tab = r->tab;
    return RSG_OK;
}

// One WavesAcoustic3D state in one allocation: P, Vx, Vy, Vz, then the 12
// strips in slot order (left/right nz*ny*lpml, front/back nz*lpml*nx,
// top/bottom lpml*ny*nx).
int a3d_state(RtmMem &dm, const rsg_grid *g, const rsg_rtm_a3d *r, rsg_state *s, float **base, size_t *total,
              hipStream_t st) {
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    const size_t face[3] = {(size_t)g->nz * g->ny * g->lpml, (size_t)g->nz * g->lpml * g->nx,
                            (size_t)g->lpml * g->ny * g->nx};
    size_t off[RSG_A3D_NSTRIP], o = 4 * n;
    for (int i = 0; i < RSG_A3D_NSTRIP; i++) {
        int f = 2;
        if (i == RSG_A3D_P_LEFT || i == RSG_A3D_P_RIGHT || i == RSG_A3D_VXX_LEFT || i == RSG_A3D_VXX_RIGHT) f = 0;
        if (i == RSG_A3D_P_FRONT || i == RSG_A3D_P_BACK || i == RSG_A3D_VYY_FRONT || i == RSG_A3D_VYY_BACK) f = 1;
        off[i] = o;
        o += face[f];
    }
    *total = o;
    if (int rc = dm.alloc(base, *total)) return rc;
    if (int rc = launch_zero(*base, *total * sizeof(float), st)) return rc;
    *s = rsg_state{};
    s->P = *base;
    s->Vx = *base + n;
    s->Vy = *base + 2 * n;
    s->Vz = *base + 3 * n;
    for (int i = 0; i < RSG_A3D_NSTRIP; i++) s->psi[i] = *base + off[i];
    s->L = r->L;
    s->Rx = r->Rx;
    s->Ry = r->Ry;
    s->Rz = r->Rz;
    s->tab = r->tab;
    return RSG_OK;
}

// The WavesAcoustic2D states of ns shots in one allocation: each array holds
// all shots back to back (P, Vx, Vz, then the 8 strips).
int a2d_batch_state(RtmMem &dm, const rsg_grid *g, const rsg_rtm_a2d *r, int ns, rsg_state *s, rsg_batch *b,
                    float **base, size_t *total, hipStream_t st) {
    const size_t n = (size_t)g->nx * g->nz, sx = (size_t)g->lpml * g->nz, sz = (size_t)g->nx * g->lpml;
    *total = (size_t)ns * (3 * n + 4 * sx + 4 * sz);
    if (int rc = dm.alloc(base, *total)) return rc;
    if (int rc = launch_zero(*base, *total * sizeof(float), st)) return rc;
    *s = rsg_state{};
    s->P = *base;
    s->Vx = *base + ns * n;
    s->Vz = *base + 2 * ns * n;
    float *p = *base + 3 * ns * n;
    const int xs[4] = {RSG_A2D_P_LEFT, RSG_A2D_VXX_LEFT, RSG_A2D_P_RIGHT, RSG_A2D_VXX_RIGHT};
    const int zs[4] = {RSG_A2D_P_TOP, RSG_A2D_VZZ_TOP, RSG_A2D_P_BOTTOM, RSG_A2D_VZZ_BOTTOM};
    for (int i = 0; i < 4; i++, p += ns * sx) s->psi[xs[i]] = p;
    for (int i = 0; i < 4; i++, p += ns * sz) s->psi[zs[i]] = p;
    s->L = r->L;
    s->Rx = r->Rx;
    s->Rz = r->Rz;
    s->tab = r->tab;
    *b = rsg_batch{};
    b->nshot = ns;
    b->field_stride = n;
    b->xstrip_stride = sx;
    b->zstrip_stride = sz;
    return RSG_OK;
}

}  // namespace

static int rtm_a2d_batch_impl(const rsg_grid *g, const rsg_rtm_a2d *r, int ns, hipStream_t st) {
    const int nt = r->src.nt;
    if (ns < 1 || ns > 65535) return fail(RSG_EINVAL, "rtm_a2d_batch: 1 <= nshot <= 65535");
    if (r->data.nt != nt || r->data.field != RSG_PRESSURE || r->src.field != RSG_PRESSURE)
        return fail(RSG_EINVAL, "rtm_a2d_batch: source and data must be PRESSURE traces with equal nt");
    if (r->nx + 2 * g->lpml != g->nx || r->nz + 2 * g->lpml != g->nz || g->ny != 1)
        return fail(RSG_EINVAL, "rtm_a2d_batch: image size must be the unpadded 2D grid");
    if (r->nhx < 1 || r->nhz < 1 || (long)r->nhx * r->nhz > 65535 || !r->image)
        return fail(RSG_EINVAL, "rtm_a2d_batch: need an image and 1 <= nhx*nhz <= 65535 subsurface offsets");
    if (r->ncheck < 1 || nt < 1) return fail(RSG_EINVAL, "rtm_a2d_batch: need nt >= 1 and ncheck >= 1");
    Revolve rev(nt, r->ncheck);
    {
        const size_t one = (size_t)ns * (3 * (size_t)g->nx * g->nz + 4 * (size_t)g->lpml * (g->nx + g->nz));
        size_t free_b = 0, total_b = 0;
        RSG_HIP(hipMemGetInfo(&free_b, &total_b));
        if (one * sizeof(float) * (2 + (size_t)rev.snaps) + ((size_t)1 << 30) > free_b)
            return fail(RSG_ENOMEM, "rtm_a2d_batch: wavefields and checkpoints do not fit in device memory");
    }
    RtmMem dm;
    rsg_state sf{}, sb{};
    rsg_batch bf{}, bb{};
    float *fwb = nullptr, *bwb = nullptr, *ck = nullptr;
    size_t total = 0;
    if (int rc = a2d_batch_state(dm, g, r, ns, &sf, &bf, &fwb, &total, st)) return rc;
    if (int rc = a2d_batch_state(dm, g, r, ns, &sb, &bb, &bwb, &total, st)) return rc;
    if (int rc = dm.alloc(&ck, total * (size_t)rev.snaps)) return rc;
    OrderedInject2D oi;
    if (r->data.ntrace > 0)
        if (int rc = oi.build(dm, r->data.map, r->data.ntrace, ns, g->nx, g->nz, map_pad(g), st)) return rc;
    const size_t nimg = (size_t)r->nhx * r->nhz * r->nx * r->nz, fs = (size_t)g->nx * g->nz;
    if (int rc = launch_zero(r->image, nimg * sizeof(float), st)) return rc;
    const dim3 xb(64, 4), xg((r->nx + 63) / 64, (r->nz + 3) / 4, r->nhx * r->nhz);

    auto fstep = [&](int it) -> int {
        if (int rc = rsg_a2d_forward_velocity_batch(g, &sf, &bf, st)) return rc;
        if (int rc = rsg_a2d_forward_stress_batch(g, &sf, &bf, st)) return rc;
        return insert_source(RSG_KIND_A2D, g, &sf, &r->src, nullptr, it, st, ns, fs);
    };
    auto image_step = [&](int capo) -> int {
        if (oi.maxcells > 0) {
            inject2d_ordered_kernel<<<dim3((oi.maxcells + 63) / 64, ns), 64, 0, st>>>(
                sb.P, r->L, g->dt, oi.cells, oi.starts, oi.ents, oi.coff, r->data.shift, r->data.data,
                r->data.ntrace, nt, capo, fs);
            RSG_LAUNCH_CHECK("inject2d_ordered_kernel");
        }
        rtm_image_a2d_batch_kernel<<<xg, xb, 0, st>>>(sf.P, sb.P, r->image, r->nx, r->nz, g->lpml, r->nhx, r->nhz,
                                                      ns, fs);
        RSG_LAUNCH_CHECK("rtm_image_a2d_batch_kernel");
        return RSG_OK;
    };
    for (;;) {
        const int oldcapo = rev.capo;
        const Revolve::Action a = rev.next();
        const int capo = rev.capo;
        if (a == Revolve::ADVANCE) {
            for (int it = oldcapo; it < capo; it++)
                if (int rc = fstep(it)) return rc;
        } else if (a == Revolve::FIRSTURN) {
            if (int rc = fstep(capo)) return rc;
            if (int rc = image_step(capo)) return rc;
        } else if (a == Revolve::YOUTURN) {
            if (int rc = rsg_a2d_forward_velocity_batch(g, &sb, &bb, st)) return rc;
            if (int rc = rsg_a2d_forward_stress_batch(g, &sb, &bb, st)) return rc;
            if (int rc = image_step(capo)) return rc;
        } else if (a == Revolve::TAKESHOT) {
            RSG_HIP(hipMemcpyAsync(ck + total * rev.check, fwb, total * sizeof(float), hipMemcpyDeviceToDevice, st));
        } else if (a == Revolve::RESTORE) {
            RSG_HIP(hipMemcpyAsync(fwb, ck + total * rev.check, total * sizeof(float), hipMemcpyDeviceToDevice, st));
        } else if (a == Revolve::TERMINATE) {
            break;
        } else {
            return fail(RSG_EINVAL, "rtm_a2d_batch: Revolve returned an error");
        }
    }
    RSG_HIP(hipStreamSynchronize(st));
    return RSG_OK;
}

static int rtm_a2d_impl(const rsg_grid *g, const rsg_rtm_a2d *r, hipStream_t st) {
    const int nt = r->src.nt;
    if (r->data.nt != nt || r->data.field != RSG_PRESSURE || r->src.field != RSG_PRESSURE)
        return fail(RSG_EINVAL, "rtm_a2d: source and data must be PRESSURE traces with equal nt");
    if (r->nx + 2 * g->lpml != g->nx || r->nz + 2 * g->lpml != g->nz || g->ny != 1)
        return fail(RSG_EINVAL, "rtm_a2d: image size must be the unpadded 2D grid");
    if (r->nhx < 1 || r->nhz < 1 || (long)r->nhx * r->nhz > 65535 || !r->image)
        return fail(RSG_EINVAL, "rtm_a2d: need an image and 1 <= nhx*nhz <= 65535 subsurface offsets");
    if (r->ncheck < 1 || nt < 1) return fail(RSG_EINVAL, "rtm_a2d: need nt >= 1 and ncheck >= 1");
    Revolve rev(nt, r->ncheck);
    RtmMem dm;
    rsg_state sf{}, sb{};
    float *fwb = nullptr, *bwb = nullptr, *ck = nullptr;
    size_t total = 0;
    if (int rc = a2d_state(dm, g, r, &sf, &fwb, &total, st)) return rc;
    if (int rc = a2d_state(dm, g, r, &sb, &bwb, &total, st)) return rc;
    {
        size_t free_b = 0, total_b = 0;
        RSG_HIP(hipMemGetInfo(&free_b, &total_b));
        if (total * 4 * (size_t)rev.snaps + ((size_t)1 << 30) > free_b)
            return fail(RSG_ENOMEM, "rtm_a2d: checkpoints do not fit in device memory");
    }
    if (int rc = dm.alloc(&ck, total * (size_t)rev.snaps)) return rc;
    OrderedInject2D oi;
    if (r->data.ntrace > 0)
        if (int rc = oi.build(dm, r->data.map, r->data.ntrace, 1, g->nx, g->nz, map_pad(g), st)) return rc;
    // RtmAcoustic2D::run_optimal allocates (zeroes) the image before the run
    const size_t nimg = (size_t)r->nhx * r->nhz * r->nx * r->nz;
    if (int rc = launch_zero(r->image, nimg * sizeof(float), st)) return rc;
    const dim3 xb(64, 4), xg((r->nx + 63) / 64, (r->nz + 3) / 4, r->nhx * r->nhz);

    auto fstep = [&](int it) -> int {
        if (int rc = rsg_a2d_forward_velocity(g, &sf, st)) return rc;
        if (int rc = rsg_a2d_forward_stress(g, &sf, st)) return rc;
        return insert_source(RSG_KIND_A2D, g, &sf, &r->src, nullptr, it, st);
    };
    auto image_step = [&](int capo) -> int {
        if (oi.maxcells > 0) {
            inject2d_ordered_kernel<<<dim3((oi.maxcells + 63) / 64, 1), 64, 0, st>>>(
                sb.P, r->L, g->dt, oi.cells, oi.starts, oi.ents, oi.coff, r->data.shift, r->data.data,
                r->data.ntrace, nt, capo, 0);
            RSG_LAUNCH_CHECK("inject2d_ordered_kernel");
        }
        rtm_image_a2d_kernel<<<xg, xb, 0, st>>>(sf.P, sb.P, r->image, r->nx, r->nz, g->lpml, r->nhx, r->nhz);
        RSG_LAUNCH_CHECK("rtm_image_a2d_kernel");
        return RSG_OK;
    };
    for (;;) {
        const int oldcapo = rev.capo;
        const Revolve::Action a = rev.next();
        const int capo = rev.capo;
        if (a == Revolve::ADVANCE) {
            for (int it = oldcapo; it < capo; it++)
                if (int rc = fstep(it)) return rc;
        } else if (a == Revolve::FIRSTURN) {
            if (int rc = fstep(capo)) return rc;
            if (int rc = image_step(capo)) return rc;
        } else if (a == Revolve::YOUTURN) {
            if (int rc = rsg_a2d_forward_velocity(g, &sb, st)) return rc;
            if (int rc = rsg_a2d_forward_stress(g, &sb, st)) return rc;
            if (int rc = image_step(capo)) return rc;
        } else if (a == Revolve::TAKESHOT) {
            RSG_HIP(hipMemcpyAsync(ck + total * rev.check, fwb, total * sizeof(float), hipMemcpyDeviceToDevice, st));
        } else if (a == Revolve::RESTORE) {
            RSG_HIP(hipMemcpyAsync(fwb, ck + total * rev.check, total * sizeof(float), hipMemcpyDeviceToDevice, st));
        } else if (a == Revolve::TERMINATE) {
            break;
        } else {
            return fail(RSG_EINVAL, "rtm_a2d: Revolve returned an error");
        }
    }
    // the checkpoints and states are freed on return: finish the queued work first
    RSG_HIP(hipStreamSynchronize(st));
    return RSG_OK;
}

static int rtm_a3d_impl(const rsg_grid *g, const rsg_rtm_a3d *r, hipStream_t st) {
    const int nt = r->src.nt;
    if (r->data.nt != nt || r->data.field != RSG_PRESSURE || r->src.field != RSG_PRESSURE)
        return fail(RSG_EINVAL, "rtm_a3d: source and data must be PRESSURE traces with equal nt");
    if (r->nx + 2 * g->lpml != g->nx || r->ny + 2 * g->lpml != g->ny || r->nz + 2 * g->lpml != g->nz)
        return fail(RSG_EINVAL, "rtm_a3d: image size must be the unpadded grid");
    if (r->nhx < 1 || r->nhy < 1 || r->nhz < 1 || !r->image || r->nz > 65535)
        return fail(RSG_EINVAL, "rtm_a3d: need an image, nhx, nhy, nhz >= 1 and nz <= 65535");
    if (r->ncheck < 1 || nt < 1) return fail(RSG_EINVAL, "rtm_a3d: need nt >= 1 and ncheck >= 1");
    Revolve rev(nt, r->ncheck);
    RtmMem dm;
    rsg_state sf{}, sb{};
    float *fwb = nullptr, *bwb = nullptr, *ck = nullptr;
    size_t total = 0;
    {
        // both wavefields and the checkpoints (4 fields and 4 strips of each
        // face orientation per state), before allocating any of them
        const size_t n = (size_t)g->nx * g->ny * g->nz;
        const size_t state = 4 * n + 4 * ((size_t)g->nz * g->ny * g->lpml + (size_t)g->nz * g->lpml * g->nx +
                                          (size_t)g->lpml * g->ny * g->nx);
        size_t free_b = 0, total_b = 0;
        RSG_HIP(hipMemGetInfo(&free_b, &total_b));
        if (state * sizeof(float) * (2 + (size_t)rev.snaps) + ((size_t)1 << 30) > free_b)
            return fail(RSG_ENOMEM, "rtm_a3d: wavefields and checkpoints do not fit in device memory");
    }
    if (int rc = a3d_state(dm, g, r, &sf, &fwb, &total, st)) return rc;
    if (int rc = a3d_state(dm, g, r, &sb, &bwb, &total, st)) return rc;
    if (int rc = dm.alloc(&ck, total * (size_t)rev.snaps)) return rc;
    OrderedInject3D oi;
    if (r->data.ntrace > 0)
        if (int rc = oi.build(dm, r->data.map, r->data.ntrace, g->nx, g->ny, g->nz, map_pad(g), st)) return rc;
    const size_t nimg = (size_t)r->nhx * r->nhy * r->nhz * r->nx * r->ny * r->nz;
    if (int rc = launch_zero(r->image, nimg * sizeof(float), st)) return rc;
    const dim3 xb(64, 4), xg((r->nx + 63) / 64, (r->ny + 3) / 4, r->nz);

    auto fstep = [&](int it) -> int {
        if (int rc = rsg_a3d_forward_velocity(g, &sf, st)) return rc;
        if (int rc = rsg_a3d_forward_stress(g, &sf, st)) return rc;
        return insert_source(RSG_KIND_A3D, g, &sf, &r->src, nullptr, it, st);
    };
    auto image_step = [&](int capo) -> int {
        if (oi.ncells > 0) {
            inject3d_ordered_kernel<<<(oi.ncells + 63) / 64, 64, 0, st>>>(sb.P, r->L, g->dt, oi.cells, oi.starts,
                                                                          oi.ents, oi.ncells, r->data.shift,
                                                                          r->data.data, nt, capo);
            RSG_LAUNCH_CHECK("inject3d_ordered_kernel");
        }
        rtm_image_a3d_kernel<<<xg, xb, 0, st>>>(sf.P, sb.P, r->image, r->nx, r->ny, r->nz, g->lpml, r->nhx, r->nhy,
                                                r->nhz);
        RSG_LAUNCH_CHECK("rtm_image_a3d_kernel");
        return RSG_OK;
    };
    for (;;) {
        const int oldcapo = rev.capo;
        const Revolve::Action a = rev.next();
        const int capo = rev.capo;
        if (a == Revolve::ADVANCE) {
            for (int it = oldcapo; it < capo; it++)
                if (int rc = fstep(it)) return rc;
        } else if (a == Revolve::FIRSTURN) {
            if (int rc = fstep(capo)) return rc;
            if (int rc = image_step(capo)) return rc;
        } else if (a == Revolve::YOUTURN) {
            if (int rc = rsg_a3d_forward_velocity(g, &sb, st)) return rc;
            if (int rc = rsg_a3d_forward_stress(g, &sb, st)) return rc;
            if (int rc = image_step(capo)) return rc;
        } else if (a == Revolve::TAKESHOT) {
            RSG_HIP(hipMemcpyAsync(ck + total * rev.check, fwb, total * sizeof(float), hipMemcpyDeviceToDevice, st));
        } else if (a == Revolve::RESTORE) {
            RSG_HIP(hipMemcpyAsync(fwb, ck + total * rev.check, total * sizeof(float), hipMemcpyDeviceToDevice, st));
        } else if (a == Revolve::TERMINATE) {
            break;
        } else {
            return fail(RSG_EINVAL, "rtm_a3d: Revolve returned an error");
        }
    }
    RSG_HIP(hipStreamSynchronize(st));
    return RSG_OK;
}

}  // namespace rsg

extern "C" int rsg_rtm_a3d_image(const rsg_grid *g, const rsg_rtm_a3d *r, void *stream) {
    if (!g || !r) return rsg::fail(RSG_EINVAL, "rsg_rtm_a3d_image: null argument");
    return rsg::rtm_a3d_impl(g, r, (hipStream_t)stream);
}

extern "C" int rsg_rtm_a2d_batch(const rsg_grid *g, const rsg_rtm_a2d *r, int nshot, void *stream) {
    if (!g || !r) return rsg::fail(RSG_EINVAL, "rsg_rtm_a2d_batch: null argument");
    return rsg::rtm_a2d_batch_impl(g, r, nshot, (hipStream_t)stream);
}

extern "C" int rsg_rtm_a2d_image(const rsg_grid *g, const rsg_rtm_a2d *r, void *stream) {
    if (!g || !r) return rsg::fail(RSG_EINVAL, "rsg_rtm_a2d_image: null argument");
    return rsg::rtm_a2d_impl(g, r, (hipStream_t)stream);
}
