// Acoustic 3D FWI gradient of one shot (FwiAcoustic3D::run_optimal,
// fwi.cpp:1991-2171) on the device.
//
// The forward wavefield is stepped with the 3D register-queue kernels; the
// Revolve schedule runs on the host and its checkpoints (pressure, the three
// velocities and the 12 PML strips) live in HBM, FP32 or FP16 with one
// power-of-two scale per array.  As in the reference, the adjoint wavefield is
// advanced with the forward operator from the DIFFERENCE residuals injected
// at the receivers; every reverse step accumulates crossCorr
// (fwi.cpp:1403-1466) into the Vp and rho gradients and records the wavelet
// gradient at the source from the adjoint pressure.
#include <algorithm>
#include <climits>
#include <utility>
#include <vector>

#include "rsg_adjoint.h"
#include "rsg_common.h"
#include "rsg_fwi.h"
#include "rsg_inject_ordered.h"
#include "rsg_srcrec.h"
#include "rsg_store16.h"

// This file is compiled with -ffp-contract=off (Makefile): every multiply and
// add is rounded separately, as in the reference's host arithmetic.

namespace rsg {

// crossCorr of one reverse step with the reference's 3D bounds ix, iy, iz in
// [1, n) (upper edge included) and its index of the second z term (wrz at
// iy-1, Rz at iz-1, fwi.cpp:1459).  Forward pressure and adjoint fields on
// the padded grid (pad lpml); Vp, Rho and the images unpadded.
__global__ __launch_bounds__(256) void crosscorr_a3d_kernel(
    const float *__restrict__ wsp, const float *__restrict__ wrx, const float *__restrict__ wry,
    const float *__restrict__ wrz, const float *__restrict__ Rx, const float *__restrict__ Ry,
    const float *__restrict__ Rz, const float *__restrict__ Vp, const float *__restrict__ Rho, int nx, int ny,
    int nz, int lpml, float dx, float dy, float dz, float *__restrict__ vpgrad, float *__restrict__ rhograd) {
    const int ix = blockIdx.x * 64 + threadIdx.x + 1;
    const int iy = blockIdx.y * 4 + threadIdx.y + 1;
    const int iz = blockIdx.z + 1;
    if (ix >= nx || iy >= ny || iz >= nz) return;
    const long nxp = nx + 2 * lpml, nxyp = nxp * (ny + 2 * lpml);
    const size_t q = (size_t)(iz + lpml) * nxyp + (size_t)(iy + lpml) * nxp + ix + lpml;
    const size_t m = ((size_t)iz * ny + iy) * nx + ix;
    const float vpscale = (float)(-2.0 / (double)Vp[m]);
    const float rhoscale1 = (float)(-1.0 / (double)Rho[m]);
    const float mrxx = (wrx[q] - wrx[q - 1]) / dx;
    const float mryy = (wry[q] - wry[q - nxp]) / dy;
    const float mrzz = (wrz[q] - wrz[q - nxyp]) / dz;
    const float ws = wsp[q];
    vpgrad[m] -= vpscale * ws * (mrxx + mryy + mrzz);
    float rg = rhograd[m] - rhoscale1 * ws * (mrxx + mryy + mrzz);
    float ux = (float)(0.5 * (double)wrx[q] * (double)Rx[q] * (double)(wsp[q + 1] - ws) / (double)dx);
    ux = (float)((double)ux + 0.5 * (double)wrx[q - 1] * (double)Rx[q - 1] * (double)(ws - wsp[q - 1]) / (double)dx);
    float uy = (float)(0.5 * (double)wry[q] * (double)Ry[q] * (double)(wsp[q + nxp] - ws) / (double)dy);
    uy = (float)((double)uy + 0.5 * (double)wry[q - nxp] * (double)Ry[q - nxp] * (double)(ws - wsp[q - nxp]) /
                                  (double)dy);
    float uz = (float)(0.5 * (double)wrz[q] * (double)Rz[q] * (double)(wsp[q + nxyp] - ws) / (double)dz);
    uz = (float)((double)uz + 0.5 * (double)wrz[q - nxp] * (double)Rz[q - nxyp] * (double)(ws - wsp[q - nxyp]) /
                                  (double)dz);
    rhograd[m] = rg + (ux + uy + uz);
}

namespace {

struct Mem3 {
    std::vector<void *> ptrs;
    ~Mem3() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **out, size_t n) {
        void *p = nullptr;
        if (n == 0) n = 1;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return fail(RSG_ENOMEM, "fwi_a3d: device allocation failed");
        ptrs.push_back(p);
        *out = (T *)p;
        return RSG_OK;
    }
};

// Strip slot i of the acoustic 3D state: 0 left/right (nz*ny*lpml),
// 1 front/back (nz*lpml*nx), 2 top/bottom (lpml*ny*nx).
inline int a3d_face(int i) {
    switch (i) {
    case RSG_A3D_P_LEFT: case RSG_A3D_P_RIGHT: case RSG_A3D_VXX_LEFT: case RSG_A3D_VXX_RIGHT: return 0;
    case RSG_A3D_P_FRONT: case RSG_A3D_P_BACK: case RSG_A3D_VYY_FRONT: case RSG_A3D_VYY_BACK: return 1;
    default: return 2;
    }
}

}  // namespace

static int fwi_a3d_impl(const rsg_grid *g, const rsg_fwi_a3d *f, hipStream_t st) {
    const int nt = f->src.nt, lpml = g->lpml;
    if (f->obs.nt != nt || f->obs.field != RSG_PRESSURE || f->src.field != RSG_PRESSURE)
        return fail(RSG_EINVAL, "fwi_a3d: source and data must be PRESSURE traces with equal nt");
    if (f->nx + 2 * lpml != g->nx || f->ny + 2 * lpml != g->ny || f->nz + 2 * lpml != g->nz)
        return fail(RSG_EINVAL, "fwi_a3d: image size must be the unpadded grid");
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    const size_t face[3] = {(size_t)g->nz * g->ny * lpml, (size_t)g->nz * lpml * g->nx, (size_t)lpml * g->ny * g->nx};
    std::vector<size_t> segs;
    size_t o = 0;
    for (int j = 0; j < 4; j++, o += n) segs.push_back(o);
    for (int i = 0; i < RSG_A3D_NSTRIP; i++) {
        segs.push_back(o);
        o += face[a3d_face(i)];
    }
    segs.push_back(o);
    const size_t total = o;
    const int nseg = (int)segs.size() - 1;
    const int store = f->fp16;
    if (store != kStoreF32 && store != kStoreF16 && store != kStoreS16)
        return fail(RSG_EINVAL, "fwi_a3d: checkpoint storage must be 0 (FP32), 1 (FP16) or 2 (S16)");
    const bool f16 = store != kStoreF32;  // 2 bytes a value
    const bool exact = f->adjoint == RSG_ADJOINT_EXACT;
    if (f->adjoint != RSG_ADJOINT_REFERENCE && !exact)
        return fail(RSG_EINVAL, "fwi_a3d: adjoint must be RSG_ADJOINT_REFERENCE or RSG_ADJOINT_EXACT");
    if (exact && f->misfit_type != 0 && f->misfit_type != 4)
        return fail(RSG_EINVAL, "fwi_a3d: the exact adjoint takes the DIFFERENCE misfit");
    if (exact && g->domdec)
        return fail(RSG_EINVAL, "fwi_a3d: the exact adjoint takes models padded by the library (no domain model)");
    Revolve rev(nt, f->ncheck);
    {
        size_t free_b = 0, total_b = 0;
        RSG_HIP(hipMemGetInfo(&free_b, &total_b));
        const size_t need = 2 * total * 4 + total * (f16 ? 2 : 4) * (size_t)rev.snaps +
                            (exact ? n * (7 * 4 + 2 * 8) + total * 4 : 0);
        if (need + ((size_t)1 << 30) > free_b)
            return fail(RSG_ENOMEM, "fwi_a3d: wavefields and checkpoints do not fit in device memory");
    }
    Mem3 dm;
    float *fwb = nullptr, *bwb = nullptr;
    if (int rc = dm.alloc(&fwb, total)) return rc;
    if (int rc = dm.alloc(&bwb, total)) return rc;
    if (int rc = launch_zero(fwb, total * sizeof(float), st)) return rc;
    if (int rc = launch_zero(bwb, total * sizeof(float), st)) return rc;
    rsg_state sf{}, sb{};
    for (int w = 0; w < 2; w++) {
        float *b = w ? bwb : fwb;
        rsg_state &s = w ? sb : sf;
        s.P = b;
        s.Vx = b + n;
        s.Vy = b + 2 * n;
        s.Vz = b + 3 * n;
        for (int i = 0; i < RSG_A3D_NSTRIP; i++) s.psi[i] = b + segs[4 + i];
        s.L = f->L;
        s.Rx = f->Rx;
        s.Ry = f->Ry;
        s.Rz = f->Rz;
        s.tab = f->tab;
    }
    double *dmis = nullptr;
    if (int rc = dm.alloc(&dmis, 1)) return rc;
    if (int rc = launch_zero(dmis, sizeof(double), st)) return rc;
    if (int rc = launch_zero(f->dmod, sizeof(float) * f->obs.ntrace * nt, st)) return rc;
    if (f->wavgrad)
        if (int rc = launch_zero(f->wavgrad, sizeof(float) * f->src.ntrace * nt, st)) return rc;
    float *ck = nullptr;
    half_t *ck16 = nullptr;
    unsigned *ckamax = nullptr;
    size_t *dsegs = nullptr;
    Chunk16 *dchunks = nullptr;
    unsigned *ckpart = nullptr;  // per-chunk maxima of the FP16 packing
    // FP16 checkpoint slots start at multiples of 4 values (8-byte aligned)
    const size_t cstride = (total + 3) & ~(size_t)3;
    const std::vector<Chunk16> chunks = store == kStoreF16 ? store16_chunks(segs) : std::vector<Chunk16>();
    const int nchunk = (int)chunks.size();
    // S16: blocks restart at every field and PML face (segs); a checkpoint is
    // nblk blocks of kS16Block values
    short *cks = nullptr, *ckexp = nullptr;
    size_t *dsblk = nullptr;
    int *dsseg = nullptr;
    const std::vector<size_t> sblkoff = store == kStoreS16 ? s16_block_offsets(segs) : std::vector<size_t>();
    const size_t nblk = store == kStoreS16 ? sblkoff.back() : 0;
    if (store == kStoreS16) {
        if (int rc = dm.alloc(&cks, nblk * kS16Block * (size_t)rev.snaps)) return rc;
        if (int rc = dm.alloc(&ckexp, nblk * rev.snaps)) return rc;
        if (int rc = dm.alloc(&dsegs, segs.size())) return rc;
        if (int rc = dm.alloc(&dsblk, sblkoff.size())) return rc;
        RSG_HIP(hipMemcpyAsync(dsegs, segs.data(), sizeof(size_t) * segs.size(), hipMemcpyHostToDevice, st));
        RSG_HIP(hipMemcpyAsync(dsblk, sblkoff.data(), sizeof(size_t) * sblkoff.size(), hipMemcpyHostToDevice, st));
        const std::vector<int> sseg = s16_block_segments(sblkoff);
        if (int rc = dm.alloc(&dsseg, sseg.size())) return rc;
        RSG_HIP(hipMemcpyAsync(dsseg, sseg.data(), sizeof(int) * sseg.size(), hipMemcpyHostToDevice, st));
        RSG_HIP(hipStreamSynchronize(st));  // sseg goes out of scope
    } else if (f16) {
        if (int rc = dm.alloc(&ck16, cstride * (size_t)rev.snaps)) return rc;
        if (int rc = dm.alloc(&ckamax, (size_t)nseg * rev.snaps)) return rc;
        if (int rc = dm.alloc(&dsegs, segs.size())) return rc;
        if (int rc = dm.alloc(&dchunks, chunks.size())) return rc;
        if (int rc = dm.alloc(&ckpart, chunks.size())) return rc;
        RSG_HIP(hipMemcpyAsync(dsegs, segs.data(), sizeof(size_t) * segs.size(), hipMemcpyHostToDevice, st));
        RSG_HIP(hipMemcpyAsync(dchunks, chunks.data(), sizeof(Chunk16) * chunks.size(), hipMemcpyHostToDevice, st));
    } else {
        if (int rc = dm.alloc(&ck, total * (size_t)rev.snaps)) return rc;
    }
    rsg_traces mod = f->obs, res = f->obs, wg = f->src;
    mod.data = f->dmod;
    res.data = f->dres;
    wg.data = f->wavgrad;
    OrderedInject3D oi;
    if (f->obs.ntrace > 0)
        if (int rc = oi.build(dm, f->obs.map, f->obs.ntrace, g->nx, g->ny, g->nz, map_pad(g), st)) return rc;
    const dim3 xb(64, 4), xg((f->nx - 1 + 63) / 64, (f->ny - 1 + 3) / 4, std::max(f->nz - 1, 1));
    // exact adjoint: work arrays, and the state before the first turn's
    // forward step (which runs first for the records)
    A3DExactBufs ex{};
    float *turn_state = nullptr;
    if (exact) {
        float *w = nullptr;
        if (int rc = dm.alloc(&w, 7 * n)) return rc;
        if (int rc = launch_zero(w, 7 * n * sizeof(float), st)) return rc;
        for (int a = 0; a < 3; a++) {
            ex.W[a] = w + a * n;
            ex.gR[a] = w + (4 + a) * n;
        }
        ex.gL = w + 3 * n;
        if (int rc = dm.alloc(&ex.gvp, n)) return rc;
        if (int rc = dm.alloc(&ex.grho, n)) return rc;
        if (int rc = dm.alloc(&turn_state, total)) return rc;
    }
    auto exact_step = [&](int capo) -> int {
        if (int rc = rsg_a3d_forward_velocity(g, &sf, st)) return rc;
        if (oi.ncells > 0) {  // the transposed gather: weight 1
            inject3d_ordered_kernel<<<(oi.ncells + 63) / 64, 64, 0, st>>>(sb.P, nullptr, 1.f, oi.cells, oi.starts,
                                                                          oi.ents, oi.ncells, res.shift, res.data,
                                                                          nt, capo);
            RSG_LAUNCH_CHECK("inject3d_ordered_kernel");
        }
        return a3d_exact_reverse(g, &sf, &sb, ex, &f->src, f->wavgrad, capo, map_pad(g), st);
    };

    auto fstep = [&](int it, bool rec) -> int {
        if (int rc = rsg_a3d_forward_velocity(g, &sf, st)) return rc;
        if (int rc = rsg_a3d_forward_stress(g, &sf, st)) return rc;
        if (int rc = insert_source(RSG_KIND_A3D, g, &sf, &f->src, nullptr, it, st)) return rc;
        if (rec)
            if (int rc = record_data(RSG_KIND_A3D, g, &sf, &mod, nullptr, it, st)) return rc;
        return RSG_OK;
    };
    auto reverse_tail = [&](int capo) -> int {
        if (oi.ncells > 0) {
            inject3d_ordered_kernel<<<(oi.ncells + 63) / 64, 64, 0, st>>>(sb.P, f->L, g->dt, oi.cells, oi.starts,
                                                                          oi.ents, oi.ncells, res.shift, res.data,
                                                                          nt, capo);
            RSG_LAUNCH_CHECK("inject3d_ordered_kernel");
        }
        crosscorr_a3d_kernel<<<xg, xb, 0, st>>>(sf.P, sb.Vx, sb.Vy, sb.Vz, f->Rx, f->Ry, f->Rz, f->Vp, f->Rho, f->nx,
                                                f->ny, f->nz, lpml, g->dx, g->dy, g->dz, f->vpgrad, f->rhograd);
        RSG_LAUNCH_CHECK("crosscorr_a3d_kernel");
        if (f->wavgrad)
            if (int rc = record_data(RSG_KIND_A3D, g, &sb, &wg, nullptr, capo, st)) return rc;
        return RSG_OK;
    };
    bool reverse = false;
    for (;;) {
        const int oldcapo = rev.capo;
        const Revolve::Action a = rev.next();
        const int capo = rev.capo;
        if (a == Revolve::ADVANCE) {
            for (int it = oldcapo; it < capo; it++)
                if (int rc = fstep(it, !reverse)) return rc;
        } else if (a == Revolve::FIRSTURN) {
            if (exact)
                RSG_HIP(hipMemcpyAsync(turn_state, fwb, total * sizeof(float), hipMemcpyDeviceToDevice, st));
            if (int rc = fstep(capo, true)) return rc;
            if (int rc = acoustic_misfit_residuals(f->misfit_type, f->dmod, f->obs.data, f->weight, f->dres,
                                                   f->obs.map, 3, f->obs.ntrace, nt, 1, 1, dmis, st))
                return rc;
            if (exact) {
                RSG_HIP(hipMemcpyAsync(fwb, turn_state, total * sizeof(float), hipMemcpyDeviceToDevice, st));
                if (int rc = exact_step(capo)) return rc;
            } else if (int rc = reverse_tail(capo)) {
                return rc;
            }
            reverse = true;
        } else if (a == Revolve::YOUTURN && exact) {
            // fw is left half stepped: Revolve restores before the next step
            if (int rc = exact_step(capo)) return rc;
        } else if (a == Revolve::YOUTURN) {
            if (int rc = rsg_a3d_forward_velocity(g, &sb, st)) return rc;
            if (int rc = rsg_a3d_forward_stress(g, &sb, st)) return rc;
            if (int rc = reverse_tail(capo)) return rc;
        } else if (a == Revolve::TAKESHOT) {
            if (store == kStoreS16) {
                if (int rc = pack_s16(fwb, cks + nblk * kS16Block * rev.check, ckexp + nblk * rev.check, dsegs, dsblk,
                                      dsseg, nblk, st))
                    return rc;
            } else if (f16) {
                if (int rc = pack16_table(fwb, ck16 + cstride * rev.check, ckamax + (size_t)nseg * rev.check, ckpart,
                                          dsegs, dchunks, nchunk, nseg, total, st))
                    return rc;
            } else {
                RSG_HIP(hipMemcpyAsync(ck + total * rev.check, fwb, total * sizeof(float), hipMemcpyDeviceToDevice, st));
            }
        } else if (a == Revolve::RESTORE) {
            if (store == kStoreS16) {
                if (int rc = unpack_s16(cks + nblk * kS16Block * rev.check, ckexp + nblk * rev.check, fwb, dsegs, dsblk,
                                      dsseg, nblk, st))
                    return rc;
            } else if (f16) {
                if (int rc = unpack16_table(ck16 + cstride * rev.check, ckamax + (size_t)nseg * rev.check, fwb, dsegs,
                                            dchunks, nchunk, total, st))
                    return rc;
            } else {
                RSG_HIP(hipMemcpyAsync(fwb, ck + total * rev.check, total * sizeof(float), hipMemcpyDeviceToDevice, st));
            }
        } else if (a == Revolve::TERMINATE) {
            break;
        } else {
            return fail(RSG_EINVAL, "fwi_a3d: Revolve returned an error");
        }
    }
    if (exact)
        if (int rc = a3d_exact_model_grad(g, f->nx, f->ny, f->nz, f->Vp, f->Rho, ex, f->vpgrad, f->rhograd, st))
            return rc;
    double m = 0.0;
    RSG_HIP(hipMemcpyAsync(&m, dmis, sizeof(double), hipMemcpyDeviceToHost, st));
    RSG_HIP(hipStreamSynchronize(st));
    if (f->misfit) *f->misfit = m;
    return RSG_OK;
}

}  // namespace rsg

extern "C" int rsg_fwi_a3d_gradient(const rsg_grid *g, const rsg_fwi_a3d *f, void *stream) {
    if (!g || !f || !f->L || !f->Rx || !f->Ry || !f->Rz || !f->tab || !f->Vp || !f->Rho || !f->vpgrad ||
        !f->rhograd || !f->dmod || !f->dres || !f->obs.data || !f->obs.map || !f->src.data)
        return rsg::fail(RSG_EINVAL, "rsg_fwi_a3d_gradient: null argument");
    return rsg::fwi_a3d_impl(g, f, (hipStream_t)stream);
}
