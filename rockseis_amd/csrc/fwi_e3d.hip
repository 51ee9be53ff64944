// FwiElastic3D::run_optimal (fwi.cpp:6770-6991) on the device: one shot's
// Vp, Vs and rho gradients with Revolve checkpoints of the nine fields and
// 36 PML strips in HBM (FP32 copies, FP16 or S16 packing), the Ux, Uy, Uz
// records, misfit and residuals (misfit.hip), the residuals injected cell by
// cell in trace order into the adjoint state of backwardstepVelocity /
// backwardstepStress (fd3d_elastic.hip), crossCorr (fwi.cpp:5046-5291) into
// lambda, mu and rho kernels, and scaleGrad (fwi.cpp:5553-5630) at the end.
//
// Layout: one contiguous block per state (forward, adjoint): Vx, Vy, Vz,
// Sxx, Syy, Szz, Syz, Sxz, Sxy of the padded grid, then the 36 strips in
// slot order, so a checkpoint is one copy (or one packing pass) of the
// forward block.  The kernels of the imaging condition run one thread per
// interior cell in blocks of 64 (x) by 4 (y), one z plane per grid row.
#include <algorithm>
#include <vector>

#include "rsg_common.h"
#include "rsg_fwi.h"
#include "rsg_inject_ordered.h"
#include "rsg_store16.h"

namespace rsg {

namespace {

struct MemE3 {
    std::vector<void *> ptrs;
    ~MemE3() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **out, size_t n) {
        void *p = nullptr;
        if (n == 0) n = 1;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return fail(RSG_ENOMEM, "fwi_e3d: device allocation failed");
        ptrs.push_back(p);
        *out = (T *)p;
        return RSG_OK;
    }
};

// Face of strip slot i: 0 left/right (nz*ny*lpml), 1 front/back
// (nz*lpml*nx), 2 top/bottom (lpml*ny*nx).
inline int e3d_face(int i) { return i < RSG_E3D_SYY_FRONT ? 0 : (i < RSG_E3D_SZZ_TOP ? 1 : 2); }

// Fields of the imaging condition: source side ws*/fs*, receiver side
// wr*/rs*, order Vx, Vy, Vz, Sxx, Syy, Szz, Syz, Sxz, Sxy.
struct XcE3 {
    const float *s[9], *r[9];
    const float *Rx, *Ry, *Rz;
    int nx, ny, nz, lpml;
    float dx, dy, dz;
    float *lam, *mu, *rho;
};

// crossCorr (fwi.cpp:5046-5291) at interior cell (ix, iy, iz), 1 <= i <= n-2:
// lambda -= div(vs) (rsxx + rsyy + rszz); mu -= 0.25 of the twelve shear
// strain terms times the receiver shear stress at the cell, then
// 2 (msxx rsxx + msyy rsyy + mszz rszz); rho -= the buoyancy-weighted
// stress divergences of the source field against the receiver velocities.
// The reference's operand order and float / double promotions are kept, and
// so are its indices: every shear term of a pair reads the receiver stress at
// the cell itself, and the second, fourth and sixth z terms of the rho
// kernel read wrz one cell back in y with Rz one cell back in z.
__global__ __launch_bounds__(256) void crosscorr_e3d_kernel(const XcE3 k) {
    const int ix = blockIdx.x * 64 + threadIdx.x + 1, iy = blockIdx.y * 4 + threadIdx.y + 1, iz = blockIdx.z + 1;
    if (ix >= k.nx - 1 || iy >= k.ny - 1) return;
    const long nxp = k.nx + 2 * k.lpml, nxyp = nxp * (k.ny + 2 * k.lpml);
    const long X = 1, Y = nxp, Z = nxyp;
    const long c = (long)(iz + k.lpml) * nxyp + (long)(iy + k.lpml) * nxp + ix + k.lpml;
    const long m = ((long)iz * k.ny + iy) * k.nx + ix;
    const float *wsx = k.s[0], *wsy = k.s[1], *wsz = k.s[2], *fsxx = k.s[3], *fsyy = k.s[4], *fszz = k.s[5],
                *fsyz = k.s[6], *fsxz = k.s[7], *fsxy = k.s[8];
    const float *wrx = k.r[0], *wry = k.r[1], *wrz = k.r[2], *rsxx = k.r[3], *rsyy = k.r[4], *rszz = k.r[5],
                *rsyz = k.r[6], *rsxz = k.r[7], *rsxy = k.r[8];
    const float dx = k.dx, dy = k.dy, dz = k.dz;
    const float msxx = (wsx[c] - wsx[c - X]) / dx;
    const float msyy = (wsy[c] - wsy[c - Y]) / dy;
    const float mszz = (wsz[c] - wsz[c - Z]) / dz;
    k.lam[m] -= (msxx + msyy + mszz) * (rsxx[c] + rsyy[c] + rszz[c]);
    // shear terms: (a - b) / h1 + (p - q) / h2, each times the receiver stress
    float v = k.mu[m];
    auto sub = [&](float t, float r) { v = (float)((double)v - 0.25 * (double)t * (double)r); };
    const float ryz = rsyz[c], rxz = rsxz[c], rxy = rsxy[c];
    sub((wsy[c - Y] - wsy[c - Y - Z]) / dz + (wsz[c - Z] - wsz[c - Y - Z]) / dy, ryz);
    sub((wsy[c] - wsy[c - Z]) / dz + (wsz[c + Y - Z] - wsz[c - Z]) / dy, ryz);
    sub((wsy[c - Y + Z] - wsy[c - Y]) / dz + (wsz[c] - wsz[c - Y]) / dy, ryz);
    sub((wsy[c + Z] - wsy[c]) / dz + (wsz[c + Y] - wsz[c]) / dy, ryz);
    sub((wsx[c - X] - wsx[c - X - Z]) / dz + (wsz[c - Z] - wsz[c - X - Z]) / dx, rxz);
    sub((wsx[c] - wsx[c - Z]) / dz + (wsz[c + X - Z] - wsz[c - Z]) / dx, rxz);
    sub((wsx[c - X + Z] - wsx[c - X]) / dz + (wsz[c] - wsz[c - X]) / dx, rxz);
    sub((wsx[c + Z] - wsx[c]) / dz + (wsz[c + X] - wsz[c]) / dx, rxz);
    sub((wsy[c - Y] - wsy[c - X - Y]) / dx + (wsx[c - X] - wsx[c - X - Y]) / dy, rxy);
    sub((wsy[c + X - Y] - wsy[c - Y]) / dx + (wsx[c] - wsx[c - Y]) / dy, rxy);
    sub((wsy[c] - wsy[c - X]) / dx + (wsx[c - X + Y] - wsx[c - X]) / dy, rxy);
    sub((wsy[c + X] - wsy[c]) / dx + (wsx[c + Y] - wsx[c]) / dy, rxy);
    k.mu[m] = (float)((double)v - (2.0 * (double)msxx * (double)rsxx[c] + 2.0 * (double)msyy * (double)rsyy[c] +
                                   2.0 * (double)mszz * (double)rszz[c]));
    // rho: u = (float)(u + 0.5 w R (a - b) / h) term by term
    auto add = [](float u, float w, float r, float d, float h) {
        return (float)((double)u + 0.5 * (double)w * (double)r * (double)d / (double)h);
    };
    const float *Rx = k.Rx, *Ry = k.Ry, *Rz = k.Rz;
    float ux = (float)(0.5 * (double)wrx[c] * (double)Rx[c] * (double)(fsxx[c + X] - fsxx[c]) / (double)dx);
    ux = add(ux, wrx[c - X], Rx[c - X], fsxx[c] - fsxx[c - X], dx);
    ux = add(ux, wrx[c], Rx[c], fsxy[c] - fsxy[c - Y], dy);
    ux = add(ux, wrx[c - X], Rx[c - X], fsxy[c - X] - fsxy[c - X - Y], dy);
    ux = add(ux, wrx[c], Rx[c], fsxz[c] - fsxz[c - Z], dz);
    ux = add(ux, wrx[c - X], Rx[c - X], fsxz[c - X] - fsxz[c - X - Z], dz);
    float uy = (float)(0.5 * (double)wry[c] * (double)Ry[c] * (double)(fsxy[c] - fsxy[c - X]) / (double)dx);
    uy = add(uy, wry[c - Y], Ry[c - Y], fsxy[c - Y] - fsxy[c - X - Y], dx);
    uy = add(uy, wry[c], Ry[c], fsyy[c + Y] - fsyy[c], dy);
    uy = add(uy, wry[c - Y], Ry[c - Y], fsyy[c] - fsyy[c - Y], dy);
    uy = add(uy, wry[c], Ry[c], fsyz[c] - fsyz[c - Z], dz);
    uy = add(uy, wry[c - Y], Ry[c - Y], fsyz[c - Y] - fsyz[c - Y - Z], dz);
    float uz = (float)(0.5 * (double)wrz[c] * (double)Rz[c] * (double)(fsxz[c] - fsxz[c - X]) / (double)dx);
    uz = add(uz, wrz[c - Y], Rz[c - Z], fsxz[c - Z] - fsxz[c - X - Z], dx);
    uz = add(uz, wrz[c], Rz[c], fsyz[c] - fsyz[c - Y], dy);
    uz = add(uz, wrz[c - Y], Rz[c - Z], fsyz[c - Z] - fsyz[c - Y - Z], dy);
    uz = add(uz, wrz[c], Rz[c], fszz[c + Z] - fszz[c], dz);
    uz = add(uz, wrz[c - Y], Rz[c - Z], fszz[c] - fszz[c - Z], dz);
    k.rho[m] -= (ux + uy + uz);
}

// Wavelet gradient of crossCorr (fwi.cpp:5143-5185): source traces mapped to
// interior cells take the receiver field of the source's kind at step it:
// Sxx + Syy + Szz, or the mean of the velocity and its neighbour behind.
__global__ void wavgrad_e3d_kernel(const rsg_traces src, const float *Sxx, const float *Syy, const float *Szz,
                                   const float *Vx, const float *Vy, const float *Vz, int nx, int ny, int nz, int lpml,
                                   float *wavgrad, int it) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= src.ntrace) return;
    const int *mp = src.map + 3 * (size_t)i;
    if (mp[0] < 1 || mp[0] > nx - 2 || mp[1] < 1 || mp[1] > ny - 2 || mp[2] < 1 || mp[2] > nz - 2) return;
    const long nxp = nx + 2 * lpml, nxyp = nxp * (ny + 2 * lpml);
    const long c = (long)(mp[2] + lpml) * nxyp + (long)(mp[1] + lpml) * nxp + mp[0] + lpml;
    float v;
    switch (src.field) {
    case RSG_VX: v = 0.5f * (Vx[c] + Vx[c - 1]); break;
    case RSG_VY: v = 0.5f * (Vy[c] + Vy[c - nxp]); break;
    case RSG_VZ: v = 0.5f * (Vz[c] + Vz[c - nxyp]); break;
    default: v = Sxx[c] + Syy[c] + Szz[c]; break;
    }
    wavgrad[(size_t)i * src.nt + it] = v;
}

// scaleGrad (fwi.cpp:5553-5630) of the lambda / mu / rho kernels, added to
// the caller's Vp / Vs / rho gradients.  The rho gradient is
// vp^2 lambda + vs^2 mu + rho, as the reference forms it.
__global__ void scalegrad_e3d_kernel(const float *Vp, const float *Vs, const float *Rho, const float *lam,
                                     const float *mu, const float *rk, size_t n, float *gvp, float *gvs,
                                     float *grho) {
    const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float vpscale = (float)(2.0 * (double)Rho[k] * (double)Vp[k]);
    const float vsscale = (float)(2.0 * (double)Rho[k] * (double)Vs[k]);
    const float rhoscale1 = Vp[k] * Vp[k], rhoscale2 = Vs[k] * Vs[k];
    const float l = lam[k], u = mu[k];
    gvp[k] += vpscale * l;
    gvs[k] += (float)((double)(vsscale * u) - 2.0 * (double)vsscale * (double)l);
    grho[k] += rhoscale1 * l + rhoscale2 * u + rk[k];
}

// Revolve checkpoint slots of one state block: FP32 copies, FP16 with one
// power-of-two scale per field / strip, or S16 blocks.
struct CheckStore {
    int store = kStoreF32, snaps = 0, nseg = 0, nchunk = 0;
    size_t total = 0, cstride = 0, nblk = 0;
    float *ck = nullptr;
    half_t *ck16 = nullptr;
    unsigned *amax = nullptr, *part = nullptr;
    size_t *dsegs = nullptr, *dsblk = nullptr;
    Chunk16 *dchunks = nullptr;
    short *cks = nullptr, *ckexp = nullptr;
    int *dsseg = nullptr;

    static size_t bytes(int store, size_t total, size_t nseg_blocks, int snaps) {
        if (store == kStoreF32) return total * 4 * (size_t)snaps;
        if (store == kStoreS16) return (nseg_blocks * kS16Block * 2 + nseg_blocks * 2) * (size_t)snaps;
        return ((total + 3) & ~(size_t)3) * 2 * (size_t)snaps;
    }

    int init(MemE3 &dm, int store_, const std::vector<size_t> &segs, int snaps_, hipStream_t st) {
        store = store_;
        snaps = snaps_;
        total = segs.back();
        nseg = (int)segs.size() - 1;
        if (store == kStoreS16) {
            const std::vector<size_t> sblk = s16_block_offsets(segs);
            nblk = sblk.back();
            const std::vector<int> sseg = s16_block_segments(sblk);
            if (int rc = dm.alloc(&cks, nblk * kS16Block * (size_t)snaps)) return rc;
            if (int rc = dm.alloc(&ckexp, nblk * (size_t)snaps)) return rc;
            if (int rc = dm.alloc(&dsegs, segs.size())) return rc;
            if (int rc = dm.alloc(&dsblk, sblk.size())) return rc;
            if (int rc = dm.alloc(&dsseg, sseg.size())) return rc;
            RSG_HIP(hipMemcpyAsync(dsegs, segs.data(), sizeof(size_t) * segs.size(), hipMemcpyHostToDevice, st));
            RSG_HIP(hipMemcpyAsync(dsblk, sblk.data(), sizeof(size_t) * sblk.size(), hipMemcpyHostToDevice, st));
            RSG_HIP(hipMemcpyAsync(dsseg, sseg.data(), sizeof(int) * sseg.size(), hipMemcpyHostToDevice, st));
            RSG_HIP(hipStreamSynchronize(st));  // the host tables go out of scope
        } else if (store == kStoreF16) {
            cstride = (total + 3) & ~(size_t)3;
            const std::vector<Chunk16> chunks = store16_chunks(segs);
            nchunk = (int)chunks.size();
            if (int rc = dm.alloc(&ck16, cstride * (size_t)snaps)) return rc;
            if (int rc = dm.alloc(&amax, (size_t)nseg * snaps)) return rc;
            if (int rc = dm.alloc(&dsegs, segs.size())) return rc;
            if (int rc = dm.alloc(&dchunks, chunks.size())) return rc;
            if (int rc = dm.alloc(&part, chunks.size())) return rc;
            RSG_HIP(hipMemcpyAsync(dsegs, segs.data(), sizeof(size_t) * segs.size(), hipMemcpyHostToDevice, st));
            RSG_HIP(hipMemcpyAsync(dchunks, chunks.data(), sizeof(Chunk16) * chunks.size(), hipMemcpyHostToDevice, st));
            RSG_HIP(hipStreamSynchronize(st));
        } else {
            if (int rc = dm.alloc(&ck, total * (size_t)snaps)) return rc;
        }
        return RSG_OK;
    }
    int save(const float *src, int slot, hipStream_t st) {
        if (store == kStoreS16)
            return pack_s16(src, cks + nblk * kS16Block * slot, ckexp + nblk * slot, dsegs, dsblk, dsseg, nblk, st);
        if (store == kStoreF16)
            return pack16_table(src, ck16 + cstride * slot, amax + (size_t)nseg * slot, part, dsegs, dchunks, nchunk,
                                nseg, total, st);
        RSG_HIP(hipMemcpyAsync(ck + total * slot, src, total * sizeof(float), hipMemcpyDeviceToDevice, st));
        return RSG_OK;
    }
    int load(float *dst, int slot, hipStream_t st) {
        if (store == kStoreS16)
            return unpack_s16(cks + nblk * kS16Block * slot, ckexp + nblk * slot, dst, dsegs, dsblk, dsseg, nblk, st);
        if (store == kStoreF16)
            return unpack16_table(ck16 + cstride * slot, amax + (size_t)nseg * slot, dst, dsegs, dchunks, nchunk,
                                  total, st);
        RSG_HIP(hipMemcpyAsync(dst, ck + total * slot, total * sizeof(float), hipMemcpyDeviceToDevice, st));
        return RSG_OK;
    }
};

bool traces_ok(const rsg_traces &t, int nt) {
    return t.ntrace >= 0 && t.nt == nt && (t.ntrace == 0 || (t.map && t.shift && t.data));
}

}  // namespace

static int fwi_e3d_impl(const rsg_grid *g, const rsg_fwi_e3d *f, hipStream_t st) {
    const int nt = f->src.nt, lpml = g->lpml, nrec = f->obs[0].ntrace;
    if (nt < 1 || !traces_ok(f->src, nt))
        return fail(RSG_EINVAL, "fwi_e3d: source traces need map, shift, wavelet and nt >= 1");
    if (f->src.field != RSG_PRESSURE && f->src.field != RSG_VX && f->src.field != RSG_VY && f->src.field != RSG_VZ)
        return fail(RSG_EINVAL, "fwi_e3d: source field must be PRESSURE, VX, VY or VZ");
    const int recfield[3] = {RSG_VX, RSG_VY, RSG_VZ};
    for (int c = 0; c < 3; c++) {
        if (!traces_ok(f->obs[c], nt) || f->obs[c].ntrace != nrec || f->obs[c].field != recfield[c])
            return fail(RSG_EINVAL, "fwi_e3d: observed Ux, Uy, Uz must be VX, VY, VZ traces of equal count and nt");
        if (nrec > 0 && (!f->dmod[c] || !f->dres[c])) return fail(RSG_EINVAL, "fwi_e3d: null record or residual");
    }
    if (g->domdec) return fail(RSG_EINVAL, "fwi_e3d: models padded by the library only (no domain model)");
    if (f->nx + 2 * lpml != g->nx || f->ny + 2 * lpml != g->ny || f->nz + 2 * lpml != g->nz)
        return fail(RSG_EINVAL, "fwi_e3d: image size must be the unpadded grid");
    if (f->nx < 1 || f->ny < 1 || f->nz < 1 || 2 * lpml >= g->nx || 2 * lpml >= g->ny || 2 * lpml >= g->nz)
        return fail(RSG_EINVAL, "fwi_e3d: grid sizes (padded, 2 lpml < n)");
    if (f->fp16 != kStoreF32 && f->fp16 != kStoreF16 && f->fp16 != kStoreS16)
        return fail(RSG_EINVAL, "fwi_e3d: checkpoint storage must be 0 (FP32), 1 (FP16) or 2 (S16)");
    if (f->misfit_type < 0 || f->misfit_type > 4)
        return fail(RSG_EINVAL, "fwi_e3d: misfit type must be 0..4 (rs_fwimisfit)");
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    const size_t face[3] = {(size_t)g->nz * g->ny * lpml, (size_t)g->nz * lpml * g->nx, (size_t)lpml * g->ny * g->nx};
    std::vector<size_t> segs;
    size_t o = 0;
    for (int j = 0; j < 9; j++, o += n) segs.push_back(o);
    for (int i = 0; i < RSG_E3D_NSTRIP; i++) {
        segs.push_back(o);
        o += face[e3d_face(i)];
    }
    segs.push_back(o);
    const size_t total = o, nm = (size_t)f->nx * f->ny * f->nz;
    Revolve rev(nt, f->ncheck);
    {
        size_t free_b = 0, total_b = 0;
        RSG_HIP(hipMemGetInfo(&free_b, &total_b));
        size_t nblk = 0;
        for (size_t s = 0; s + 1 < segs.size(); s++) nblk += s16_blocks(segs[s + 1] - segs[s]);
        const size_t need = 2 * total * 4 + CheckStore::bytes(f->fp16, total, nblk, rev.snaps) + 3 * nm * 4 + 6 * n * 4;
        if (need + ((size_t)1 << 30) > free_b)
            return fail(RSG_ENOMEM, "fwi_e3d: wavefields and checkpoints do not fit in device memory");
    }
    MemE3 dm;
    float *fwb = nullptr, *bwb = nullptr, *acc = nullptr;
    double *dmis = nullptr;
    if (int rc = dm.alloc(&fwb, total)) return rc;
    if (int rc = dm.alloc(&bwb, total)) return rc;
    if (int rc = dm.alloc(&acc, 3 * nm)) return rc;
    float *wrk = nullptr;  // the backward velocity step's weighted stresses
    if (int rc = dm.alloc(&wrk, 6 * n)) return rc;
    if (int rc = dm.alloc(&dmis, 1)) return rc;
    if (int rc = launch_zero(fwb, total * sizeof(float), st)) return rc;
    if (int rc = launch_zero(bwb, total * sizeof(float), st)) return rc;
    if (int rc = launch_zero(acc, 3 * nm * sizeof(float), st)) return rc;
    if (int rc = launch_zero(dmis, sizeof(double), st)) return rc;
    for (int c = 0; c < 3; c++)
        if (nrec > 0)
            if (int rc = launch_zero(f->dmod[c], sizeof(float) * nrec * (size_t)nt, st)) return rc;
    if (f->wavgrad)
        if (int rc = launch_zero(f->wavgrad, sizeof(float) * f->src.ntrace * (size_t)nt, st)) return rc;
    rsg_state_e3d sf{}, sb{};
    for (int w = 0; w < 2; w++) {
        float *b = w ? bwb : fwb;
        rsg_state_e3d &s = w ? sb : sf;
        float **fld[9] = {&s.Vx, &s.Vy, &s.Vz, &s.Sxx, &s.Syy, &s.Szz, &s.Syz, &s.Sxz, &s.Sxy};
        for (int j = 0; j < 9; j++) *fld[j] = b + segs[j];
        for (int i = 0; i < RSG_E3D_NSTRIP; i++) s.psi[i] = b + segs[9 + i];
        s.L = f->L;
        s.L2M = f->L2M;
        s.Mxz = f->Mxz;
        s.Myz = f->Myz;
        s.Mxy = f->Mxy;
        s.Rx = f->Rx;
        s.Ry = f->Ry;
        s.Rz = f->Rz;
        s.tab = f->tab;
    }
    CheckStore cs;
    if (int rc = cs.init(dm, f->fp16, segs, rev.snaps, st)) return rc;
    // source and residual injection, cell by cell in trace order
    OrderedInject3D osrc, ores[3];
    const int pad = map_pad(g);
    if (f->src.ntrace > 0)
        if (int rc = osrc.build(dm, f->src.map, f->src.ntrace, g->nx, g->ny, g->nz, pad, st)) return rc;
    for (int c = 0; c < 3; c++)
        if (nrec > 0)
            if (int rc = ores[c].build(dm, f->obs[c].map, nrec, g->nx, g->ny, g->nz, pad, st)) return rc;
    E3DRecs recs{};
    recs.nrec = nrec;
    for (int c = 0; c < 3; c++) {
        recs.mod[c] = f->dmod[c];
        recs.obs[c] = f->obs[c].data;
        recs.wei[c] = f->weight[c];
        recs.res[c] = f->dres[c];
    }
    XcE3 xc{};
    for (int w = 0; w < 2; w++) {
        const rsg_state_e3d &s = w ? sb : sf;
        const float **d = w ? xc.r : xc.s;
        const float *p[9] = {s.Vx, s.Vy, s.Vz, s.Sxx, s.Syy, s.Szz, s.Syz, s.Sxz, s.Sxy};
        for (int j = 0; j < 9; j++) d[j] = p[j];
    }
    xc.Rx = f->Rx;
    xc.Ry = f->Ry;
    xc.Rz = f->Rz;
    xc.nx = f->nx;
    xc.ny = f->ny;
    xc.nz = f->nz;
    xc.lpml = lpml;
    xc.dx = g->dx;
    xc.dy = g->dy;
    xc.dz = g->dz;
    xc.lam = acc;
    xc.mu = acc + nm;
    xc.rho = acc + 2 * nm;
    const dim3 xb(64, 4), xg((std::max(f->nx - 2, 1) + 63) / 64, (std::max(f->ny - 2, 1) + 3) / 4,
                             std::max(f->nz - 2, 1));
    const bool interior = f->nx > 2 && f->ny > 2 && f->nz > 2;

    auto src_inject = [&](int it) -> int {
        if (osrc.ncells == 0) return RSG_OK;
        float *F[3] = {nullptr, nullptr, nullptr};
        const float *M = nullptr;
        float scale = g->dt;
        switch (f->src.field) {
        case RSG_VX: F[0] = sf.Vx; M = f->Rx; break;
        case RSG_VY: F[0] = sf.Vy; M = f->Ry; break;
        case RSG_VZ: F[0] = sf.Vz; M = f->Rz; break;
        default: F[0] = sf.Sxx; F[1] = sf.Syy; F[2] = sf.Szz; scale = -g->dt; break;  // explosive: -dt wav w
        }
        for (int j = 0; j < 3 && F[j]; j++) {
            inject3d_ordered_kernel<<<(osrc.ncells + 63) / 64, 64, 0, st>>>(F[j], M, scale, osrc.cells, osrc.starts,
                                                                            osrc.ents, osrc.ncells, f->src.shift,
                                                                            f->src.data, nt, it);
            RSG_LAUNCH_CHECK("inject3d_ordered_kernel");
        }
        return RSG_OK;
    };
    auto fstep = [&](int it, bool rec) -> int {
        if (int rc = rsg_e3d_forward_velocity(g, &sf, st)) return rc;
        if (int rc = rsg_e3d_forward_stress(g, &sf, st)) return rc;
        if (int rc = src_inject(it)) return rc;
        if (rec && nrec > 0)
            for (int c = 0; c < 3; c++) {
                rsg_traces r = f->obs[c];
                r.data = f->dmod[c];
                if (int rc = rsg_e3d_record_data(g, &sf, &r, nullptr, nullptr, nullptr, it, st)) return rc;
            }
        return RSG_OK;
    };
    auto reverse_tail = [&](int capo) -> int {
        float *F[3] = {sb.Vx, sb.Vy, sb.Vz};
        const float *M[3] = {f->Rx, f->Ry, f->Rz};
        for (int c = 0; c < 3; c++)
            if (ores[c].ncells > 0) {
                inject3d_ordered_kernel<<<(ores[c].ncells + 63) / 64, 64, 0, st>>>(
                    F[c], M[c], g->dt, ores[c].cells, ores[c].starts, ores[c].ents, ores[c].ncells, f->obs[c].shift,
                    f->dres[c], nt, capo);
                RSG_LAUNCH_CHECK("inject3d_ordered_kernel");
            }
        if (interior) {
            crosscorr_e3d_kernel<<<xg, xb, 0, st>>>(xc);
            RSG_LAUNCH_CHECK("crosscorr_e3d_kernel");
        }
        if (f->wavgrad && f->src.ntrace > 0) {
            wavgrad_e3d_kernel<<<(f->src.ntrace + 63) / 64, 64, 0, st>>>(f->src, sb.Sxx, sb.Syy, sb.Szz, sb.Vx, sb.Vy,
                                                                          sb.Vz, f->nx, f->ny, f->nz, lpml, f->wavgrad,
                                                                          capo);
            RSG_LAUNCH_CHECK("wavgrad_e3d_kernel");
        }
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
            if (int rc = fstep(capo, true)) return rc;
            if (int rc = elastic3d_misfit_residuals(f->misfit_type, recs, nt, dmis, st)) return rc;
            if (int rc = reverse_tail(capo)) return rc;
            reverse = true;
        } else if (a == Revolve::YOUTURN) {
            if (int rc = e3d_backward_velocity_premul(g, &sb, wrk, st)) return rc;
            if (int rc = rsg_e3d_backward_stress(g, &sb, st)) return rc;
            if (int rc = reverse_tail(capo)) return rc;
        } else if (a == Revolve::TAKESHOT) {
            if (int rc = cs.save(fwb, rev.check, st)) return rc;
        } else if (a == Revolve::RESTORE) {
            if (int rc = cs.load(fwb, rev.check, st)) return rc;
        } else if (a == Revolve::TERMINATE) {
            break;
        } else {
            return fail(RSG_EINVAL, "fwi_e3d: Revolve returned an error");
        }
    }
    scalegrad_e3d_kernel<<<(unsigned)((nm + 255) / 256), 256, 0, st>>>(f->Vp, f->Vs, f->Rho, acc, acc + nm,
                                                                        acc + 2 * nm, nm, f->vpgrad, f->vsgrad,
                                                                        f->rhograd);
    RSG_LAUNCH_CHECK("scalegrad_e3d_kernel");
    double m = 0.0;
    RSG_HIP(hipMemcpyAsync(&m, dmis, sizeof(double), hipMemcpyDeviceToHost, st));
    RSG_HIP(hipStreamSynchronize(st));
    if (f->misfit) *f->misfit = m;
    return RSG_OK;
}

}  // namespace rsg

extern "C" int rsg_fwi_e3d_gradient(const rsg_grid *g, const rsg_fwi_e3d *f, void *stream) {
    if (!g || !f || !f->L || !f->L2M || !f->Mxz || !f->Myz || !f->Mxy || !f->Rx || !f->Ry || !f->Rz || !f->tab ||
        !f->Vp || !f->Vs || !f->Rho || !f->vpgrad || !f->vsgrad || !f->rhograd || !f->src.data)
        return rsg::fail(RSG_EINVAL, "rsg_fwi_e3d_gradient: null argument");
    return rsg::fwi_e3d_impl(g, f, (hipStream_t)stream);
}
