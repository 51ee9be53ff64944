// Acoustic 2D FWI gradient (FwiAcoustic2D::run_optimal, fwi.cpp:1159-1332,
// and FwiAcoustic2D::run, fwi.cpp:1030-1156) on the device, for one shot or a
// batch of shots on one grid.
//
// The forward wavefield is stepped with the fused acoustic kernels.  As in
// the reference, the adjoint wavefield is advanced with the forward operator
// and driven by the DIFFERENCE residuals injected at the receivers; reverse
// steps accumulate the zero-lag cross-correlation of crossCorr
// (fwi.cpp:511-571) into the Vp and rho gradients.  Three ways to get the
// forward pressure back in reverse time:
//
//  RSG_FWI_REVOLVE  run_optimal: the Revolve schedule (revolve.cpp:150-305)
//                   runs on the host; its checkpoints (pressure, velocities,
//                   the 8 PML strips: the payload of revolve.cpp:648-700)
//                   live in HBM.
//  RSG_FWI_STORE    same gradient as run_optimal, forward sweep once: the
//                   unpadded pressure after every step is kept in HBM (the
//                   reference correlates, at reverse step capo, the forward
//                   pressure after capo-1 steps; after nt-1 at the first
//                   turn).  Falls back to REVOLVE when it does not fit.
//  RSG_FWI_FULL     run: the unpadded pressure every snapinc-th step (the
//                   Snapshot2D cadence, snap.cpp:289-346, 638-689) kept in
//                   HBM instead of a snapshot file; the adjoint correlates at
//                   reverse times t with (t - enddiff) % snapinc == 0.
//
// With fp16 set, snapshots and checkpoints are stored as FP16 with one
// power-of-two scale per stored array (rsg_store16.h).
#include <algorithm>
#include <cstring>
#include <utility>
#include <vector>

#include "rsg_common.h"
#include "rsg_adjoint.h"
#include "rsg_fwi.h"
#include "rsg_inject_ordered.h"
#include "rsg_srcrec.h"
#include "rsg_store16.h"

// This file is compiled with -ffp-contract=off (Makefile): every multiply and
// add is rounded separately, as in the reference's host arithmetic.

namespace rsg {

// ---------------------------------------------------------------------------
// Revolve schedule (binomial checkpointing, Griewank & Walther, ACM TOMS
// Algorithm 799): the integer state machine of revolve.cpp:150-305.
// ---------------------------------------------------------------------------

int Revolve::maxrange(int ss, int tt) {
    if (tt < 0 || ss < 0) return -1;
    double res = 1.0;
    for (int i = 1; i <= tt; i++) {
        res *= (ss + i);
        res /= i;
        if (res > 2147483647.0) return 2147483647;
    }
    return (int)res;
}

int Revolve::adjust(int steps) {
    int snaps = 1, reps = 1, s = 0;
    while (maxrange(snaps + s, reps + s) > steps) s--;
    while (maxrange(snaps + s, reps + s) < steps) s++;
    snaps += s;
    reps += s;
    s = -1;
    while (maxrange(snaps, reps) >= steps) {
        if (snaps > reps) {
            snaps -= 1;
            s = 0;
        } else {
            reps -= 1;
            s = 1;
        }
    }
    if (s == 0) snaps += 1;
    if (s == 1) reps += 1;
    return snaps;
}

int Revolve::numforw(int steps, int snaps) {
    if (snaps < 1 || snaps > kCheckUp) return -1;
    int reps = 0, range = 1;
    while (range < steps) {
        reps += 1;
        range = range * (reps + snaps) / reps;
    }
    if (reps > kRepsUp) return -1;
    return reps * steps - range * reps / (snaps + 1);
}

Revolve::Revolve(int steps_, int snaps_) : steps(steps_), snaps(snaps_) {
    if (snaps > steps || snaps < 2) snaps = adjust(steps);
    capo = 0;
    fine = steps;
    check = -1;
    turn = 0;
    oldfine = 0;
    ch.assign(kCheckUp, 0);
}

Revolve::Action Revolve::next() {
    if (check < -1 || capo > fine) return ERROR;
    if (check == -1 && capo < fine) {
        turn = 0;
        ch[0] = capo - 1;
    }
    const int span = fine - capo;
    if (span == 0) {
        if (check == -1 || capo == ch[0]) {
            check -= 1;
            return TERMINATE;
        }
        capo = ch[check];
        oldfine = fine;
        return RESTORE;
    }
    if (span == 1) {
        fine -= 1;
        if (check >= 0 && ch[check] == capo) check -= 1;
        oldfine = fine;
        if (turn == 0) {
            turn = 1;
            return FIRSTURN;
        }
        return YOUTURN;
    }
    if (check == -1 || ch[check] != capo) {
        check += 1;
        if (check >= kCheckUp || check + 1 > snaps) return ERROR;
        ch[check] = capo;
        oldfine = fine;
        return TAKESHOT;
    }
    if (oldfine < fine && snaps == check + 1) return ERROR;
    const int ds = snaps - check;
    if (ds < 1) return ERROR;
    int reps = 0, range = 1;
    while (range < span) {
        reps += 1;
        range = range * (reps + ds) / reps;
    }
    if (reps > kRepsUp) return ERROR;
    const int b1 = range * reps / (ds + reps);
    const int b2 = ds > 1 ? b1 * ds / (ds + reps - 1) : 1;
    const int b3 = ds == 1 ? 0 : (ds > 2 ? b2 * (ds - 1) / (ds + reps - 2) : 1);
    const int b4 = b2 * (reps - 1) / ds;
    const int b5 = ds < 3 ? 0 : (ds > 3 ? b3 * (ds - 2) / reps : 1);
    const int old = capo;
    if (span <= b1 + b3)
        capo = capo + b4;
    else if (span >= range - b5)
        capo = capo + b1;
    else
        capo = fine - b2 - b3;
    if (capo == old) capo = old + 1;
    oldfine = fine;
    return ADVANCE;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// crossCorr for one reverse step, interior ix, iz in [1, n-1) of the
// nx x nz image (fwi.cpp:550-569); Vx_a, Vz_a, Rx, Rz padded by padr (lpml,
// or 0 on a domain model, whose image is the whole padded domain:
// fwi.cpp:535-544).  The rho-gradient velocity terms keep the reference's
// double temporaries.  A batch of nshot shots (fields fstride apart) is
// summed into the one image, shot 0 first.
// Pf is either the padded forward pressure (pf_pad = padr) or a stored
// unpadded snapshot (pf_pad = 0): FP32, FP16 scaled by aux = amax[shot], or
// S16 with aux = the block exponents of the flat [nshot][nz][nx] snapshot;
// pf_stride separates its shots.  srcilum (optional) takes the source
// illumination -(-2/Vp) Pf^2 (fwi.cpp:567-569).
template <int STORE>
__global__ void crosscorr_a2d_kernel(const void *__restrict__ Pf_, const void *__restrict__ aux,
                                     const float *__restrict__ Vxa, const float *__restrict__ Vza,
                                     const float *__restrict__ Rx, const float *__restrict__ Rz,
                                     const float *__restrict__ Vp, const float *__restrict__ Rho, int nx,
                                     int nz, int padr, float dx, float dz, float *__restrict__ vpgrad,
                                     float *__restrict__ rhograd, float *__restrict__ srcilum, int nshot,
                                     size_t fstride, int pf_pad, size_t pf_stride) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int iz = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (ix >= nx - 1 || iz >= nz - 1) return;
    const int nxp = nx + 2 * padr;
    const size_t m = (size_t)iz * nx + ix;
    const size_t q = (size_t)(iz + padr) * nxp + ix + padr;
    const int pld = nx + 2 * pf_pad;
    const size_t qp = (size_t)(iz + pf_pad) * pld + ix + pf_pad;
    const float vpscale = (float)(-2.0 / (double)Vp[m]);
    const float rhoscale = (float)(-1.0 / (double)Rho[m]);
    const double rx0 = Rx[q], rx1 = Rx[q - 1], rz0 = Rz[q], rz1 = Rz[q - nxp];
    // the reference's x / (double)dx as x * (1 / (double)dx): within one
    // double ulp, rounded to float afterwards
    const double idx = 1.0 / (double)dx, idz = 1.0 / (double)dz;
    float vg = vpgrad[m], rg = rhograd[m], si = srcilum ? srcilum[m] : 0.f;
    // unrolled so that several shots' loads are in flight at once (the S16
    // decode would spill if unrolled); the gradients still add shot by shot
    constexpr int kUnroll = STORE == kStoreS16 ? 1 : 4;
#pragma unroll kUnroll
    for (int k = 0; k < nshot; k++) {
        float pc, pe, pw, ps, pn;
        if constexpr (STORE == kStoreS16) {
            const short *Q = (const short *)Pf_;
            const short *E = (const short *)aux;
            const size_t o = k * pf_stride + qp;
            pc = s16_value(Q, E, o);
            pe = s16_value(Q, E, o + 1);
            pw = s16_value(Q, E, o - 1);
            ps = s16_value(Q, E, o + pld);
            pn = s16_value(Q, E, o - pld);
        } else if constexpr (STORE == kStoreF16) {
            const half_t *P = (const half_t *)Pf_ + k * pf_stride;
            const float us = store16_unscale(((const unsigned *)aux)[k]);
            pc = (float)P[qp] * us;
            pe = (float)P[qp + 1] * us;
            pw = (float)P[qp - 1] * us;
            ps = (float)P[qp + pld] * us;
            pn = (float)P[qp - pld] * us;
        } else {
            const float *P = (const float *)Pf_ + k * pf_stride;
            pc = P[qp];
            pe = P[qp + 1];
            pw = P[qp - 1];
            ps = P[qp + pld];
            pn = P[qp - pld];
        }
        const float *Vx = Vxa + k * fstride, *Vz = Vza + k * fstride;
        const float mrxx = (Vx[q] - Vx[q - 1]) / dx;
        const float mrzz = (Vz[q] - Vz[q - nxp]) / dz;
        vg -= vpscale * pc * (mrxx + mrzz);
        rg -= rhoscale * pc * (mrxx + mrzz);
        float ux = (float)(0.5 * (double)Vx[q] * rx0 * (double)(pe - pc) * idx);
        ux = (float)((double)ux + 0.5 * (double)Vx[q - 1] * rx1 * (double)(pc - pw) * idx);
        float uz = (float)(0.5 * (double)Vz[q] * rz0 * (double)(ps - pc) * idz);
        uz = (float)((double)uz + 0.5 * (double)Vz[q - nxp] * rz1 * (double)(pc - pn) * idz);
        rg += ux + uz;
        if (srcilum) si -= vpscale * pc * pc;
    }
    vpgrad[m] = vg;
    rhograd[m] = rg;
    if (srcilum) srcilum[m] = si;
}

// ---------------------------------------------------------------------------
// Driver
// ---------------------------------------------------------------------------

namespace {

struct DevBuf {
    std::vector<void *> ptrs;
    ~DevBuf() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **out, size_t n) {
        void *p = nullptr;
        if (n == 0) n = 1;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return fail(RSG_ENOMEM, "fwi: device allocation failed");
        ptrs.push_back(p);
        *out = (T *)p;
        return RSG_OK;
    }
};

// WavesAcoustic2D states of a batch of shots: P, Vx, Vz and the 8 PML
// strips, one allocation; each array holds all shots back to back.
struct A2DWaves {
    float *base = nullptr;
    size_t n = 0, nstrip_x = 0, nstrip_z = 0, total = 0;
    int ns = 1;
    rsg_state st{};
    rsg_batch b{};
    int init(DevBuf &db, const rsg_grid *g, const rsg_fwi_a2d *f, int nshot, hipStream_t s) {
        ns = nshot;
        n = (size_t)g->nx * g->nz;
        nstrip_x = (size_t)g->lpml * g->nz;
        nstrip_z = (size_t)g->nx * g->lpml;
        total = ns * (3 * n + 4 * nstrip_x + 4 * nstrip_z);
        if (int rc = db.alloc(&base, total)) return rc;
        if (int rc = launch_zero(base, total * sizeof(float), s)) return rc;
        st = rsg_state{};
        st.L = f->L;
        st.Rx = f->Rx;
        st.Rz = f->Rz;
        st.tab = f->tab;
        st = at(base);
        b.nshot = ns;
        b.field_stride = n;
        b.xstrip_stride = nstrip_x;
        b.zstrip_stride = nstrip_z;
        return RSG_OK;
    }
    // The state laid out over any buffer of `total` values (the model and
    // PML tables of st): P, Vx, Vz, then the x strips and the z strips.
    rsg_state at(float *buf) const {
        rsg_state r = st;
        r.P = buf;
        r.Vx = buf + ns * n;
        r.Vz = buf + 2 * ns * n;
        float *p = buf + 3 * ns * n;
        const int xs[4] = {RSG_A2D_P_LEFT, RSG_A2D_VXX_LEFT, RSG_A2D_P_RIGHT, RSG_A2D_VXX_RIGHT};
        const int zs[4] = {RSG_A2D_P_TOP, RSG_A2D_VZZ_TOP, RSG_A2D_P_BOTTOM, RSG_A2D_VZZ_BOTTOM};
        for (int i = 0; i < 4; i++, p += ns * nstrip_x) r.psi[xs[i]] = p;
        for (int i = 0; i < 4; i++, p += ns * nstrip_z) r.psi[zs[i]] = p;
        return r;
    }
};

}  // namespace

static int fwi_a2d_impl(const rsg_grid *g, const rsg_fwi_a2d *f, int ns, hipStream_t st) {
    const int nt = f->src.nt;
    if (f->obs.nt != nt || f->obs.field != RSG_PRESSURE || f->src.field != RSG_PRESSURE)
        return fail(RSG_EINVAL, "fwi: source and data must be PRESSURE traces with equal nt");
    // image offset on the grid: lpml, or none on a domain model (fwi.cpp:535-544)
    const int pad = g->domdec ? 0 : g->lpml;
    if (f->nx + 2 * pad != g->nx || f->nz + 2 * pad != g->nz)
        return fail(RSG_EINVAL, g->domdec ? "fwi: image size must be the grid of the domain model"
                                          : "fwi: image size must be the unpadded grid");
    if (g->domdec && f->adjoint == RSG_ADJOINT_EXACT)
        return fail(RSG_EINVAL, "fwi: the exact adjoint takes models padded by the library (no domain model)");
    if (f->srcilum && f->adjoint == RSG_ADJOINT_EXACT)
        return fail(RSG_EINVAL, "fwi: source illumination is computed by the reference imaging condition only");
    if (f->mode == RSG_FWI_FULL && f->snapinc < 1) return fail(RSG_EINVAL, "fwi: snapinc must be >= 1");
    if (f->adjoint != RSG_ADJOINT_REFERENCE && f->adjoint != RSG_ADJOINT_EXACT)
        return fail(RSG_EINVAL, "fwi: adjoint must be RSG_ADJOINT_REFERENCE or RSG_ADJOINT_EXACT");
    const bool exact = f->adjoint == RSG_ADJOINT_EXACT;
    if (exact && (f->mode != RSG_FWI_REVOLVE || (f->misfit_type != 0 && f->misfit_type != 4)))
        return fail(RSG_EINVAL, "fwi: the exact adjoint takes RSG_FWI_REVOLVE and the DIFFERENCE misfit");
    const int store = f->fp16;
    if (store != kStoreF32 && store != kStoreF16 && store != kStoreS16)
        return fail(RSG_EINVAL, "fwi: storage must be 0 (FP32), 1 (FP16) or 2 (S16)");
    const bool f16 = store != kStoreF32;  // 2 bytes a stored value
    if (store == kStoreF16 && 4 * ns > 65535) return fail(RSG_EINVAL, "fwi: FP16 storage takes at most 16383 shots");
    DevBuf db;
    A2DWaves fw, bw;
    if (int rc = fw.init(db, g, f, ns, st)) return rc;
    if (int rc = bw.init(db, g, f, ns, st)) return rc;
    const size_t ndat = (size_t)ns * f->obs.ntrace * nt;
    double *misfit_d = nullptr;
    if (int rc = db.alloc(&misfit_d, 1)) return rc;
    if (int rc = launch_zero(misfit_d, sizeof(double), st)) return rc;
    if (int rc = launch_zero(f->dmod, sizeof(float) * ndat, st)) return rc;
    if (f->wavgrad)
        if (int rc = launch_zero(f->wavgrad, sizeof(float) * ns * f->src.ntrace * nt, st)) return rc;

    rsg_traces mod = f->obs, res = f->obs, wg = f->src;
    mod.data = f->dmod;
    res.data = f->dres;
    wg.data = f->wavgrad;
    OrderedInject2D oi;
    if (f->obs.ntrace > 0)
        if (int rc = oi.build(db, f->obs.map, f->obs.ntrace, ns, g->nx, g->nz, map_pad(g), st)) return rc;
    const dim3 xb(64, 4), xg((f->nx + 63) / 64, (f->nz + 3) / 4);
    const size_t fs = fw.n;

    auto fstep = [&](int it, bool rec) -> int {
        if (int rc = rsg_a2d_forward_velocity_batch(g, &fw.st, &fw.b, st)) return rc;
        if (int rc = rsg_a2d_forward_stress_batch(g, &fw.st, &fw.b, st)) return rc;
        if (int rc = insert_source(RSG_KIND_A2D, g, &fw.st, &f->src, nullptr, it, st, ns, fs)) return rc;
        if (rec)
            if (int rc = record_data(RSG_KIND_A2D, g, &fw.st, &mod, nullptr, it, st, ns, fs)) return rc;
        return RSG_OK;
    };
    auto bstep = [&]() -> int {
        if (int rc = rsg_a2d_forward_velocity_batch(g, &bw.st, &bw.b, st)) return rc;
        return rsg_a2d_forward_stress_batch(g, &bw.st, &bw.b, st);
    };
    auto first_turn = [&]() -> int {
        return acoustic_misfit_residuals(f->misfit_type, f->dmod, f->obs.data, f->weight, f->dres, f->obs.map, 2,
                                         f->obs.ntrace, nt, ns, 1, misfit_d, st);
    };
    // Residual injection at time t, crossCorr with Pf (skipped when null;
    // FP16 or S16 as `store` when aux is set) and the wavelet gradient at t.
    auto reverse_tail = [&](int t, const void *Pf, int pf_pad, size_t pf_stride, const void *aux) -> int {
        if (oi.maxcells > 0) {
            inject2d_ordered_kernel<<<dim3((oi.maxcells + 63) / 64, ns), 64, 0, st>>>(
                bw.st.P, f->L, g->dt, oi.cells, oi.starts, oi.ents, oi.coff, res.shift, res.data, res.ntrace, nt, t,
                fs);
            RSG_LAUNCH_CHECK("inject2d_ordered_kernel");
        }
        if (Pf) {
            if (aux && store == kStoreS16)
                crosscorr_a2d_kernel<kStoreS16><<<xg, xb, 0, st>>>(Pf, aux, bw.st.Vx, bw.st.Vz, f->Rx, f->Rz, f->Vp,
                                                                   f->Rho, f->nx, f->nz, pad, g->dx, g->dz,
                                                                   f->vpgrad, f->rhograd, f->srcilum, ns, fs, pf_pad,
                                                                   pf_stride);
            else if (aux)
                crosscorr_a2d_kernel<kStoreF16><<<xg, xb, 0, st>>>(Pf, aux, bw.st.Vx, bw.st.Vz, f->Rx, f->Rz, f->Vp,
                                                                   f->Rho, f->nx, f->nz, pad, g->dx, g->dz,
                                                                   f->vpgrad, f->rhograd, f->srcilum, ns, fs, pf_pad,
                                                                   pf_stride);
            else
                crosscorr_a2d_kernel<kStoreF32><<<xg, xb, 0, st>>>(Pf, nullptr, bw.st.Vx, bw.st.Vz, f->Rx, f->Rz,
                                                                   f->Vp, f->Rho, f->nx, f->nz, pad, g->dx,
                                                                   g->dz, f->vpgrad, f->rhograd, f->srcilum, ns, fs,
                                                                   pf_pad, pf_stride);
            RSG_LAUNCH_CHECK("crosscorr_a2d_kernel");
        }
        if (f->wavgrad)
            if (int rc = record_data(RSG_KIND_A2D, g, &bw.st, &wg, nullptr, t, st, ns, fs)) return rc;
        return RSG_OK;
    };

    // Snapshot history of the STORE and FULL modes: nslot slots of the
    // unpadded pressure of every shot, FP32, FP16 (+ one amax per shot) or
    // S16 (+ the exponents of the slot's blocks).
    const size_t nint = (size_t)f->nx * f->nz, slot = (size_t)ns * nint;
    const size_t esz = f16 ? sizeof(half_t) : sizeof(float);
    int mode = f->mode, snapnt = nt, enddiff = 0;
    if (mode == RSG_FWI_FULL)
        if (int rc = rsg_snap_cadence(nt, g->dt, f->snapinc, &snapnt, &enddiff)) return rc;
    if (mode != RSG_FWI_REVOLVE) {
        size_t free_b = 0, total_b = 0;
        RSG_HIP(hipMemGetInfo(&free_b, &total_b));
        const bool fits = slot * (size_t)snapnt * esz + ((size_t)2 << 30) <= free_b;
        if (!fits && mode == RSG_FWI_FULL)
            return fail(RSG_ENOMEM, "fwi: the snapshots of RSG_FWI_FULL do not fit in device memory");
        if (!fits) mode = RSG_FWI_REVOLVE;
    }
    if (f->mode_used) *f->mode_used = mode;
    char *hist = nullptr;
    unsigned *hamax = nullptr;
    short *hexp = nullptr;
    const size_t sblk = s16_blocks(slot);
    if (mode != RSG_FWI_REVOLVE) {
        if (int rc = db.alloc(&hist, slot * (size_t)snapnt * esz)) return rc;
        if (store == kStoreF16)
            if (int rc = db.alloc(&hamax, (size_t)ns * snapnt)) return rc;
        if (store == kStoreS16)
            if (int rc = db.alloc(&hexp, sblk * snapnt)) return rc;
    }
    auto snap_at = [&](int k) -> void * { return hist + slot * (size_t)k * esz; };
    auto amax_at = [&](int k) -> const void * {
        return store == kStoreF16 ? (const void *)(hamax + (size_t)ns * k)
                                  : store == kStoreS16 ? (const void *)(hexp + sblk * k) : nullptr;
    };
    auto take_snap = [&](int k) -> int {
        return pack_interior(fw.st.P, snap_at(k), (void *)amax_at(k), f->nx, f->nz, pad, fw.n, ns, store, st);
    };

    if (mode == RSG_FWI_STORE) {
        for (int it = 0; it < nt; it++) {
            if (int rc = fstep(it, true)) return rc;
            if (int rc = take_snap(it)) return rc;
        }
        if (int rc = first_turn()) return rc;
        if (int rc = reverse_tail(nt - 1, snap_at(nt - 1), 0, nint, amax_at(nt - 1))) return rc;
        for (int capo = nt - 2; capo >= 0; capo--) {
            if (int rc = bstep()) return rc;
            // state after capo-1 steps; before the first step it is zero
            if (capo >= 1) {
                if (int rc = reverse_tail(capo, snap_at(capo - 1), 0, nint, amax_at(capo - 1))) return rc;
            } else {
                if (int rc = launch_zero(fw.st.P, (size_t)ns * fw.n * sizeof(float), st)) return rc;
                if (int rc = reverse_tail(capo, fw.st.P, pad, fw.n, nullptr)) return rc;
            }
        }
    } else if (mode == RSG_FWI_FULL) {
        for (int it = 0; it < nt; it++) {
            if (int rc = fstep(it, true)) return rc;
            if (it % f->snapinc == 0)
                if (int rc = take_snap(it / f->snapinc)) return rc;
        }
        if (int rc = first_turn()) return rc;
        for (int t = nt - 1; t >= 0; t--) {
            if (int rc = bstep()) return rc;
            const bool corr = (t - enddiff) % f->snapinc == 0;
            const int k = (t - enddiff) / f->snapinc;
            if (int rc = reverse_tail(t, corr ? snap_at(k) : nullptr, 0, nint, corr ? amax_at(k) : nullptr))
                return rc;
        }
    } else {
        // exact adjoint: work arrays and a copy of the state before the last
        // forward step (firsturn runs that step for the records first)
        A2DExactBufs ex{};
        float *turn_state = nullptr;
        if (exact) {
            const size_t nn = (size_t)ns * fw.n;
            float *w = nullptr;
            if (int rc = db.alloc(&w, 5 * nn)) return rc;
            if (int rc = launch_zero(w, 5 * nn * sizeof(float), st)) return rc;
            ex.W1 = w;
            ex.W2 = w + nn;
            ex.gL = w + 2 * nn;
            ex.gRx = w + 3 * nn;
            ex.gRz = w + 4 * nn;
            if (int rc = db.alloc(&ex.gvp, fw.n)) return rc;
            if (int rc = db.alloc(&ex.grho, fw.n)) return rc;
            if (int rc = db.alloc(&turn_state, fw.total)) return rc;
        }
        auto exact_step = [&](int t) -> int {
            return a2d_exact_step(g, &fw.st, &bw.st, &fw.b, ex, &res, &f->src, f->wavgrad, t, st);
        };
        Revolve rev(nt, f->ncheck);
        // checkpoint = P, Vx, Vz (3*ns arrays of n), 4*ns x strips, 4*ns z strips
        float *ck = nullptr;
        half_t *ck16 = nullptr;
        unsigned *ckamax = nullptr;
        const int nseg = 11 * ns;
        // S16: one segment per array of each shot (blocks restart at every
        // field and strip); a checkpoint is cblk blocks of kS16Block values
        std::vector<size_t> soff(1, 0), sblkoff;
        short *cks = nullptr, *ckexp = nullptr;
        size_t *dsoff = nullptr, *dsblk = nullptr, cblk = 0;
        int *dsseg = nullptr;
        std::vector<int> sseg;
        if (store == kStoreS16) {
            const size_t lens[3] = {fw.n, fw.nstrip_x, fw.nstrip_z};
            const int narr[3] = {3, 4, 4};
            for (int g3 = 0; g3 < 3; g3++)
                for (int a = 0; a < narr[g3] * ns; a++) soff.push_back(soff.back() + lens[g3]);
            sblkoff = s16_block_offsets(soff);
            cblk = sblkoff.back();
            if (int rc = db.alloc(&cks, cblk * kS16Block * (size_t)rev.snaps)) return rc;
            if (int rc = db.alloc(&ckexp, cblk * rev.snaps)) return rc;
            if (int rc = db.alloc(&dsoff, soff.size())) return rc;
            if (int rc = db.alloc(&dsblk, sblkoff.size())) return rc;
            RSG_HIP(hipMemcpyAsync(dsoff, soff.data(), sizeof(size_t) * soff.size(), hipMemcpyHostToDevice, st));
            RSG_HIP(hipMemcpyAsync(dsblk, sblkoff.data(), sizeof(size_t) * sblkoff.size(), hipMemcpyHostToDevice, st));
            sseg = s16_block_segments(sblkoff);
            if (int rc = db.alloc(&dsseg, sseg.size())) return rc;
            RSG_HIP(hipMemcpyAsync(dsseg, sseg.data(), sizeof(int) * sseg.size(), hipMemcpyHostToDevice, st));
        } else if (f16) {
            if (int rc = db.alloc(&ck16, fw.total * (size_t)rev.snaps)) return rc;
            if (int rc = db.alloc(&ckamax, (size_t)nseg * rev.snaps)) return rc;
        } else {
            if (int rc = db.alloc(&ck, fw.total * (size_t)rev.snaps)) return rc;
        }
        const size_t xoff = 3 * fw.n * ns, zoff = xoff + 4 * fw.nstrip_x * ns;
        auto ck_pack = [&](int c) -> int {
            if (store == kStoreS16)
                return pack_s16(fw.base, cks + cblk * kS16Block * c, ckexp + cblk * c, dsoff, dsblk, dsseg, cblk, st);
            half_t *d = ck16 + fw.total * c;
            unsigned *am = ckamax + (size_t)nseg * c;
            if (int rc = pack16_segments(fw.base, d, am, fw.n, 3 * ns, st)) return rc;
            if (int rc = pack16_segments(fw.base + xoff, d + xoff, am + 3 * ns, fw.nstrip_x, 4 * ns, st))
                return rc;
            return pack16_segments(fw.base + zoff, d + zoff, am + 7 * ns, fw.nstrip_z, 4 * ns, st);
        };
        auto ck_unpack = [&](int c) -> int {
            if (store == kStoreS16)
                return unpack_s16(cks + cblk * kS16Block * c, ckexp + cblk * c, fw.base, dsoff, dsblk, dsseg, cblk, st);
            const half_t *d = ck16 + fw.total * c;
            const unsigned *am = ckamax + (size_t)nseg * c;
            if (int rc = unpack16_segments(d, am, fw.base, fw.n, 3 * ns, st)) return rc;
            if (int rc = unpack16_segments(d + xoff, am + 3 * ns, fw.base + xoff, fw.nstrip_x, 4 * ns, st))
                return rc;
            return unpack16_segments(d + zoff, am + 7 * ns, fw.base + zoff, fw.nstrip_z, 4 * ns, st);
        };
        // FP32 checkpoints of the reference adjoint are kept by reference, not
        // copied: the forward state lives in one of rev.snaps + 2 buffers
        // (fw.base, a spare, the checkpoint slots).  TAKESHOT hands the
        // current buffer to the checkpoint, RESTORE points the current state
        // at it, and the next forward step after either runs out of place
        // from that buffer into a free one, so the checkpoint is never
        // written.  A turn only reads the forward pressure, straight from the
        // checkpoint (1674 of the 2999 restores at nt 3000 / 50 snapshots
        // are followed by a turn, the rest by a forward step).
        const bool zc = store == kStoreF32 && !exact;
        std::vector<float *> pool, slot(rev.snaps, nullptr);
        float *cur = fw.base;
        bool cur_ro = false;  // cur is a checkpoint's buffer
        if (zc) {
            float *spare = nullptr;
            if (int rc = db.alloc(&spare, fw.total)) return rc;
            pool.push_back(spare);
            for (int c = rev.snaps - 1; c >= 0; c--) pool.push_back(ck + fw.total * c);
        }
        auto take_free = [&]() -> float * {
            if (pool.empty()) return nullptr;
            float *p = pool.back();
            pool.pop_back();
            return p;
        };
        auto zstep = [&](int it, bool rec) -> int {
            rsg_state cs = fw.at(cur);
            if (cur_ro) {
                float *nb = take_free();
                if (!nb) return fail(RSG_EINVAL, "fwi: no free checkpoint buffer");
                const rsg_state out = fw.at(nb);
                if (int rc = a2d_forward_step_from(g, &cs, &out, &fw.b, st)) return rc;
                cur = nb;
                cur_ro = false;
                cs = out;
            } else {
                if (int rc = rsg_a2d_forward_velocity_batch(g, &cs, &fw.b, st)) return rc;
                if (int rc = rsg_a2d_forward_stress_batch(g, &cs, &fw.b, st)) return rc;
            }
            if (int rc = insert_source(RSG_KIND_A2D, g, &cs, &f->src, nullptr, it, st, ns, fs)) return rc;
            if (rec)
                if (int rc = record_data(RSG_KIND_A2D, g, &cs, &mod, nullptr, it, st, ns, fs)) return rc;
            return RSG_OK;
        };
        auto step = [&](int it, bool rec) -> int { return zc ? zstep(it, rec) : fstep(it, rec); };
        const float *const *Pcur = zc ? (const float *const *)&cur : (const float *const *)&fw.st.P;
        bool reverse = false;
        for (;;) {
            const int oldcapo = rev.capo;
            const Revolve::Action a = rev.next();
            const int capo = rev.capo;
            if (a == Revolve::ADVANCE) {
                for (int it = oldcapo; it < capo; it++)
                    if (int rc = step(it, !reverse)) return rc;
            } else if (a == Revolve::FIRSTURN) {
                if (exact)
                    RSG_HIP(hipMemcpyAsync(turn_state, fw.base, fw.total * sizeof(float), hipMemcpyDeviceToDevice,
                                           st));
                if (int rc = step(capo, true)) return rc;
                if (int rc = first_turn()) return rc;
                if (exact) {
                    RSG_HIP(hipMemcpyAsync(fw.base, turn_state, fw.total * sizeof(float), hipMemcpyDeviceToDevice,
                                           st));
                    if (int rc = exact_step(capo)) return rc;
                } else if (int rc = reverse_tail(capo, *Pcur, pad, fw.n, nullptr)) {
                    return rc;
                }
                reverse = true;
            } else if (a == Revolve::YOUTURN) {
                // the state is discarded after a turn: Revolve restores next
                if (exact) {
                    if (int rc = exact_step(capo)) return rc;
                    continue;
                }
                if (int rc = bstep()) return rc;
                if (int rc = reverse_tail(capo, *Pcur, pad, fw.n, nullptr)) return rc;
            } else if (a == Revolve::TAKESHOT && zc) {
                float *&sl = slot[rev.check];
                if (!cur_ro) {
                    if (sl) pool.push_back(sl);
                    sl = cur;
                    cur_ro = true;
                } else if (sl != cur) {
                    float *nb = take_free();
                    if (!nb) return fail(RSG_EINVAL, "fwi: no free checkpoint buffer");
                    RSG_HIP(hipMemcpyAsync(nb, cur, fw.total * sizeof(float), hipMemcpyDeviceToDevice, st));
                    if (sl) pool.push_back(sl);
                    sl = nb;
                }
            } else if (a == Revolve::RESTORE && zc) {
                if (!slot[rev.check]) return fail(RSG_EINVAL, "fwi: Revolve restored an empty checkpoint");
                if (!cur_ro) pool.push_back(cur);
                cur = slot[rev.check];
                cur_ro = true;
            } else if (a == Revolve::TAKESHOT) {
                if (f16) {
                    if (int rc = ck_pack(rev.check)) return rc;
                } else {
                    RSG_HIP(hipMemcpyAsync(ck + fw.total * rev.check, fw.base, fw.total * sizeof(float),
                                           hipMemcpyDeviceToDevice, st));
                }
            } else if (a == Revolve::RESTORE) {
                if (f16) {
                    if (int rc = ck_unpack(rev.check)) return rc;
                } else {
                    RSG_HIP(hipMemcpyAsync(fw.base, ck + fw.total * rev.check, fw.total * sizeof(float),
                                           hipMemcpyDeviceToDevice, st));
                }
            } else if (a == Revolve::TERMINATE) {
                break;
            } else {
                return fail(RSG_EINVAL, "fwi: Revolve returned an error");
            }
        }
        if (exact)
            if (int rc = a2d_exact_model_grad(g, f->nx, f->nz, f->Vp, f->Rho, ex, ns, fw.n, f->vpgrad, f->rhograd,
                                              st))
                return rc;
    }
    double m = 0.0;
    RSG_HIP(hipMemcpyAsync(&m, misfit_d, sizeof(double), hipMemcpyDeviceToHost, st));
    RSG_HIP(hipStreamSynchronize(st));
    if (f->misfit) *f->misfit = m;
    return RSG_OK;
}

}  // namespace rsg

extern "C" int rsg_fwi_a2d_gradient(const rsg_grid *g, const rsg_fwi_a2d *f, void *stream) {
    if (!g || !f || !f->L || !f->Rx || !f->Rz || !f->tab || !f->Vp || !f->Rho || !f->vpgrad ||
        !f->rhograd || !f->dmod || !f->dres || !f->obs.data)
        return rsg::fail(RSG_EINVAL, "rsg_fwi_a2d_gradient: null argument");
    if (f->mode != RSG_FWI_REVOLVE && f->mode != RSG_FWI_STORE && f->mode != RSG_FWI_FULL)
        return rsg::fail(RSG_EINVAL, "rsg_fwi_a2d_gradient: unknown mode");
    return rsg::fwi_a2d_impl(g, f, 1, (hipStream_t)stream);
}

extern "C" int rsg_fwi_a2d_gradient_batch(const rsg_grid *g, const rsg_fwi_a2d *f, int nshot,
                                          void *stream) {
    if (!g || !f || !f->L || !f->Rx || !f->Rz || !f->tab || !f->Vp || !f->Rho || !f->vpgrad ||
        !f->rhograd || !f->dmod || !f->dres || !f->obs.data)
        return rsg::fail(RSG_EINVAL, "rsg_fwi_a2d_gradient_batch: null argument");
    if (nshot < 1 || nshot > 65535) return rsg::fail(RSG_EINVAL, "rsg_fwi_a2d_gradient_batch: nshot 1..65535");
    if (f->mode != RSG_FWI_REVOLVE && f->mode != RSG_FWI_STORE && f->mode != RSG_FWI_FULL)
        return rsg::fail(RSG_EINVAL, "rsg_fwi_a2d_gradient_batch: unknown mode");
    return rsg::fwi_a2d_impl(g, f, nshot, (hipStream_t)stream);
}

extern "C" int rsg_crosscorr_a2d(const float *wsp, int pads, const float *wrx, const float *wrz, int padr,
                                 const float *Vp, const float *Rho, const float *Rx, const float *Rz, int nx, int nz,
                                 float dx, float dz, int domdec, float *vpgrad, float *rhograd, float *srcilum,
                                 void *stream) {
    if (!wsp || !wrx || !wrz || !Vp || !Rho || !Rx || !Rz || !vpgrad || !rhograd || nx < 1 || nz < 1 || pads < 0 ||
        padr < 0)
        return rsg::fail(RSG_EINVAL, "rsg_crosscorr_a2d: bad arguments");
    if (domdec) pads = padr = 0;
    if (nx < 3 || nz < 3) return RSG_OK;  // no interior cell
    const dim3 xb(64, 4), xg((nx + 63) / 64, (nz + 3) / 4);
    rsg::crosscorr_a2d_kernel<rsg::kStoreF32><<<xg, xb, 0, (hipStream_t)stream>>>(
        wsp, nullptr, wrx, wrz, Rx, Rz, Vp, Rho, nx, nz, padr, dx, dz, vpgrad, rhograd, srcilum, 1, 0, pads, 0);
    RSG_LAUNCH_CHECK("crosscorr_a2d_kernel");
    return RSG_OK;
}

extern "C" int rsg_revolve_schedule(int nt, int snaps, int *actions, int *capos, int *checks, int maxn,
                                    int *n) {
    if (!actions || !capos || !checks || !n || nt < 1) return rsg::fail(RSG_EINVAL, "bad arguments");
    rsg::Revolve r(nt, snaps);
    int k = 0;
    for (;;) {
        const rsg::Revolve::Action a = r.next();
        if (k >= maxn) return rsg::fail(RSG_EINVAL, "rsg_revolve_schedule: maxn too small");
        actions[k] = (int)a;
        capos[k] = r.capo;
        checks[k] = r.check;
        k++;
        if (a == rsg::Revolve::TERMINATE || a == rsg::Revolve::ERROR) break;
    }
    *n = k;
    return RSG_OK;
}

extern "C" int rsg_revolve_numforw(int nt, int snaps) { return rsg::Revolve::numforw(nt, snaps); }

extern "C" int rsg_revolve_adjust(int nt) { return rsg::Revolve::adjust(nt); }

extern "C" int rsg_revolve_init(rsg_revolve_state *st, int nt, int snaps, int *adjusted) {
    if (!st || nt < 1) return rsg::fail(RSG_EINVAL, "rsg_revolve_init: bad arguments");
    rsg::Revolve r(nt, snaps);
    if (adjusted) *adjusted = r.snaps != snaps;
    st->steps = r.steps;
    st->snaps = r.snaps;
    st->capo = r.capo;
    st->fine = r.fine;
    st->check = r.check;
    st->turn = r.turn;
    st->oldfine = r.oldfine;
    for (int i = 0; i < rsg::Revolve::kCheckUp; i++) st->ch[i] = r.ch[i];
    return RSG_OK;
}

extern "C" int rsg_revolve_next(rsg_revolve_state *st) {
    if (!st) return rsg::Revolve::ERROR;
    rsg::Revolve r(st->steps > 1 ? st->steps : 2, 2);
    r.steps = st->steps;
    r.snaps = st->snaps;
    r.capo = st->capo;
    r.fine = st->fine;
    r.check = st->check;
    r.turn = st->turn;
    r.oldfine = st->oldfine;
    r.ch.assign(st->ch, st->ch + rsg::Revolve::kCheckUp);
    const int a = (int)r.next();
    st->capo = r.capo;
    st->fine = r.fine;
    st->check = r.check;
    st->turn = r.turn;
    st->oldfine = r.oldfine;
    for (int i = 0; i < rsg::Revolve::kCheckUp; i++) st->ch[i] = r.ch[i];
    return a;
}
