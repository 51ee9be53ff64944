// Exact discrete adjoint of the acoustic 2D FWI gradient (the gradient mode
// RSG_ADJOINT_EXACT of rsg_fwi_a2d_gradient*).  The reference has no such
// mode: FwiAcoustic2D::run_optimal (fwi.cpp:1159-1332) advances its adjoint
// wavefield with the forward step and images with crossCorr.  Here every
// linear map of the forward step (WavesAcoustic2D::forwardstepVelocity /
// forwardstepStress, waves.cpp:217-404; Der, der.cpp:155-540; the CFS-PML
// strip updates) is transposed as it stands, so the gradient is that of the
// DIFFERENCE misfit of the discrete forward map.
//
// Reverse step t, with the forward state before step t in `fw` (Revolve's
// youturn) and the adjoint state in `bw` (P, Vx, Vz and the eight strips
// hold the adjoints of the same quantities after step t):
//   1. forward velocity half step on fw: V of step t, P strips after it;
//   2. bw.P += R^T res(t): residuals at the receivers, bilinear weights;
//   3. insertSource: wavgrad(t) = dt sum w L bw.P, gL += dt wav(t) w bw.P;
//   4. reverse stress half step: per cell, with d = D-(V) and the strip
//      update psi' = B psi + A d of the x / z terms,
//        W = dt L (1-A-C) lP + A lpsi   (dt L lP outside the strips)
//        lpsi <- B lpsi - dt L B lP,    gL += lP (dt d - dt (psi' + C d)),
//      then lVx += Dx-^T Wx, lVz += Dz-^T Wz;
//   5. reverse velocity half step, the same with (P, Vx, Vz, Rx, Rz, D+):
//      gRx += lVx (dt d - dt (psi' + C d)), lP += Dx+^T Ux + Dz+^T Uz.
// Steps 4 and 5 are two kernels each: a per-cell pass that needs the forward
// derivative (LDS tile of the forward field) and writes W, then the
// transposed-stencil gather over W (LDS tile of W).  In the interior the
// transposed staggered stencils are -D+ and -D- with the same coefficients;
// tiles whose stencils reach the reduced-order edge rows sum the transpose
// term by term with each neighbour's own order.
// After the last reverse step gL, gRx, gRz (FP32 time sums per shot) go
// through the transposed model staggering (ModelAcoustic2D::staggerModels,
// model.cpp:1350-1411: L = rho vp^2, Rx, Rz = 1 / mean of the edge-padded
// density, free surface L = 0 and Rx doubled on row lpml) in FP64.
#include "rsg_adjoint.h"
#include "rsg_adjoint_dev.h"
#include "rsg_fd2d.h"
#include "rsg_srcrec_dev.h"

namespace rsg {

namespace {

constexpr int RZA = 2, TZA = BZ * RZA;

struct KAdj {
    K2D f;  // forward state, model, tables and coefficients
    float *lP, *lVx, *lVz;
    float *lpsi[RSG_MAX_STRIPS];
    float *W1, *W2;  // adjoints of the x and z derivatives of the half step
    float *gL, *gRx, *gRz;
};

// Per-cell pass of the reverse stress half step.
template <int ORDER, bool FAST>
__device__ __forceinline__ void stress_pre_body(const KAdj &k, int x0, int z0, float *sVx, float *sVz) {
    const K2D &a = k.f;
    constexpr int H = ORDER, LDX = BX + 2 * H, LDZ = BX, TZ = TZA;
    const int nx = a.nx, nz = a.nz, tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    const size_t fo = blockIdx.y * (size_t)a.sf, xo = blockIdx.y * (size_t)a.sxs,
                 zo = blockIdx.y * (size_t)a.szs;
    const unsigned nb = (unsigned)(nx * nz * 4);
    const int rowb = nx * 4;
    const unsigned cb = (unsigned)(((z0 + ty * RZA) * nx + x) * 4);
    Tile<TZ, H, 0> tvx;
    Tile<TZ, 0, H> tvz;
    tile_fetch<TZ, H, 0, FAST>(tvx, Ld1{mkbuf(a.Vx + fo, nb)}, nx, nz, x0, z0);
    tile_fetch<TZ, 0, H, FAST>(tvz, Ld1{mkbuf(a.Vz + fo, nb)}, nx, nz, x0, z0);
    tile_store(tvx, sVx);
    tile_store(tvz, sVz);
    __syncthreads();
    if (!FAST && x >= nx) return;
    const Buf bLP = mkbuf(k.lP + fo, nb), bL = mkbuf(a.L, nb), bG = mkbuf(k.gL + fo, nb),
              bW1 = mkbuf(k.W1 + fo, nb), bW2 = mkbuf(k.W2 + fo, nb);
    const Tab tb{a.tab, a.lpml};
    const PmlPos px = FAST ? PmlPos{-1, 0} : pml_pos(x, nx, a.lpml);
    const int kx = FAST ? ORDER : kbw(x, nx, ORDER);
    const float dt = a.dt;
#pragma unroll
    for (int r = 0; r < RZA; r++) {
        const int zl = ty * RZA + r, z = z0 + zl;
        if (!FAST && z >= nz) break;
        const float dVx = ddbw<ORDER, FAST>(sVx + zl * LDX + tx + H, 1, a.cx.c, kx, a.invdx);
        const float dVz = ddbw<ORDER, FAST>(sVz + (zl + H) * LDZ + tx, LDZ, a.cz.c,
                                            FAST ? ORDER : kbw(z, nz, ORDER), a.invdz);
        const float lp = cell_ld<FAST>(bLP, cb, r, rowb), dtL = dt * cell_ld<FAST>(bL, cb, r, rowb);
        float wx = dtL * lp, wz = dtL * lp, fx = dt * dVx, fz = dt * dVz;
        if (!FAST) {
            if (px.side >= 0) {
                const size_t q = xo + (size_t)z * a.lpml + px.i;
                const float A = tb.A(px.side, 0, px.i), B = tb.B(px.side, 0, px.i), C = tb.C(px.side, 0, px.i);
                const float po = a.psi[px.side ? RSG_A2D_VXX_RIGHT : RSG_A2D_VXX_LEFT][q];
                strip_term(wx, fx, dt, dtL, lp, dVx, A, B, C, B * po + A * dVx,
                           k.lpsi[px.side ? RSG_A2D_VXX_RIGHT : RSG_A2D_VXX_LEFT] + q);
            }
            const PmlPos pz = pml_pos(z, nz, a.lpml);
            if (pz.side >= 0) {
                const size_t q = zo + (size_t)pz.i * nx + x;
                const float A = tb.A(pz.side, 0, pz.i), B = tb.B(pz.side, 0, pz.i), C = tb.C(pz.side, 0, pz.i);
                const float po = a.psi[pz.side ? RSG_A2D_VZZ_BOTTOM : RSG_A2D_VZZ_TOP][q];
                strip_term(wz, fz, dt, dtL, lp, dVz, A, B, C, B * po + A * dVz,
                           k.lpsi[pz.side ? RSG_A2D_VZZ_BOTTOM : RSG_A2D_VZZ_TOP] + q);
            }
        }
        cell_st<FAST>(bG, cell_ld<FAST>(bG, cb, r, rowb) + lp * (fx + fz), cb, r, rowb);
        cell_st<FAST>(bW1, wx, cb, r, rowb);
        cell_st<FAST>(bW2, wz, cb, r, rowb);
    }
}

// Per-cell pass of the reverse velocity half step.
template <int ORDER, bool FAST>
__device__ __forceinline__ void velocity_pre_body(const KAdj &k, int x0, int z0, float *sP) {
    const K2D &a = k.f;
    constexpr int H = ORDER, TZ = TZA, LD = BX + 2 * H;
    const int nx = a.nx, nz = a.nz, tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    const size_t fo = blockIdx.y * (size_t)a.sf, xo = blockIdx.y * (size_t)a.sxs,
                 zo = blockIdx.y * (size_t)a.szs;
    const unsigned nb = (unsigned)(nx * nz * 4);
    const int rowb = nx * 4;
    const unsigned cb = (unsigned)(((z0 + ty * RZA) * nx + x) * 4);
    Tile<TZ, H, H> tp;
    tile_fetch<TZ, H, H, FAST>(tp, Ld1{mkbuf(a.P + fo, nb)}, nx, nz, x0, z0);
    tile_store(tp, sP);
    __syncthreads();
    if (!FAST && x >= nx) return;
    const Buf bLX = mkbuf(k.lVx + fo, nb), bLZ = mkbuf(k.lVz + fo, nb), bRx = mkbuf(a.Rx, nb),
              bRz = mkbuf(a.Rz, nb), bGx = mkbuf(k.gRx + fo, nb), bGz = mkbuf(k.gRz + fo, nb),
              bW1 = mkbuf(k.W1 + fo, nb), bW2 = mkbuf(k.W2 + fo, nb);
    const Tab tb{a.tab, a.lpml};
    const PmlPos px = FAST ? PmlPos{-1, 0} : pml_pos(x, nx, a.lpml);
    const int kx = FAST ? ORDER : kfw(x, nx, ORDER);
    const float dt = a.dt;
#pragma unroll
    for (int r = 0; r < RZA; r++) {
        const int zl = ty * RZA + r, z = z0 + zl;
        if (!FAST && z >= nz) break;
        const float *s = sP + (zl + H) * LD + tx + H;
        const float dPx = ddfw<ORDER, FAST>(s, 1, a.cx.c, kx, a.invdx);
        const float dPz = ddfw<ORDER, FAST>(s, LD, a.cz.c, FAST ? ORDER : kfw(z, nz, ORDER), a.invdz);
        const float lvx = cell_ld<FAST>(bLX, cb, r, rowb), lvz = cell_ld<FAST>(bLZ, cb, r, rowb);
        const float dtRx = dt * cell_ld<FAST>(bRx, cb, r, rowb), dtRz = dt * cell_ld<FAST>(bRz, cb, r, rowb);
        float ux = dtRx * lvx, uz = dtRz * lvz, fx = dt * dPx, fz = dt * dPz;
        if (!FAST) {
            if (px.side >= 0) {
                const size_t q = xo + (size_t)z * a.lpml + px.i;
                const float A = tb.A(px.side, 1, px.i), B = tb.B(px.side, 1, px.i), C = tb.C(px.side, 1, px.i);
                strip_term(ux, fx, dt, dtRx, lvx, dPx, A, B, C,
                           a.psi[px.side ? RSG_A2D_P_RIGHT : RSG_A2D_P_LEFT][q],
                           k.lpsi[px.side ? RSG_A2D_P_RIGHT : RSG_A2D_P_LEFT] + q);
            }
            const PmlPos pz = pml_pos(z, nz, a.lpml);
            if (pz.side >= 0) {
                const size_t q = zo + (size_t)pz.i * nx + x;
                const float A = tb.A(pz.side, 1, pz.i), B = tb.B(pz.side, 1, pz.i), C = tb.C(pz.side, 1, pz.i);
                strip_term(uz, fz, dt, dtRz, lvz, dPz, A, B, C,
                           a.psi[pz.side ? RSG_A2D_P_BOTTOM : RSG_A2D_P_TOP][q],
                           k.lpsi[pz.side ? RSG_A2D_P_BOTTOM : RSG_A2D_P_TOP] + q);
            }
        }
        cell_st<FAST>(bGx, cell_ld<FAST>(bGx, cb, r, rowb) + lvx * fx, cb, r, rowb);
        cell_st<FAST>(bGz, cell_ld<FAST>(bGz, cb, r, rowb) + lvz * fz, cb, r, rowb);
        cell_st<FAST>(bW1, ux, cb, r, rowb);
        cell_st<FAST>(bW2, uz, cb, r, rowb);
    }
}

// Transposed-stencil gather: FWD (velocity half step, D+): lP += Dx+^T W1 +
// Dz+^T W2; else (stress half step, D-): lVx += Dx-^T W1, lVz += Dz-^T W2.
template <int ORDER, bool FWD, bool FAST>
__device__ __forceinline__ void gather_body(const KAdj &k, int x0, int z0, float *sW1, float *sW2) {
    const K2D &a = k.f;
    constexpr int H = ORDER, LDX = BX + 2 * H, LDZ = BX, TZ = TZA;
    const int nx = a.nx, nz = a.nz, tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    const size_t fo = blockIdx.y * (size_t)a.sf;
    const unsigned nb = (unsigned)(nx * nz * 4);
    const int rowb = nx * 4;
    const unsigned cb = (unsigned)(((z0 + ty * RZA) * nx + x) * 4);
    Tile<TZ, H, 0> t1;
    Tile<TZ, 0, H> t2;
    tile_fetch<TZ, H, 0, FAST>(t1, Ld1{mkbuf(k.W1 + fo, nb)}, nx, nz, x0, z0);
    tile_fetch<TZ, 0, H, FAST>(t2, Ld1{mkbuf(k.W2 + fo, nb)}, nx, nz, x0, z0);
    tile_store(t1, sW1);
    tile_store(t2, sW2);
    __syncthreads();
    if (!FAST && x >= nx) return;
    const Buf bA = mkbuf((FWD ? k.lP : k.lVx) + fo, nb), bB = mkbuf(k.lVz + fo, nb);
#pragma unroll
    for (int r = 0; r < RZA; r++) {
        const int zl = ty * RZA + r, z = z0 + zl;
        if (!FAST && z >= nz) break;
        const float *s1 = sW1 + zl * LDX + tx + H, *s2 = sW2 + (zl + H) * LDZ + tx;
        float ax, az;
        if (FAST) {
            ax = FWD ? -dbw<ORDER>(s1, 1, a.cx.c + cind(ORDER)) : -dfw<ORDER>(s1, 1, a.cx.c + cind(ORDER));
            az = FWD ? -dbw<ORDER>(s2, LDZ, a.cz.c + cind(ORDER)) : -dfw<ORDER>(s2, LDZ, a.cz.c + cind(ORDER));
        } else {
            ax = dtr_edge<FWD>(s1, 1, x, nx, ORDER, a.invdx);
            az = dtr_edge<FWD>(s2, LDZ, z, nz, ORDER, a.invdz);
        }
        if (FWD) {
            cell_st<FAST>(bA, cell_ld<FAST>(bA, cb, r, rowb) + (ax + az), cb, r, rowb);
        } else {
            cell_st<FAST>(bA, cell_ld<FAST>(bA, cb, r, rowb) + ax, cb, r, rowb);
            cell_st<FAST>(bB, cell_ld<FAST>(bB, cb, r, rowb) + az, cb, r, rowb);
        }
    }
}

template <int ORDER>
__global__ __launch_bounds__(BX *BZ) void adj_stress_pre_kernel(const KAdj k) {
    constexpr int H = ORDER;
    __shared__ float sVx[TZA * (BX + 2 * H)];
    __shared__ float sVz[(TZA + 2 * H) * BX];
    int x0, z0;
    if (!tile_of(k.f, TZA, x0, z0)) return;
    if (interior_tile(k.f, x0, z0, TZA, H))
        stress_pre_body<ORDER, true>(k, x0, z0, sVx, sVz);
    else
        stress_pre_body<ORDER, false>(k, x0, z0, sVx, sVz);
}

template <int ORDER>
__global__ __launch_bounds__(BX *BZ) void adj_velocity_pre_kernel(const KAdj k) {
    constexpr int H = ORDER;
    __shared__ float sP[(TZA + 2 * H) * (BX + 2 * H)];
    int x0, z0;
    if (!tile_of(k.f, TZA, x0, z0)) return;
    if (interior_tile(k.f, x0, z0, TZA, H))
        velocity_pre_body<ORDER, true>(k, x0, z0, sP);
    else
        velocity_pre_body<ORDER, false>(k, x0, z0, sP);
}

// The gather's fast path needs every row i within ORDER of its cells at the
// full order: tiles 2*ORDER away from the array edges (interior_tile's halo
// test with 2*ORDER).
template <int ORDER, bool FWD>
__global__ __launch_bounds__(BX *BZ) void adj_gather_kernel(const KAdj k) {
    constexpr int H = ORDER;
    __shared__ float sW1[TZA * (BX + 2 * H)];
    __shared__ float sW2[(TZA + 2 * H) * BX];
    int x0, z0;
    if (!tile_of(k.f, TZA, x0, z0)) return;
    if (interior_tile(k.f, x0, z0, TZA, 2 * H))
        gather_body<ORDER, FWD, true>(k, x0, z0, sW1, sW2);
    else
        gather_body<ORDER, FWD, false>(k, x0, z0, sW1, sW2);
}

// Adjoint of recordData at time t: lP += res(t) * w at each valid trace's
// 2x2 cells (atomic: traces may share cells).
__global__ void adj_record_kernel(float *lP, int nx, int lpml, int ntrace, const int *map, const float *shift,
                                  const float *res, int nt, int t, size_t fstride) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntrace) return;
    const size_t k = blockIdx.y;
    inject2d_trace(i, lP + k * fstride, nullptr, 1.f, nx, lpml, map + k * 2 * ntrace, shift + k * 2 * ntrace,
                   res + k * (size_t)ntrace * nt, nt, t);
}

// Adjoint of insertSource at time t (P += dt L wav(t) w): the wavelet
// gradient dt sum w L lP and the L gradient dt wav(t) w lP.
__global__ void adj_source_kernel(const float *lP, const float *L, float *gL, float dt, int nx, int lpml,
                                  int ntrace, const int *map, const float *shift, const float *wav,
                                  float *wavgrad, int nt, int t, size_t fstride) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntrace) return;
    const size_t k = blockIdx.y;
    const int *m = map + k * 2 * ntrace + 2 * i;
    const float *sh = shift + k * 2 * ntrace + 2 * i;
    if (m[0] < 0 || m[1] < 0) return;
    const float xtri[2] = {1 - sh[0], sh[0]}, ytri[2] = {1 - sh[1], sh[1]};
    const size_t tr = k * (size_t)ntrace + i;
    const float w = wav[tr * nt + t];
    float acc = 0.f;
    for (int i1 = 0; i1 < 2; i1++)
        for (int i2 = 0; i2 < 2; i2++) {
            const size_t p = (size_t)(lpml + m[1] + i1) * nx + (lpml + m[0] + i2);
            const float wt = xtri[i2] * ytri[i1], lp = lP[k * fstride + p];
            acc += dt * wt * L[p] * lp;
            unsafeAtomicAdd(gL + k * fstride + p, dt * w * wt * lp);
        }
    if (wavgrad) wavgrad[tr * nt + t] += acc;
}

// Transposed staggering, per padded cell, in FP64, gradients summed over the
// shots: gvp = 2 rho vp gL, grho = vp^2 gL + the Rx / Rz terms (Rx = s / Ax,
// Ax(i) = (rho(i) + rho(i+1)) / 2, Ax(n-1) = Ax(n-2), s = 2 on the free
// surface row): dJ/dAx = -gRx s / Ax^2, folded from the copied last plane,
// shared half and half by the two densities of the mean.
__global__ void adj_model_kernel(const float *gL, const float *gRx, const float *gRz, const float *Vp,
                                 const float *Rho, int nx, int nz, int lpml, int fs, int nshot, size_t fstride,
                                 double *gvp, double *grho) {
    const int nxp = nx + 2 * lpml, nzp = nz + 2 * lpml;
    const int px = blockIdx.x * blockDim.x + threadIdx.x, pz = blockIdx.y * blockDim.y + threadIdx.y;
    if (px >= nxp || pz >= nzp) return;
    auto rho = [&](int qx, int qz) -> double {
        const int ix = min(max(qx - lpml, 0), nx - 1), iz = min(max(qz - lpml, 0), nz - 1);
        return (double)Rho[(size_t)iz * nx + ix];
    };
    auto sum = [&](const float *g, int qx, int qz) -> double {
        double s = 0.0;
        for (int k = 0; k < nshot; k++) s += (double)g[k * fstride + (size_t)qz * nxp + qx];
        return s;
    };
    const size_t p = (size_t)pz * nxp + px;
    const int ix = min(max(px - lpml, 0), nx - 1), iz = min(max(pz - lpml, 0), nz - 1);
    const double vp = (double)Vp[(size_t)iz * nx + ix], rp = rho(px, pz);
    const bool fsrow = fs && pz == lpml;
    const double gl = fsrow ? 0.0 : sum(gL, px, pz);
    double gr = gl * vp * vp;
    const double sx = fsrow ? 2.0 : 1.0;
    // dJ/dAx at i (i <= nxp-2), with the copied last plane folded into nxp-2
    auto gax = [&](int i) -> double {
        const double a = 0.5 * (rho(i, pz) + rho(i + 1, pz));
        double g = sum(gRx, i, pz);
        if (i == nxp - 2) g += sum(gRx, nxp - 1, pz);
        return -g * sx / (a * a);
    };
    auto gaz = [&](int i) -> double {
        const double a = 0.5 * (rho(px, i) + rho(px, i + 1));
        double g = sum(gRz, px, i);
        if (i == nzp - 2) g += sum(gRz, px, nzp - 1);
        return -g / (a * a);
    };
    if (px <= nxp - 2) gr += 0.5 * gax(px);
    if (px >= 1) gr += 0.5 * gax(px - 1);
    if (pz <= nzp - 2) gr += 0.5 * gaz(pz);
    if (pz >= 1) gr += 0.5 * gaz(pz - 1);
    gvp[p] = 2.0 * gl * rp * vp;
    grho[p] = gr;
}

// Transposed edge padding: each model cell sums the padded cells that
// replicate it; added to the FP32 images.
__global__ void adj_unpad_kernel(const double *gvp, const double *grho, int nx, int nz, int lpml,
                                 float *vpgrad, float *rhograd) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iz = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= nx || iz >= nz) return;
    const int nxp = nx + 2 * lpml;
    const int x0 = ix == 0 ? 0 : ix + lpml, x1 = ix == nx - 1 ? nx + 2 * lpml - 1 : ix + lpml;
    const int z0 = iz == 0 ? 0 : iz + lpml, z1 = iz == nz - 1 ? nz + 2 * lpml - 1 : iz + lpml;
    double a = 0.0, b = 0.0;
    for (int pz = z0; pz <= z1; pz++)
        for (int px = x0; px <= x1; px++) {
            a += gvp[(size_t)pz * nxp + px];
            b += grho[(size_t)pz * nxp + px];
        }
    const size_t m = (size_t)iz * nx + ix;
    vpgrad[m] = (float)((double)vpgrad[m] + a);
    rhograd[m] = (float)((double)rhograd[m] + b);
}

KAdj make_kadj(const rsg_grid *g, const rsg_state *fw, const rsg_state *bw, const rsg_batch *b,
               const A2DExactBufs &e) {
    KAdj k;
    k.f = make_k2d(g, fw, TZA);
    k.f.sf = (long)b->field_stride;
    k.f.sxs = (long)b->xstrip_stride;
    k.f.szs = (long)b->zstrip_stride;
    k.lP = bw->P;
    k.lVx = bw->Vx;
    k.lVz = bw->Vz;
    for (int i = 0; i < RSG_MAX_STRIPS; i++) k.lpsi[i] = bw->psi[i];
    k.W1 = e.W1;
    k.W2 = e.W2;
    k.gL = e.gL;
    k.gRx = e.gRx;
    k.gRz = e.gRz;
    return k;
}

#define ADJ_ORDER_CASES(LAUNCH) \
    switch (g->order) {         \
    case 1: LAUNCH(1); break;   \
    case 2: LAUNCH(2); break;   \
    case 3: LAUNCH(3); break;   \
    case 4: LAUNCH(4); break;   \
    case 5: LAUNCH(5); break;   \
    case 6: LAUNCH(6); break;   \
    case 7: LAUNCH(7); break;   \
    default: LAUNCH(8); break;  \
    }

}  // namespace

int a2d_exact_step(const rsg_grid *g, const rsg_state *fw, const rsg_state *bw, const rsg_batch *b,
                   const A2DExactBufs &e, const rsg_traces *res, const rsg_traces *src, float *wavgrad, int t,
                   hipStream_t st) {
    if (int rc = rsg_a2d_forward_velocity_batch(g, fw, b, st)) return rc;
    const size_t fs = b->field_stride;
    const int ns = b->nshot;
    if (res->ntrace > 0) {
        adj_record_kernel<<<dim3((res->ntrace + 63) / 64, ns), 64, 0, st>>>(
            bw->P, g->nx, g->lpml, res->ntrace, res->map, res->shift, res->data, res->nt, t, fs);
        RSG_LAUNCH_CHECK("adj_record_kernel");
    }
    if (src->ntrace > 0) {
        adj_source_kernel<<<dim3((src->ntrace + 63) / 64, ns), 64, 0, st>>>(
            bw->P, fw->L, e.gL, g->dt, g->nx, g->lpml, src->ntrace, src->map, src->shift, src->data, wavgrad,
            src->nt, t, fs);
        RSG_LAUNCH_CHECK("adj_source_kernel");
    }
    const KAdj k = make_kadj(g, fw, bw, b, e);
    const dim3 grid(nblocks(k.f), ns), block(BX, BZ);
#define L_(O) adj_stress_pre_kernel<O><<<grid, block, 0, st>>>(k)
    ADJ_ORDER_CASES(L_);
#undef L_
    RSG_LAUNCH_CHECK("adj_stress_pre_kernel");
#define L_(O) adj_gather_kernel<O, false><<<grid, block, 0, st>>>(k)
    ADJ_ORDER_CASES(L_);
#undef L_
    RSG_LAUNCH_CHECK("adj_gather_kernel<D->");
#define L_(O) adj_velocity_pre_kernel<O><<<grid, block, 0, st>>>(k)
    ADJ_ORDER_CASES(L_);
#undef L_
    RSG_LAUNCH_CHECK("adj_velocity_pre_kernel");
#define L_(O) adj_gather_kernel<O, true><<<grid, block, 0, st>>>(k)
    ADJ_ORDER_CASES(L_);
#undef L_
    RSG_LAUNCH_CHECK("adj_gather_kernel<D+>");
    return RSG_OK;
}

int a2d_exact_model_grad(const rsg_grid *g, int nx, int nz, const float *Vp, const float *Rho,
                         const A2DExactBufs &e, int nshot, size_t fstride, float *vpgrad, float *rhograd,
                         hipStream_t st) {
    const dim3 b(64, 4);
    adj_model_kernel<<<dim3((g->nx + 63) / 64, (g->nz + 3) / 4), b, 0, st>>>(
        e.gL, e.gRx, e.gRz, Vp, Rho, nx, nz, g->lpml, g->fs, nshot, fstride, e.gvp, e.grho);
    RSG_LAUNCH_CHECK("adj_model_kernel");
    adj_unpad_kernel<<<dim3((nx + 63) / 64, (nz + 3) / 4), b, 0, st>>>(e.gvp, e.grho, nx, nz, g->lpml, vpgrad,
                                                                        rhograd);
    RSG_LAUNCH_CHECK("adj_unpad_kernel");
    return RSG_OK;
}

}  // namespace rsg
