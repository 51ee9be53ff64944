// Exact discrete adjoint of the elastic 2D FWI gradient (the gradient mode
// RSG_ADJOINT_EXACT of rsg_fwi_e2d_gradient_batch and the elastic plan).
// The reference has no such mode: FwiElastic2D::run_optimal
// (fwi.cpp:3937-4143) back-propagates with the adjoint-state steps
// WavesElastic2D::backwardstepVelocity / backwardstepStress
// (waves.cpp:1738-2080) and images with crossCorr.  Here every linear map of
// the forward step (forwardstepVelocity / forwardstepStress,
// waves.cpp:1416-1735; Der, der.cpp:155-540; the CFS-PML strip updates; the
// free-surface row Szz = 0) is transposed as it stands, so the gradient is
// that of the DIFFERENCE misfit of the discrete forward map.
//
// Reverse step t of one shot, with the forward state before step t in fw
// and the adjoint state (fields and the sixteen strips) in bw:
//   1. forward velocity half step on fw (V of step t, stress strips after);
//   2. bw += R^T res(t): P residuals into Sxx and Szz with 1/2, Ux into Vx,
//      Uz into Vz, bilinear weights (fwi_e2d.hip);
//   3. insertSource: the wavelet gradient, and dt wav w lV into gRx / gRz
//      for a force source (fwi_e2d.hip);
//   4. reverse stress half step: the free surface drops lSzz on row lpml;
//      per cell, with the four derivatives d of the step and S the
//      modulus-weighted adjoint stress of each term (L2M lSxx + L lSzz for
//      Dx- Vx, L lSxx + L2M lSzz for Dz- Vz, M lSxz for Dx+ Vz and Dz+ Vx),
//        W = dt S (1-A-C) + A lpsi   (dt S outside the strips)
//        lpsi <- B lpsi - dt S B,  f = dt d - dt (psi' + C d),
//      gL2M += lSxx f_x + lSzz f_z, gL += lSzz f_x + lSxx f_z,
//      gM += lSxz (f_zx + f_xz); then lVx += Dx-^T W_x + Dz+^T W_xz,
//      lVz += Dz-^T W_z + Dx+^T W_zx;
//   5. reverse velocity half step, the same with (Rx, Rz) and the stress
//      derivatives: lSxx += Dx+^T U1, lSxz += Dz-^T U2 + Dx-^T U3,
//      lSzz += Dz+^T U4; gRx, gRz += lV f.
// Steps 4 and 5 are two kernels each (a per-cell pass on LDS tiles of the
// forward fields, then the transposed-stencil gather on LDS tiles of W), for
// all shots of a ragged batch in one launch.  After the sweep the
// time-summed moduli / buoyancy gradients go through the transposed
// ModelElastic2D::staggerModels (model.cpp:2720-2801) in FP64.
#include <algorithm>
#include <vector>

#include "rsg_adjoint.h"
#include "rsg_adjoint_dev.h"
#include "rsg_fd2d.h"

namespace rsg {

namespace {

constexpr int RZE = 2, TZE = BZ * RZE;

// Everything the reverse-step kernels need of one shot.
struct KAdjE {
    K2D f;  // forward state (fw), model, tables, coefficients
    float *lVx, *lVz, *lSxx, *lSzz, *lSxz;
    float *lpsi[RSG_E2D_NSTRIP];
    float *W[4];
    float *gL, *gL2M, *gM, *gRx, *gRz;
};

// Shot and tile of this block: the shots' tiles are numbered consecutively
// (tile0[k] = first tile of shot k, tile0[nshot] = total).
__device__ __forceinline__ const KAdjE *exact_tile(const KAdjE *__restrict__ desc, const int *__restrict__ tile0,
                                                   int nshot, int &x0, int &z0) {
    const int t = blockIdx.x;
    if (t >= tile0[nshot]) return nullptr;
    int lo = 0, hi = nshot - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile0[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    const KAdjE *k = desc + lo;
    const int tl = t - tile0[lo];
    x0 = (tl % k->f.ntx) * BX;
    z0 = (tl / k->f.ntx) * TZE;
    return k;
}

// Per-cell pass of the reverse stress half step.
template <int ORDER, bool FAST>
__device__ __forceinline__ void e2d_stress_pre_body(const KAdjE &k, int x0, int z0, float *sVx, float *sVz) {
    const K2D &a = k.f;
    constexpr int H = ORDER, TZ = TZE, LD = BX + 2 * H;
    const int nx = a.nx, nz = a.nz, tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    const unsigned nb = (unsigned)(nx * nz * 4);
    const int rowb = nx * 4;
    const unsigned cb = (unsigned)(((z0 + ty * RZE) * nx + x) * 4);
    Tile<TZ, H, H> tvx, tvz;
    tile_fetch<TZ, H, H, FAST>(tvx, Ld1{mkbuf(a.Vx, nb)}, nx, nz, x0, z0);
    tile_fetch<TZ, H, H, FAST>(tvz, Ld1{mkbuf(a.Vz, nb)}, nx, nz, x0, z0);
    tile_store(tvx, sVx);
    tile_store(tvz, sVz);
    __syncthreads();
    if (!FAST && x >= nx) return;
    const Buf bXX = mkbuf(k.lSxx, nb), bZZ = mkbuf(k.lSzz, nb), bXZ = mkbuf(k.lSxz, nb), bL = mkbuf(a.L, nb),
              bL2M = mkbuf(a.L2M, nb), bM = mkbuf(a.M, nb), bGL = mkbuf(k.gL, nb), bGL2M = mkbuf(k.gL2M, nb),
              bGM = mkbuf(k.gM, nb), bW0 = mkbuf(k.W[0], nb), bW1 = mkbuf(k.W[1], nb), bW2 = mkbuf(k.W[2], nb),
              bW3 = mkbuf(k.W[3], nb);
    const Tab tb{a.tab, a.lpml};
    const PmlPos px = FAST ? PmlPos{-1, 0} : pml_pos(x, nx, a.lpml);
    const int kxf = FAST ? ORDER : kfw(x, nx, ORDER), kxb = FAST ? ORDER : kbw(x, nx, ORDER);
    const float dt = a.dt;
#pragma unroll
    for (int r = 0; r < RZE; r++) {
        const int zl = ty * RZE + r, z = z0 + zl;
        if (!FAST && z >= nz) break;
        const float *svx = sVx + (zl + H) * LD + tx + H, *svz = sVz + (zl + H) * LD + tx + H;
        const int kzf = FAST ? ORDER : kfw(z, nz, ORDER), kzb = FAST ? ORDER : kbw(z, nz, ORDER);
        const float dVx_x = ddbw<ORDER, FAST>(svx, 1, a.cx.c, kxb, a.invdx);
        const float dVz_z = ddbw<ORDER, FAST>(svz, LD, a.cz.c, kzb, a.invdz);
        const float dVz_x = ddfw<ORDER, FAST>(svz, 1, a.cx.c, kxf, a.invdx);
        const float dVx_z = ddfw<ORDER, FAST>(svx, LD, a.cz.c, kzf, a.invdz);
        const float lxx = cell_ld<FAST>(bXX, cb, r, rowb), lxz = cell_ld<FAST>(bXZ, cb, r, rowb);
        float lzz = cell_ld<FAST>(bZZ, cb, r, rowb);
        if (!FAST && a.fs && z == a.lpml) {  // Szz' = 0 there: its adjoint ends here
            lzz = 0.f;
            cell_st<FAST>(bZZ, 0.f, cb, r, rowb);
        }
        const float ml = cell_ld<FAST>(bL, cb, r, rowb), ml2m = cell_ld<FAST>(bL2M, cb, r, rowb),
                    mm = cell_ld<FAST>(bM, cb, r, rowb);
        const float dSx = dt * (ml2m * lxx + ml * lzz), dSz = dt * (ml * lxx + ml2m * lzz), dSm = dt * (mm * lxz);
        float w0 = dSx, w1 = dSz, w2 = dSm, w3 = dSm;
        float fx = dt * dVx_x, fz = dt * dVz_z, fzx = dt * dVz_x, fxz = dt * dVx_z;
        if (!FAST) {
            if (px.side >= 0) {
                const size_t q = (size_t)z * a.lpml + px.i;
                const int s0 = px.side ? RSG_E2D_VXX_RIGHT : RSG_E2D_VXX_LEFT;
                const int s1 = px.side ? RSG_E2D_VZX_RIGHT : RSG_E2D_VZX_LEFT;
                float A = tb.A(px.side, 0, px.i), B = tb.B(px.side, 0, px.i), C = tb.C(px.side, 0, px.i);
                strip_term(w0, fx, dt, dSx, 1.f, dVx_x, A, B, C, B * a.psi[s0][q] + A * dVx_x, k.lpsi[s0] + q);
                A = tb.A(px.side, 1, px.i);
                B = tb.B(px.side, 1, px.i);
                C = tb.C(px.side, 1, px.i);
                strip_term(w2, fzx, dt, dSm, 1.f, dVz_x, A, B, C, B * a.psi[s1][q] + A * dVz_x, k.lpsi[s1] + q);
            }
            const PmlPos pz = pml_pos(z, nz, a.lpml);
            if (pz.side >= 0) {
                const size_t q = (size_t)pz.i * nx + x;
                const int s0 = pz.side ? RSG_E2D_VZZ_BOTTOM : RSG_E2D_VZZ_TOP;
                const int s1 = pz.side ? RSG_E2D_VXZ_BOTTOM : RSG_E2D_VXZ_TOP;
                float A = tb.A(pz.side, 0, pz.i), B = tb.B(pz.side, 0, pz.i), C = tb.C(pz.side, 0, pz.i);
                strip_term(w1, fz, dt, dSz, 1.f, dVz_z, A, B, C, B * a.psi[s0][q] + A * dVz_z, k.lpsi[s0] + q);
                A = tb.A(pz.side, 1, pz.i);
                B = tb.B(pz.side, 1, pz.i);
                C = tb.C(pz.side, 1, pz.i);
                strip_term(w3, fxz, dt, dSm, 1.f, dVx_z, A, B, C, B * a.psi[s1][q] + A * dVx_z, k.lpsi[s1] + q);
            }
        }
        cell_st<FAST>(bGL2M, cell_ld<FAST>(bGL2M, cb, r, rowb) + (lxx * fx + lzz * fz), cb, r, rowb);
        cell_st<FAST>(bGL, cell_ld<FAST>(bGL, cb, r, rowb) + (lzz * fx + lxx * fz), cb, r, rowb);
        cell_st<FAST>(bGM, cell_ld<FAST>(bGM, cb, r, rowb) + lxz * (fzx + fxz), cb, r, rowb);
        cell_st<FAST>(bW0, w0, cb, r, rowb);
        cell_st<FAST>(bW1, w1, cb, r, rowb);
        cell_st<FAST>(bW2, w2, cb, r, rowb);
        cell_st<FAST>(bW3, w3, cb, r, rowb);
    }
}

// Per-cell pass of the reverse velocity half step.
template <int ORDER, bool FAST>
__device__ __forceinline__ void e2d_velocity_pre_body(const KAdjE &k, int x0, int z0, float *sXX, float *sZZ,
                                                      float *sXZ) {
    const K2D &a = k.f;
    constexpr int H = ORDER, TZ = TZE, LDXX = BX + 2 * H, LDZZ = BX, LDXZ = BX + 2 * H;
    const int nx = a.nx, nz = a.nz, tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    const unsigned nb = (unsigned)(nx * nz * 4);
    const int rowb = nx * 4;
    const unsigned cb = (unsigned)(((z0 + ty * RZE) * nx + x) * 4);
    Tile<TZ, H, 0> txx;
    Tile<TZ, 0, H> tzz;
    Tile<TZ, H, H> txz;
    tile_fetch<TZ, H, 0, FAST>(txx, Ld1{mkbuf(a.Sxx, nb)}, nx, nz, x0, z0);
    tile_fetch<TZ, 0, H, FAST>(tzz, Ld1{mkbuf(a.Szz, nb)}, nx, nz, x0, z0);
    tile_fetch<TZ, H, H, FAST>(txz, Ld1{mkbuf(a.Sxz, nb)}, nx, nz, x0, z0);
    tile_store(txx, sXX);
    tile_store(tzz, sZZ);
    tile_store(txz, sXZ);
    __syncthreads();
    if (!FAST && x >= nx) return;
    const Buf bLX = mkbuf(k.lVx, nb), bLZ = mkbuf(k.lVz, nb), bRx = mkbuf(a.Rx, nb), bRz = mkbuf(a.Rz, nb),
              bGx = mkbuf(k.gRx, nb), bGz = mkbuf(k.gRz, nb), bW0 = mkbuf(k.W[0], nb), bW1 = mkbuf(k.W[1], nb),
              bW2 = mkbuf(k.W[2], nb), bW3 = mkbuf(k.W[3], nb);
    const Tab tb{a.tab, a.lpml};
    const PmlPos px = FAST ? PmlPos{-1, 0} : pml_pos(x, nx, a.lpml);
    const int kxf = FAST ? ORDER : kfw(x, nx, ORDER), kxb = FAST ? ORDER : kbw(x, nx, ORDER);
    const float dt = a.dt;
#pragma unroll
    for (int r = 0; r < RZE; r++) {
        const int zl = ty * RZE + r, z = z0 + zl;
        if (!FAST && z >= nz) break;
        const float d1 = ddfw<ORDER, FAST>(sXX + zl * LDXX + tx + H, 1, a.cx.c, kxf, a.invdx);
        const float *sxz = sXZ + (zl + H) * LDXZ + tx + H;
        const float d2 = ddbw<ORDER, FAST>(sxz, LDXZ, a.cz.c, FAST ? ORDER : kbw(z, nz, ORDER), a.invdz);
        const float d3 = ddbw<ORDER, FAST>(sxz, 1, a.cx.c, kxb, a.invdx);
        const float d4 = ddfw<ORDER, FAST>(sZZ + (zl + H) * LDZZ + tx, LDZZ, a.cz.c,
                                           FAST ? ORDER : kfw(z, nz, ORDER), a.invdz);
        const float lvx = cell_ld<FAST>(bLX, cb, r, rowb), lvz = cell_ld<FAST>(bLZ, cb, r, rowb);
        const float dSx = dt * cell_ld<FAST>(bRx, cb, r, rowb) * lvx, dSz = dt * cell_ld<FAST>(bRz, cb, r, rowb) * lvz;
        float u0 = dSx, u1 = dSx, u2 = dSz, u3 = dSz;
        float f1 = dt * d1, f2 = dt * d2, f3 = dt * d3, f4 = dt * d4;
        if (!FAST) {
            if (px.side >= 0) {
                const size_t q = (size_t)z * a.lpml + px.i;
                const int s0 = px.side ? RSG_E2D_SXX_RIGHT : RSG_E2D_SXX_LEFT;
                const int s1 = px.side ? RSG_E2D_SXZX_RIGHT : RSG_E2D_SXZX_LEFT;
                strip_term(u0, f1, dt, dSx, 1.f, d1, tb.A(px.side, 1, px.i), tb.B(px.side, 1, px.i),
                           tb.C(px.side, 1, px.i), a.psi[s0][q], k.lpsi[s0] + q);
                strip_term(u2, f3, dt, dSz, 1.f, d3, tb.A(px.side, 0, px.i), tb.B(px.side, 0, px.i),
                           tb.C(px.side, 0, px.i), a.psi[s1][q], k.lpsi[s1] + q);
            }
            const PmlPos pz = pml_pos(z, nz, a.lpml);
            if (pz.side >= 0) {
                const size_t q = (size_t)pz.i * nx + x;
                const int s0 = pz.side ? RSG_E2D_SXZZ_BOTTOM : RSG_E2D_SXZZ_TOP;
                const int s1 = pz.side ? RSG_E2D_SZZ_BOTTOM : RSG_E2D_SZZ_TOP;
                strip_term(u1, f2, dt, dSx, 1.f, d2, tb.A(pz.side, 0, pz.i), tb.B(pz.side, 0, pz.i),
                           tb.C(pz.side, 0, pz.i), a.psi[s0][q], k.lpsi[s0] + q);
                strip_term(u3, f4, dt, dSz, 1.f, d4, tb.A(pz.side, 1, pz.i), tb.B(pz.side, 1, pz.i),
                           tb.C(pz.side, 1, pz.i), a.psi[s1][q], k.lpsi[s1] + q);
            }
        }
        cell_st<FAST>(bGx, cell_ld<FAST>(bGx, cb, r, rowb) + lvx * (f1 + f2), cb, r, rowb);
        cell_st<FAST>(bGz, cell_ld<FAST>(bGz, cb, r, rowb) + lvz * (f3 + f4), cb, r, rowb);
        cell_st<FAST>(bW0, u0, cb, r, rowb);
        cell_st<FAST>(bW1, u1, cb, r, rowb);
        cell_st<FAST>(bW2, u2, cb, r, rowb);
        cell_st<FAST>(bW3, u3, cb, r, rowb);
    }
}

// Transposed derivative of an LDS tile at a cell: FWD transposes D+ (-D- of
// the same coefficients in the interior), else D- (-D+).
template <int ORDER, bool FWD, bool FAST>
__device__ __forceinline__ float dtr(const float *s, int st, const float *c, int j, int n, float invd) {
    if (FAST) return FWD ? -dbw<ORDER>(s, st, c + cind(ORDER)) : -dfw<ORDER>(s, st, c + cind(ORDER));
    return dtr_edge<FWD>(s, st, j, n, ORDER, invd);
}

// Transposed-stencil gather.  VEL = false (after the stress pass):
// lVx += Dx-^T W0 + Dz+^T W3, lVz += Dz-^T W1 + Dx+^T W2.  VEL = true
// (after the velocity pass): lSxx += Dx+^T W0, lSxz += Dz-^T W1 + Dx-^T W2,
// lSzz += Dz+^T W3.
template <int ORDER, bool VEL, bool FAST>
__device__ __forceinline__ void e2d_gather_body(const KAdjE &k, int x0, int z0, float *s0, float *s1, float *s2,
                                                float *s3) {
    const K2D &a = k.f;
    constexpr int H = ORDER, TZ = TZE, LDX = BX + 2 * H, LDZ = BX;
    const int nx = a.nx, nz = a.nz, tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    const unsigned nb = (unsigned)(nx * nz * 4);
    const int rowb = nx * 4;
    const unsigned cb = (unsigned)(((z0 + ty * RZE) * nx + x) * 4);
    // W0, W2 are x derivatives' adjoints (x halo), W1, W3 z ones (z halo)
    Tile<TZ, H, 0> t0, t2;
    Tile<TZ, 0, H> t1, t3;
    tile_fetch<TZ, H, 0, FAST>(t0, Ld1{mkbuf(k.W[0], nb)}, nx, nz, x0, z0);
    tile_fetch<TZ, 0, H, FAST>(t1, Ld1{mkbuf(k.W[1], nb)}, nx, nz, x0, z0);
    tile_fetch<TZ, H, 0, FAST>(t2, Ld1{mkbuf(k.W[2], nb)}, nx, nz, x0, z0);
    tile_fetch<TZ, 0, H, FAST>(t3, Ld1{mkbuf(k.W[3], nb)}, nx, nz, x0, z0);
    tile_store(t0, s0);
    tile_store(t1, s1);
    tile_store(t2, s2);
    tile_store(t3, s3);
    __syncthreads();
    if (!FAST && x >= nx) return;
#pragma unroll
    for (int r = 0; r < RZE; r++) {
        const int zl = ty * RZE + r, z = z0 + zl;
        if (!FAST && z >= nz) break;
        const float *p0 = s0 + zl * LDX + tx + H, *p2 = s2 + zl * LDX + tx + H;
        const float *p1 = s1 + (zl + H) * LDZ + tx, *p3 = s3 + (zl + H) * LDZ + tx;
        if (!VEL) {
            const float ax = dtr<ORDER, false, FAST>(p0, 1, a.cx.c, x, nx, a.invdx) +
                             dtr<ORDER, true, FAST>(p3, LDZ, a.cz.c, z, nz, a.invdz);
            const float az = dtr<ORDER, false, FAST>(p1, LDZ, a.cz.c, z, nz, a.invdz) +
                             dtr<ORDER, true, FAST>(p2, 1, a.cx.c, x, nx, a.invdx);
            const Buf bX = mkbuf(k.lVx, nb), bZ = mkbuf(k.lVz, nb);
            cell_st<FAST>(bX, cell_ld<FAST>(bX, cb, r, rowb) + ax, cb, r, rowb);
            cell_st<FAST>(bZ, cell_ld<FAST>(bZ, cb, r, rowb) + az, cb, r, rowb);
        } else {
            const float axx = dtr<ORDER, true, FAST>(p0, 1, a.cx.c, x, nx, a.invdx);
            const float axz = dtr<ORDER, false, FAST>(p1, LDZ, a.cz.c, z, nz, a.invdz) +
                              dtr<ORDER, false, FAST>(p2, 1, a.cx.c, x, nx, a.invdx);
            const float azz = dtr<ORDER, true, FAST>(p3, LDZ, a.cz.c, z, nz, a.invdz);
            const Buf bXX = mkbuf(k.lSxx, nb), bXZ = mkbuf(k.lSxz, nb), bZZ = mkbuf(k.lSzz, nb);
            cell_st<FAST>(bXX, cell_ld<FAST>(bXX, cb, r, rowb) + axx, cb, r, rowb);
            cell_st<FAST>(bXZ, cell_ld<FAST>(bXZ, cb, r, rowb) + axz, cb, r, rowb);
            cell_st<FAST>(bZZ, cell_ld<FAST>(bZZ, cb, r, rowb) + azz, cb, r, rowb);
        }
    }
}

// The per-cell passes read the forward fields' derivatives at the full order
// on tiles ORDER away from the edges (interior_tile); the gathers need every
// row i within ORDER of their cells at the full order: 2*ORDER away.
template <int ORDER>
__global__ __launch_bounds__(BX *BZ) void e2d_adj_stress_pre_kernel(const KAdjE *__restrict__ desc,
                                                                    const int *__restrict__ tile0, int nshot) {
    constexpr int H = ORDER, LD = BX + 2 * H;
    __shared__ float sVx[(TZE + 2 * H) * LD];
    __shared__ float sVz[(TZE + 2 * H) * LD];
    int x0, z0;
    const KAdjE *k = exact_tile(desc, tile0, nshot, x0, z0);
    if (!k) return;
    if (interior_tile(k->f, x0, z0, TZE, H))
        e2d_stress_pre_body<ORDER, true>(*k, x0, z0, sVx, sVz);
    else
        e2d_stress_pre_body<ORDER, false>(*k, x0, z0, sVx, sVz);
}

template <int ORDER>
__global__ __launch_bounds__(BX *BZ) void e2d_adj_velocity_pre_kernel(const KAdjE *__restrict__ desc,
                                                                      const int *__restrict__ tile0, int nshot) {
    constexpr int H = ORDER;
    __shared__ float sXX[TZE * (BX + 2 * H)];
    __shared__ float sZZ[(TZE + 2 * H) * BX];
    __shared__ float sXZ[(TZE + 2 * H) * (BX + 2 * H)];
    int x0, z0;
    const KAdjE *k = exact_tile(desc, tile0, nshot, x0, z0);
    if (!k) return;
    if (interior_tile(k->f, x0, z0, TZE, H))
        e2d_velocity_pre_body<ORDER, true>(*k, x0, z0, sXX, sZZ, sXZ);
    else
        e2d_velocity_pre_body<ORDER, false>(*k, x0, z0, sXX, sZZ, sXZ);
}

template <int ORDER, bool VEL>
__global__ __launch_bounds__(BX *BZ) void e2d_adj_gather_kernel(const KAdjE *__restrict__ desc,
                                                                const int *__restrict__ tile0, int nshot) {
    constexpr int H = ORDER;
    __shared__ float s0[TZE * (BX + 2 * H)];
    __shared__ float s1[(TZE + 2 * H) * BX];
    __shared__ float s2[TZE * (BX + 2 * H)];
    __shared__ float s3[(TZE + 2 * H) * BX];
    int x0, z0;
    const KAdjE *k = exact_tile(desc, tile0, nshot, x0, z0);
    if (!k) return;
    if (interior_tile(k->f, x0, z0, TZE, 2 * H))
        e2d_gather_body<ORDER, VEL, true>(*k, x0, z0, s0, s1, s2, s3);
    else
        e2d_gather_body<ORDER, VEL, false>(*k, x0, z0, s0, s1, s2, s3);
}

// Transposed ModelElastic2D::staggerModels for one window, per padded cell
// in FP64: L2M = rho vp^2, M0 = rho vs^2, L = L2M - 2 M0 (free surface row
// lpml: L = 0, L2M = M0), M = M0 staggered along z then x, Rx / Rz = 1 / mean
// of the density along x / z (Rx doubled on the free-surface row); the
// staggering's last plane copies its neighbour.  scale = 1: gradients of
// (vp, vs, rho); 0: of (lambda, mu, rho) as independent parameters
// (lambda = L, mu = M0).
__global__ void e2d_adj_model_kernel(const float *gL, const float *gL2M, const float *gM, const float *gRx,
                                     const float *gRz, const float *Vp, const float *Vs, const float *Rho, int nx,
                                     int nz, int lpml, int fs, int scale, double *g0, double *g1, double *g2) {
    const int nxp = nx + 2 * lpml, nzp = nz + 2 * lpml;
    const int px = blockIdx.x * blockDim.x + threadIdx.x, pz = blockIdx.y * blockDim.y + threadIdx.y;
    if (px >= nxp || pz >= nzp) return;
    auto cell = [&](const float *m, int qx, int qz) -> double {
        const int ix = min(max(qx - lpml, 0), nx - 1), iz = min(max(qz - lpml, 0), nz - 1);
        return (double)m[(size_t)iz * nx + ix];
    };
    auto at = [&](const float *g, int qx, int qz) -> double { return (double)g[(size_t)qz * nxp + qx]; };
    // staggering transposes: folded last plane, then half to each of the two cells of the mean
    auto gx_fold = [&](const float *g, int i, int qz) -> double {
        return at(g, i, qz) + (i == nxp - 2 ? at(g, nxp - 1, qz) : 0.0);
    };
    auto gM1 = [&](int qz) -> double {  // stagger_x^T of gM at (px, qz)
        double s = 0.0;
        if (px <= nxp - 2) s += 0.5 * gx_fold(gM, px, qz);
        if (px >= 1) s += 0.5 * gx_fold(gM, px - 1, qz);
        return s;
    };
    auto gM1_fold = [&](int j) -> double { return gM1(j) + (j == nzp - 2 ? gM1(nzp - 1) : 0.0); };
    double gm0 = 0.0;  // stagger_z^T of gM1
    if (pz <= nzp - 2) gm0 += 0.5 * gM1_fold(pz);
    if (pz >= 1) gm0 += 0.5 * gM1_fold(pz - 1);
    const double v = cell(Vp, px, pz), w = cell(Vs, px, pz), r = cell(Rho, px, pz);
    const bool fsrow = fs && pz == lpml;
    double gl = at(gL, px, pz), gl2m = at(gL2M, px, pz);
    if (fsrow) {
        gm0 += gl2m;  // L2M = M0 there
        gl = 0.0;
        gl2m = 0.0;
    }
    const double sx = fsrow ? 2.0 : 1.0;
    auto gax = [&](int i) -> double {
        const double a = 0.5 * (cell(Rho, i, pz) + cell(Rho, i + 1, pz));
        double g = at(gRx, i, pz);
        if (i == nxp - 2) g += at(gRx, nxp - 1, pz);
        return -g * sx / (a * a);
    };
    auto gaz = [&](int i) -> double {
        const double a = 0.5 * (cell(Rho, px, i) + cell(Rho, px, i + 1));
        double g = at(gRz, px, i);
        if (i == nzp - 2) g += at(gRz, px, nzp - 1);
        return -g / (a * a);
    };
    double gr = 0.0;
    if (px <= nxp - 2) gr += 0.5 * gax(px);
    if (px >= 1) gr += 0.5 * gax(px - 1);
    if (pz <= nzp - 2) gr += 0.5 * gaz(pz);
    if (pz >= 1) gr += 0.5 * gaz(pz - 1);
    const size_t p = (size_t)pz * nxp + px;
    if (scale) {
        g0[p] = 2.0 * r * v * (gl2m + gl);
        g1[p] = 2.0 * r * w * (gm0 - 2.0 * gl);
        g2[p] = gr + v * v * gl2m + (v * v - 2.0 * w * w) * gl + w * w * gm0;
    } else {
        g0[p] = gl + gl2m;
        g1[p] = 2.0 * gl2m + gm0;
        g2[p] = gr;
    }
}

// Transposed edge padding: each window cell sums the padded cells that
// replicate it, into the window's three FP32 images.
__global__ void e2d_adj_unpad_kernel(const double *g0, const double *g1, const double *g2, int nx, int nz, int lpml,
                                     float *o0, float *o1, float *o2) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iz = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= nx || iz >= nz) return;
    const int nxp = nx + 2 * lpml;
    const int x0 = ix == 0 ? 0 : ix + lpml, x1 = ix == nx - 1 ? nx + 2 * lpml - 1 : ix + lpml;
    const int z0 = iz == 0 ? 0 : iz + lpml, z1 = iz == nz - 1 ? nz + 2 * lpml - 1 : iz + lpml;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int pz = z0; pz <= z1; pz++)
        for (int px = x0; px <= x1; px++) {
            const size_t p = (size_t)pz * nxp + px;
            a += g0[p];
            b += g1[p];
            c += g2[p];
        }
    const size_t m = (size_t)iz * nx + ix;
    o0[m] = (float)((double)o0[m] + a);
    o1[m] = (float)((double)o1[m] + b);
    o2[m] = (float)((double)o2[m] + c);
}

#define ADJE_ORDER_CASES(LAUNCH) \
    switch (p.order) {           \
    case 1: LAUNCH(1); break;    \
    case 2: LAUNCH(2); break;    \
    case 3: LAUNCH(3); break;    \
    case 4: LAUNCH(4); break;    \
    case 5: LAUNCH(5); break;    \
    case 6: LAUNCH(6); break;    \
    case 7: LAUNCH(7); break;    \
    default: LAUNCH(8); break;   \
    }

}  // namespace

int e2d_exact_create(int nshot, const rsg_grid *grids, const rsg_state *fw, const rsg_state *bw,
                     const E2DExactWork *work, E2DExact *out, hipStream_t st) {
    if (nshot < 1 || !grids || !fw || !bw || !work || !out) return fail(RSG_EINVAL, "e2d exact: bad arguments");
    std::vector<KAdjE> desc(nshot);
    std::vector<int> tile0(nshot + 1, 0);
    for (int k = 0; k < nshot; k++) {
        if (grids[k].order != grids[0].order) return fail(RSG_EINVAL, "e2d exact: one stencil order for all shots");
        KAdjE &d = desc[k];
        d.f = make_k2d(grids + k, fw + k, TZE);
        d.lVx = bw[k].Vx;
        d.lVz = bw[k].Vz;
        d.lSxx = bw[k].Sxx;
        d.lSzz = bw[k].Szz;
        d.lSxz = bw[k].Sxz;
        for (int i = 0; i < RSG_E2D_NSTRIP; i++) d.lpsi[i] = bw[k].psi[i];
        for (int i = 0; i < 4; i++) d.W[i] = work[k].W[i];
        d.gL = work[k].gL;
        d.gL2M = work[k].gL2M;
        d.gM = work[k].gM;
        d.gRx = work[k].gRx;
        d.gRz = work[k].gRz;
        const long next = (long)tile0[k] + d.f.ntiles;
        if (next > (1L << 30)) return fail(RSG_EINVAL, "e2d exact: too many tiles");
        tile0[k + 1] = (int)next;
    }
    E2DExact p;
    p.nshot = nshot;
    p.order = grids[0].order;
    p.total = tile0[nshot];
    if (hipMalloc((void **)&p.d_desc, sizeof(KAdjE) * nshot) != hipSuccess)
        return fail(RSG_ENOMEM, "e2d exact: device allocation failed");
    if (hipMalloc((void **)&p.d_tile0, sizeof(int) * (nshot + 1)) != hipSuccess) {
        (void)hipFree(p.d_desc);
        return fail(RSG_ENOMEM, "e2d exact: device allocation failed");
    }
    hipError_t e = hipMemcpyAsync(p.d_desc, desc.data(), sizeof(KAdjE) * nshot, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p.d_tile0, tile0.data(), sizeof(int) * (nshot + 1), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // the host tables go out of scope
    if (e != hipSuccess) {
        (void)hipFree(p.d_desc);
        (void)hipFree(p.d_tile0);
        return check_hip(e, "e2d exact: upload");
    }
    *out = p;
    return RSG_OK;
}

void e2d_exact_destroy(E2DExact *p) {
    if (!p) return;
    if (p->d_desc) (void)hipFree(p->d_desc);
    if (p->d_tile0) (void)hipFree(p->d_tile0);
    p->d_desc = nullptr;
    p->d_tile0 = nullptr;
}

int e2d_exact_reverse(const E2DExact &p, hipStream_t st) {
    if (!p.d_desc) return fail(RSG_EINVAL, "e2d exact: no plan");
    const KAdjE *d = static_cast<const KAdjE *>(p.d_desc);
    const dim3 grid((unsigned)p.total), block(BX, BZ);
#define L_(O) e2d_adj_stress_pre_kernel<O><<<grid, block, 0, st>>>(d, p.d_tile0, p.nshot)
    ADJE_ORDER_CASES(L_);
#undef L_
    RSG_LAUNCH_CHECK("e2d_adj_stress_pre_kernel");
#define L_(O) e2d_adj_gather_kernel<O, false><<<grid, block, 0, st>>>(d, p.d_tile0, p.nshot)
    ADJE_ORDER_CASES(L_);
#undef L_
    RSG_LAUNCH_CHECK("e2d_adj_gather_kernel<stress>");
#define L_(O) e2d_adj_velocity_pre_kernel<O><<<grid, block, 0, st>>>(d, p.d_tile0, p.nshot)
    ADJE_ORDER_CASES(L_);
#undef L_
    RSG_LAUNCH_CHECK("e2d_adj_velocity_pre_kernel");
#define L_(O) e2d_adj_gather_kernel<O, true><<<grid, block, 0, st>>>(d, p.d_tile0, p.nshot)
    ADJE_ORDER_CASES(L_);
#undef L_
    RSG_LAUNCH_CHECK("e2d_adj_gather_kernel<velocity>");
    return RSG_OK;
}

int e2d_exact_model_grad(const E2DExactWork &w, int nx, int nz, int lpml, int fs, int scale, const float *Vp,
                         const float *Vs, const float *Rho, double *pad3, float *o0, float *o1, float *o2,
                         hipStream_t st) {
    const size_t np = (size_t)(nx + 2 * lpml) * (nz + 2 * lpml);
    const dim3 b(64, 4);
    e2d_adj_model_kernel<<<dim3((nx + 2 * lpml + 63) / 64, (nz + 2 * lpml + 3) / 4), b, 0, st>>>(
        w.gL, w.gL2M, w.gM, w.gRx, w.gRz, Vp, Vs, Rho, nx, nz, lpml, fs, scale, pad3, pad3 + np, pad3 + 2 * np);
    RSG_LAUNCH_CHECK("e2d_adj_model_kernel");
    e2d_adj_unpad_kernel<<<dim3((nx + 63) / 64, (nz + 3) / 4), b, 0, st>>>(pad3, pad3 + np, pad3 + 2 * np, nx, nz,
                                                                           lpml, o0, o1, o2);
    RSG_LAUNCH_CHECK("e2d_adj_unpad_kernel");
    return RSG_OK;
}

}  // namespace rsg
