// Exact discrete adjoint of the acoustic 3D FWI gradient (the gradient mode
// RSG_ADJOINT_EXACT of rsg_fwi_a3d_gradient).  The reference has no such
// mode: FwiAcoustic3D::run_optimal (fwi.cpp:1991-2171) advances its adjoint
// wavefield with the forward step and images with crossCorr.  Here every
// linear map of the forward step (WavesAcoustic3D::forwardstepVelocity /
// forwardstepStress, waves.cpp:710-996; Der; the CFS-PML strips of the six
// faces) is transposed as it stands.
//
// Reverse step t, forward state before step t in fw, adjoint state in bw:
//   1. forward velocity half step on fw (V of step t, P strips after it);
//   2. bw.P += R^T res(t), trilinear weights, cells in trace order;
//   3. insertSource: wavgrad(t) = dt sum w L lP, gL += dt wav(t) w lP;
//   4. reverse stress half step, per cell and axis a with d_a = D-_a V_a:
//        W_a = dt L (1-A-C) lP + A lpsi   (dt L lP outside the a strips)
//        lpsi <- B lpsi - dt L B lP,  gL += lP (dt d_a - dt (psi' + C d_a)),
//      then lV_a += D-_a^T W_a;
//   5. reverse velocity half step, the same with (R_a, D+_a P):
//      gR_a += lV_a f_a, then lP += sum_a D+_a^T U_a.
// One thread per cell, stencils read straight from HBM (the L2 keeps the
// neighbouring planes), run-time stencil orders so that the reduced edge
// stencils and their transposes need no separate path.  This mode is for
// checking and for gradients that must be exact; the production path is the
// reference adjoint with the register-queue kernels.
#include "rsg_adjoint.h"
#include "rsg_adjoint_dev.h"
#include "rsg_fd2d.h"

namespace rsg {

namespace {

struct KAdj3 {
    int nx, ny, nz, lpml, order;
    float dt, inv[3];
    const float *tab;
    const float *P, *V[3], *L, *R[3];
    const float *psiP[6], *psiV[6];  // forward strips: P and V_a, (low, high) of axis a at [2a], [2a+1]
    float *lP, *lV[3];
    float *lpsiP[6], *lpsiV[6];
    float *W[3];
    float *gL, *gR[3];
};

// Strip index of cell (x, y, z) on face side s of axis a, with its table
// index i; -1 when the cell is in no strip of that axis.
__device__ __forceinline__ long strip_q(const KAdj3 &k, int a, int x, int y, int z, int &side, int &i) {
    const int L = k.lpml;
    const int c = a == 0 ? x : (a == 1 ? y : z), n = a == 0 ? k.nx : (a == 1 ? k.ny : k.nz);
    if (c < L) {
        side = 0;
        i = c;
    } else if (c >= n - L) {
        side = 1;
        i = c - (n - L);
    } else {
        return -1;
    }
    if (a == 0) return ((long)z * k.ny + y) * L + i;
    if (a == 1) return ((long)z * L + i) * k.nx + x;
    return ((long)i * k.ny + y) * k.nx + x;
}

__device__ __forceinline__ void cell_of(const KAdj3 &k, int &x, int &y, int &z, bool &ok) {
    x = blockIdx.x * blockDim.x + threadIdx.x;
    y = blockIdx.y * blockDim.y + threadIdx.y;
    z = blockIdx.z;
    ok = x < k.nx && y < k.ny && z < k.nz;
}

__global__ __launch_bounds__(256) void a3d_adj_stress_pre_kernel(const KAdj3 k) {
    int x, y, z;
    bool ok;
    cell_of(k, x, y, z, ok);
    if (!ok) return;
    const long sxy = (long)k.nx * k.ny, p = z * sxy + (long)y * k.nx + x;
    const long st[3] = {1, k.nx, sxy};
    const int c[3] = {x, y, z}, n[3] = {k.nx, k.ny, k.nz};
    const Tab tb{k.tab, k.lpml};
    const float lp = k.lP[p], dtL = k.dt * k.L[p];
    float g = 0.f;
    for (int a = 0; a < 3; a++) {
        const float d = dbw_k(k.V[a] + p, (int)st[a], kbw(c[a], n[a], k.order), k.inv[a]);
        float w = dtL * lp, f = k.dt * d;
        int side, i;
        const long q = strip_q(k, a, x, y, z, side, i);
        if (q >= 0) {
            const float A = tb.A(side, 0, i), B = tb.B(side, 0, i), C = tb.C(side, 0, i);
            strip_term(w, f, k.dt, dtL, lp, d, A, B, C, B * k.psiV[2 * a + side][q] + A * d,
                       k.lpsiV[2 * a + side] + q);
        }
        k.W[a][p] = w;
        g += f;
    }
    k.gL[p] += lp * g;
}

__global__ __launch_bounds__(256) void a3d_adj_velocity_pre_kernel(const KAdj3 k) {
    int x, y, z;
    bool ok;
    cell_of(k, x, y, z, ok);
    if (!ok) return;
    const long sxy = (long)k.nx * k.ny, p = z * sxy + (long)y * k.nx + x;
    const long st[3] = {1, k.nx, sxy};
    const int c[3] = {x, y, z}, n[3] = {k.nx, k.ny, k.nz};
    const Tab tb{k.tab, k.lpml};
    for (int a = 0; a < 3; a++) {
        const float d = dfw_k(k.P + p, (int)st[a], kfw(c[a], n[a], k.order), k.inv[a]);
        const float lv = k.lV[a][p], dtR = k.dt * k.R[a][p];
        float u = dtR * lv, f = k.dt * d;
        int side, i;
        const long q = strip_q(k, a, x, y, z, side, i);
        if (q >= 0)
            strip_term(u, f, k.dt, dtR, lv, d, tb.A(side, 1, i), tb.B(side, 1, i), tb.C(side, 1, i),
                       k.psiP[2 * a + side][q], k.lpsiP[2 * a + side] + q);
        k.W[a][p] = u;
        k.gR[a][p] += lv * f;
    }
}

// VEL = false: lV_a += D-_a^T W_a; true: lP += sum_a D+_a^T W_a.
template <bool VEL>
__global__ __launch_bounds__(256) void a3d_adj_gather_kernel(const KAdj3 k) {
    int x, y, z;
    bool ok;
    cell_of(k, x, y, z, ok);
    if (!ok) return;
    const long sxy = (long)k.nx * k.ny, p = z * sxy + (long)y * k.nx + x;
    const long st[3] = {1, k.nx, sxy};
    const int c[3] = {x, y, z}, n[3] = {k.nx, k.ny, k.nz};
    float acc = 0.f;
    for (int a = 0; a < 3; a++) {
        const float t = dtr_edge<VEL>(k.W[a] + p, (int)st[a], c[a], n[a], k.order, k.inv[a]);
        if (VEL)
            acc += t;
        else
            k.lV[a][p] += t;
    }
    if (VEL) k.lP[p] += acc;
}

// Exact adjoint of insertSource (P += dt L wav w): the wavelet gradient
// dt sum w L lP and the L gradient dt wav w lP (atomic: traces may share
// cells).
__global__ void a3d_adj_source_kernel(const float *lP, const float *L, float *gL, float dt, int nx, int ny, int nz,
                                      int pad, int ntrace, const int *map, const float *shift, const float *wav,
                                      float *wavgrad, int nt, int t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntrace) return;
    const int *m = map + 3 * (size_t)i;
    if (m[0] < 0 || m[1] < 0 || m[2] < 0) return;
    const float *sh = shift + 3 * (size_t)i;
    const float xt[2] = {1 - sh[0], sh[0]}, yt[2] = {1 - sh[1], sh[1]}, zt[2] = {1 - sh[2], sh[2]};
    const float w = wav[(size_t)i * nt + t];
    const size_t ncell = (size_t)nx * ny * nz;
    float acc = 0.f;
    for (int i1 = 0; i1 < 2; i1++)
        for (int i2 = 0; i2 < 2; i2++)
            for (int i3 = 0; i3 < 2; i3++) {
                const size_t p = ((size_t)(pad + m[2] + i1) * ny + (pad + m[1] + i2)) * nx + (pad + m[0] + i3);
                if (p >= ncell) continue;
                const float wt = xt[i3] * yt[i2] * zt[i1], lp = lP[p];
                acc += dt * wt * L[p] * lp;
                unsafeAtomicAdd(gL + p, dt * w * wt * lp);
            }
    if (wavgrad) wavgrad[(size_t)i * nt + t] += acc;
}

// Transposed ModelAcoustic3D::staggerModels per padded cell, in FP64:
// L = rho vp^2 (0 on the free-surface plane z = lpml), R_a = 1 / mean of the
// density along axis a, the last plane copying its neighbour, Rx and Ry
// doubled on the free-surface plane.
__global__ void a3d_adj_model_kernel(const float *gL, const float *gRx, const float *gRy, const float *gRz,
                                     const float *Vp, const float *Rho, int nx, int ny, int nz, int lpml, int fs,
                                     double *gvp, double *grho) {
    const int nxp = nx + 2 * lpml, nyp = ny + 2 * lpml, nzp = nz + 2 * lpml;
    const int px = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y * blockDim.y + threadIdx.y;
    const int pz = blockIdx.z;
    if (px >= nxp || py >= nyp || pz >= nzp) return;
    auto rho = [&](int qx, int qy, int qz) -> double {
        const int ix = min(max(qx - lpml, 0), nx - 1), iy = min(max(qy - lpml, 0), ny - 1),
                  iz = min(max(qz - lpml, 0), nz - 1);
        return (double)Rho[((size_t)iz * ny + iy) * nx + ix];
    };
    auto at = [&](const float *g, int qx, int qy, int qz) -> double {
        return (double)g[((size_t)qz * nyp + qy) * nxp + qx];
    };
    const bool fsrow = fs && pz == lpml;
    const double sxy = fsrow ? 2.0 : 1.0;
    const int ix = min(max(px - lpml, 0), nx - 1), iy = min(max(py - lpml, 0), ny - 1),
              iz = min(max(pz - lpml, 0), nz - 1);
    const double vp = (double)Vp[((size_t)iz * ny + iy) * nx + ix], rp = rho(px, py, pz);
    const double gl = fsrow ? 0.0 : at(gL, px, py, pz);
    double gr = gl * vp * vp;
    // dJ/dA at i (the mean along the axis between i and i+1), with the copied
    // last plane folded into i = n-2
    auto ga = [&](int axis, int i) -> double {
        const int qx = axis == 0 ? i : px, qy = axis == 1 ? i : py, qz = axis == 2 ? i : pz;
        const int np = axis == 0 ? nxp : (axis == 1 ? nyp : nzp);
        const double a = 0.5 * (rho(qx, qy, qz) + rho(qx + (axis == 0), qy + (axis == 1), qz + (axis == 2)));
        const float *g = axis == 0 ? gRx : (axis == 1 ? gRy : gRz);
        double s = at(g, qx, qy, qz);
        if (i == np - 2) s += at(g, qx + (axis == 0), qy + (axis == 1), qz + (axis == 2));
        return -s * (axis < 2 ? sxy : 1.0) / (a * a);
    };
    const int c[3] = {px, py, pz}, np[3] = {nxp, nyp, nzp};
    for (int a = 0; a < 3; a++) {
        if (c[a] <= np[a] - 2) gr += 0.5 * ga(a, c[a]);
        if (c[a] >= 1) gr += 0.5 * ga(a, c[a] - 1);
    }
    const size_t p = ((size_t)pz * nyp + py) * nxp + px;
    gvp[p] = 2.0 * gl * rp * vp;
    grho[p] = gr;
}

// Transposed edge padding: each model cell sums the padded cells that
// replicate it; added to the FP32 images.
__global__ void a3d_adj_unpad_kernel(const double *gvp, const double *grho, int nx, int ny, int nz, int lpml,
                                     float *vpgrad, float *rhograd) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y * blockDim.y + threadIdx.y;
    const int iz = blockIdx.z;
    if (ix >= nx || iy >= ny || iz >= nz) return;
    const int nxp = nx + 2 * lpml, nyp = ny + 2 * lpml;
    auto lo = [&](int i) { return i == 0 ? 0 : i + lpml; };
    auto hi = [&](int i, int n) { return i == n - 1 ? n + 2 * lpml - 1 : i + lpml; };
    double a = 0.0, b = 0.0;
    for (int pz = lo(iz); pz <= hi(iz, nz); pz++)
        for (int py = lo(iy); py <= hi(iy, ny); py++)
            for (int px = lo(ix); px <= hi(ix, nx); px++) {
                const size_t q = ((size_t)pz * nyp + py) * nxp + px;
                a += gvp[q];
                b += grho[q];
            }
    const size_t m = ((size_t)iz * ny + iy) * nx + ix;
    vpgrad[m] = (float)((double)vpgrad[m] + a);
    rhograd[m] = (float)((double)rhograd[m] + b);
}

KAdj3 make_kadj3(const rsg_grid *g, const rsg_state *fw, const rsg_state *bw, const A3DExactBufs &e) {
    KAdj3 k;
    k.nx = g->nx;
    k.ny = g->ny;
    k.nz = g->nz;
    k.lpml = g->lpml;
    k.order = g->order < 1 ? 1 : (g->order > 8 ? 8 : g->order);
    k.dt = g->dt;
    k.inv[0] = 1.0f / g->dx;
    k.inv[1] = 1.0f / g->dy;
    k.inv[2] = 1.0f / g->dz;
    k.tab = fw->tab;
    k.P = fw->P;
    k.V[0] = fw->Vx;
    k.V[1] = fw->Vy;
    k.V[2] = fw->Vz;
    k.L = fw->L;
    k.R[0] = fw->Rx;
    k.R[1] = fw->Ry;
    k.R[2] = fw->Rz;
    const int sp[6] = {RSG_A3D_P_LEFT, RSG_A3D_P_RIGHT, RSG_A3D_P_FRONT, RSG_A3D_P_BACK, RSG_A3D_P_TOP,
                       RSG_A3D_P_BOTTOM};
    const int sv[6] = {RSG_A3D_VXX_LEFT, RSG_A3D_VXX_RIGHT, RSG_A3D_VYY_FRONT, RSG_A3D_VYY_BACK, RSG_A3D_VZZ_TOP,
                       RSG_A3D_VZZ_BOTTOM};
    for (int i = 0; i < 6; i++) {
        k.psiP[i] = fw->psi[sp[i]];
        k.psiV[i] = fw->psi[sv[i]];
        k.lpsiP[i] = bw->psi[sp[i]];
        k.lpsiV[i] = bw->psi[sv[i]];
    }
    k.lP = bw->P;
    k.lV[0] = bw->Vx;
    k.lV[1] = bw->Vy;
    k.lV[2] = bw->Vz;
    for (int a = 0; a < 3; a++) {
        k.W[a] = e.W[a];
        k.gR[a] = e.gR[a];
    }
    k.gL = e.gL;
    return k;
}

}  // namespace

int a3d_exact_reverse(const rsg_grid *g, const rsg_state *fw, const rsg_state *bw, const A3DExactBufs &e,
                      const rsg_traces *src, float *wavgrad, int t, int pad, hipStream_t st) {
    const KAdj3 k = make_kadj3(g, fw, bw, e);
    if (src->ntrace > 0) {
        a3d_adj_source_kernel<<<(src->ntrace + 63) / 64, 64, 0, st>>>(bw->P, fw->L, e.gL, g->dt, g->nx, g->ny, g->nz,
                                                                     pad, src->ntrace, src->map, src->shift,
                                                                     src->data, wavgrad, src->nt, t);
        RSG_LAUNCH_CHECK("a3d_adj_source_kernel");
    }
    const dim3 b(64, 4), grid((g->nx + 63) / 64, (g->ny + 3) / 4, g->nz);
    a3d_adj_stress_pre_kernel<<<grid, b, 0, st>>>(k);
    RSG_LAUNCH_CHECK("a3d_adj_stress_pre_kernel");
    a3d_adj_gather_kernel<false><<<grid, b, 0, st>>>(k);
    RSG_LAUNCH_CHECK("a3d_adj_gather_kernel<D->");
    a3d_adj_velocity_pre_kernel<<<grid, b, 0, st>>>(k);
    RSG_LAUNCH_CHECK("a3d_adj_velocity_pre_kernel");
    a3d_adj_gather_kernel<true><<<grid, b, 0, st>>>(k);
    RSG_LAUNCH_CHECK("a3d_adj_gather_kernel<D+>");
    return RSG_OK;
}

int a3d_exact_model_grad(const rsg_grid *g, int nx, int ny, int nz, const float *Vp, const float *Rho,
                         const A3DExactBufs &e, float *vpgrad, float *rhograd, hipStream_t st) {
    const dim3 b(64, 4);
    a3d_adj_model_kernel<<<dim3((g->nx + 63) / 64, (g->ny + 3) / 4, g->nz), b, 0, st>>>(
        e.gL, e.gR[0], e.gR[1], e.gR[2], Vp, Rho, nx, ny, nz, g->lpml, g->fs, e.gvp, e.grho);
    RSG_LAUNCH_CHECK("a3d_adj_model_kernel");
    a3d_adj_unpad_kernel<<<dim3((nx + 63) / 64, (ny + 3) / 4, nz), b, 0, st>>>(e.gvp, e.grho, nx, ny, nz, g->lpml,
                                                                              vpgrad, rhograd);
    RSG_LAUNCH_CHECK("a3d_adj_unpad_kernel");
    return RSG_OK;
}

}  // namespace rsg
