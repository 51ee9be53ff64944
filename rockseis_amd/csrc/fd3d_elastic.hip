// WavesElastic3D (waves.cpp:2310-4403) on the device: the velocity and
// stress half steps of the 3D velocity-stress system with the CFS-PML
// memory strips of all six faces (36 strips, PmlElastic3D, pml.h:197-234)
// fused into the updates, the free surface (Szz = 0 on plane lpml), source
// injection and receiver gathers, and ModelElastic3D::staggerModels on the
// host (bit-exact, model.cpp:3371-3473).
//
// Kernels: one thread per cell, blocks of 64 (x) by 4 (y) cells, one z plane
// per block row of the grid.  Every derivative reads its stencil straight
// from HBM: the x stencil is coalesced within the wave, the y and z stencils
// hit the L2 lines the neighbouring blocks and planes brought in.  Cells whose
// stencils are at the full order take the compile-time stencil with the
// coefficients pre-divided by the step; edge cells take the reference's
// reduced orders.  Each derivative term is applied as the reference applies
// it: F += dt M d, then in a strip psi = B psi + A d, F -= dt M (psi + C d).
// Results agree with the reference to rounding (contraction is on here).
//
// The adjoint-state steps of FwiElastic3D (backwardstepVelocity / Stress,
// waves.cpp:3260-4103) share the kernels: the velocity step differentiates
// the moduli-weighted stresses (L2M Sxx + L Syy + L Szz, M_xy Sxy, ...),
// formed at each stencil point where the reference first fills its wrk
// array, and the stress step adds bare velocity derivatives.
#include <vector>

#include "rsg_common.h"
#include "rsg_fwi.h"

namespace rsg {

namespace {

struct K3E {
    int nx, ny, nz, lpml, fs, order;
    float dt, invd[3];
    const float *tab;
    float *V[3], *Sxx, *Syy, *Szz, *Syz, *Sxz, *Sxy;
    const float *L, *L2M, *Mxz, *Myz, *Mxy, *R[3];
    float *psi[RSG_E3D_NSTRIP];
    Coefs c[3];
};

// Position of a cell against the two strips of axis a: side 0 / 1 and
// profile index, or side -1.
__device__ __forceinline__ int side_of(int c, int n, int L, int &i) {
    i = 0;
    if (c < L) {
        i = c;
        return 0;
    }
    if (c >= n - L) {
        i = c - (n - L);
        return 1;
    }
    return -1;
}

struct Cell {
    int c[3];
    long p;
    long st[3];
    int side[3], ti[3];
    long q[3];  // strip index on each axis' face
};

__device__ __forceinline__ bool cell_of(const K3E &k, Cell &cl) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
    if (x >= k.nx || y >= k.ny) return false;
    cl.c[0] = x;
    cl.c[1] = y;
    cl.c[2] = z;
    cl.st[0] = 1;
    cl.st[1] = k.nx;
    cl.st[2] = (long)k.nx * k.ny;
    cl.p = z * cl.st[2] + (long)y * k.nx + x;
    const int n[3] = {k.nx, k.ny, k.nz};
    for (int a = 0; a < 3; a++) cl.side[a] = side_of(cl.c[a], n[a], k.lpml, cl.ti[a]);
    const int L = k.lpml;
    cl.q[0] = ((long)z * k.ny + y) * L + cl.ti[0];
    cl.q[1] = ((long)z * L + cl.ti[1]) * k.nx + x;
    cl.q[2] = ((long)cl.ti[2] * k.ny + y) * k.nx + x;
    return true;
}

// Derivative of f along axis a at the cell: forward (D+) or backward (D-).
template <int ORDER, bool FWD>
__device__ __forceinline__ float deriv(const K3E &k, const Cell &cl, const float *f, int a) {
    const int n = a == 0 ? k.nx : (a == 1 ? k.ny : k.nz);
    const int kk = FWD ? kfw(cl.c[a], n, ORDER) : kbw(cl.c[a], n, ORDER);
    const float *s = f + cl.p;
    const int st = (int)cl.st[a];
    if (kk == ORDER) return FWD ? dfw<ORDER>(s, st, k.c[a].c + cind(ORDER)) : dbw<ORDER>(s, st, k.c[a].c + cind(ORDER));
    return FWD ? dfw_k(s, st, kk, k.invd[a]) : dbw_k(s, st, kk, k.invd[a]);
}

// Moduli-weighted sum of N stress fields at a cell: the wrk array of
// backwardstepVelocity, m[0] f[0] + m[1] f[1] + m[2] f[2] in that order.
template <int N>
struct Mix {
    const float *f[N], *m[N];
    __device__ __forceinline__ float operator()(long q) const {
        float v = m[0][q] * f[0][q];
#pragma unroll
        for (int i = 1; i < N; i++) v += m[i][q] * f[i][q];
        return v;
    }
};

// deriv for a field given by a functor of the cell index.
template <int ORDER, bool FWD, class G>
__device__ __forceinline__ float derivm(const K3E &k, const Cell &cl, const G &g, int a) {
    const int n = a == 0 ? k.nx : (a == 1 ? k.ny : k.nz);
    const int kk = FWD ? kfw(cl.c[a], n, ORDER) : kbw(cl.c[a], n, ORDER);
    if (kk <= 0) return 0.f;
    const long p = cl.p, st = cl.st[a];
    if (kk == ORDER) {  // coefficients pre-divided by the step
        const float *c = k.c[a].c + cind(ORDER);
        float acc = 0.f;
#pragma unroll
        for (int io = 0; io < ORDER; io++)
            acc += FWD ? c[io] * (g(p + (io + 1) * st) - g(p - io * st)) : c[io] * (g(p + io * st) - g(p - (io + 1) * st));
        return acc;
    }
    const float *c = kCoefC + cind(kk);
    float acc = 0.f;
    for (int io = 0; io < kk; io++)
        acc += FWD ? c[io] * (g(p + (io + 1) * st) - g(p - io * st)) : c[io] * (g(p + io * st) - g(p - (io + 1) * st));
    return acc * k.invd[a];
}

// One derivative term into up to three targets sharing one strip.
__device__ __forceinline__ void term3(const K3E &k, const Cell &cl, int a, int stag, int slot, float d, float &f1,
                                      float m1, float *f2, float m2, float *f3, float m3) {
    const float dt = k.dt;
    f1 += dt * m1 * d;
    if (f2) *f2 += dt * m2 * d;
    if (f3) *f3 += dt * m3 * d;
    const int s = cl.side[a];
    if (s >= 0) {
        const int i = cl.ti[a];
        const float A = k.tab[((s ? T_RBB : 0) + stag * T_STAG + T_A) * k.lpml + i];
        const float B = k.tab[((s ? T_RBB : 0) + stag * T_STAG + T_B) * k.lpml + i];
        const float C = k.tab[((s ? T_RBB : 0) + stag * T_STAG + T_C) * k.lpml + i];
        float *psi = k.psi[slot + s] + cl.q[a];
        const float ps = B * (*psi) + A * d;
        *psi = ps;
        f1 -= dt * m1 * (ps + C * d);
        if (f2) *f2 -= dt * m2 * (ps + C * d);
        if (f3) *f3 -= dt * m3 * (ps + C * d);
    }
}

template <int ORDER>
__global__ __launch_bounds__(256) void e3d_velocity_kernel(const K3E k) {
    Cell cl;
    if (!cell_of(k, cl)) return;
    const long p = cl.p;
    const float rx = k.R[0][p], ry = k.R[1][p], rz = k.R[2][p];
    float vx = k.V[0][p], vy = k.V[1][p], vz = k.V[2][p];
    term3(k, cl, 0, 1, RSG_E3D_SXX_LEFT, deriv<ORDER, true>(k, cl, k.Sxx, 0), vx, rx, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 1, 0, RSG_E3D_SXYY_FRONT, deriv<ORDER, false>(k, cl, k.Sxy, 1), vx, rx, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 2, 0, RSG_E3D_SXZZ_TOP, deriv<ORDER, false>(k, cl, k.Sxz, 2), vx, rx, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 0, 0, RSG_E3D_SXYX_LEFT, deriv<ORDER, false>(k, cl, k.Sxy, 0), vy, ry, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 1, 1, RSG_E3D_SYY_FRONT, deriv<ORDER, true>(k, cl, k.Syy, 1), vy, ry, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 2, 0, RSG_E3D_SYZZ_TOP, deriv<ORDER, false>(k, cl, k.Syz, 2), vy, ry, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 0, 0, RSG_E3D_SXZX_LEFT, deriv<ORDER, false>(k, cl, k.Sxz, 0), vz, rz, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 1, 0, RSG_E3D_SYZY_FRONT, deriv<ORDER, false>(k, cl, k.Syz, 1), vz, rz, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 2, 1, RSG_E3D_SZZ_TOP, deriv<ORDER, true>(k, cl, k.Szz, 2), vz, rz, nullptr, 0.f, nullptr, 0.f);
    k.V[0][p] = vx;
    k.V[1][p] = vy;
    k.V[2][p] = vz;
}

// backwardstepVelocity (waves.cpp:3260-3713): the forward velocity step's
// terms and strips on the moduli-weighted stresses.
template <int ORDER>
__global__ __launch_bounds__(256) void e3d_velocity_bw_kernel(const K3E k) {
    Cell cl;
    if (!cell_of(k, cl)) return;
    const long p = cl.p;
    const float rx = k.R[0][p], ry = k.R[1][p], rz = k.R[2][p];
    float vx = k.V[0][p], vy = k.V[1][p], vz = k.V[2][p];
    const Mix<3> nx{{k.Sxx, k.Syy, k.Szz}, {k.L2M, k.L, k.L}};
    const Mix<3> ny{{k.Sxx, k.Syy, k.Szz}, {k.L, k.L2M, k.L}};
    const Mix<3> nz{{k.Sxx, k.Syy, k.Szz}, {k.L, k.L, k.L2M}};
    const Mix<1> sxy{{k.Sxy}, {k.Mxy}}, sxz{{k.Sxz}, {k.Mxz}}, syz{{k.Syz}, {k.Myz}};
    term3(k, cl, 0, 1, RSG_E3D_SXX_LEFT, derivm<ORDER, true>(k, cl, nx, 0), vx, rx, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 1, 0, RSG_E3D_SXYY_FRONT, derivm<ORDER, false>(k, cl, sxy, 1), vx, rx, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 2, 0, RSG_E3D_SXZZ_TOP, derivm<ORDER, false>(k, cl, sxz, 2), vx, rx, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 0, 0, RSG_E3D_SXYX_LEFT, derivm<ORDER, false>(k, cl, sxy, 0), vy, ry, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 1, 1, RSG_E3D_SYY_FRONT, derivm<ORDER, true>(k, cl, ny, 1), vy, ry, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 2, 0, RSG_E3D_SYZZ_TOP, derivm<ORDER, false>(k, cl, syz, 2), vy, ry, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 0, 0, RSG_E3D_SXZX_LEFT, derivm<ORDER, false>(k, cl, sxz, 0), vz, rz, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 1, 0, RSG_E3D_SYZY_FRONT, derivm<ORDER, false>(k, cl, syz, 1), vz, rz, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 2, 1, RSG_E3D_SZZ_TOP, derivm<ORDER, true>(k, cl, nz, 2), vz, rz, nullptr, 0.f, nullptr, 0.f);
    k.V[0][p] = vx;
    k.V[1][p] = vy;
    k.V[2][p] = vz;
}

// The moduli-weighted stresses of backwardstepVelocity as six arrays: the
// three normal combinations for Dx+, Dy+, Dz+ and M_xy Sxy, M_xz Sxz,
// M_yz Syz, formed as e3d_velocity_bw_kernel forms them per stencil point.
// The forward velocity kernel on these arrays is the backward velocity step.
__global__ __launch_bounds__(256) void e3d_premul_kernel(const K3E k, float *w0, float *w1, float *w2, float *w3,
                                                         float *w4, float *w5, size_t n) {
    const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const Mix<3> nx{{k.Sxx, k.Syy, k.Szz}, {k.L2M, k.L, k.L}};
    const Mix<3> ny{{k.Sxx, k.Syy, k.Szz}, {k.L, k.L2M, k.L}};
    const Mix<3> nz{{k.Sxx, k.Syy, k.Szz}, {k.L, k.L, k.L2M}};
    const Mix<1> sxy{{k.Sxy}, {k.Mxy}}, sxz{{k.Sxz}, {k.Mxz}}, syz{{k.Syz}, {k.Myz}};
    w0[p] = nx((long)p);
    w1[p] = ny((long)p);
    w2[p] = nz((long)p);
    w3[p] = sxy((long)p);
    w4[p] = sxz((long)p);
    w5[p] = syz((long)p);
}

// forwardstepStress, or with BW backwardstepStress (waves.cpp:3715-4103):
// the same terms and strips with unit moduli.
template <int ORDER, bool BW>
__global__ __launch_bounds__(256) void e3d_stress_kernel(const K3E k) {
    Cell cl;
    if (!cell_of(k, cl)) return;
    const long p = cl.p;
    float sxx = k.Sxx[p], syy = k.Syy[p], szz = k.Szz[p];
    if (BW) {
        term3(k, cl, 0, 0, RSG_E3D_VXX_LEFT, deriv<ORDER, false>(k, cl, k.V[0], 0), sxx, 1.f, nullptr, 0.f, nullptr, 0.f);
        term3(k, cl, 1, 0, RSG_E3D_VYY_FRONT, deriv<ORDER, false>(k, cl, k.V[1], 1), syy, 1.f, nullptr, 0.f, nullptr,
              0.f);
        term3(k, cl, 2, 0, RSG_E3D_VZZ_TOP, deriv<ORDER, false>(k, cl, k.V[2], 2), szz, 1.f, nullptr, 0.f, nullptr, 0.f);
    } else {
        const float l = k.L[p], l2m = k.L2M[p];
        term3(k, cl, 0, 0, RSG_E3D_VXX_LEFT, deriv<ORDER, false>(k, cl, k.V[0], 0), sxx, l2m, &syy, l, &szz, l);
        term3(k, cl, 1, 0, RSG_E3D_VYY_FRONT, deriv<ORDER, false>(k, cl, k.V[1], 1), sxx, l, &syy, l2m, &szz, l);
        term3(k, cl, 2, 0, RSG_E3D_VZZ_TOP, deriv<ORDER, false>(k, cl, k.V[2], 2), sxx, l, &syy, l, &szz, l2m);
    }
    // free surface: the plane lpml lies in no z strip, so zeroing after the
    // strips is the reference's zeroing before them
    if (k.fs && cl.c[2] == k.lpml) szz = 0.f;
    k.Sxx[p] = sxx;
    k.Syy[p] = syy;
    k.Szz[p] = szz;
    const float mxz = BW ? 1.f : k.Mxz[p], myz = BW ? 1.f : k.Myz[p], mxy = BW ? 1.f : k.Mxy[p];
    float sxz = k.Sxz[p], syz = k.Syz[p], sxy = k.Sxy[p];
    term3(k, cl, 0, 1, RSG_E3D_VZX_LEFT, deriv<ORDER, true>(k, cl, k.V[2], 0), sxz, mxz, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 2, 1, RSG_E3D_VXZ_TOP, deriv<ORDER, true>(k, cl, k.V[0], 2), sxz, mxz, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 2, 1, RSG_E3D_VYZ_TOP, deriv<ORDER, true>(k, cl, k.V[1], 2), syz, myz, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 1, 1, RSG_E3D_VZY_FRONT, deriv<ORDER, true>(k, cl, k.V[2], 1), syz, myz, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 1, 1, RSG_E3D_VXY_FRONT, deriv<ORDER, true>(k, cl, k.V[0], 1), sxy, mxy, nullptr, 0.f, nullptr, 0.f);
    term3(k, cl, 0, 1, RSG_E3D_VYX_LEFT, deriv<ORDER, true>(k, cl, k.V[1], 0), sxy, mxy, nullptr, 0.f, nullptr, 0.f);
    k.Sxz[p] = sxz;
    k.Syz[p] = syz;
    k.Sxy[p] = sxy;
}

// insertSource (waves.cpp:4104-4246): PRESSURE is explosive (Sxx, Syy, Szz
// -= dt wav w), VX / VY / VZ add dt R wav w; trilinear over the 8 corners
// (atomic: traces may share cells).
__global__ void e3d_inject_kernel(float *F1, float *F2, float *F3, const float *Mod, float scale, int nx, int ny,
                                  int nz, int pad, int ntrace, const int *map, const float *shift, const float *wav,
                                  int nt, int it) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntrace) return;
    const int *m = map + 3 * (size_t)i;
    if (m[0] < 0 || m[1] < 0 || m[2] < 0) return;
    const float *sh = shift + 3 * (size_t)i;
    const float xt[2] = {1 - sh[0], sh[0]}, yt[2] = {1 - sh[1], sh[1]}, zt[2] = {1 - sh[2], sh[2]};
    const float w = wav[(size_t)i * nt + it];
    const size_t ncell = (size_t)nx * ny * nz;
    for (int i1 = 0; i1 < 2; i1++)
        for (int i2 = 0; i2 < 2; i2++)
            for (int i3 = 0; i3 < 2; i3++) {
                const size_t p = ((size_t)(pad + m[2] + i1) * ny + (pad + m[1] + i2)) * nx + (pad + m[0] + i3);
                if (p >= ncell) continue;
                const float v = (Mod ? scale * Mod[p] : scale) * w * xt[i3] * yt[i2] * zt[i1];
                unsafeAtomicAdd(F1 + p, v);
                if (F2) unsafeAtomicAdd(F2 + p, v);
                if (F3) unsafeAtomicAdd(F3 + p, v);
            }
}

// recordData (waves.cpp:4247-4403): velocity gathers, or the pressure
// gather (Sxx + Syy + Szz) / 3 divided by (3 rho vp^2 - 4 rho vs^2) / 3 at
// the trace's model cell, in double as the reference forms it; one thread
// per trace, corners in the reference's order.
__global__ void e3d_record_kernel(const float *F1, const float *F2, const float *F3, const float *Vp, const float *Vs,
                                  const float *Rho, int mnx, int mny, int nx, int ny, int nz, int pad, int ntrace,
                                  const int *map, const float *shift, float *data, int nt, int it) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntrace) return;
    const int *m = map + 3 * (size_t)i;
    if (m[0] < 0 || m[1] < 0 || m[2] < 0) return;
    const float *sh = shift + 3 * (size_t)i;
    const float xt[2] = {1 - sh[0], sh[0]}, yt[2] = {1 - sh[1], sh[1]}, zt[2] = {1 - sh[2], sh[2]};
    const size_t ncell = (size_t)nx * ny * nz;
    float *d = data + (size_t)i * nt + it;
    float acc = *d;
    if (F2) {
        const size_t c = ((size_t)m[2] * mny + m[1]) * mnx + m[0];
        const double r = Rho[c], vp = Vp[c], vs = Vs[c];
        const float factor = (float)((3. * r * vp * vp - 4. * r * vs * vs) / 3.);
        for (int i1 = 0; i1 < 2; i1++)
            for (int i2 = 0; i2 < 2; i2++)
                for (int i3 = 0; i3 < 2; i3++) {
                    const size_t p = ((size_t)(pad + m[2] + i1) * ny + (pad + m[1] + i2)) * nx + (pad + m[0] + i3);
                    if (p >= ncell) continue;
                    const double w = (1. / 3.) * (double)xt[i3] * (double)yt[i2] * (double)zt[i1];
                    acc = (float)((double)acc + w * (double)F1[p] / (double)factor);
                    acc = (float)((double)acc + w * (double)F2[p] / (double)factor);
                    acc = (float)((double)acc + w * (double)F3[p] / (double)factor);
                }
    } else {
        for (int i1 = 0; i1 < 2; i1++)
            for (int i2 = 0; i2 < 2; i2++)
                for (int i3 = 0; i3 < 2; i3++) {
                    const size_t p = ((size_t)(pad + m[2] + i1) * ny + (pad + m[1] + i2)) * nx + (pad + m[0] + i3);
                    if (p >= ncell) continue;
                    acc += xt[i3] * yt[i2] * zt[i1] * F1[p];
                }
    }
    *d = acc;
}

int check_e3d(const rsg_grid *g, const rsg_state_e3d *s) {
    if (!g || !s) return fail(RSG_EINVAL, "e3d: null argument");
    if (g->nx < 2 || g->ny < 2 || g->nz < 2 || g->nz > 65535 || g->lpml < 1 || 2 * g->lpml >= g->nx ||
        2 * g->lpml >= g->ny || 2 * g->lpml >= g->nz || g->dx <= 0 || g->dy <= 0 || g->dz <= 0)
        return fail(RSG_EINVAL, "e3d: grid sizes (padded, 2 lpml < n, nz <= 65535) or steps");
    if (!s->Vx || !s->Vy || !s->Vz || !s->Sxx || !s->Syy || !s->Szz || !s->Syz || !s->Sxz || !s->Sxy || !s->L ||
        !s->L2M || !s->Mxz || !s->Myz || !s->Mxy || !s->Rx || !s->Ry || !s->Rz || !s->tab)
        return fail(RSG_EINVAL, "e3d: null field, model or table");
    for (int i = 0; i < RSG_E3D_NSTRIP; i++)
        if (!s->psi[i]) return fail(RSG_EINVAL, "e3d: null PML strip");
    return RSG_OK;
}

K3E make_k3e(const rsg_grid *g, const rsg_state_e3d *s) {
    K3E k;
    k.nx = g->nx;
    k.ny = g->ny;
    k.nz = g->nz;
    k.lpml = g->lpml;
    k.fs = g->fs;
    k.order = g->order < 1 ? 1 : (g->order > 8 ? 8 : g->order);
    k.dt = g->dt;
    const float d[3] = {g->dx, g->dy, g->dz};
    for (int a = 0; a < 3; a++) {
        k.invd[a] = 1.0f / d[a];
        k.c[a] = scaled_coefs(d[a]);
    }
    k.tab = s->tab;
    k.V[0] = s->Vx;
    k.V[1] = s->Vy;
    k.V[2] = s->Vz;
    k.Sxx = s->Sxx;
    k.Syy = s->Syy;
    k.Szz = s->Szz;
    k.Syz = s->Syz;
    k.Sxz = s->Sxz;
    k.Sxy = s->Sxy;
    k.L = s->L;
    k.L2M = s->L2M;
    k.Mxz = s->Mxz;
    k.Myz = s->Myz;
    k.Mxy = s->Mxy;
    k.R[0] = s->Rx;
    k.R[1] = s->Ry;
    k.R[2] = s->Rz;
    for (int i = 0; i < RSG_E3D_NSTRIP; i++) k.psi[i] = s->psi[i];
    return k;
}

#define E3D_ORDER_CASES(LAUNCH) \
    switch (k.order) {          \
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

}  // namespace rsg

namespace rsg {

// backwardstepVelocity with a scratch of six padded arrays: the weighted
// stresses once per cell, then the forward velocity kernel on them.  The
// same values as rsg_e3d_backward_velocity, which forms them per stencil
// point; this path reads 6 instead of up to 96 values per cell's taps.
int e3d_backward_velocity_premul(const rsg_grid *g, const rsg_state_e3d *s, float *scratch, hipStream_t st) {
    if (int rc = check_e3d(g, s)) return rc;
    if (!scratch) return fail(RSG_EINVAL, "e3d: null scratch");
    K3E k = make_k3e(g, s);
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    float *w[6];
    for (int i = 0; i < 6; i++) w[i] = scratch + i * n;
    e3d_premul_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(k, w[0], w[1], w[2], w[3], w[4], w[5], n);
    RSG_LAUNCH_CHECK("e3d_premul_kernel");
    k.Sxx = w[0];
    k.Syy = w[1];
    k.Szz = w[2];
    k.Sxy = w[3];
    k.Sxz = w[4];
    k.Syz = w[5];
    const dim3 grid((g->nx + 63) / 64, (g->ny + 3) / 4, g->nz), block(64, 4);
#define L_(O) e3d_velocity_kernel<O><<<grid, block, 0, st>>>(k)
    E3D_ORDER_CASES(L_);
#undef L_
    RSG_LAUNCH_CHECK("e3d_velocity_kernel");
    return RSG_OK;
}

}  // namespace rsg

extern "C" {

int rsg_e3d_forward_velocity(const rsg_grid *g, const rsg_state_e3d *s, void *stream) {
    if (int rc = rsg::check_e3d(g, s)) return rc;
    const rsg::K3E k = rsg::make_k3e(g, s);
    const dim3 grid((g->nx + 63) / 64, (g->ny + 3) / 4, g->nz), block(64, 4);
#define L_(O) rsg::e3d_velocity_kernel<O><<<grid, block, 0, (hipStream_t)stream>>>(k)
    E3D_ORDER_CASES(L_);
#undef L_
    RSG_LAUNCH_CHECK("e3d_velocity_kernel");
    return RSG_OK;
}

int rsg_e3d_forward_stress(const rsg_grid *g, const rsg_state_e3d *s, void *stream) {
    if (int rc = rsg::check_e3d(g, s)) return rc;
    const rsg::K3E k = rsg::make_k3e(g, s);
    const dim3 grid((g->nx + 63) / 64, (g->ny + 3) / 4, g->nz), block(64, 4);
#define L_(O) rsg::e3d_stress_kernel<O, false><<<grid, block, 0, (hipStream_t)stream>>>(k)
    E3D_ORDER_CASES(L_);
#undef L_
    RSG_LAUNCH_CHECK("e3d_stress_kernel");
    return RSG_OK;
}

int rsg_e3d_backward_velocity(const rsg_grid *g, const rsg_state_e3d *s, void *stream) {
    if (int rc = rsg::check_e3d(g, s)) return rc;
    const rsg::K3E k = rsg::make_k3e(g, s);
    const dim3 grid((g->nx + 63) / 64, (g->ny + 3) / 4, g->nz), block(64, 4);
#define L_(O) rsg::e3d_velocity_bw_kernel<O><<<grid, block, 0, (hipStream_t)stream>>>(k)
    E3D_ORDER_CASES(L_);
#undef L_
    RSG_LAUNCH_CHECK("e3d_velocity_bw_kernel");
    return RSG_OK;
}

int rsg_e3d_backward_velocity_ws(const rsg_grid *g, const rsg_state_e3d *s, float *workspace, void *stream) {
    return rsg::e3d_backward_velocity_premul(g, s, workspace, (hipStream_t)stream);
}

int rsg_e3d_backward_stress(const rsg_grid *g, const rsg_state_e3d *s, void *stream) {
    if (int rc = rsg::check_e3d(g, s)) return rc;
    const rsg::K3E k = rsg::make_k3e(g, s);
    const dim3 grid((g->nx + 63) / 64, (g->ny + 3) / 4, g->nz), block(64, 4);
#define L_(O) rsg::e3d_stress_kernel<O, true><<<grid, block, 0, (hipStream_t)stream>>>(k)
    E3D_ORDER_CASES(L_);
#undef L_
    RSG_LAUNCH_CHECK("e3d_stress_kernel");
    return RSG_OK;
}

int rsg_e3d_insert_source(const rsg_grid *g, const rsg_state_e3d *s, const rsg_traces *src, int it, void *stream) {
    if (int rc = rsg::check_e3d(g, s)) return rc;
    if (!src || src->ntrace < 0 || (src->ntrace > 0 && (!src->map || !src->shift || !src->data)) || it < 0 ||
        it >= src->nt)
        return rsg::fail(RSG_EINVAL, "rsg_e3d_insert_source: bad traces or time index");
    if (src->ntrace == 0) return RSG_OK;
    float *F1, *F2 = nullptr, *F3 = nullptr;
    const float *Mod = nullptr;
    float scale = g->dt;
    switch (src->field) {
    case RSG_PRESSURE: F1 = s->Sxx; F2 = s->Syy; F3 = s->Szz; scale = -g->dt; break;
    case RSG_VX: F1 = s->Vx; Mod = s->Rx; break;
    case RSG_VY: F1 = s->Vy; Mod = s->Ry; break;
    case RSG_VZ: F1 = s->Vz; Mod = s->Rz; break;
    default: return rsg::fail(RSG_EINVAL, "rsg_e3d_insert_source: field must be P, VX, VY or VZ");
    }
    rsg::e3d_inject_kernel<<<(src->ntrace + 63) / 64, 64, 0, (hipStream_t)stream>>>(
        F1, F2, F3, Mod, scale, g->nx, g->ny, g->nz, rsg::map_pad(g), src->ntrace, src->map, src->shift, src->data, src->nt,
        it);
    RSG_LAUNCH_CHECK("e3d_inject_kernel");
    return RSG_OK;
}

int rsg_e3d_record_data(const rsg_grid *g, const rsg_state_e3d *s, const rsg_traces *rec, const float *Vp,
                        const float *Vs, const float *Rho, int it, void *stream) {
    if (int rc = rsg::check_e3d(g, s)) return rc;
    if (!rec || rec->ntrace < 0 || (rec->ntrace > 0 && (!rec->map || !rec->shift || !rec->data)) || it < 0 ||
        it >= rec->nt)
        return rsg::fail(RSG_EINVAL, "rsg_e3d_record_data: bad traces or time index");
    if (rec->ntrace == 0) return RSG_OK;
    const float *F1, *F2 = nullptr, *F3 = nullptr;
    switch (rec->field) {
    case RSG_PRESSURE:
        if (!Vp || !Vs || !Rho) return rsg::fail(RSG_EINVAL, "rsg_e3d_record_data: pressure needs Vp, Vs, Rho");
        F1 = s->Sxx; F2 = s->Syy; F3 = s->Szz;
        break;
    case RSG_VX: F1 = s->Vx; break;
    case RSG_VY: F1 = s->Vy; break;
    case RSG_VZ: F1 = s->Vz; break;
    default: return rsg::fail(RSG_EINVAL, "rsg_e3d_record_data: field must be P, VX, VY or VZ");
    }
    const int pad = rsg::map_pad(g);
    rsg::e3d_record_kernel<<<(rec->ntrace + 63) / 64, 64, 0, (hipStream_t)stream>>>(
        F1, F2, F3, Vp, Vs, Rho, g->nx - 2 * pad, g->ny - 2 * pad, g->nx, g->ny, g->nz, pad, rec->ntrace, rec->map,
        rec->shift, rec->data, rec->nt, it);
    RSG_LAUNCH_CHECK("e3d_record_kernel");
    return RSG_OK;
}

}  // extern "C"
