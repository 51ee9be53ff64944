// 2D staggered-grid velocity-stress kernels for gfx950: acoustic
// (WavesAcoustic2D, waves.cpp:217-404) and elastic forward and adjoint-state
// steps (WavesElastic2D, waves.cpp:1416-2080).
//
// One fused kernel per half step.  The reference makes, per derivative, a
// full-grid Der pass into a scratch array, a full-grid update pass and a PML
// strip pass (der.cpp, waves.cpp); here each block stages the stencil inputs
// of a TX x TZ tile (plus halo) in LDS once, evaluates every derivative of the
// half step from LDS, and applies the field update and the CFS-PML memory
// update of every cell in registers, so each array crosses HBM once per half
// step.  The per-cell operation sequence is the reference's: += dt*M*d for
// the first derivative, its PML correction, then the next derivative.
//
// Block: 64 x 4 threads (four wave64 rows); each thread owns RZ consecutive
// cells of one x column, so z-derivative taps are reused from registers and
// x-derivative taps are conflict-free LDS row reads.  Cells whose stencil
// reaches the array edge use the reference's reduced-order coefficient sets.
#include "rsg_common.h"

namespace rsg {

constexpr int BX = 64;
constexpr int BZ = 4;
constexpr int RZ2D = 8;

struct K2D {
    int nx, nz, lpml, fs;
    float dt, invdx, invdz;
    const float *tab;
    float *P, *Vx, *Vz, *Sxx, *Szz, *Sxz;
    const float *L, *L2M, *M, *Rx, *Rz;
    float *psi[RSG_MAX_STRIPS];
    Coefs cx, cz;
};

// Stage rows [z0-HZ, z0+TZ+HZ) x cols [x0-HX, x0+TX+HX) of g into LDS,
// zero outside the grid (those entries are never read by the edge stencils).
template <int TX, int TZ, int HX, int HZ>
__device__ __forceinline__ void load_tile(float *lds, const float *__restrict__ g, int nx, int nz,
                                          int x0, int z0) {
    constexpr int LD = TX + 2 * HX, N = LD * (TZ + 2 * HZ);
    for (int i = threadIdx.x + threadIdx.y * BX; i < N; i += BX * BZ) {
        const int r = i / LD, c = i - r * LD;
        const int z = z0 - HZ + r, x = x0 - HX + c;
        float v = 0.f;
        if (x >= 0 && x < nx && z >= 0 && z < nz) v = g[(size_t)z * nx + x];
        lds[i] = v;
    }
}

// Same, staging a*b + c*d (c may be null: a*b) -- the moduli-weighted
// stresses of WavesElastic2D::backwardstepVelocity (waves.cpp:1771-1775).
template <int TX, int TZ, int HX, int HZ>
__device__ __forceinline__ void load_tile_mix(float *lds, const float *__restrict__ a,
                                              const float *__restrict__ b,
                                              const float *__restrict__ c,
                                              const float *__restrict__ d, int nx, int nz, int x0,
                                              int z0) {
    constexpr int LD = TX + 2 * HX, N = LD * (TZ + 2 * HZ);
    for (int i = threadIdx.x + threadIdx.y * BX; i < N; i += BX * BZ) {
        const int r = i / LD, cc = i - r * LD;
        const int z = z0 - HZ + r, x = x0 - HX + cc;
        float v = 0.f;
        if (x >= 0 && x < nx && z >= 0 && z < nz) {
            const size_t p = (size_t)z * nx + x;
            v = c ? a[p] * b[p] + c[p] * d[p] : a[p] * b[p];
        }
        lds[i] = v;
    }
}

// Derivative of an LDS tile at a cell: compile-time order in the interior,
// reduced order where the reference's Der shortens the stencil.
template <int ORDER>
__device__ __forceinline__ float ddfw(const float *s, int st, const float *c, int k, float invd) {
    float d = dfw<ORDER>(s, st, c + cind(ORDER));
    if (k != ORDER) d = dfw_k(s, st, k, invd);
    return d;
}
template <int ORDER>
__device__ __forceinline__ float ddbw(const float *s, int st, const float *c, int k, float invd) {
    float d = dbw<ORDER>(s, st, c + cind(ORDER));
    if (k != ORDER) d = dbw_k(s, st, k, invd);
    return d;
}

// PML table accessors: stag selects the staggered set, right the rbb side.
struct Tab {
    const float *t;
    int L;
    __device__ __forceinline__ float A(int right, int stag, int i) const {
        return t[(right * T_RBB + stag * T_STAG + T_A) * L + i];
    }
    __device__ __forceinline__ float B(int right, int stag, int i) const {
        return t[(right * T_RBB + stag * T_STAG + T_B) * L + i];
    }
    __device__ __forceinline__ float C(int right, int stag, int i) const {
        return t[(right * T_RBB + stag * T_STAG + T_C) * L + i];
    }
};

// Position of a cell relative to the x-side and z-side PML strips.
struct PmlPos {
    int side;  // -1 none, 0 low (left/top), 1 high (right/bottom)
    int i;     // index into the strip profile (table index)
};
__device__ __forceinline__ PmlPos pml_pos(int x, int n, int L) {
    PmlPos p{-1, 0};
    if (x < L) {
        p.side = 0;
        p.i = x;
    } else if (x >= n - L) {
        p.side = 1;
        p.i = x - (n - L);
    }
    return p;
}

// Apply dt*M*d then the strip correction of one derivative term.
// xs: strip is an x-side strip (index z*L+i), else a z-side strip
// (index i*nx+x).  lo/hi: the two memory arrays of that term.
__device__ __forceinline__ void term(float &F, float dtM, float d, const PmlPos &pp, const Tab &tb,
                                     int stag, float *lo, float *hi, size_t sidx) {
    F += dtM * d;
    if (pp.side >= 0) {
        float *psi = (pp.side ? hi : lo) + sidx;
        pml_corr(F, dtM, psi, tb.A(pp.side, stag, pp.i), tb.B(pp.side, stag, pp.i),
                 tb.C(pp.side, stag, pp.i), d);
    }
}

// ---------------------------------------------------------------------------
// Acoustic 2D
// ---------------------------------------------------------------------------

template <int ORDER, int RZ>
__global__ __launch_bounds__(BX *BZ) void a2d_velocity_kernel(const K2D a) {
    constexpr int H = ORDER, TZ = BZ * RZ, LD = BX + 2 * H;
    __shared__ float sP[(TZ + 2 * H) * LD];
    const int x0 = blockIdx.x * BX, z0 = blockIdx.y * TZ;
    load_tile<BX, TZ, H, H>(sP, a.P, a.nx, a.nz, x0, z0);
    __syncthreads();
    const int tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    if (x >= a.nx) return;
    const Tab tb{a.tab, a.lpml};
    const PmlPos px = pml_pos(x, a.nx, a.lpml);
    const int kx = kfw(x, a.nx, ORDER);
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int zl = ty * RZ + r, z = z0 + zl;
        if (z >= a.nz) break;
        const float *s = sP + (zl + H) * LD + tx + H;
        const float dPx = ddfw<ORDER>(s, 1, a.cx.c, kx, a.invdx);
        const float dPz = ddfw<ORDER>(s, LD, a.cz.c, kfw(z, a.nz, ORDER), a.invdz);
        const size_t p = (size_t)z * a.nx + x;
        const PmlPos pz = pml_pos(z, a.nz, a.lpml);
        float vx = a.Vx[p];
        term(vx, a.dt * a.Rx[p], dPx, px, tb, 1, a.psi[RSG_A2D_P_LEFT], a.psi[RSG_A2D_P_RIGHT],
             (size_t)z * a.lpml + px.i);
        a.Vx[p] = vx;
        float vz = a.Vz[p];
        term(vz, a.dt * a.Rz[p], dPz, pz, tb, 1, a.psi[RSG_A2D_P_TOP], a.psi[RSG_A2D_P_BOTTOM],
             (size_t)pz.i * a.nx + x);
        a.Vz[p] = vz;
    }
}

template <int ORDER, int RZ>
__global__ __launch_bounds__(BX *BZ) void a2d_stress_kernel(const K2D a) {
    constexpr int H = ORDER, TZ = BZ * RZ, LDX = BX + 2 * H, LDZ = BX;
    __shared__ float sVx[TZ * LDX];
    __shared__ float sVz[(TZ + 2 * H) * LDZ];
    const int x0 = blockIdx.x * BX, z0 = blockIdx.y * TZ;
    load_tile<BX, TZ, H, 0>(sVx, a.Vx, a.nx, a.nz, x0, z0);
    load_tile<BX, TZ, 0, H>(sVz, a.Vz, a.nx, a.nz, x0, z0);
    __syncthreads();
    const int tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    if (x >= a.nx) return;
    const Tab tb{a.tab, a.lpml};
    const PmlPos px = pml_pos(x, a.nx, a.lpml);
    const int kx = kbw(x, a.nx, ORDER);
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int zl = ty * RZ + r, z = z0 + zl;
        if (z >= a.nz) break;
        const float dVx = ddbw<ORDER>(sVx + zl * LDX + tx + H, 1, a.cx.c, kx, a.invdx);
        const float dVz = ddbw<ORDER>(sVz + (zl + H) * LDZ + tx, LDZ, a.cz.c, kbw(z, a.nz, ORDER),
                                      a.invdz);
        const size_t p = (size_t)z * a.nx + x;
        const PmlPos pz = pml_pos(z, a.nz, a.lpml);
        const float dtL = a.dt * a.L[p];
        float P = a.P[p];
        term(P, dtL, dVx, px, tb, 0, a.psi[RSG_A2D_VXX_LEFT], a.psi[RSG_A2D_VXX_RIGHT],
             (size_t)z * a.lpml + px.i);
        term(P, dtL, dVz, pz, tb, 0, a.psi[RSG_A2D_VZZ_TOP], a.psi[RSG_A2D_VZZ_BOTTOM],
             (size_t)pz.i * a.nx + x);
        a.P[p] = P;
    }
}

// ---------------------------------------------------------------------------
// Elastic 2D
// ---------------------------------------------------------------------------

// Velocity update shared by the forward (BW=false, plain stresses) and
// adjoint-state (BW=true, moduli-weighted stresses) steps:
//   Vx += dt*Rx*(ddx_fw(Wxx) + ddz_bw(Wxz)),  Vz += dt*Rz*(ddx_bw(Wxz) + ddz_fw(Wzz))
// with Wxx=Sxx, Wxz=Sxz, Wzz=Szz (forward) or Wxx=L2M*Sxx+L*Szz, Wxz=M*Sxz,
// Wzz=L*Sxx+L2M*Szz (adjoint state, waves.cpp:1771-1878).
template <int ORDER, int RZ, bool BW>
__global__ __launch_bounds__(BX *BZ) void e2d_velocity_kernel(const K2D a) {
    constexpr int H = ORDER, TZ = BZ * RZ;
    constexpr int LDXX = BX + 2 * H, LDZZ = BX, LDXZ = BX + 2 * H;
    __shared__ float sXX[TZ * LDXX];
    __shared__ float sZZ[(TZ + 2 * H) * LDZZ];
    __shared__ float sXZ[(TZ + 2 * H) * LDXZ];
    const int x0 = blockIdx.x * BX, z0 = blockIdx.y * TZ;
    if (!BW) {
        load_tile<BX, TZ, H, 0>(sXX, a.Sxx, a.nx, a.nz, x0, z0);
        load_tile<BX, TZ, 0, H>(sZZ, a.Szz, a.nx, a.nz, x0, z0);
        load_tile<BX, TZ, H, H>(sXZ, a.Sxz, a.nx, a.nz, x0, z0);
    } else {
        load_tile_mix<BX, TZ, H, 0>(sXX, a.L2M, a.Sxx, a.L, a.Szz, a.nx, a.nz, x0, z0);
        load_tile_mix<BX, TZ, 0, H>(sZZ, a.L, a.Sxx, a.L2M, a.Szz, a.nx, a.nz, x0, z0);
        load_tile_mix<BX, TZ, H, H>(sXZ, a.M, a.Sxz, nullptr, nullptr, a.nx, a.nz, x0, z0);
    }
    __syncthreads();
    const int tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    if (x >= a.nx) return;
    const Tab tb{a.tab, a.lpml};
    const PmlPos px = pml_pos(x, a.nx, a.lpml);
    const int kxf = kfw(x, a.nx, ORDER), kxb = kbw(x, a.nx, ORDER);
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int zl = ty * RZ + r, z = z0 + zl;
        if (z >= a.nz) break;
        const float dSxx = ddfw<ORDER>(sXX + zl * LDXX + tx + H, 1, a.cx.c, kxf, a.invdx);
        const float *sxz = sXZ + (zl + H) * LDXZ + tx + H;
        const float dSxz_z = ddbw<ORDER>(sxz, LDXZ, a.cz.c, kbw(z, a.nz, ORDER), a.invdz);
        const float dSxz_x = ddbw<ORDER>(sxz, 1, a.cx.c, kxb, a.invdx);
        const float dSzz = ddfw<ORDER>(sZZ + (zl + H) * LDZZ + tx, LDZZ, a.cz.c,
                                       kfw(z, a.nz, ORDER), a.invdz);
        const size_t p = (size_t)z * a.nx + x;
        const PmlPos pz = pml_pos(z, a.nz, a.lpml);
        const size_t ix = (size_t)z * a.lpml + px.i, iz = (size_t)pz.i * a.nx + x;
        const float dtRx = a.dt * a.Rx[p];
        float vx = a.Vx[p];
        term(vx, dtRx, dSxx, px, tb, 1, a.psi[RSG_E2D_SXX_LEFT], a.psi[RSG_E2D_SXX_RIGHT], ix);
        term(vx, dtRx, dSxz_z, pz, tb, 0, a.psi[RSG_E2D_SXZZ_TOP], a.psi[RSG_E2D_SXZZ_BOTTOM], iz);
        a.Vx[p] = vx;
        const float dtRz = a.dt * a.Rz[p];
        float vz = a.Vz[p];
        term(vz, dtRz, dSxz_x, px, tb, 0, a.psi[RSG_E2D_SXZX_LEFT], a.psi[RSG_E2D_SXZX_RIGHT], ix);
        term(vz, dtRz, dSzz, pz, tb, 1, a.psi[RSG_E2D_SZZ_TOP], a.psi[RSG_E2D_SZZ_BOTTOM], iz);
        a.Vz[p] = vz;
    }
}

// Stress update.  Forward (BW=false, waves.cpp:1571-1735):
//   Sxx += dt*(L2M ddx_bw(Vx) + L ddz_bw(Vz)), Szz += dt*(L ddx_bw(Vx) + L2M ddz_bw(Vz)),
//   free surface Szz(z=lpml)=0, Sxz += dt*M*(ddx_fw(Vz) + ddz_fw(Vx)).
// Adjoint state (BW=true, waves.cpp:1907-2080): the same derivatives with
// unit moduli, Sxx only from ddx_bw(Vx) and Szz only from ddz_bw(Vz).
template <int ORDER, int RZ, bool BW>
__global__ __launch_bounds__(BX *BZ) void e2d_stress_kernel(const K2D a) {
    constexpr int H = ORDER, TZ = BZ * RZ, LD = BX + 2 * H;
    __shared__ float sVx[(TZ + 2 * H) * LD];
    __shared__ float sVz[(TZ + 2 * H) * LD];
    const int x0 = blockIdx.x * BX, z0 = blockIdx.y * TZ;
    load_tile<BX, TZ, H, H>(sVx, a.Vx, a.nx, a.nz, x0, z0);
    load_tile<BX, TZ, H, H>(sVz, a.Vz, a.nx, a.nz, x0, z0);
    __syncthreads();
    const int tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    if (x >= a.nx) return;
    const Tab tb{a.tab, a.lpml};
    const PmlPos px = pml_pos(x, a.nx, a.lpml);
    const int kxf = kfw(x, a.nx, ORDER), kxb = kbw(x, a.nx, ORDER);
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int zl = ty * RZ + r, z = z0 + zl;
        if (z >= a.nz) break;
        const float *svx = sVx + (zl + H) * LD + tx + H;
        const float *svz = sVz + (zl + H) * LD + tx + H;
        const int kzf = kfw(z, a.nz, ORDER), kzb = kbw(z, a.nz, ORDER);
        const float dVx_x = ddbw<ORDER>(svx, 1, a.cx.c, kxb, a.invdx);
        const float dVz_z = ddbw<ORDER>(svz, LD, a.cz.c, kzb, a.invdz);
        const float dVz_x = ddfw<ORDER>(svz, 1, a.cx.c, kxf, a.invdx);
        const float dVx_z = ddfw<ORDER>(svx, LD, a.cz.c, kzf, a.invdz);
        const size_t p = (size_t)z * a.nx + x;
        const PmlPos pz = pml_pos(z, a.nz, a.lpml);
        const size_t ix = (size_t)z * a.lpml + px.i, iz = (size_t)pz.i * a.nx + x;
        float sxx = a.Sxx[p], szz = a.Szz[p], sxz = a.Sxz[p];
        const float dt = a.dt;
        if (!BW) {
            const float dtL2M = dt * a.L2M[p], dtL = dt * a.L[p];
            // d/dx Vx feeds Sxx (L2M) and Szz (L) through one memory variable
            sxx += dtL2M * dVx_x;
            szz += dtL * dVx_x;
            if (px.side >= 0) {
                float *psi = a.psi[px.side ? RSG_E2D_VXX_RIGHT : RSG_E2D_VXX_LEFT] + ix;
                const float A = tb.A(px.side, 0, px.i), B = tb.B(px.side, 0, px.i),
                            C = tb.C(px.side, 0, px.i);
                const float q = B * (*psi) + A * dVx_x;
                *psi = q;
                sxx -= dtL2M * (q + C * dVx_x);
                szz -= dtL * (q + C * dVx_x);
            }
            sxx += dtL * dVz_z;
            szz += dtL2M * dVz_z;
            if (pz.side >= 0) {
                float *psi = a.psi[pz.side ? RSG_E2D_VZZ_BOTTOM : RSG_E2D_VZZ_TOP] + iz;
                const float A = tb.A(pz.side, 0, pz.i), B = tb.B(pz.side, 0, pz.i),
                            C = tb.C(pz.side, 0, pz.i);
                const float q = B * (*psi) + A * dVz_z;
                *psi = q;
                sxx -= dtL * (q + C * dVz_z);
                szz -= dtL2M * (q + C * dVz_z);
            }
            if (a.fs && z == a.lpml) szz = 0.f;
            const float dtM = dt * a.M[p];
            term(sxz, dtM, dVz_x, px, tb, 1, a.psi[RSG_E2D_VZX_LEFT], a.psi[RSG_E2D_VZX_RIGHT], ix);
            term(sxz, dtM, dVx_z, pz, tb, 1, a.psi[RSG_E2D_VXZ_TOP], a.psi[RSG_E2D_VXZ_BOTTOM], iz);
        } else {
            term(sxx, dt, dVx_x, px, tb, 0, a.psi[RSG_E2D_VXX_LEFT], a.psi[RSG_E2D_VXX_RIGHT], ix);
            term(szz, dt, dVz_z, pz, tb, 0, a.psi[RSG_E2D_VZZ_TOP], a.psi[RSG_E2D_VZZ_BOTTOM], iz);
            if (a.fs && z == a.lpml) szz = 0.f;
            term(sxz, dt, dVz_x, px, tb, 1, a.psi[RSG_E2D_VZX_LEFT], a.psi[RSG_E2D_VZX_RIGHT], ix);
            term(sxz, dt, dVx_z, pz, tb, 1, a.psi[RSG_E2D_VXZ_TOP], a.psi[RSG_E2D_VXZ_BOTTOM], iz);
        }
        a.Sxx[p] = sxx;
        a.Szz[p] = szz;
        a.Sxz[p] = sxz;
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

static int check_grid2d(const rsg_grid *g, const rsg_state *s) {
    if (!g || !s) return fail(RSG_EINVAL, "null grid/state");
    if (g->order < 1 || g->order > 8) return fail(RSG_EINVAL, "order must be 1..8");
    if (g->ny != 1) return fail(RSG_EINVAL, "2D kernels need ny == 1");
    if (g->lpml < 1 || g->nx <= 2 * g->lpml || g->nz <= 2 * g->lpml)
        return fail(RSG_EINVAL, "grid must exceed 2*lpml in x and z");
    if (g->nx < 2 * g->order + 2 || g->nz < 2 * g->order + 2)
        return fail(RSG_EINVAL, "grid too small for the stencil order");
    if (!s->tab) return fail(RSG_EINVAL, "PML tables not set");
    return RSG_OK;
}

static K2D make_k2d(const rsg_grid *g, const rsg_state *s) {
    K2D a;
    a.nx = g->nx;
    a.nz = g->nz;
    a.lpml = g->lpml;
    a.fs = g->fs;
    a.dt = g->dt;
    a.invdx = 1.0f / g->dx;
    a.invdz = 1.0f / g->dz;
    a.tab = s->tab;
    a.P = s->P;
    a.Vx = s->Vx;
    a.Vz = s->Vz;
    a.Sxx = s->Sxx;
    a.Szz = s->Szz;
    a.Sxz = s->Sxz;
    a.L = s->L;
    a.L2M = s->L2M;
    a.M = s->M;
    a.Rx = s->Rx;
    a.Rz = s->Rz;
    for (int i = 0; i < RSG_MAX_STRIPS; i++) a.psi[i] = s->psi[i];
    a.cx = scaled_coefs(g->dx);
    a.cz = scaled_coefs(g->dz);
    return a;
}

#define RSG_ORDER_CASES(LAUNCH) \
    switch (g->order) {          \
    case 1: LAUNCH(1); break;    \
    case 2: LAUNCH(2); break;    \
    case 3: LAUNCH(3); break;    \
    case 4: LAUNCH(4); break;    \
    case 5: LAUNCH(5); break;    \
    case 6: LAUNCH(6); break;    \
    case 7: LAUNCH(7); break;    \
    default: LAUNCH(8); break;   \
    }

static dim3 grid2d(const rsg_grid *g) {
    return dim3((g->nx + BX - 1) / BX, (g->nz + BZ * RZ2D - 1) / (BZ * RZ2D), 1);
}

}  // namespace rsg

using namespace rsg;

extern "C" int rsg_a2d_forward_velocity(const rsg_grid *g, const rsg_state *s, void *stream) {
    if (int rc = check_grid2d(g, s)) return rc;
    if (!s->P || !s->Vx || !s->Vz || !s->Rx || !s->Rz) return fail(RSG_EINVAL, "a2d: null array");
    const K2D a = make_k2d(g, s);
#define L_(O) a2d_velocity_kernel<O, RZ2D><<<grid2d(g), dim3(BX, BZ), 0, (hipStream_t)stream>>>(a)
    RSG_ORDER_CASES(L_)
#undef L_
    RSG_LAUNCH_CHECK("a2d_velocity_kernel");
    return RSG_OK;
}

extern "C" int rsg_a2d_forward_stress(const rsg_grid *g, const rsg_state *s, void *stream) {
    if (int rc = check_grid2d(g, s)) return rc;
    if (!s->P || !s->Vx || !s->Vz || !s->L) return fail(RSG_EINVAL, "a2d: null array");
    const K2D a = make_k2d(g, s);
#define L_(O) a2d_stress_kernel<O, RZ2D><<<grid2d(g), dim3(BX, BZ), 0, (hipStream_t)stream>>>(a)
    RSG_ORDER_CASES(L_)
#undef L_
    RSG_LAUNCH_CHECK("a2d_stress_kernel");
    return RSG_OK;
}

static int e2d_check(const rsg_grid *g, const rsg_state *s) {
    if (int rc = check_grid2d(g, s)) return rc;
    if (!s->Vx || !s->Vz || !s->Sxx || !s->Szz || !s->Sxz || !s->L || !s->L2M || !s->M || !s->Rx ||
        !s->Rz)
        return fail(RSG_EINVAL, "e2d: null array");
    return RSG_OK;
}

extern "C" int rsg_e2d_forward_velocity(const rsg_grid *g, const rsg_state *s, void *stream) {
    if (int rc = e2d_check(g, s)) return rc;
    const K2D a = make_k2d(g, s);
#define L_(O) \
    e2d_velocity_kernel<O, RZ2D, false><<<grid2d(g), dim3(BX, BZ), 0, (hipStream_t)stream>>>(a)
    RSG_ORDER_CASES(L_)
#undef L_
    RSG_LAUNCH_CHECK("e2d_velocity_kernel");
    return RSG_OK;
}

extern "C" int rsg_e2d_forward_stress(const rsg_grid *g, const rsg_state *s, void *stream) {
    if (int rc = e2d_check(g, s)) return rc;
    const K2D a = make_k2d(g, s);
#define L_(O) \
    e2d_stress_kernel<O, RZ2D, false><<<grid2d(g), dim3(BX, BZ), 0, (hipStream_t)stream>>>(a)
    RSG_ORDER_CASES(L_)
#undef L_
    RSG_LAUNCH_CHECK("e2d_stress_kernel");
    return RSG_OK;
}

extern "C" int rsg_e2d_backward_velocity(const rsg_grid *g, const rsg_state *s, void *stream) {
    if (int rc = e2d_check(g, s)) return rc;
    const K2D a = make_k2d(g, s);
#define L_(O) \
    e2d_velocity_kernel<O, RZ2D, true><<<grid2d(g), dim3(BX, BZ), 0, (hipStream_t)stream>>>(a)
    RSG_ORDER_CASES(L_)
#undef L_
    RSG_LAUNCH_CHECK("e2d_velocity_kernel(bw)");
    return RSG_OK;
}

extern "C" int rsg_e2d_backward_stress(const rsg_grid *g, const rsg_state *s, void *stream) {
    if (int rc = e2d_check(g, s)) return rc;
    const K2D a = make_k2d(g, s);
#define L_(O) \
    e2d_stress_kernel<O, RZ2D, true><<<grid2d(g), dim3(BX, BZ), 0, (hipStream_t)stream>>>(a)
    RSG_ORDER_CASES(L_)
#undef L_
    RSG_LAUNCH_CHECK("e2d_stress_kernel(bw)");
    return RSG_OK;
}
