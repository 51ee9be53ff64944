// 3D acoustic staggered-grid kernels for gfx950 (WavesAcoustic3D,
// waves.cpp:710-996): fused velocity and stress half steps with the CFS-PML
// memory updates of all six faces.
//
// Each block owns a 64 x 8 column tile in (x, y) and marches through a chunk
// of z planes.  The z-derivative taps of every column live in a register
// queue of 2*ORDER values that shifts by one plane per step (one new global
// load per cell), the x- and y-derivative taps come from the current plane
// staged in LDS with its halo.
#include "rsg_common.h"

namespace rsg {

constexpr int B3X = 64, B3Y = 4, R3Y = 2, T3Y = B3Y * R3Y;

struct K3D {
    int nx, ny, nz, lpml;
    float dt, invdx, invdy, invdz;
    const float *tab;
    float *P, *Vx, *Vy, *Vz;
    const float *L, *Rx, *Ry, *Rz;
    float *psi[RSG_MAX_STRIPS];
    Coefs cx, cy, cz;
};

struct Tab3 {
    const float *t;
    int L;
    __device__ __forceinline__ float get(int right, int stag, int which, int i) const {
        return t[(right * T_RBB + stag * T_STAG + which) * L + i];
    }
};

__device__ __forceinline__ void side_of(int i, int n, int L, int &side, int &idx) {
    side = -1;
    idx = 0;
    if (i < L) {
        side = 0;
        idx = i;
    } else if (i >= n - L) {
        side = 1;
        idx = i - (n - L);
    }
}

__device__ __forceinline__ void term3(float &F, float dtM, float d, int side, int idx, const Tab3 &tb,
                                      int stag, float *lo, float *hi, size_t sidx) {
    F += dtM * d;
    if (side >= 0) {
        float *psi = (side ? hi : lo) + sidx;
        pml_corr(F, dtM, psi, tb.get(side, stag, T_A, idx), tb.get(side, stag, T_B, idx),
                 tb.get(side, stag, T_C, idx), d);
    }
}

// Stage plane z rows [y0-HY, y0+T3Y+HY) x cols [x0-HX, x0+B3X+HX) of g.
template <int HX, int HY>
__device__ __forceinline__ void load_plane(float *lds, const float *__restrict__ g, int nx, int ny,
                                           int nz, int x0, int y0, int z) {
    constexpr int LD = B3X + 2 * HX, N = LD * (T3Y + 2 * HY);
    const bool zin = z >= 0 && z < nz;
    const size_t base = (size_t)z * nx * ny;
    for (int i = threadIdx.x + threadIdx.y * B3X; i < N; i += B3X * B3Y) {
        const int r = i / LD, c = i - r * LD;
        const int y = y0 - HY + r, x = x0 - HX + c;
        float v = 0.f;
        if (zin && x >= 0 && x < nx && y >= 0 && y < ny) v = g[base + (size_t)y * nx + x];
        lds[i] = v;
    }
}

// z-derivative from a queue q[j] = f(z + off + j): both the forward
// (off = 1-H) and the backward (off = -H) staggered forms reduce to
// sum_io c_io (q[H+io] - q[H-1-io]).
template <int H>
__device__ __forceinline__ float dq(const float (&q)[2 * H], const float *c) {
    float acc = c[0] * (q[H] - q[H - 1]);
#pragma unroll
    for (int io = 1; io < H; io++) acc += c[io] * (q[H + io] - q[H - 1 - io]);
    return acc;
}
template <int H>
__device__ __forceinline__ float dq_k(const float (&q)[2 * H], int k, float invd) {
    if (k <= 0) return 0.f;
    const float *c = kCoefC + cind(k);
    float acc = c[0] * (q[H] - q[H - 1]);
#pragma unroll
    for (int io = 1; io < H; io++)
        if (io < k) acc += c[io] * (q[H + io] - q[H - 1 - io]);
    return acc * invd;
}

__device__ __forceinline__ float ld_or0(const float *g, int x, int y, int z, int nx, int ny, int nz) {
    return (z >= 0 && z < nz) ? g[((size_t)z * ny + y) * nx + x] : 0.f;
}

template <int ORDER>
__global__ __launch_bounds__(B3X *B3Y) void a3d_velocity_kernel(const K3D a, int zc) {
    constexpr int H = ORDER, LD = B3X + 2 * H;
    __shared__ float sP[(T3Y + 2 * H) * LD];
    const int x0 = blockIdx.x * B3X, y0 = blockIdx.y * T3Y;
    const int zb = blockIdx.z * zc, ze = min(zb + zc, a.nz);
    const int tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    const bool xin = x < a.nx;
    const Tab3 tb{a.tab, a.lpml};
    int sx, ix;
    side_of(x, a.nx, a.lpml, sx, ix);
    const int kx = kfw(x, a.nx, ORDER);
    float q[R3Y][2 * H];
#pragma unroll
    for (int r = 0; r < R3Y; r++) {
        const int y = y0 + ty * R3Y + r;
#pragma unroll
        for (int j = 0; j < 2 * H; j++)
            q[r][j] = (xin && y < a.ny) ? ld_or0(a.P, x, y, zb + 1 - H + j, a.nx, a.ny, a.nz) : 0.f;
    }
    for (int z = zb; z < ze; z++) {
        __syncthreads();
        load_plane<H, H>(sP, a.P, a.nx, a.ny, a.nz, x0, y0, z);
        __syncthreads();
        int sz, iz;
        side_of(z, a.nz, a.lpml, sz, iz);
        const int kz = kfw(z, a.nz, ORDER);
#pragma unroll
        for (int r = 0; r < R3Y; r++) {
            const int yl = ty * R3Y + r, y = y0 + yl;
            if (xin && y < a.ny) {
                const float *s = sP + (yl + H) * LD + tx + H;
                float dPx = dfw<ORDER>(s, 1, a.cx.c + cind(ORDER));
                if (kx != ORDER) dPx = dfw_k(s, 1, kx, a.invdx);
                const int ky = kfw(y, a.ny, ORDER);
                float dPy = dfw<ORDER>(s, LD, a.cy.c + cind(ORDER));
                if (ky != ORDER) dPy = dfw_k(s, LD, ky, a.invdy);
                float dPz = dq<H>(q[r], a.cz.c + cind(ORDER));
                if (kz != ORDER) dPz = dq_k<H>(q[r], kz, a.invdz);
                int sy, iy;
                side_of(y, a.ny, a.lpml, sy, iy);
                const size_t p = ((size_t)z * a.ny + y) * a.nx + x;
                float vx = a.Vx[p];
                term3(vx, a.dt * a.Rx[p], dPx, sx, ix, tb, 1, a.psi[RSG_A3D_P_LEFT],
                      a.psi[RSG_A3D_P_RIGHT], ((size_t)z * a.ny + y) * a.lpml + ix);
                a.Vx[p] = vx;
                float vy = a.Vy[p];
                term3(vy, a.dt * a.Ry[p], dPy, sy, iy, tb, 1, a.psi[RSG_A3D_P_FRONT],
                      a.psi[RSG_A3D_P_BACK], ((size_t)z * a.lpml + iy) * a.nx + x);
                a.Vy[p] = vy;
                float vz = a.Vz[p];
                term3(vz, a.dt * a.Rz[p], dPz, sz, iz, tb, 1, a.psi[RSG_A3D_P_TOP],
                      a.psi[RSG_A3D_P_BOTTOM], ((size_t)iz * a.ny + y) * a.nx + x);
                a.Vz[p] = vz;
            }
#pragma unroll
            for (int j = 0; j < 2 * H - 1; j++) q[r][j] = q[r][j + 1];
            q[r][2 * H - 1] = (xin && y < a.ny) ? ld_or0(a.P, x, y, z + 1 + H, a.nx, a.ny, a.nz) : 0.f;
        }
    }
}

template <int ORDER>
__global__ __launch_bounds__(B3X *B3Y) void a3d_stress_kernel(const K3D a, int zc) {
    constexpr int H = ORDER, LDX = B3X + 2 * H, LDY = B3X;
    __shared__ float sVx[T3Y * LDX];
    __shared__ float sVy[(T3Y + 2 * H) * LDY];
    const int x0 = blockIdx.x * B3X, y0 = blockIdx.y * T3Y;
    const int zb = blockIdx.z * zc, ze = min(zb + zc, a.nz);
    const int tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    const bool xin = x < a.nx;
    const Tab3 tb{a.tab, a.lpml};
    int sx, ix;
    side_of(x, a.nx, a.lpml, sx, ix);
    const int kx = kbw(x, a.nx, ORDER);
    float q[R3Y][2 * H];
#pragma unroll
    for (int r = 0; r < R3Y; r++) {
        const int y = y0 + ty * R3Y + r;
#pragma unroll
        for (int j = 0; j < 2 * H; j++)
            q[r][j] = (xin && y < a.ny) ? ld_or0(a.Vz, x, y, zb - H + j, a.nx, a.ny, a.nz) : 0.f;
    }
    for (int z = zb; z < ze; z++) {
        __syncthreads();
        load_plane<H, 0>(sVx, a.Vx, a.nx, a.ny, a.nz, x0, y0, z);
        load_plane<0, H>(sVy, a.Vy, a.nx, a.ny, a.nz, x0, y0, z);
        __syncthreads();
        int sz, iz;
        side_of(z, a.nz, a.lpml, sz, iz);
        const int kz = kbw(z, a.nz, ORDER);
#pragma unroll
        for (int r = 0; r < R3Y; r++) {
            const int yl = ty * R3Y + r, y = y0 + yl;
            if (xin && y < a.ny) {
                const float *sx_ = sVx + yl * LDX + tx + H;
                float dVx = dbw<ORDER>(sx_, 1, a.cx.c + cind(ORDER));
                if (kx != ORDER) dVx = dbw_k(sx_, 1, kx, a.invdx);
                const float *sy_ = sVy + (yl + H) * LDY + tx;
                const int ky = kbw(y, a.ny, ORDER);
                float dVy = dbw<ORDER>(sy_, LDY, a.cy.c + cind(ORDER));
                if (ky != ORDER) dVy = dbw_k(sy_, LDY, ky, a.invdy);
                float dVz = dq<H>(q[r], a.cz.c + cind(ORDER));
                if (kz != ORDER) dVz = dq_k<H>(q[r], kz, a.invdz);
                int sy, iy;
                side_of(y, a.ny, a.lpml, sy, iy);
                const size_t p = ((size_t)z * a.ny + y) * a.nx + x;
                const float dtL = a.dt * a.L[p];
                float P = a.P[p];
                term3(P, dtL, dVx, sx, ix, tb, 0, a.psi[RSG_A3D_VXX_LEFT], a.psi[RSG_A3D_VXX_RIGHT],
                      ((size_t)z * a.ny + y) * a.lpml + ix);
                term3(P, dtL, dVy, sy, iy, tb, 0, a.psi[RSG_A3D_VYY_FRONT], a.psi[RSG_A3D_VYY_BACK],
                      ((size_t)z * a.lpml + iy) * a.nx + x);
                term3(P, dtL, dVz, sz, iz, tb, 0, a.psi[RSG_A3D_VZZ_TOP], a.psi[RSG_A3D_VZZ_BOTTOM],
                      ((size_t)iz * a.ny + y) * a.nx + x);
                a.P[p] = P;
            }
#pragma unroll
            for (int j = 0; j < 2 * H - 1; j++) q[r][j] = q[r][j + 1];
            q[r][2 * H - 1] = (xin && y < a.ny) ? ld_or0(a.Vz, x, y, z + H, a.nx, a.ny, a.nz) : 0.f;
        }
    }
}

static int check_grid3d(const rsg_grid *g, const rsg_state *s) {
    if (!g || !s) return fail(RSG_EINVAL, "null grid/state");
    if (g->order < 1 || g->order > 8) return fail(RSG_EINVAL, "order must be 1..8");
    if (g->lpml < 1 || g->nx <= 2 * g->lpml || g->ny <= 2 * g->lpml || g->nz <= 2 * g->lpml)
        return fail(RSG_EINVAL, "grid must exceed 2*lpml on every axis");
    if (g->nx < 2 * g->order + 2 || g->ny < 2 * g->order + 2 || g->nz < 2 * g->order + 2)
        return fail(RSG_EINVAL, "grid too small for the stencil order");
    if (!s->tab || !s->P || !s->Vx || !s->Vy || !s->Vz) return fail(RSG_EINVAL, "a3d: null array");
    return RSG_OK;
}

static K3D make_k3d(const rsg_grid *g, const rsg_state *s) {
    K3D a;
    a.nx = g->nx;
    a.ny = g->ny;
    a.nz = g->nz;
    a.lpml = g->lpml;
    a.dt = g->dt;
    a.invdx = 1.0f / g->dx;
    a.invdy = 1.0f / g->dy;
    a.invdz = 1.0f / g->dz;
    a.tab = s->tab;
    a.P = s->P;
    a.Vx = s->Vx;
    a.Vy = s->Vy;
    a.Vz = s->Vz;
    a.L = s->L;
    a.Rx = s->Rx;
    a.Ry = s->Ry;
    a.Rz = s->Rz;
    for (int i = 0; i < RSG_MAX_STRIPS; i++) a.psi[i] = s->psi[i];
    a.cx = scaled_coefs(g->dx);
    a.cy = scaled_coefs(g->dy);
    a.cz = scaled_coefs(g->dz);
    return a;
}

// z-chunk length: enough blocks to fill 256 CUs several times over while
// keeping the queue warm-up (2*ORDER planes) small against the chunk.
static int zchunk(const rsg_grid *g) {
    if (tuning().zc3d > 0) return tuning().zc3d;
    const long cols = (long)((g->nx + B3X - 1) / B3X) * ((g->ny + T3Y - 1) / T3Y);
    int zc = 64;
    while (zc > 16 && cols * ((g->nz + zc - 1) / zc) < 2048) zc /= 2;
    return zc;
}

}  // namespace rsg

using namespace rsg;

#define RSG3_ORDER_CASES(LAUNCH) \
    switch (g->order) {           \
    case 1: LAUNCH(1); break;     \
    case 2: LAUNCH(2); break;     \
    case 3: LAUNCH(3); break;     \
    case 4: LAUNCH(4); break;     \
    case 5: LAUNCH(5); break;     \
    case 6: LAUNCH(6); break;     \
    case 7: LAUNCH(7); break;     \
    default: LAUNCH(8); break;    \
    }

extern "C" int rsg_a3d_forward_velocity(const rsg_grid *g, const rsg_state *s, void *stream) {
    if (int rc = check_grid3d(g, s)) return rc;
    if (!s->Rx || !s->Ry || !s->Rz) return fail(RSG_EINVAL, "a3d: null model");
    const K3D a = make_k3d(g, s);
    const int zc = zchunk(g);
    const dim3 grid((g->nx + B3X - 1) / B3X, (g->ny + T3Y - 1) / T3Y, (g->nz + zc - 1) / zc);
#define L_(O) a3d_velocity_kernel<O><<<grid, dim3(B3X, B3Y), 0, (hipStream_t)stream>>>(a, zc)
    RSG3_ORDER_CASES(L_)
#undef L_
    RSG_LAUNCH_CHECK("a3d_velocity_kernel");
    return RSG_OK;
}

extern "C" int rsg_a3d_forward_stress(const rsg_grid *g, const rsg_state *s, void *stream) {
    if (int rc = check_grid3d(g, s)) return rc;
    if (!s->L) return fail(RSG_EINVAL, "a3d: null model");
    const K3D a = make_k3d(g, s);
    const int zc = zchunk(g);
    const dim3 grid((g->nx + B3X - 1) / B3X, (g->ny + T3Y - 1) / T3Y, (g->nz + zc - 1) / zc);
#define L_(O) a3d_stress_kernel<O><<<grid, dim3(B3X, B3Y), 0, (hipStream_t)stream>>>(a, zc)
    RSG3_ORDER_CASES(L_)
#undef L_
    RSG_LAUNCH_CHECK("a3d_stress_kernel");
    return RSG_OK;
}
