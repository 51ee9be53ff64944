// 3D acoustic staggered-grid kernels for gfx950 (WavesAcoustic3D,
// waves.cpp:710-996): fused velocity and stress half steps with the CFS-PML
// memory updates of all six faces.
//
// Each block owns a 64 x 8 column tile in (x, y) and marches through a chunk
// of z planes.  The z-derivative taps of every column live in a register
// queue of 2*ORDER values that shifts by one plane per step (one new global
// load per cell), the x- and y-derivative taps come from the current plane
// staged in LDS with its halo.  The march is software-pipelined: the next
// plane's staging loads, queue entry and own cells are issued before the
// current plane is computed, so memory latency is paid once per chunk rather
// than once per plane.  Raw buffer loads with per-plane descriptors and fixed
// 32-bit in-plane offsets keep the address arithmetic out of the loop.
#include <algorithm>
#include <vector>

#include "rsg_common.h"

namespace rsg {

constexpr int B3X = 64, B3Y = 4, R3Y = 2, T3Y = B3Y * R3Y;

struct K3D {
    int nx, ny, nz, lpml;
    // Tile order.  ncx x ncy columns of B3X x T3Y cells, nzc z chunks.  With
    // xgroups set the launch is one-dimensional: blocks are dealt round-robin
    // to the 8 XCDs (XCD k runs blocks k, k+8, ...), and XCD k marches its own
    // contiguous group of columns cg[k] .. cg[k+1]-1 (x fastest, then y)
    // through the z chunks, columns fastest.  The blocks an XCD holds at once
    // are then neighbours in x and y at the same depth: the cache lines their
    // tile rows share (rows of a padded grid are rarely 128-byte aligned) and
    // their y halos are fetched into that XCD's L2 once, not once per XCD.
    int xgroups, ncx, ncy, nzc, cgmax;
    int cg[9];
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

// Byte offset past every plane: raw buffer loads there return 0 and stores
// are dropped, so out-of-grid lanes and halo cells need no branches.
constexpr unsigned kOut = 0xFFFFFF00u;

// One plane of a field as a raw buffer (32-bit in-plane byte offsets).
__device__ __forceinline__ Buf plane_buf(const float *f, int z, size_t nxy) {
    return mkbuf(f + (size_t)z * nxy, (unsigned)(nxy * 4));
}

// Staging of the (T3Y+2HY) x (64+2HX) window of one plane.  Each thread's
// element list (in-plane offsets, kOut outside the grid) is fixed for the
// whole z march; fetch issues the loads into registers, store writes them to
// LDS, so the next plane's loads are in flight while this plane computes.
template <int HX, int HY>
struct Plane {
    static constexpr int LD = B3X + 2 * HX, N = LD * (T3Y + 2 * HY), K = (N + B3X * B3Y - 1) / (B3X * B3Y);
    unsigned off[K];
    float v[K];
    __device__ __forceinline__ void init(int nx, int ny, int x0, int y0) {
        const int tid = threadIdx.x + threadIdx.y * B3X;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = tid + k * B3X * B3Y;
            const int r = i / LD, c = i - r * LD;
            const int y = y0 - HY + r, x = x0 - HX + c;
            off[k] = (i < N && x >= 0 && x < nx && y >= 0 && y < ny) ? (unsigned)((y * nx + x) * 4) : kOut;
        }
    }
    __device__ __forceinline__ void fetch(const Buf &b) {
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = b.ld(off[k], 0);
    }
    __device__ __forceinline__ void store(float *lds) const {
        const int tid = threadIdx.x + threadIdx.y * B3X;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (N % (B3X * B3Y) == 0 || tid + k * B3X * B3Y < N) lds[tid + k * B3X * B3Y] = v[k];
    }
};

// Tile (x0, y0, chunk) of this block; false for the surplus blocks of an XCD
// whose column group is one column shorter.
__device__ __forceinline__ bool tile3_of(const K3D &a, int &x0, int &y0, int &zchunk) {
    int cx, cy;
    if (!a.xgroups) {
        cx = blockIdx.x;
        cy = blockIdx.y;
        zchunk = blockIdx.z;
    } else {
        const int k = blockIdx.x % 8, j = blockIdx.x / 8;
        const int ncol = a.cg[k + 1] - a.cg[k];
        if (ncol <= 0 || j >= ncol * a.nzc) return false;
        zchunk = j / ncol;
        const int col = a.cg[k] + (j - zchunk * ncol);
        cy = col / a.ncx;
        cx = col - cy * a.ncx;
    }
    x0 = cx * B3X;
    y0 = cy * T3Y;
    return true;
}

// Per-thread column set-up shared by both half steps: in-plane offsets of the
// thread's R3Y cells (kOut outside the grid) and their PML positions and
// edge stencil orders in x and y, fixed for the whole z march.
struct Cols {
    unsigned oc[R3Y];
    bool ok[R3Y];
    int sy[R3Y], iy[R3Y], ky[R3Y];
    int sx, ix, kx;
};
template <bool FW>
__device__ __forceinline__ Cols make_cols(const K3D &a, int x, int y0, int order) {
    Cols c;
    side_of(x, a.nx, a.lpml, c.sx, c.ix);
    c.kx = FW ? kfw(x, a.nx, order) : kbw(x, a.nx, order);
#pragma unroll
    for (int r = 0; r < R3Y; r++) {
        const int y = y0 + threadIdx.y * R3Y + r;
        c.ok[r] = x < a.nx && y < a.ny;
        c.oc[r] = c.ok[r] ? (unsigned)((y * a.nx + x) * 4) : kOut;
        side_of(y, a.ny, a.lpml, c.sy[r], c.iy[r]);
        c.ky[r] = FW ? kfw(y, a.ny, order) : kbw(y, a.ny, order);
    }
    return c;
}

// Velocity half step.  Per plane z of the block's z chunk: the staged plane
// of P (x and y taps) and the queue of P along z for the thread's own cells
// (z taps); the next plane of P, the next queue entry and the next own cells
// (Vx, Vy, Vz, Rx, Ry, Rz) are loaded while this plane computes.
template <int ORDER>
__global__ __launch_bounds__(B3X *B3Y) void a3d_velocity_kernel(const K3D a, int zc) {
    constexpr int H = ORDER, LD = B3X + 2 * H;
    __shared__ float sP[(T3Y + 2 * H) * LD];
    int x0, y0, zck;
    if (!tile3_of(a, x0, y0, zck)) return;
    const int zb = zck * zc, ze = min(zb + zc, a.nz);
    const int tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    const size_t nxy = (size_t)a.nx * a.ny;
    const Tab3 tb{a.tab, a.lpml};
    const Cols c = make_cols<true>(a, x, y0, ORDER);
    Plane<H, H> pl;
    pl.init(a.nx, a.ny, x0, y0);
    float q[R3Y][2 * H];
#pragma unroll
    for (int j = 0; j < 2 * H; j++) {
        const int zz = zb + 1 - H + j;
        const bool in = zz >= 0 && zz < a.nz;
        const Buf b = plane_buf(a.P, in ? zz : 0, nxy);
#pragma unroll
        for (int r = 0; r < R3Y; r++) q[r][j] = in ? b.ld(c.oc[r], 0) : 0.f;
    }
    float vx[R3Y], vy[R3Y], vz[R3Y], rx[R3Y], ry[R3Y], rz[R3Y];
    auto own = [&](int z) {
        const Buf bVx = plane_buf(a.Vx, z, nxy), bVy = plane_buf(a.Vy, z, nxy), bVz = plane_buf(a.Vz, z, nxy),
                  bRx = plane_buf(a.Rx, z, nxy), bRy = plane_buf(a.Ry, z, nxy), bRz = plane_buf(a.Rz, z, nxy);
#pragma unroll
        for (int r = 0; r < R3Y; r++) {
            vx[r] = bVx.ld(c.oc[r], 0);
            vy[r] = bVy.ld(c.oc[r], 0);
            vz[r] = bVz.ld(c.oc[r], 0);
            rx[r] = bRx.ld(c.oc[r], 0);
            ry[r] = bRy.ld(c.oc[r], 0);
            rz[r] = bRz.ld(c.oc[r], 0);
        }
    };
    if (zb < ze) {
        pl.fetch(plane_buf(a.P, zb, nxy));
        own(zb);
    }
    for (int z = zb; z < ze; z++) {
        // zero, opaque to the compiler: keeps the plane's LDS addresses inside
        // the loop instead of hoisted, each holding a VGPR for the whole march
        int o = 0;
        asm volatile("" : "+v"(o));
        float *const sp = sP + o;
        pl.store(sp);
        float cvx[R3Y], cvy[R3Y], cvz[R3Y], crx[R3Y], cry[R3Y], crz[R3Y], qn[R3Y];
#pragma unroll
        for (int r = 0; r < R3Y; r++) {
            cvx[r] = vx[r];
            cvy[r] = vy[r];
            cvz[r] = vz[r];
            crx[r] = rx[r];
            cry[r] = ry[r];
            crz[r] = rz[r];
        }
        __syncthreads();
        // loads of the next plane, in flight during this plane's compute
        if (z + 1 < ze) {
            pl.fetch(plane_buf(a.P, z + 1, nxy));
            own(z + 1);
        }
        {
            const int zq = z + 1 + H;
            const bool in = zq < a.nz;
            const Buf b = plane_buf(a.P, in ? zq : 0, nxy);
#pragma unroll
            for (int r = 0; r < R3Y; r++) qn[r] = in ? b.ld(c.oc[r], 0) : 0.f;
        }
        int sz, iz;
        side_of(z, a.nz, a.lpml, sz, iz);
        const int kz = kfw(z, a.nz, ORDER);
        const Buf bVx = plane_buf(a.Vx, z, nxy), bVy = plane_buf(a.Vy, z, nxy), bVz = plane_buf(a.Vz, z, nxy);
#pragma unroll
        for (int r = 0; r < R3Y; r++) {
            const int yl = ty * R3Y + r, y = y0 + yl;
            const float *s = sp + (yl + H) * LD + tx + H;
            float dPx = dfw<ORDER>(s, 1, a.cx.c + cind(ORDER));
            if (c.kx != ORDER) dPx = dfw_k(s, 1, c.kx, a.invdx);
            float dPy = dfw<ORDER>(s, LD, a.cy.c + cind(ORDER));
            if (c.ky[r] != ORDER) dPy = dfw_k(s, LD, c.ky[r], a.invdy);
            float dPz = dq<H>(q[r], a.cz.c + cind(ORDER));
            if (kz != ORDER) dPz = dq_k<H>(q[r], kz, a.invdz);
            if (c.ok[r]) {
                const int sx = c.sx, ix = c.ix, sy = c.sy[r], iy = c.iy[r];
                term3(cvx[r], a.dt * crx[r], dPx, sx, ix, tb, 1, a.psi[RSG_A3D_P_LEFT], a.psi[RSG_A3D_P_RIGHT],
                      ((size_t)z * a.ny + y) * a.lpml + ix);
                term3(cvy[r], a.dt * cry[r], dPy, sy, iy, tb, 1, a.psi[RSG_A3D_P_FRONT], a.psi[RSG_A3D_P_BACK],
                      ((size_t)z * a.lpml + iy) * a.nx + x);
                term3(cvz[r], a.dt * crz[r], dPz, sz, iz, tb, 1, a.psi[RSG_A3D_P_TOP], a.psi[RSG_A3D_P_BOTTOM],
                      ((size_t)iz * a.ny + y) * a.nx + x);
            }
            bVx.st(cvx[r], c.oc[r], 0);
            bVy.st(cvy[r], c.oc[r], 0);
            bVz.st(cvz[r], c.oc[r], 0);
#pragma unroll
            for (int j = 0; j < 2 * H - 1; j++) q[r][j] = q[r][j + 1];
            q[r][2 * H - 1] = qn[r];
        }
        __syncthreads();  // the next plane's staging overwrites the LDS
    }
}

// Stress half step: staged planes of Vx (x taps) and Vy (y taps), the queue
// of Vz along z, own cells P and L; the next planes, queue entry and own
// cells are loaded while this plane computes.
template <int ORDER>
__global__ __launch_bounds__(B3X *B3Y) void a3d_stress_kernel(const K3D a, int zc) {
    constexpr int H = ORDER, LDX = B3X + 2 * H, LDY = B3X;
    __shared__ float sVx[T3Y * LDX];
    __shared__ float sVy[(T3Y + 2 * H) * LDY];
    int x0, y0, zck;
    if (!tile3_of(a, x0, y0, zck)) return;
    const int zb = zck * zc, ze = min(zb + zc, a.nz);
    const int tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    const size_t nxy = (size_t)a.nx * a.ny;
    const Tab3 tb{a.tab, a.lpml};
    const Cols c = make_cols<false>(a, x, y0, ORDER);
    Plane<H, 0> px;
    Plane<0, H> py;
    px.init(a.nx, a.ny, x0, y0);
    py.init(a.nx, a.ny, x0, y0);
    float q[R3Y][2 * H];
#pragma unroll
    for (int j = 0; j < 2 * H; j++) {
        const int zz = zb - H + j;
        const bool in = zz >= 0 && zz < a.nz;
        const Buf b = plane_buf(a.Vz, in ? zz : 0, nxy);
#pragma unroll
        for (int r = 0; r < R3Y; r++) q[r][j] = in ? b.ld(c.oc[r], 0) : 0.f;
    }
    float pv[R3Y], lv[R3Y];
    auto own = [&](int z) {
        const Buf bP = plane_buf(a.P, z, nxy), bL = plane_buf(a.L, z, nxy);
#pragma unroll
        for (int r = 0; r < R3Y; r++) {
            pv[r] = bP.ld(c.oc[r], 0);
            lv[r] = bL.ld(c.oc[r], 0);
        }
    };
    if (zb < ze) {
        px.fetch(plane_buf(a.Vx, zb, nxy));
        py.fetch(plane_buf(a.Vy, zb, nxy));
        own(zb);
    }
    for (int z = zb; z < ze; z++) {
        int o = 0;  // as in the velocity kernel
        asm volatile("" : "+v"(o));
        float *const svx = sVx + o, *const svy = sVy + o;
        px.store(svx);
        py.store(svy);
        float cp[R3Y], cl[R3Y], qn[R3Y];
#pragma unroll
        for (int r = 0; r < R3Y; r++) {
            cp[r] = pv[r];
            cl[r] = lv[r];
        }
        __syncthreads();
        if (z + 1 < ze) {
            px.fetch(plane_buf(a.Vx, z + 1, nxy));
            py.fetch(plane_buf(a.Vy, z + 1, nxy));
            own(z + 1);
        }
        {
            const int zq = z + H;
            const bool in = zq < a.nz;
            const Buf b = plane_buf(a.Vz, in ? zq : 0, nxy);
#pragma unroll
            for (int r = 0; r < R3Y; r++) qn[r] = in ? b.ld(c.oc[r], 0) : 0.f;
        }
        int sz, iz;
        side_of(z, a.nz, a.lpml, sz, iz);
        const int kz = kbw(z, a.nz, ORDER);
        const Buf bP = plane_buf(a.P, z, nxy);
#pragma unroll
        for (int r = 0; r < R3Y; r++) {
            const int yl = ty * R3Y + r, y = y0 + yl;
            const float *sx_ = svx + yl * LDX + tx + H;
            float dVx = dbw<ORDER>(sx_, 1, a.cx.c + cind(ORDER));
            if (c.kx != ORDER) dVx = dbw_k(sx_, 1, c.kx, a.invdx);
            const float *sy_ = svy + (yl + H) * LDY + tx;
            float dVy = dbw<ORDER>(sy_, LDY, a.cy.c + cind(ORDER));
            if (c.ky[r] != ORDER) dVy = dbw_k(sy_, LDY, c.ky[r], a.invdy);
            float dVz = dq<H>(q[r], a.cz.c + cind(ORDER));
            if (kz != ORDER) dVz = dq_k<H>(q[r], kz, a.invdz);
            if (c.ok[r]) {
                const int sx = c.sx, ix = c.ix, sy = c.sy[r], iy = c.iy[r];
                const float dtL = a.dt * cl[r];
                term3(cp[r], dtL, dVx, sx, ix, tb, 0, a.psi[RSG_A3D_VXX_LEFT], a.psi[RSG_A3D_VXX_RIGHT],
                      ((size_t)z * a.ny + y) * a.lpml + ix);
                term3(cp[r], dtL, dVy, sy, iy, tb, 0, a.psi[RSG_A3D_VYY_FRONT], a.psi[RSG_A3D_VYY_BACK],
                      ((size_t)z * a.lpml + iy) * a.nx + x);
                term3(cp[r], dtL, dVz, sz, iz, tb, 0, a.psi[RSG_A3D_VZZ_TOP], a.psi[RSG_A3D_VZZ_BOTTOM],
                      ((size_t)iz * a.ny + y) * a.nx + x);
            }
            bP.st(cp[r], c.oc[r], 0);
#pragma unroll
            for (int j = 0; j < 2 * H - 1; j++) q[r][j] = q[r][j + 1];
            q[r][2 * H - 1] = qn[r];
        }
        __syncthreads();  // the next planes' staging overwrites the LDS
    }
}

static int check_grid3d(const rsg_grid *g, const rsg_state *s) {
    if (!g || !s) return fail(RSG_EINVAL, "null grid/state");
    if (g->order < 1 || g->order > 8) return fail(RSG_EINVAL, "order must be 1..8");
    if (g->lpml < 1 || g->nx <= 2 * g->lpml || g->ny <= 2 * g->lpml || g->nz <= 2 * g->lpml)
        return fail(RSG_EINVAL, "grid must exceed 2*lpml on every axis");
    if (g->nx < 2 * g->order + 2 || g->ny < 2 * g->order + 2 || g->nz < 2 * g->order + 2)
        return fail(RSG_EINVAL, "grid too small for the stencil order");
    if ((long long)g->nx * g->ny * 4 >= (1LL << 31))
        return fail(RSG_EINVAL, "3D plane larger than 2^29 cells (32-bit in-plane offsets)");
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
    a.xgroups = 0;
    a.ncx = (g->nx + B3X - 1) / B3X;
    a.ncy = (g->ny + T3Y - 1) / T3Y;
    a.nzc = a.cgmax = 0;
    for (int k = 0; k < 9; k++) a.cg[k] = 0;
    return a;
}

// Launch grid of a 3D half step with z chunks of zc planes, and the tile
// order it implies (K3D.xgroups, cg).
static dim3 grid3(K3D &a, int zc) {
    a.nzc = (a.nz + zc - 1) / zc;
    if (!tuning().xcd3d) {
        a.xgroups = 0;
        return dim3(a.ncx, a.ncy, a.nzc);
    }
    a.xgroups = 1;
    // groups of equal estimated cost: a cell in an x or y PML column counts
    // pml_cost percent of an interior one (it updates memory variables on
    // every plane), so the PML-heavy groups at the ends of the y range are
    // shorter
    const int ncol = a.ncx * a.ncy;
    const long pc = tuning().pml_cost;
    std::vector<long long> pre(ncol + 1, 0);
    for (int c = 0; c < ncol; c++) {
        const int x0 = (c % a.ncx) * B3X, y0 = (c / a.ncx) * T3Y;
        const int x1 = std::min(x0 + B3X, a.nx), y1 = std::min(y0 + T3Y, a.ny);
        long cost = 0;
        for (int y = y0; y < y1; y++)
            for (int x = x0; x < x1; x++) {
                const bool pml = x < a.lpml || x >= a.nx - a.lpml || y < a.lpml || y >= a.ny - a.lpml;
                cost += pml ? pc : 100;
            }
        pre[c + 1] = pre[c] + cost;
    }
    a.cg[0] = 0;
    a.cgmax = 0;
    int c = 0;
    for (int k = 1; k <= 8; k++) {
        const long long target = pre[ncol] * k / 8;
        while (c < ncol && pre[c + 1] <= target) c++;
        if (k == 8) c = ncol;
        a.cg[k] = c;
        a.cgmax = std::max(a.cgmax, a.cg[k] - a.cg[k - 1]);
    }
    return dim3((unsigned)(8L * a.cgmax * a.nzc), 1, 1);
}

// z-chunk length: short chunks give many blocks (the pipelined march keeps
// 2-3 blocks per CU resident, so the grid must fill the CUs many times over
// to balance), long ones amortize the queue warm-up (2*ORDER planes).
// Measured on MI355X at 440^3: 8 and 16 planes are best, 64 costs 25%.
static int zchunk(const rsg_grid *g) {
    if (tuning().zc3d > 0) return tuning().zc3d;
    const long cols = (long)((g->nx + B3X - 1) / B3X) * ((g->ny + T3Y - 1) / T3Y);
    int zc = 16;
    while (zc > 8 && cols * ((g->nz + zc - 1) / zc) < 16384) zc /= 2;
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
    K3D a = make_k3d(g, s);
    const int zc = zchunk(g);
    const dim3 grid = grid3(a, zc);
#define L_(O) a3d_velocity_kernel<O><<<grid, dim3(B3X, B3Y), 0, (hipStream_t)stream>>>(a, zc)
    RSG3_ORDER_CASES(L_)
#undef L_
    RSG_LAUNCH_CHECK("a3d_velocity_kernel");
    return RSG_OK;
}

extern "C" int rsg_a3d_forward_stress(const rsg_grid *g, const rsg_state *s, void *stream) {
    if (int rc = check_grid3d(g, s)) return rc;
    if (!s->L) return fail(RSG_EINVAL, "a3d: null model");
    K3D a = make_k3d(g, s);
    const int zc = zchunk(g);
    const dim3 grid = grid3(a, zc);
#define L_(O) a3d_stress_kernel<O><<<grid, dim3(B3X, B3Y), 0, (hipStream_t)stream>>>(a, zc)
    RSG3_ORDER_CASES(L_)
#undef L_
    RSG_LAUNCH_CHECK("a3d_stress_kernel");
    return RSG_OK;
}
