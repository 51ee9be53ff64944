// Device side of the 2D staggered-grid kernels (shared by the uniform
// batch launchers in fd2d.hip and the ragged-batch launchers in
// fd2d_ragged.hip): kernel arguments, LDS tile staging, per-cell update
// bodies, the kernels, and the host helpers that build their arguments.
#pragma once
#include <algorithm>
#include <vector>

#include "rsg_common.h"

// 2D staggered-grid velocity-stress kernels for gfx950: acoustic
// (WavesAcoustic2D, waves.cpp:217-404) and elastic forward and adjoint-state
// steps (WavesElastic2D, waves.cpp:1416-2080).
//
// One fused kernel per half step.  The reference makes, per derivative, a
// full-grid Der pass into a scratch array, a full-grid update pass and a PML
// strip pass (der.cpp, waves.cpp); here each block stages the stencil inputs
// of a 64 x TZ tile (plus halo) in LDS once, evaluates every derivative of
// the half step from LDS, and applies the field update and the CFS-PML
// memory update of every cell in registers, so each array crosses HBM once
// per half step.  The per-cell operation sequence is the reference's: += dt*M*d
// for the first derivative, its PML correction, then the next derivative.
//
// Block: 64 x 4 threads (four wave64 rows); each thread owns RZ consecutive
// cells of one x column.  The thread's own field and model values are loaded
// into registers before the tile is staged so both sets of loads are in
// flight together; x-derivative taps are conflict-free LDS row reads.  Cells
// whose stencil reaches the array edge use the reference's reduced-order
// coefficient sets.  Tiles are numbered so that consecutive tiles, which
// share halos, run on the same XCD (blocks are dealt round-robin to the 8
// XCDs, each with its own L2).

namespace rsg {

// Row r of a thread's RZ cells, whose first row sits at byte offset cb.
template <bool FAST>
__device__ __forceinline__ float cell_ld(const Buf &b, unsigned cb, int r, int rowb) {
    return FAST ? b.ld(cb, r * rowb) : b.ld(cb + (unsigned)(r * rowb), 0);
}
template <bool FAST>
__device__ __forceinline__ void cell_st(const Buf &b, float v, unsigned cb, int r, int rowb) {
    if (FAST)
        b.st(v, cb, r * rowb);
    else
        b.st(v, cb + (unsigned)(r * rowb), 0);
}

constexpr int BX = 64;
constexpr int BZ = 4;

// Source injection and velocity gathers done by the stress kernel itself
// (insertSource / recordData of the same time step, waves.cpp:407-595,
// 2085-2288), so a modelling step is two launches: each stress-type source
// cell is updated by the thread that owns it, right after its stress update,
// traces in order (the reference's serial order); each velocity receiver is
// gathered by one thread of the block whose tile holds its base cell, from
// that block's LDS tiles of Vx and Vz (the velocities of the step are final
// and the tiles carry a halo on both axes).
constexpr int kFusedSrc = 4, kFusedRec = 2;
struct FusedIO {
    int nsrc;                           // traces injected (0: none)
    int sx[kFusedSrc], sz[kFusedSrc];   // padded base cell; sx < 0: trace skipped
    float wx[kFusedSrc][2], wz[kFusedSrc][2];
    const float *wav;                   // wavelets, src_nt samples per trace
    int src_nt;
    float scale;                        // dt (acoustic: times L) or -dt (elastic explosive)
    int nrec;                           // velocity receiver sets gathered (elastic only)
    int rvz[kFusedRec];                 // 0: Vx, 1: Vz
    const int *rmap[kFusedRec];
    const float *rshift[kFusedRec];
    float *rdata[kFusedRec];
    int rnt[kFusedRec];
    const int *rtile[kFusedRec];        // receivers of tile t: rlist[rtile[t] .. rtile[t+1])
    const int *rlist[kFusedRec];
    const int *it_dev;                  // time index *it_dev + it, or it when null
    int it;
};

struct K2D {
    int nx, nz, lpml, fs;
    int ntx, ntiles;  // tiles along x, total tiles
    int xs[9], xper;  // tiles of XCD k: [xs[k], xs[k+1]); the most any XCD holds
    long sf, sxs, szs;  // per-shot strides of fields, x-side and z-side strips (blockIdx.y = shot)
    float dt, invdx, invdz;
    const float *tab;
    float *P, *Vx, *Vz, *Sxx, *Szz, *Sxz;
    const float *L, *L2M, *M, *Rx, *Rz;
    float *psi[RSG_MAX_STRIPS];
    // Acoustic kernels read the updated arrays' previous values from these
    // (the fields and strips above, unless a step runs out of place: from a
    // checkpoint buffer into the working state).  The velocity step reads
    // P (stencil), Vx, Vz and the P strips here; the stress step reads P and
    // the Vxx / Vzz strips here and the new Vx, Vz from the outputs.
    const float *iP, *iVx, *iVz;
    const float *ipsi[RSG_A2D_NSTRIP];
    Coefs cx, cz;
    FusedIO io;  // stress kernels of single-shot modelling runs only
};

// Injection of the fused sources, run by each thread after it has stored
// its updated stress cells (column x, rows zt .. zt+RZ-1 at byte offset cb,
// rows rowb apart): F += ((scale * modulus) * w) * xtri * ytri at every
// source corner among them, traces in order (the cell was written by this
// same thread, so reading it back is ordered).  Mod null: modulus 1; F2 set:
// the same term into F2 (elastic explosive source, Sxx and Szz).
template <int RZ, bool FAST>
__device__ __forceinline__ void fused_inject(const FusedIO &io, const Buf &bF, const Buf *bF2, const Buf *bMod,
                                          int x, int zt, int nz, unsigned cb, int rowb, int it) {
    for (int i = 0; i < io.nsrc; i++) {
        const int i2 = x - io.sx[i];
        if (io.sx[i] < 0 || i2 < 0 || i2 > 1) continue;
        const float w = io.wav[(size_t)i * io.src_nt + it];
        for (int i1 = 0; i1 < 2; i1++) {
            const int r = io.sz[i] + i1 - zt;
            if (r < 0 || r >= RZ || zt + r >= nz) continue;
            const float m = bMod ? io.scale * cell_ld<FAST>(*bMod, cb, r, rowb) : io.scale;
            const float v = m * w * io.wx[i][i2] * io.wz[i][i1];
            cell_st<FAST>(bF, cell_ld<FAST>(bF, cb, r, rowb) + v, cb, r, rowb);
            if (bF2) cell_st<FAST>(*bF2, cell_ld<FAST>(*bF2, cb, r, rowb) + v, cb, r, rowb);
        }
    }
}

// Velocity gathers of the tile (tile index t, origin x0, z0) from the LDS
// tiles sVx, sVz (origin x0 - H, z0 - H, row length LD): acc = data, then
// acc += xtri * ytri * V over the 2x2 cells in the reference's corner order.
__device__ __forceinline__ void fused_record(const FusedIO &io, int t, int x0, int z0, int H, int LD, int lpml,
                                             const float *sVx, const float *sVz, int it) {
    const int tid = threadIdx.x + threadIdx.y * BX;
    for (int r = 0; r < io.nrec; r++) {
        const int b = io.rtile[r][t], e = io.rtile[r][t + 1];
        const float *V = io.rvz[r] ? sVz : sVx;
        for (int k = b + tid; k < e; k += BX * BZ) {
            const int i = io.rlist[r][k];
            const int px = lpml + io.rmap[r][2 * i] - x0 + H, pz = lpml + io.rmap[r][2 * i + 1] - z0 + H;
            const float sx = io.rshift[r][2 * i], sz = io.rshift[r][2 * i + 1];
            const float xtri[2] = {1 - sx, sx}, ytri[2] = {1 - sz, sz};
            float *d = io.rdata[r] + (size_t)i * io.rnt[r] + it;
            float acc = *d;
#pragma unroll
            for (int i1 = 0; i1 < 2; i1++)
#pragma unroll
                for (int i2 = 0; i2 < 2; i2++) acc += xtri[i2] * ytri[i1] * V[(pz + i1) * LD + px + i2];
            *d = acc;
        }
    }
}

// Tile handled by this block: blocks are dealt round-robin to the 8 XCDs,
// so XCD k runs blocks k, k+8, ...; it gets the contiguous tile range
// [xs[k], xs[k+1]) (tiles that share halos stay in one L2), ranges of equal
// estimated cost (make_k2d), and blocks past its range exit.
__device__ __forceinline__ bool tile_of(const K2D &a, int TZ, int &x0, int &z0) {
    const int k = blockIdx.x % 8;
    const int t = a.xs[k] + (int)(blockIdx.x / 8);
    if (t >= a.xs[k + 1]) return false;
    x0 = (t % a.ntx) * BX;
    z0 = (t / a.ntx) * TZ;
    return true;
}

// A tile is "interior" when it holds no PML cell and no free-surface row and
// its whole halo lies inside the grid: such tiles (most of the grid) take a
// straight-line path with no bounds, edge-stencil or strip tests.
__host__ __device__ inline bool interior_tile(const K2D &a, int x0, int z0, int TZ, int H) {
    const int lo = a.lpml > H ? a.lpml : H;
    return x0 >= lo && x0 + BX + H <= a.nx - a.lpml && z0 >= lo && z0 + TZ + H <= a.nz - a.lpml &&
           !(a.fs && z0 <= a.lpml);
}

// Staging of a (TZ+2HZ) x (64+2HX) window into LDS, in two phases so that
// every global load of a block is in flight before the first LDS write: fetch
// issues all loads into registers (thread row ty takes window rows ty, ty+4,
// ...; columns tx and, for halo columns, tx+64), store writes them out.
// Boundary tiles clamp the address into the grid and select 0 outside it, so
// the loads stay unconditional; zeros outside the grid are never read by the
// reduced edge stencils.
template <int TZ, int HX, int HZ>
struct Tile {
    static constexpr int LD = BX + 2 * HX, ROWS = TZ + 2 * HZ, KR = (ROWS + BZ - 1) / BZ;
    float a[KR], b[KR];
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <int TZ, int HX, int HZ, bool FAST, class F>
__device__ __forceinline__ void tile_fetch(Tile<TZ, HX, HZ> &t, F &&ld, int nx, int nz, int x0, int z0) {
    using T = Tile<TZ, HX, HZ>;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x = x0 - HX + tx;
    if (FAST) {
        // one VGPR offset; rows k*BZ apart as uniform SGPR offsets
        const unsigned v0 = (unsigned)(((z0 - HZ + ty) * nx + x) * 4);
#pragma unroll
        for (int k = 0; k < T::KR; k++) {
            const int soff = k * BZ * nx * 4;
            t.a[k] = ld(v0, soff);
            if (2 * HX > 0) t.b[k] = tx < 2 * HX ? ld(v0 + BX * 4, soff) : 0.f;
        }
    } else {
        // rows outside the grid fall outside the buffer (loads return 0);
        // columns outside it would wrap to a neighbouring row: select 0
        const bool xa = x >= 0 && x < nx, xb = x + BX < nx;
#pragma unroll
        for (int k = 0; k < T::KR; k++) {
            const int z = z0 - HZ + ty + k * BZ;
            const unsigned v = (unsigned)((z * nx + x) * 4);
            const float va = ld(v, 0);
            t.a[k] = xa ? va : 0.f;
            if (2 * HX > 0) {
                const float vb = tx < 2 * HX ? ld(v + BX * 4, 0) : 0.f;
                t.b[k] = xb ? vb : 0.f;
            }
        }
    }
}

template <int TZ, int HX, int HZ>
__device__ __forceinline__ void tile_store(const Tile<TZ, HX, HZ> &t, float *lds) {
    using T = Tile<TZ, HX, HZ>;
    const int tx = threadIdx.x, ty = threadIdx.y;
#pragma unroll
    for (int k = 0; k < T::KR; k++) {
        const int r = ty + k * BZ;
        if (T::ROWS % BZ == 0 || r < T::ROWS) {
            lds[r * T::LD + tx] = t.a[k];
            if (2 * HX > 0 && tx < 2 * HX) lds[r * T::LD + tx + BX] = t.b[k];
        }
    }
}

// Loaders: a plain array, or the products a*b + c*d / a*b that
// WavesElastic2D::backwardstepVelocity differentiates (waves.cpp:1771-1775).
struct Ld1 {
    Buf g;
    __device__ __forceinline__ float operator()(unsigned o, int so) const { return g.ld(o, so); }
};
struct Ld2 {
    Buf a, b;
    __device__ __forceinline__ float operator()(unsigned o, int so) const { return a.ld(o, so) * b.ld(o, so); }
};
struct Ld4 {
    Buf a, b, c, d;
    __device__ __forceinline__ float operator()(unsigned o, int so) const {
        return a.ld(o, so) * b.ld(o, so) + c.ld(o, so) * d.ld(o, so);
    }
};

// Derivative of an LDS tile at a cell: compile-time order in the interior,
// reduced order where the reference's Der shortens the stencil.
template <int ORDER, bool FAST>
__device__ __forceinline__ float ddfw(const float *s, int st, const float *c, int k, float invd) {
    float d = dfw<ORDER>(s, st, c + cind(ORDER));
    if (!FAST && k != ORDER) d = dfw_k(s, st, k, invd);
    return d;
}
template <int ORDER, bool FAST>
__device__ __forceinline__ float ddbw(const float *s, int st, const float *c, int k, float invd) {
    float d = dbw<ORDER>(s, st, c + cind(ORDER));
    if (!FAST && k != ORDER) d = dbw_k(s, st, k, invd);
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

// Position of a cell relative to the strips of one axis.
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
template <bool FAST>
__device__ __forceinline__ void term(float &F, float dtM, float d, const PmlPos &pp, const Tab &tb,
                                     int stag, float *lo, float *hi, size_t sidx) {
    F += dtM * d;
    if (!FAST && pp.side >= 0) {
        float *psi = (pp.side ? hi : lo) + sidx;
        pml_corr(F, dtM, psi, tb.A(pp.side, stag, pp.i), tb.B(pp.side, stag, pp.i),
                 tb.C(pp.side, stag, pp.i), d);
    }
}

// The same with the strip's previous value read from ilo / ihi (equal to
// lo / hi in place).
template <bool FAST>
__device__ __forceinline__ void term(float &F, float dtM, float d, const PmlPos &pp, const Tab &tb,
                                     int stag, float *lo, float *hi, const float *ilo, const float *ihi,
                                     size_t sidx) {
    F += dtM * d;
    if (!FAST && pp.side >= 0) {
        const float A = tb.A(pp.side, stag, pp.i), B = tb.B(pp.side, stag, pp.i), C = tb.C(pp.side, stag, pp.i);
        const float p = B * (pp.side ? ihi : ilo)[sidx] + A * d;
        (pp.side ? hi : lo)[sidx] = p;
        F -= dtM * (p + C * d);
    }
}

// ---------------------------------------------------------------------------
// Acoustic 2D
// ---------------------------------------------------------------------------

// Kernel bodies take the LDS buffers from the kernel so the interior (FAST)
// and boundary instantiations share one allocation.
template <int ORDER, int RZ, bool FAST>
__device__ __forceinline__ void a2d_velocity_body(const K2D &a, int x0, int z0, float *sP) {
    constexpr int H = ORDER, TZ = BZ * RZ, LD = BX + 2 * H;
    const int nx = a.nx, nz = a.nz, tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    const bool xin = FAST || x < nx;
    const size_t fo = blockIdx.y * (size_t)a.sf, xo = blockIdx.y * (size_t)a.sxs,
                 zo = blockIdx.y * (size_t)a.szs;
    const unsigned nb = (unsigned)(nx * nz * 4);
    const int rowb = nx * 4;
    const unsigned cb = (unsigned)(((z0 + ty * RZ) * nx + x) * 4);
    const Buf bVx = mkbuf(a.Vx + fo, nb), bVz = mkbuf(a.Vz + fo, nb), bRx = mkbuf(a.Rx, nb),
              bRz = mkbuf(a.Rz, nb);
    const Buf iVx = mkbuf(a.iVx + fo, nb), iVz = mkbuf(a.iVz + fo, nb);
    float vx[RZ], vz[RZ], rx[RZ], rz[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        vx[r] = cell_ld<FAST>(iVx, cb, r, rowb);
        vz[r] = cell_ld<FAST>(iVz, cb, r, rowb);
        rx[r] = cell_ld<FAST>(bRx, cb, r, rowb);
        rz[r] = cell_ld<FAST>(bRz, cb, r, rowb);
    }
    Tile<TZ, H, H> tp;
    tile_fetch<TZ, H, H, FAST>(tp, Ld1{mkbuf(a.iP + fo, nb)}, nx, nz, x0, z0);
    tile_store(tp, sP);
    __syncthreads();
    if (!xin) return;
    const Tab tb{a.tab, a.lpml};
    const PmlPos px = FAST ? PmlPos{-1, 0} : pml_pos(x, nx, a.lpml);
    const int kx = FAST ? ORDER : kfw(x, nx, ORDER);
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int zl = ty * RZ + r, z = z0 + zl;
        if (!FAST && z >= nz) break;
        const float *s = sP + (zl + H) * LD + tx + H;
        const float dPx = ddfw<ORDER, FAST>(s, 1, a.cx.c, kx, a.invdx);
        const float dPz = ddfw<ORDER, FAST>(s, LD, a.cz.c, FAST ? ORDER : kfw(z, nz, ORDER), a.invdz);
        const PmlPos pz = FAST ? PmlPos{-1, 0} : pml_pos(z, nz, a.lpml);
        term<FAST>(vx[r], a.dt * rx[r], dPx, px, tb, 1, (a.psi[RSG_A2D_P_LEFT] + xo),
                   (a.psi[RSG_A2D_P_RIGHT] + xo), (a.ipsi[RSG_A2D_P_LEFT] + xo), (a.ipsi[RSG_A2D_P_RIGHT] + xo),
                   (size_t)z * a.lpml + px.i);
        term<FAST>(vz[r], a.dt * rz[r], dPz, pz, tb, 1, (a.psi[RSG_A2D_P_TOP] + zo),
                   (a.psi[RSG_A2D_P_BOTTOM] + zo), (a.ipsi[RSG_A2D_P_TOP] + zo), (a.ipsi[RSG_A2D_P_BOTTOM] + zo),
                   (size_t)pz.i * nx + x);
        cell_st<FAST>(bVx, vx[r], cb, r, rowb);
        cell_st<FAST>(bVz, vz[r], cb, r, rowb);
    }
}

template <int ORDER, int RZ>
__global__ __launch_bounds__(BX *BZ) void a2d_velocity_kernel(const K2D a) {
    constexpr int H = ORDER, TZ = BZ * RZ, LD = BX + 2 * H;
    __shared__ float sP[(TZ + 2 * H) * LD];
    int x0, z0;
    if (!tile_of(a, TZ, x0, z0)) return;
    if (interior_tile(a, x0, z0, TZ, H))
        a2d_velocity_body<ORDER, RZ, true>(a, x0, z0, sP);
    else
        a2d_velocity_body<ORDER, RZ, false>(a, x0, z0, sP);
}

template <int ORDER, int RZ, bool FAST, bool IO = true>
__device__ __forceinline__ void a2d_stress_body(const K2D &a, int x0, int z0, float *sVx, float *sVz) {
    constexpr int H = ORDER, LDX = BX + 2 * H, LDZ = BX, TZ = BZ * RZ;
    const int nx = a.nx, nz = a.nz, tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    const bool xin = FAST || x < nx;
    const size_t fo = blockIdx.y * (size_t)a.sf, xo = blockIdx.y * (size_t)a.sxs,
                 zo = blockIdx.y * (size_t)a.szs;
    const unsigned nb = (unsigned)(nx * nz * 4);
    const int rowb = nx * 4;
    const unsigned cb = (unsigned)(((z0 + ty * RZ) * nx + x) * 4);
    const Buf bP = mkbuf(a.P + fo, nb), bL = mkbuf(a.L, nb), iP = mkbuf(a.iP + fo, nb);
    float pv[RZ], lv[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        pv[r] = cell_ld<FAST>(iP, cb, r, rowb);
        lv[r] = cell_ld<FAST>(bL, cb, r, rowb);
    }
    Tile<TZ, H, 0> tvx;
    Tile<TZ, 0, H> tvz;
    tile_fetch<TZ, H, 0, FAST>(tvx, Ld1{mkbuf(a.Vx + fo, nb)}, nx, nz, x0, z0);
    tile_fetch<TZ, 0, H, FAST>(tvz, Ld1{mkbuf(a.Vz + fo, nb)}, nx, nz, x0, z0);
    tile_store(tvx, sVx);
    tile_store(tvz, sVz);
    __syncthreads();
    if (!xin) return;
    const int it = (IO && a.io.nsrc) ? (a.io.it_dev ? *a.io.it_dev + a.io.it : a.io.it) : 0;
    const Tab tb{a.tab, a.lpml};
    const PmlPos px = FAST ? PmlPos{-1, 0} : pml_pos(x, nx, a.lpml);
    const int kx = FAST ? ORDER : kbw(x, nx, ORDER);
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int zl = ty * RZ + r, z = z0 + zl;
        if (!FAST && z >= nz) break;
        const float dVx = ddbw<ORDER, FAST>(sVx + zl * LDX + tx + H, 1, a.cx.c, kx, a.invdx);
        const float dVz = ddbw<ORDER, FAST>(sVz + (zl + H) * LDZ + tx, LDZ, a.cz.c,
                                            FAST ? ORDER : kbw(z, nz, ORDER), a.invdz);
        const PmlPos pz = FAST ? PmlPos{-1, 0} : pml_pos(z, nz, a.lpml);
        const float dtL = a.dt * lv[r];
        term<FAST>(pv[r], dtL, dVx, px, tb, 0, (a.psi[RSG_A2D_VXX_LEFT] + xo), (a.psi[RSG_A2D_VXX_RIGHT] + xo),
                   (a.ipsi[RSG_A2D_VXX_LEFT] + xo), (a.ipsi[RSG_A2D_VXX_RIGHT] + xo), (size_t)z * a.lpml + px.i);
        term<FAST>(pv[r], dtL, dVz, pz, tb, 0, (a.psi[RSG_A2D_VZZ_TOP] + zo), (a.psi[RSG_A2D_VZZ_BOTTOM] + zo),
                   (a.ipsi[RSG_A2D_VZZ_TOP] + zo), (a.ipsi[RSG_A2D_VZZ_BOTTOM] + zo), (size_t)pz.i * nx + x);
        cell_st<FAST>(bP, pv[r], cb, r, rowb);
    }
    if (IO && a.io.nsrc) fused_inject<RZ, FAST>(a.io, bP, nullptr, &bL, x, z0 + ty * RZ, nz, cb, rowb, it);
}

template <int ORDER, int RZ>
__global__ __launch_bounds__(BX *BZ) void a2d_stress_kernel(const K2D a) {
    constexpr int H = ORDER, TZ = BZ * RZ, LDX = BX + 2 * H, LDZ = BX;
    __shared__ float sVx[TZ * LDX];
    __shared__ float sVz[(TZ + 2 * H) * LDZ];
    int x0, z0;
    if (!tile_of(a, TZ, x0, z0)) return;
    if (interior_tile(a, x0, z0, TZ, H))
        a2d_stress_body<ORDER, RZ, true>(a, x0, z0, sVx, sVz);
    else
        a2d_stress_body<ORDER, RZ, false>(a, x0, z0, sVx, sVz);
}

// ---------------------------------------------------------------------------
// Elastic 2D
// ---------------------------------------------------------------------------

// Velocity update shared by the forward (BW=false, plain stresses) and
// adjoint-state (BW=true, moduli-weighted stresses) steps:
//   Vx += dt*Rx*(ddx_fw(Wxx) + ddz_bw(Wxz)),  Vz += dt*Rz*(ddx_bw(Wxz) + ddz_fw(Wzz))
// with Wxx=Sxx, Wxz=Sxz, Wzz=Szz (forward) or Wxx=L2M*Sxx+L*Szz, Wxz=M*Sxz,
// Wzz=L*Sxx+L2M*Szz (adjoint state, waves.cpp:1771-1878).
template <int ORDER, int RZ, bool BW, bool FAST>
__device__ __forceinline__ void e2d_velocity_body(const K2D &a, int x0, int z0, float *sXX,
                                                  float *sZZ, float *sXZ) {
    constexpr int H = ORDER, TZ = BZ * RZ;
    constexpr int LDXX = BX + 2 * H, LDZZ = BX, LDXZ = BX + 2 * H;
    const int nx = a.nx, nz = a.nz, tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    const bool xin = FAST || x < nx;
    const size_t fo = blockIdx.y * (size_t)a.sf, xo = blockIdx.y * (size_t)a.sxs,
                 zo = blockIdx.y * (size_t)a.szs;
    const unsigned nb = (unsigned)(nx * nz * 4);
    const int rowb = nx * 4;
    const unsigned cb = (unsigned)(((z0 + ty * RZ) * nx + x) * 4);
    const Buf bVx = mkbuf(a.Vx + fo, nb), bVz = mkbuf(a.Vz + fo, nb), bRx = mkbuf(a.Rx, nb),
              bRz = mkbuf(a.Rz, nb);
    float vx[RZ], vz[RZ], rx[RZ], rz[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        vx[r] = cell_ld<FAST>(bVx, cb, r, rowb);
        vz[r] = cell_ld<FAST>(bVz, cb, r, rowb);
        rx[r] = cell_ld<FAST>(bRx, cb, r, rowb);
        rz[r] = cell_ld<FAST>(bRz, cb, r, rowb);
    }
    Tile<TZ, H, 0> txx;
    Tile<TZ, 0, H> tzz;
    Tile<TZ, H, H> txz;
    const Buf bSxx = mkbuf(a.Sxx + fo, nb), bSzz = mkbuf(a.Szz + fo, nb), bSxz = mkbuf(a.Sxz + fo, nb);
    if (!BW) {
        tile_fetch<TZ, H, 0, FAST>(txx, Ld1{bSxx}, nx, nz, x0, z0);
        tile_fetch<TZ, 0, H, FAST>(tzz, Ld1{bSzz}, nx, nz, x0, z0);
        tile_fetch<TZ, H, H, FAST>(txz, Ld1{bSxz}, nx, nz, x0, z0);
    } else {
        const Buf bL = mkbuf(a.L, nb), bL2M = mkbuf(a.L2M, nb), bM = mkbuf(a.M, nb);
        tile_fetch<TZ, H, 0, FAST>(txx, Ld4{bL2M, bSxx, bL, bSzz}, nx, nz, x0, z0);
        tile_fetch<TZ, 0, H, FAST>(tzz, Ld4{bL, bSxx, bL2M, bSzz}, nx, nz, x0, z0);
        tile_fetch<TZ, H, H, FAST>(txz, Ld2{bM, bSxz}, nx, nz, x0, z0);
    }
    tile_store(txx, sXX);
    tile_store(tzz, sZZ);
    tile_store(txz, sXZ);
    __syncthreads();
    if (!xin) return;
    const Tab tb{a.tab, a.lpml};
    const PmlPos px = FAST ? PmlPos{-1, 0} : pml_pos(x, nx, a.lpml);
    const int kxf = FAST ? ORDER : kfw(x, nx, ORDER), kxb = FAST ? ORDER : kbw(x, nx, ORDER);
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int zl = ty * RZ + r, z = z0 + zl;
        if (!FAST && z >= nz) break;
        const float dSxx = ddfw<ORDER, FAST>(sXX + zl * LDXX + tx + H, 1, a.cx.c, kxf, a.invdx);
        const float *sxz = sXZ + (zl + H) * LDXZ + tx + H;
        const float dSxz_z = ddbw<ORDER, FAST>(sxz, LDXZ, a.cz.c, FAST ? ORDER : kbw(z, nz, ORDER),
                                               a.invdz);
        const float dSxz_x = ddbw<ORDER, FAST>(sxz, 1, a.cx.c, kxb, a.invdx);
        const float dSzz = ddfw<ORDER, FAST>(sZZ + (zl + H) * LDZZ + tx, LDZZ, a.cz.c,
                                             FAST ? ORDER : kfw(z, nz, ORDER), a.invdz);
        const PmlPos pz = FAST ? PmlPos{-1, 0} : pml_pos(z, nz, a.lpml);
        const size_t ix = (size_t)z * a.lpml + px.i, iz = (size_t)pz.i * nx + x;
        const float dtRx = a.dt * rx[r];
        term<FAST>(vx[r], dtRx, dSxx, px, tb, 1, (a.psi[RSG_E2D_SXX_LEFT] + xo), (a.psi[RSG_E2D_SXX_RIGHT] + xo), ix);
        term<FAST>(vx[r], dtRx, dSxz_z, pz, tb, 0, (a.psi[RSG_E2D_SXZZ_TOP] + zo),
                   (a.psi[RSG_E2D_SXZZ_BOTTOM] + zo), iz);
        cell_st<FAST>(bVx, vx[r], cb, r, rowb);
        const float dtRz = a.dt * rz[r];
        term<FAST>(vz[r], dtRz, dSxz_x, px, tb, 0, (a.psi[RSG_E2D_SXZX_LEFT] + xo),
                   (a.psi[RSG_E2D_SXZX_RIGHT] + xo), ix);
        term<FAST>(vz[r], dtRz, dSzz, pz, tb, 1, (a.psi[RSG_E2D_SZZ_TOP] + zo), (a.psi[RSG_E2D_SZZ_BOTTOM] + zo), iz);
        cell_st<FAST>(bVz, vz[r], cb, r, rowb);
    }
}

template <int ORDER, int RZ, bool BW>
__global__ __launch_bounds__(BX *BZ) void e2d_velocity_kernel(const K2D a) {
    constexpr int H = ORDER, TZ = BZ * RZ;
    __shared__ float sXX[TZ * (BX + 2 * H)];
    __shared__ float sZZ[(TZ + 2 * H) * BX];
    __shared__ float sXZ[(TZ + 2 * H) * (BX + 2 * H)];
    int x0, z0;
    if (!tile_of(a, TZ, x0, z0)) return;
    if (interior_tile(a, x0, z0, TZ, H))
        e2d_velocity_body<ORDER, RZ, BW, true>(a, x0, z0, sXX, sZZ, sXZ);
    else
        e2d_velocity_body<ORDER, RZ, BW, false>(a, x0, z0, sXX, sZZ, sXZ);
}

// Stress update.  Forward (BW=false, waves.cpp:1571-1735):
//   Sxx += dt*(L2M ddx_bw(Vx) + L ddz_bw(Vz)), Szz += dt*(L ddx_bw(Vx) + L2M ddz_bw(Vz)),
//   free surface Szz(z=lpml)=0, Sxz += dt*M*(ddx_fw(Vz) + ddz_fw(Vx)).
// Adjoint state (BW=true, waves.cpp:1907-2080): the same derivatives with
// unit moduli, Sxx only from ddx_bw(Vx) and Szz only from ddz_bw(Vz).
template <int ORDER, int RZ, bool BW, bool FAST, bool IO = true>
__device__ __forceinline__ void e2d_stress_body(const K2D &a, int x0, int z0, float *sVx, float *sVz) {
    constexpr int H = ORDER, TZ = BZ * RZ, LD = BX + 2 * H;
    const int nx = a.nx, nz = a.nz, tx = threadIdx.x, ty = threadIdx.y, x = x0 + tx;
    const bool xin = FAST || x < nx;
    const size_t fo = blockIdx.y * (size_t)a.sf, xo = blockIdx.y * (size_t)a.sxs,
                 zo = blockIdx.y * (size_t)a.szs;
    const unsigned nb = (unsigned)(nx * nz * 4);
    const int rowb = nx * 4;
    const unsigned cb = (unsigned)(((z0 + ty * RZ) * nx + x) * 4);
    const Buf bSxx = mkbuf(a.Sxx + fo, nb), bSzz = mkbuf(a.Szz + fo, nb), bSxz = mkbuf(a.Sxz + fo, nb);
    float sxx[RZ], szz[RZ], sxz[RZ], ml[RZ], ml2m[RZ], mm[RZ];
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        sxx[r] = cell_ld<FAST>(bSxx, cb, r, rowb);
        szz[r] = cell_ld<FAST>(bSzz, cb, r, rowb);
        sxz[r] = cell_ld<FAST>(bSxz, cb, r, rowb);
        if (!BW) {
            ml[r] = cell_ld<FAST>(mkbuf(a.L, nb), cb, r, rowb);
            ml2m[r] = cell_ld<FAST>(mkbuf(a.L2M, nb), cb, r, rowb);
            mm[r] = cell_ld<FAST>(mkbuf(a.M, nb), cb, r, rowb);
        }
    }
    Tile<TZ, H, H> tvx, tvz;
    tile_fetch<TZ, H, H, FAST>(tvx, Ld1{mkbuf(a.Vx + fo, nb)}, nx, nz, x0, z0);
    tile_fetch<TZ, H, H, FAST>(tvz, Ld1{mkbuf(a.Vz + fo, nb)}, nx, nz, x0, z0);
    tile_store(tvx, sVx);
    tile_store(tvz, sVz);
    __syncthreads();
    const int it = (IO && (a.io.nsrc || a.io.nrec)) ? (a.io.it_dev ? *a.io.it_dev + a.io.it : a.io.it) : 0;
    if (IO && !BW && a.io.nrec)
        fused_record(a.io, (z0 / TZ) * a.ntx + x0 / BX, x0, z0, H, LD, a.lpml, sVx, sVz, it);
    if (!xin) return;
    const Tab tb{a.tab, a.lpml};
    const PmlPos px = FAST ? PmlPos{-1, 0} : pml_pos(x, nx, a.lpml);
    const int kxf = FAST ? ORDER : kfw(x, nx, ORDER), kxb = FAST ? ORDER : kbw(x, nx, ORDER);
    const float dt = a.dt;
#pragma unroll
    for (int r = 0; r < RZ; r++) {
        const int zl = ty * RZ + r, z = z0 + zl;
        if (!FAST && z >= nz) break;
        const float *svx = sVx + (zl + H) * LD + tx + H;
        const float *svz = sVz + (zl + H) * LD + tx + H;
        const int kzf = FAST ? ORDER : kfw(z, nz, ORDER), kzb = FAST ? ORDER : kbw(z, nz, ORDER);
        const float dVx_x = ddbw<ORDER, FAST>(svx, 1, a.cx.c, kxb, a.invdx);
        const float dVz_z = ddbw<ORDER, FAST>(svz, LD, a.cz.c, kzb, a.invdz);
        const float dVz_x = ddfw<ORDER, FAST>(svz, 1, a.cx.c, kxf, a.invdx);
        const float dVx_z = ddfw<ORDER, FAST>(svx, LD, a.cz.c, kzf, a.invdz);
        const PmlPos pz = FAST ? PmlPos{-1, 0} : pml_pos(z, nz, a.lpml);
        const size_t ix = (size_t)z * a.lpml + px.i, iz = (size_t)pz.i * nx + x;
        if (!BW) {
            const float dtL2M = dt * ml2m[r], dtL = dt * ml[r];
            // d/dx Vx feeds Sxx (L2M) and Szz (L) through one memory variable
            sxx[r] += dtL2M * dVx_x;
            szz[r] += dtL * dVx_x;
            if (!FAST && px.side >= 0) {
                float *psi = a.psi[px.side ? RSG_E2D_VXX_RIGHT : RSG_E2D_VXX_LEFT] + xo + ix;
                const float A = tb.A(px.side, 0, px.i), B = tb.B(px.side, 0, px.i),
                            C = tb.C(px.side, 0, px.i);
                const float q = B * (*psi) + A * dVx_x;
                *psi = q;
                sxx[r] -= dtL2M * (q + C * dVx_x);
                szz[r] -= dtL * (q + C * dVx_x);
            }
            sxx[r] += dtL * dVz_z;
            szz[r] += dtL2M * dVz_z;
            if (!FAST && pz.side >= 0) {
                float *psi = a.psi[pz.side ? RSG_E2D_VZZ_BOTTOM : RSG_E2D_VZZ_TOP] + zo + iz;
                const float A = tb.A(pz.side, 0, pz.i), B = tb.B(pz.side, 0, pz.i),
                            C = tb.C(pz.side, 0, pz.i);
                const float q = B * (*psi) + A * dVz_z;
                *psi = q;
                sxx[r] -= dtL * (q + C * dVz_z);
                szz[r] -= dtL2M * (q + C * dVz_z);
            }
            if (!FAST && a.fs && z == a.lpml) szz[r] = 0.f;
            const float dtM = dt * mm[r];
            term<FAST>(sxz[r], dtM, dVz_x, px, tb, 1, (a.psi[RSG_E2D_VZX_LEFT] + xo), (a.psi[RSG_E2D_VZX_RIGHT] + xo), ix);
            term<FAST>(sxz[r], dtM, dVx_z, pz, tb, 1, (a.psi[RSG_E2D_VXZ_TOP] + zo), (a.psi[RSG_E2D_VXZ_BOTTOM] + zo), iz);
        } else {
            term<FAST>(sxx[r], dt, dVx_x, px, tb, 0, (a.psi[RSG_E2D_VXX_LEFT] + xo), (a.psi[RSG_E2D_VXX_RIGHT] + xo), ix);
            term<FAST>(szz[r], dt, dVz_z, pz, tb, 0, (a.psi[RSG_E2D_VZZ_TOP] + zo), (a.psi[RSG_E2D_VZZ_BOTTOM] + zo), iz);
            if (!FAST && a.fs && z == a.lpml) szz[r] = 0.f;
            term<FAST>(sxz[r], dt, dVz_x, px, tb, 1, (a.psi[RSG_E2D_VZX_LEFT] + xo), (a.psi[RSG_E2D_VZX_RIGHT] + xo), ix);
            term<FAST>(sxz[r], dt, dVx_z, pz, tb, 1, (a.psi[RSG_E2D_VXZ_TOP] + zo), (a.psi[RSG_E2D_VXZ_BOTTOM] + zo), iz);
        }
        cell_st<FAST>(bSxx, sxx[r], cb, r, rowb);
        cell_st<FAST>(bSzz, szz[r], cb, r, rowb);
        cell_st<FAST>(bSxz, sxz[r], cb, r, rowb);
    }
    // explosive source: Sxx, Szz -= dt * w * xtri * ytri
    if (IO && !BW && a.io.nsrc) fused_inject<RZ, FAST>(a.io, bSxx, &bSzz, nullptr, x, z0 + ty * RZ, nz, cb, rowb, it);
}

template <int ORDER, int RZ, bool BW>
__global__ __launch_bounds__(BX *BZ) void e2d_stress_kernel(const K2D a) {
    constexpr int H = ORDER, TZ = BZ * RZ, LD = BX + 2 * H;
    __shared__ float sVx[(TZ + 2 * H) * LD];
    __shared__ float sVz[(TZ + 2 * H) * LD];
    int x0, z0;
    if (!tile_of(a, TZ, x0, z0)) return;
    if (interior_tile(a, x0, z0, TZ, H))
        e2d_stress_body<ORDER, RZ, BW, true>(a, x0, z0, sVx, sVz);
    else
        e2d_stress_body<ORDER, RZ, BW, false>(a, x0, z0, sVx, sVz);
}

// ---------------------------------------------------------------------------
// Ragged batches: shots on grids of different sizes in one launch
// ---------------------------------------------------------------------------

// The tiles of all shots are numbered consecutively, shot by shot (tile0[k] =
// first tile of shot k for the launch's tile height, tile0[nshot] = total);
// the XCD-contiguous block order of tile_of is kept over that numbering.  A
// block finds its shot by binary search and copies that shot's kernel
// arguments from desc[] (uniform across the block: scalar loads).  The fused
// source/receiver pass is compiled out of these kernels (IO = false): its
// dynamically indexed arrays would put the copied descriptor in private
// (scratch) memory.
__device__ __forceinline__ bool ragged_tile(const K2D *__restrict__ desc, const int *__restrict__ tile0,
                                            int nshot, int TZ, K2D &a, int &x0, int &z0) {
    const int nb = gridDim.x;
    const int per = (nb + 7) / 8;
    const int t = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (t >= tile0[nshot]) return false;
    int lo = 0, hi = nshot - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile0[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    a = desc[lo];
    const int tl = t - tile0[lo];
    x0 = (tl % a.ntx) * BX;
    z0 = (tl / a.ntx) * TZ;
    return true;
}

template <int ORDER, int RZ>
__global__ __launch_bounds__(BX *BZ) void a2d_velocity_ragged_kernel(const K2D *__restrict__ desc,
                                                                      const int *__restrict__ tile0, int nshot) {
    constexpr int H = ORDER, TZ = BZ * RZ, LD = BX + 2 * H;
    __shared__ float sP[(TZ + 2 * H) * LD];
    K2D a;
    int x0, z0;
    if (!ragged_tile(desc, tile0, nshot, TZ, a, x0, z0)) return;
    if (interior_tile(a, x0, z0, TZ, H))
        a2d_velocity_body<ORDER, RZ, true>(a, x0, z0, sP);
    else
        a2d_velocity_body<ORDER, RZ, false>(a, x0, z0, sP);
}

template <int ORDER, int RZ>
__global__ __launch_bounds__(BX *BZ) void a2d_stress_ragged_kernel(const K2D *__restrict__ desc,
                                                                    const int *__restrict__ tile0, int nshot) {
    constexpr int H = ORDER, TZ = BZ * RZ, LD = BX + 2 * H;
    __shared__ float sVx[TZ * LD];
    __shared__ float sVz[(TZ + 2 * H) * BX];
    K2D a;
    int x0, z0;
    if (!ragged_tile(desc, tile0, nshot, TZ, a, x0, z0)) return;
    if (interior_tile(a, x0, z0, TZ, H))
        a2d_stress_body<ORDER, RZ, true, false>(a, x0, z0, sVx, sVz);
    else
        a2d_stress_body<ORDER, RZ, false, false>(a, x0, z0, sVx, sVz);
}

template <int ORDER, int RZ, bool BW>
__global__ __launch_bounds__(BX *BZ) void e2d_velocity_ragged_kernel(const K2D *__restrict__ desc,
                                                                      const int *__restrict__ tile0, int nshot) {
    constexpr int H = ORDER, TZ = BZ * RZ;
    __shared__ float sXX[TZ * (BX + 2 * H)];
    __shared__ float sZZ[(TZ + 2 * H) * BX];
    __shared__ float sXZ[(TZ + 2 * H) * (BX + 2 * H)];
    K2D a;
    int x0, z0;
    if (!ragged_tile(desc, tile0, nshot, TZ, a, x0, z0)) return;
    if (interior_tile(a, x0, z0, TZ, H))
        e2d_velocity_body<ORDER, RZ, BW, true>(a, x0, z0, sXX, sZZ, sXZ);
    else
        e2d_velocity_body<ORDER, RZ, BW, false>(a, x0, z0, sXX, sZZ, sXZ);
}

template <int ORDER, int RZ, bool BW>
__global__ __launch_bounds__(BX *BZ) void e2d_stress_ragged_kernel(const K2D *__restrict__ desc,
                                                                    const int *__restrict__ tile0, int nshot) {
    constexpr int H = ORDER, TZ = BZ * RZ, LD = BX + 2 * H;
    __shared__ float sVx[(TZ + 2 * H) * LD];
    __shared__ float sVz[(TZ + 2 * H) * LD];
    K2D a;
    int x0, z0;
    if (!ragged_tile(desc, tile0, nshot, TZ, a, x0, z0)) return;
    if (interior_tile(a, x0, z0, TZ, H))
        e2d_stress_body<ORDER, RZ, BW, true, false>(a, x0, z0, sVx, sVz);
    else
        e2d_stress_body<ORDER, RZ, BW, false, false>(a, x0, z0, sVx, sVz);
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
    if ((long long)g->nx * g->nz * 4 >= (1LL << 31))
        return fail(RSG_EINVAL, "2D grid larger than 2^29 cells (32-bit buffer offsets)");
    if (!s->tab) return fail(RSG_EINVAL, "PML tables not set");
    return RSG_OK;
}

static K2D make_k2d(const rsg_grid *g, const rsg_state *s, int tz) {
    K2D a;
    a.nx = g->nx;
    a.nz = g->nz;
    a.lpml = g->lpml;
    a.fs = g->fs;
    a.ntx = (g->nx + BX - 1) / BX;
    a.ntiles = a.ntx * ((g->nz + tz - 1) / tz);
    a.sf = a.sxs = a.szs = 0;
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
    a.iP = s->P;
    a.iVx = s->Vx;
    a.iVz = s->Vz;
    for (int i = 0; i < RSG_A2D_NSTRIP; i++) a.ipsi[i] = s->psi[i];
    a.cx = scaled_coefs(g->dx);
    a.cz = scaled_coefs(g->dz);
    a.io = FusedIO{};
    // XCD ranges of equal cost: a tile off the straight-line path (PML, free
    // surface, grid edge) counts pml_cost percent of an interior tile.  The
    // PML rows of the top and bottom all fall to the first and last XCD, so
    // equal tile counts leave those two XCDs the longest.
    {
        const int pc = tuning().pml_cost;
        std::vector<long long> pre(a.ntiles + 1, 0);
        for (int t = 0; t < a.ntiles; t++) {
            const int x0 = (t % a.ntx) * BX, z0 = (t / a.ntx) * tz;
            pre[t + 1] = pre[t] + (interior_tile(a, x0, z0, tz, g->order) ? 100 : pc);
        }
        a.xs[0] = 0;
        a.xper = 0;
        int t = 0;
        for (int k = 1; k <= 8; k++) {
            const long long target = pre[a.ntiles] * k / 8;
            while (t < a.ntiles && pre[t + 1] <= target) t++;
            if (k == 8) t = a.ntiles;
            a.xs[k] = t;
            a.xper = std::max(a.xper, a.xs[k] - a.xs[k - 1]);
        }
    }
    return a;
}

// Rows per thread from the tuning table (1, 2, 4 or 8).
static int rz_of(int v) { return v <= 1 ? 1 : (v <= 2 ? 2 : (v <= 4 ? 4 : 8)); }
static int rz2d() { return rz_of(tuning().rz2d); }
// Ragged batches have their own knob: measured on MI355X over 16 elastic FWI
// windows of 600-2000 x 600 cells, 2 rows per thread is fastest for all four
// half steps (the adjoint-state velocity step, which forms the moduli-weighted
// stresses in its tile loader, 240 us at 2 vs 294 us at 1), while the
// single-shot modelling kernel prefers 1.
static int rz2d_ragged() { return rz_of(tuning().rz2d_ragged); }

static unsigned nblocks(const K2D &a) { return 8u * (unsigned)std::max(a.xper, 1); }

#define RSG_ORDER_CASES(LAUNCH, RZ) \
    switch (g->order) {              \
    case 1: LAUNCH(1, RZ); break;    \
    case 2: LAUNCH(2, RZ); break;    \
    case 3: LAUNCH(3, RZ); break;    \
    case 4: LAUNCH(4, RZ); break;    \
    case 5: LAUNCH(5, RZ); break;    \
    case 6: LAUNCH(6, RZ); break;    \
    case 7: LAUNCH(7, RZ); break;    \
    default: LAUNCH(8, RZ); break;   \
    }

// Launch with the RZ template of the value rzv (1, 2, 4 or 8): the launcher
// must size its grid and tile table for the same value.
#define RSG_DISPATCH_RZ(LAUNCH, rzv)              \
    do {                                          \
        const int rz_ = (rzv);                    \
        if (rz_ == 1) {                           \
            RSG_ORDER_CASES(LAUNCH, 1)            \
        } else if (rz_ == 2) {                    \
            RSG_ORDER_CASES(LAUNCH, 2)            \
        } else if (rz_ == 4) {                    \
            RSG_ORDER_CASES(LAUNCH, 4)            \
        } else {                                  \
            RSG_ORDER_CASES(LAUNCH, 8)            \
        }                                         \
    } while (0)
#define RSG_DISPATCH(LAUNCH) RSG_DISPATCH_RZ(LAUNCH, rz2d())

}  // namespace rsg
