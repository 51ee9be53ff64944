// Shared internals of librsgpu: error state, coefficient sets, stencil
// helpers and LDS tile loaders used by the 2D and 3D kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "../../include/rsgpu.h"

namespace rsg {

// Offset of source/receiver maps on the grid: lpml, or 0 for a domain model
// whose arrays already hold the padding (Waves*::insertSource/recordData
// take lpml = 0 when getDomdec(), waves.cpp:411-418).
inline int map_pad(const rsg_grid *g) { return g->domdec ? 0 : g->lpml; }

// ---- error state ----------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int check_hip(hipError_t e, const char *what);

#define RSG_HIP(call)                                                           \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) return ::rsg::check_hip(e_, #call);               \
    } while (0)

#define RSG_LAUNCH_CHECK(what)                                                  \
    do {                                                                        \
        hipError_t e_ = hipGetLastError();                                      \
        if (e_ != hipSuccess) return ::rsg::check_hip(e_, what);                \
    } while (0)

// Raw buffer access to one grid array: 32-bit byte offsets in a VGPR plus an
// optional uniform offset in an SGPR, no 64-bit address arithmetic per load;
// loads whose VGPR offset lies outside the array return 0 and never fault
// (the out-of-range check covers the VGPR offset, so boundary tiles put the
// whole offset there; interior tiles, always in range, use the SGPR part).
struct Buf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ float ld(unsigned off, int soff) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, soff, 0));
    }
    __device__ __forceinline__ void st(float v, unsigned off, int soff) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, soff, 0);
    }
};
// Descriptor of `bytes` bytes at p (dword 3 as gfx9 raw buffers take it).
__device__ __forceinline__ Buf mkbuf(const float *p, unsigned bytes) {
    return Buf{__builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), (short)0, (int)bytes, 0x00020000)};
}

// ---- optimised staggered coefficients (der.cpp:3-53) -----------------------
// Set for stencil order k starts at k(k-1)/2.  The reference stores double
// literals into T, except w11 which is the float literal 1.0029f.
#define RSG_DER_TABLE                                                            \
    1.0029f, 1.188401e+0, -7.046382e-2, 1.216624e+0, -9.197724e-2, 1.300041e-2,  \
        1.230862e+0, -1.034123e-1, 2.011671e-2, -3.245760e-3, 1.239407e+0,        \
        -1.105315e-1, 2.496329e-2, -5.804879e-3, 9.358680e-4, 1.245095e+0,        \
        -1.153979e-1, 2.848442e-2, -7.899473e-3, 1.898222e-3, -2.935304e-4,       \
        1.249150e+0, -1.189375e-1, 3.116091e-2, -9.623004e-3, 2.812069e-3,        \
        -6.655106e-4, 9.728191e-5, 1.252186e+0, -1.216289e-1, 3.326513e-2,        \
        -1.105777e-2, 3.647865e-3, -1.066144e-3, 2.437491e-4, -3.351959e-5
constexpr int cind(int k) { return (k * (k - 1)) / 2; }

// Unscaled table in constant memory for the run-time-order edge stencils.
static __constant__ float kCoefC[36] = {RSG_DER_TABLE};

// Coefficients divided by the grid step, one set per axis, passed by value
// in the kernel arguments (read with scalar loads at compile-time indices).
struct Coefs {
    float c[36];
};
Coefs scaled_coefs(float d);

// Stencil order of Der at index i of an n-long axis (forward / backward).
__host__ __device__ inline int kfw(int i, int n, int order) {
    int k = order;
    if (i + 1 < k) k = i + 1;
    if (n - 1 - i < k) k = n - 1 - i;
    return k;
}
__host__ __device__ inline int kbw(int i, int n, int order) {
    int k = order;
    if (i < k) k = i;
    if (n - i < k) k = n - i;
    return k;
}

// Forward staggered difference sum_io c[io]*(s[(io+1)*st] - s[-io*st]) with
// a compile-time order (interior cells) ...
template <int K>
__device__ __forceinline__ float dfw(const float *s, int st, const float *c) {
    float acc = c[0] * (s[st] - s[0]);
#pragma unroll
    for (int io = 1; io < K; io++) acc += c[io] * (s[(io + 1) * st] - s[-io * st]);
    return acc;
}
// ... backward sum_io c[io]*(s[io*st] - s[-(io+1)*st]) ...
template <int K>
__device__ __forceinline__ float dbw(const float *s, int st, const float *c) {
    float acc = c[0] * (s[0] - s[-st]);
#pragma unroll
    for (int io = 1; io < K; io++) acc += c[io] * (s[io * st] - s[-(io + 1) * st]);
    return acc;
}
// ... and with a run-time order k for the reduced edge stencils (k=0 -> 0),
// scaled by invd after the sum (der.cpp divides the sum by the step).
__device__ __forceinline__ float dfw_k(const float *s, int st, int k, float invd) {
    if (k <= 0) return 0.f;
    const float *c = kCoefC + cind(k);
    float acc = c[0] * (s[st] - s[0]);
    for (int io = 1; io < k; io++) acc += c[io] * (s[(io + 1) * st] - s[-io * st]);
    return acc * invd;
}
__device__ __forceinline__ float dbw_k(const float *s, int st, int k, float invd) {
    if (k <= 0) return 0.f;
    const float *c = kCoefC + cind(k);
    float acc = c[0] * (s[0] - s[-st]);
    for (int io = 1; io < k; io++) acc += c[io] * (s[io * st] - s[-(io + 1) * st]);
    return acc * invd;
}

// CFS-PML memory update and field correction (waves.cpp:258-263 pattern):
// psi = B*psi + A*d; F -= dt*M*(psi + C*d)
__device__ __forceinline__ void pml_corr(float &F, float dtM, float *psi, float A, float B, float C,
                                         float d) {
    float p = B * (*psi) + A * d;
    *psi = p;
    F -= dtM * (p + C * d);
}

// PML table rows (pml.h:60-74)
enum { T_A = 0, T_B = 1, T_C = 2, T_STAG = 3, T_RBB = 6 };

// Launch-shape knobs (set through rsg_set_tuning; picked by the autotuner).
struct Tuning {
    int rz2d = 2;   // rows per thread of the 2D tiles: 1, 2, 4 or 8
    int rz2d_ragged = 2;  // the same for the ragged batches (elastic FWI windows)
    int zc3d = 0;   // z planes per 3D block (0: automatic)
    int fuse_io = 1;  // modelling runs: source injection / velocity gathers in the stress kernel
    int pml_cost = 200;  // weight (percent of an interior tile) of a PML / edge tile in the XCD split
    int xcd3d = 1;  // 3D tiles: 1 = each XCD marches its own contiguous group of (x, y) columns, 0 = grid order
};
Tuning &tuning();

// Grid-stride fill used instead of hipMemsetAsync (graph-capture safe).
__global__ void fill_zero_kernel(float *p, size_t n);

int launch_zero(void *ptr, size_t bytes, hipStream_t st);

}  // namespace rsg
