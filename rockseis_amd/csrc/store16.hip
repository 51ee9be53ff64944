// FP16 snapshot / checkpoint storage with per-snapshot power-of-two scales
// (declarations and the scaling rule in rsg_store16.h), and the FP32 copy of
// the unpadded window that crossCorr reads.
#include "rsg_common.h"
#include "rsg_store16.h"

namespace rsg {

namespace {

constexpr int kThreads = 256;

// Largest magnitude of this thread's values, reduced over the block, then
// one atomicMax per block.
__device__ inline void block_amax(float v, unsigned *out) {
    __shared__ float part[kThreads / 64];
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = part[0];
        for (int i = 1; i < kThreads / 64; i++) m = fmaxf(m, part[i]);
        atomicMax(out, __float_as_uint(m));
    }
}

__global__ __launch_bounds__(kThreads) void fill_u32_kernel(unsigned *p, int n) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ __launch_bounds__(kThreads) void amax_segments_kernel(const float *__restrict__ src,
                                                                  unsigned *amax, size_t seglen) {
    const float *s = src + blockIdx.y * seglen;
    float m = 0.f;
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < seglen; i += (size_t)gridDim.x * kThreads)
        m = fmaxf(m, fabsf(s[i]));
    block_amax(m, amax + blockIdx.y);
}

__global__ __launch_bounds__(kThreads) void pack16_segments_kernel(const float *__restrict__ src,
                                                                    half_t *__restrict__ dst,
                                                                    const unsigned *__restrict__ amax,
                                                                    size_t seglen) {
    const size_t o = blockIdx.y * seglen;
    const float sc = store16_scale(amax[blockIdx.y]);
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < seglen; i += (size_t)gridDim.x * kThreads)
        dst[o + i] = (half_t)(src[o + i] * sc);
}

__global__ __launch_bounds__(kThreads) void unpack16_segments_kernel(const half_t *__restrict__ src,
                                                                      const unsigned *__restrict__ amax,
                                                                      float *__restrict__ dst, size_t seglen) {
    const size_t o = blockIdx.y * seglen;
    const float us = store16_unscale(amax[blockIdx.y]);
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < seglen; i += (size_t)gridDim.x * kThreads)
        dst[o + i] = (float)src[o + i] * us;
}

// Segments of different lengths: segment s = [off[s], off[s+1]).
__global__ __launch_bounds__(kThreads) void amax_table_kernel(const float *__restrict__ src, unsigned *amax,
                                                               const size_t *__restrict__ off) {
    const size_t o = off[blockIdx.y], n = off[blockIdx.y + 1] - o;
    float m = 0.f;
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads)
        m = fmaxf(m, fabsf(src[o + i]));
    block_amax(m, amax + blockIdx.y);
}

__global__ __launch_bounds__(kThreads) void pack16_table_kernel(const float *__restrict__ src,
                                                                 half_t *__restrict__ dst,
                                                                 const unsigned *__restrict__ amax,
                                                                 const size_t *__restrict__ off) {
    const size_t o = off[blockIdx.y], n = off[blockIdx.y + 1] - o;
    const float sc = store16_scale(amax[blockIdx.y]);
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads)
        dst[o + i] = (half_t)(src[o + i] * sc);
}

__global__ __launch_bounds__(kThreads) void unpack16_table_kernel(const half_t *__restrict__ src,
                                                                   const unsigned *__restrict__ amax,
                                                                   float *__restrict__ dst,
                                                                   const size_t *__restrict__ off) {
    const size_t o = off[blockIdx.y], n = off[blockIdx.y + 1] - o;
    const float us = store16_unscale(amax[blockIdx.y]);
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads)
        dst[o + i] = (float)src[o + i] * us;
}

// Interior window: blockIdx.x covers 256 columns, blockIdx.y 8 rows,
// blockIdx.z the shot.
constexpr int kRows = 8;

__global__ __launch_bounds__(kThreads) void amax_interior_kernel(const float *__restrict__ src,
                                                                  unsigned *amax, int nx, int nz, int lpml,
                                                                  size_t sstride) {
    const int ix = blockIdx.x * kThreads + threadIdx.x, k = blockIdx.z;
    const int nxp = nx + 2 * lpml;
    float m = 0.f;
    if (ix < nx) {
        const float *s = src + k * sstride + (size_t)lpml * nxp + lpml + ix;
        const int z0 = blockIdx.y * kRows, z1 = min(z0 + kRows, nz);
        for (int iz = z0; iz < z1; iz++) m = fmaxf(m, fabsf(s[(size_t)iz * nxp]));
    }
    block_amax(m, amax + k);
}

template <bool FP16>
__global__ __launch_bounds__(kThreads) void pack_interior_kernel(const float *__restrict__ src,
                                                                  void *__restrict__ dst_,
                                                                  const unsigned *__restrict__ amax,
                                                                  int nx, int nz, int lpml, size_t sstride) {
    const int ix = blockIdx.x * kThreads + threadIdx.x, k = blockIdx.z;
    if (ix >= nx) return;
    const int nxp = nx + 2 * lpml;
    const float *s = src + k * sstride + (size_t)lpml * nxp + lpml + ix;
    const size_t d0 = (size_t)k * nz * nx + ix;
    const int z0 = blockIdx.y * kRows, z1 = min(z0 + kRows, nz);
    if constexpr (FP16) {
        half_t *dst = (half_t *)dst_;
        const float sc = store16_scale(amax[k]);
        for (int iz = z0; iz < z1; iz++) dst[d0 + (size_t)iz * nx] = (half_t)(s[(size_t)iz * nxp] * sc);
    } else {
        float *dst = (float *)dst_;
        for (int iz = z0; iz < z1; iz++) dst[d0 + (size_t)iz * nx] = s[(size_t)iz * nxp];
    }
}

int zero_u32(unsigned *p, int n, hipStream_t st) {
    fill_u32_kernel<<<(n + kThreads - 1) / kThreads, kThreads, 0, st>>>(p, n);
    RSG_LAUNCH_CHECK("fill_u32_kernel");
    return RSG_OK;
}

dim3 seg_grid(size_t seglen, int nseg) {
    size_t bx = (seglen + kThreads * 4 - 1) / (kThreads * 4);
    if (bx > 1024) bx = 1024;
    if (bx < 1) bx = 1;
    return dim3((unsigned)bx, (unsigned)nseg);
}

}  // namespace

int pack16_segments(const float *src, half_t *dst, unsigned *amax, size_t seglen, int nseg, hipStream_t st) {
    if (nseg < 1 || nseg > 65535) return fail(RSG_EINVAL, "pack16: 1..65535 segments");
    if (int rc = zero_u32(amax, nseg, st)) return rc;
    const dim3 g = seg_grid(seglen, nseg);
    amax_segments_kernel<<<g, kThreads, 0, st>>>(src, amax, seglen);
    RSG_LAUNCH_CHECK("amax_segments_kernel");
    pack16_segments_kernel<<<g, kThreads, 0, st>>>(src, dst, amax, seglen);
    RSG_LAUNCH_CHECK("pack16_segments_kernel");
    return RSG_OK;
}

int unpack16_segments(const half_t *src, const unsigned *amax, float *dst, size_t seglen, int nseg,
                      hipStream_t st) {
    if (nseg < 1 || nseg > 65535) return fail(RSG_EINVAL, "unpack16: 1..65535 segments");
    unpack16_segments_kernel<<<seg_grid(seglen, nseg), kThreads, 0, st>>>(src, amax, dst, seglen);
    RSG_LAUNCH_CHECK("unpack16_segments_kernel");
    return RSG_OK;
}

int pack16_table(const float *src, half_t *dst, unsigned *amax, const size_t *off, int nseg, size_t maxlen,
                 hipStream_t st) {
    if (nseg < 1 || nseg > 65535) return fail(RSG_EINVAL, "pack16: 1..65535 segments");
    if (int rc = zero_u32(amax, nseg, st)) return rc;
    const dim3 g = seg_grid(maxlen, nseg);
    amax_table_kernel<<<g, kThreads, 0, st>>>(src, amax, off);
    RSG_LAUNCH_CHECK("amax_table_kernel");
    pack16_table_kernel<<<g, kThreads, 0, st>>>(src, dst, amax, off);
    RSG_LAUNCH_CHECK("pack16_table_kernel");
    return RSG_OK;
}

int unpack16_table(const half_t *src, const unsigned *amax, float *dst, const size_t *off, int nseg,
                   size_t maxlen, hipStream_t st) {
    if (nseg < 1 || nseg > 65535) return fail(RSG_EINVAL, "unpack16: 1..65535 segments");
    unpack16_table_kernel<<<seg_grid(maxlen, nseg), kThreads, 0, st>>>(src, amax, dst, off);
    RSG_LAUNCH_CHECK("unpack16_table_kernel");
    return RSG_OK;
}

int pack_interior(const float *src, void *dst, unsigned *amax, int nx, int nz, int lpml, size_t sstride,
                  int nshot, bool fp16, hipStream_t st) {
    if (nshot < 1 || nshot > 65535) return fail(RSG_EINVAL, "pack_interior: 1..65535 shots");
    const dim3 g((nx + kThreads - 1) / kThreads, (nz + kRows - 1) / kRows, nshot);
    if (fp16) {
        if (int rc = zero_u32(amax, nshot, st)) return rc;
        amax_interior_kernel<<<g, kThreads, 0, st>>>(src, amax, nx, nz, lpml, sstride);
        RSG_LAUNCH_CHECK("amax_interior_kernel");
        pack_interior_kernel<true><<<g, kThreads, 0, st>>>(src, dst, amax, nx, nz, lpml, sstride);
        RSG_LAUNCH_CHECK("pack_interior_kernel<fp16>");
    } else {
        pack_interior_kernel<false><<<g, kThreads, 0, st>>>(src, dst, nullptr, nx, nz, lpml, sstride);
        RSG_LAUNCH_CHECK("pack_interior_kernel<fp32>");
    }
    return RSG_OK;
}

}  // namespace rsg
