// FP16 snapshot / checkpoint storage with per-snapshot power-of-two scales
// (declarations and the scaling rule in rsg_store16.h), and the FP32 copy of
// the unpadded window that crossCorr reads.
#include <algorithm>
#include <cstdint>

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

// Segments of different lengths, segment s = [off[s], off[s+1]) of one flat
// buffer of `total` values, swept in chunks of kChunk16 values: block c
// covers [c*kChunk16, min((c+1)*kChunk16, total)), which touches segments
// cs[c].first .. cs[c].last.  A chunk inside one segment (nearly all of
// them) is read with float4 loads by every thread and reduced to one
// atomicMax; a chunk across segment boundaries walks the boundaries value by
// value.  The grid therefore scales with the buffer, not with the number of
// segments times the longest segment.
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
constexpr int kVec = kChunk16 / (4 * kThreads);  // float4 loads per thread

// Largest magnitude of a block's values (no atomics): the result is in
// thread 0.
__device__ inline float block_max(float v) {
    __shared__ float part[kThreads / 64];
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    float m = part[0];
    for (int i = 1; i < kThreads / 64; i++) m = fmaxf(m, part[i]);
    return m;
}

// First pass: a chunk inside one segment stores its maximum in part[chunk]
// (reduced per segment by amax_finish_kernel: no contended atomics, a
// segment may span tens of thousands of chunks); a chunk across boundaries
// folds each segment's share into amax[] with atomicMax.
__global__ __launch_bounds__(kThreads) void amax_chunks_kernel(const float *__restrict__ src, unsigned *amax,
                                                                unsigned *__restrict__ part,
                                                                const size_t *__restrict__ off,
                                                                const Chunk16 *__restrict__ cs, size_t total) {
    const size_t c0 = (size_t)blockIdx.x * kChunk16, c1 = min(c0 + (size_t)kChunk16, total);
    const Chunk16 q = cs[blockIdx.x];
    if (q.first == q.last) {
        float m = 0.f;
        if (c1 - c0 == (size_t)kChunk16) {
            float4 v[kVec];
#pragma unroll
            for (int j = 0; j < kVec; j++) v[j] = *(const float4 *)(src + c0 + 4 * (threadIdx.x + j * kThreads));
#pragma unroll
            for (int j = 0; j < kVec; j++)
                m = fmaxf(m, fmaxf(fmaxf(fabsf(v[j].x), fabsf(v[j].y)), fmaxf(fabsf(v[j].z), fabsf(v[j].w))));
        } else {
            for (size_t i = c0 + threadIdx.x; i < c1; i += kThreads) m = fmaxf(m, fabsf(src[i]));
        }
        m = block_max(m);
        if (threadIdx.x == 0) part[blockIdx.x] = __float_as_uint(m);
        return;
    }
    for (int k = q.first; k <= q.last; k++) {
        const size_t b = max(off[k], c0), e = min(off[k + 1], c1);
        float m = 0.f;
        for (size_t i = b + threadIdx.x; i < e; i += kThreads) m = fmaxf(m, fabsf(src[i]));
        block_amax(m, amax + k);
        __syncthreads();  // part[] of block_amax is reused by the next segment
    }
}

// Second pass, one block per segment: the chunks lying wholly inside segment
// s are [ceil(off[s]/kChunk16), floor(off[s+1]/kChunk16)), or up to the last
// (partial) chunk when the segment ends the buffer.
__global__ __launch_bounds__(kThreads) void amax_finish_kernel(unsigned *amax, const unsigned *__restrict__ part,
                                                                const size_t *__restrict__ off, size_t total) {
    const int s = blockIdx.x;
    const size_t nchunk = (total + kChunk16 - 1) / kChunk16;
    const size_t cf = (off[s] + kChunk16 - 1) / kChunk16;
    const size_t cl = off[s + 1] == total ? nchunk : off[s + 1] / kChunk16;
    float m = 0.f;
    for (size_t c = cf + threadIdx.x; c < cl; c += kThreads) m = fmaxf(m, __uint_as_float(part[c]));
    block_amax(m, amax + s);
}

__global__ __launch_bounds__(kThreads) void pack16_chunks_kernel(const float *__restrict__ src,
                                                                  half_t *__restrict__ dst,
                                                                  const unsigned *__restrict__ amax,
                                                                  const size_t *__restrict__ off,
                                                                  const Chunk16 *__restrict__ cs, size_t total) {
    const size_t c0 = (size_t)blockIdx.x * kChunk16, c1 = min(c0 + (size_t)kChunk16, total);
    const Chunk16 q = cs[blockIdx.x];
    if (q.first == q.last && c1 - c0 == (size_t)kChunk16) {
        const float sc = store16_scale(amax[q.first]);
        float4 v[kVec];
#pragma unroll
        for (int j = 0; j < kVec; j++) v[j] = *(const float4 *)(src + c0 + 4 * (threadIdx.x + j * kThreads));
#pragma unroll
        for (int j = 0; j < kVec; j++) {
            half4_t h;
            h.x = (half_t)(v[j].x * sc);
            h.y = (half_t)(v[j].y * sc);
            h.z = (half_t)(v[j].z * sc);
            h.w = (half_t)(v[j].w * sc);
            *(half4_t *)(dst + c0 + 4 * (threadIdx.x + j * kThreads)) = h;
        }
        return;
    }
    int k = q.first;
    for (size_t i = c0 + threadIdx.x; i < c1; i += kThreads) {
        while (off[k + 1] <= i) k++;
        dst[i] = (half_t)(src[i] * store16_scale(amax[k]));
    }
}

__global__ __launch_bounds__(kThreads) void unpack16_chunks_kernel(const half_t *__restrict__ src,
                                                                    const unsigned *__restrict__ amax,
                                                                    float *__restrict__ dst,
                                                                    const size_t *__restrict__ off,
                                                                    const Chunk16 *__restrict__ cs, size_t total) {
    const size_t c0 = (size_t)blockIdx.x * kChunk16, c1 = min(c0 + (size_t)kChunk16, total);
    const Chunk16 q = cs[blockIdx.x];
    if (q.first == q.last && c1 - c0 == (size_t)kChunk16) {
        const float us = store16_unscale(amax[q.first]);
        half4_t h[kVec];
#pragma unroll
        for (int j = 0; j < kVec; j++) h[j] = *(const half4_t *)(src + c0 + 4 * (threadIdx.x + j * kThreads));
#pragma unroll
        for (int j = 0; j < kVec; j++)
            *(float4 *)(dst + c0 + 4 * (threadIdx.x + j * kThreads)) =
                make_float4((float)h[j].x * us, (float)h[j].y * us, (float)h[j].z * us, (float)h[j].w * us);
        return;
    }
    int k = q.first;
    for (size_t i = c0 + threadIdx.x; i < c1; i += kThreads) {
        while (off[k + 1] <= i) k++;
        dst[i] = (float)src[i] * store16_unscale(amax[k]);
    }
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

// S16 blocks: one wave per block of kS16Block values, four values a lane.
constexpr int kS16PerWg = kThreads / 64;  // blocks per workgroup

__device__ inline float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// S16 interior window: one wave per block of kS16Block values of the flat
// [nshot][nz][nx] snapshot; lane l takes values l, l+64, l+128, l+192 of
// the block, so each store instruction writes 64 consecutive values.
__global__ __launch_bounds__(kThreads) void pack_interior_s16_kernel(const float *__restrict__ src,
                                                                      short *__restrict__ q, short *__restrict__ ex,
                                                                      int nx, int nz, int lpml, size_t sstride,
                                                                      size_t n, size_t nblk) {
    const size_t b = (size_t)blockIdx.x * kS16PerWg + (threadIdx.x >> 6);
    if (b >= nblk) return;
    const int lane = threadIdx.x & 63;
    const int nxp = nx + 2 * lpml;
    const size_t nint = (size_t)nx * nz;
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const size_t i = b * kS16Block + lane + 64 * j;
        v[j] = 0.f;
        if (i < n) {
            const size_t k = i / nint, r = i - k * nint;
            const size_t iz = r / nx, ix = r - iz * nx;
            v[j] = src[k * sstride + (iz + lpml) * nxp + ix + lpml];
        }
    }
    float m = fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3])));
    m = wave_max(m);
    int e = 0;
    const bool zero = !(m > 0.f);
    if (!zero) (void)frexpf(m, &e);
    if (lane == 0) ex[b] = zero ? kS16Zero : (short)e;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const size_t i = b * kS16Block + lane + 64 * j;
        if (i < n) {
            const float t = zero ? 0.f : rintf(ldexpf(v[j], 15 - e));
            q[i] = (short)fminf(fmaxf(t, -32767.f), 32767.f);
        }
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

// Block b of a segmented S16 buffer: its segment (from the per-block table),
// first source value and length.
__device__ __forceinline__ void s16_block(const size_t *__restrict__ off, const size_t *__restrict__ blk,
                                          const int *__restrict__ bseg, size_t b, size_t &first, int &len) {
    const int sg = bseg[b];
    first = off[sg] + (b - blk[sg]) * kS16Block;
    const size_t rest = off[sg + 1] - first;
    len = rest < (size_t)kS16Block ? (int)rest : kS16Block;
}

// One wave per block; lane l takes values l, l+64, l+128, l+192 of the block,
// so each load and store instruction covers 64 consecutive values.
__global__ __launch_bounds__(kThreads) void pack_s16_kernel(const float *__restrict__ src, short *__restrict__ q,
                                                             short *__restrict__ ex, const size_t *__restrict__ off,
                                                             const size_t *__restrict__ blk,
                                                             const int *__restrict__ bseg, size_t nblk) {
    const size_t b = (size_t)blockIdx.x * kS16PerWg + (threadIdx.x >> 6);
    if (b >= nblk) return;  // whole waves leave together
    const int lane = threadIdx.x & 63;
    size_t first;
    int len;
    s16_block(off, blk, bseg, b, first, len);
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = lane + 64 * j < len ? src[first + lane + 64 * j] : 0.f;
    float m = fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3])));
    m = wave_max(m);
    int e = 0;
    const bool zero = !(m > 0.f);
    if (!zero) (void)frexpf(m, &e);
    if (lane == 0) ex[b] = zero ? kS16Zero : (short)e;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float t = zero ? 0.f : rintf(ldexpf(v[j], 15 - e));
        q[b * kS16Block + lane + 64 * j] = (short)fminf(fmaxf(t, -32767.f), 32767.f);
    }
}

__global__ __launch_bounds__(kThreads) void unpack_s16_kernel(const short *__restrict__ q,
                                                               const short *__restrict__ ex, float *__restrict__ dst,
                                                               const size_t *__restrict__ off,
                                                               const size_t *__restrict__ blk,
                                                               const int *__restrict__ bseg, size_t nblk) {
    const size_t b = (size_t)blockIdx.x * kS16PerWg + (threadIdx.x >> 6);
    if (b >= nblk) return;
    const int lane = threadIdx.x & 63;
    size_t first;
    int len;
    s16_block(off, blk, bseg, b, first, len);
    const short e = ex[b];
    const int sh = e == kS16Zero ? 0 : e - 15;
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (lane + 64 * j < len)
            dst[first + lane + 64 * j] = e == kS16Zero ? 0.f : ldexpf((float)q[b * kS16Block + lane + 64 * j], sh);
}

}  // namespace

std::vector<size_t> s16_block_offsets(const std::vector<size_t> &off) {
    std::vector<size_t> blk(off.size(), 0);
    for (size_t s = 0; s + 1 < off.size(); s++) blk[s + 1] = blk[s] + s16_blocks(off[s + 1] - off[s]);
    return blk;
}

std::vector<int> s16_block_segments(const std::vector<size_t> &blk) {
    std::vector<int> seg(blk.empty() ? 0 : blk.back());
    for (size_t s = 0; s + 1 < blk.size(); s++)
        for (size_t b = blk[s]; b < blk[s + 1]; b++) seg[b] = (int)s;
    return seg;
}

int pack_s16(const float *src, short *q, short *e, const size_t *off, const size_t *blk, const int *bseg,
             size_t nblk, hipStream_t st) {
    if (nblk == 0) return RSG_OK;
    pack_s16_kernel<<<(unsigned)((nblk + kS16PerWg - 1) / kS16PerWg), kThreads, 0, st>>>(src, q, e, off, blk, bseg,
                                                                                          nblk);
    RSG_LAUNCH_CHECK("pack_s16_kernel");
    return RSG_OK;
}

int unpack_s16(const short *q, const short *e, float *dst, const size_t *off, const size_t *blk, const int *bseg,
               size_t nblk, hipStream_t st) {
    if (nblk == 0) return RSG_OK;
    unpack_s16_kernel<<<(unsigned)((nblk + kS16PerWg - 1) / kS16PerWg), kThreads, 0, st>>>(q, e, dst, off, blk,
                                                                                            bseg, nblk);
    RSG_LAUNCH_CHECK("unpack_s16_kernel");
    return RSG_OK;
}

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

std::vector<Chunk16> store16_chunks(const std::vector<size_t> &off) {
    std::vector<Chunk16> cs;
    if (off.size() < 2) return cs;
    const size_t total = off.back();
    const size_t nchunk = (total + kChunk16 - 1) / kChunk16;
    cs.resize(nchunk);
    int k = 0;
    for (size_t c = 0; c < nchunk; c++) {
        const size_t c0 = c * kChunk16, c1 = std::min(c0 + (size_t)kChunk16, total);
        while (off[k + 1] <= c0) k++;  // segment of the chunk's first value
        int l = k;
        while (off[l + 1] < c1) l++;  // segment of its last value
        cs[c] = Chunk16{k, l};
    }
    return cs;
}

int pack16_table(const float *src, half_t *dst, unsigned *amax, unsigned *part, const size_t *off,
                 const Chunk16 *cs, int nchunk, int nseg, size_t total, hipStream_t st) {
    if (nseg < 1 || nseg > 65535 || nchunk < 1) return fail(RSG_EINVAL, "pack16: bad segment table");
    if (((uintptr_t)src & 15) || ((uintptr_t)dst & 7)) return fail(RSG_EINVAL, "pack16: misaligned buffers");
    if (int rc = zero_u32(amax, nseg, st)) return rc;
    amax_chunks_kernel<<<nchunk, kThreads, 0, st>>>(src, amax, part, off, cs, total);
    RSG_LAUNCH_CHECK("amax_chunks_kernel");
    amax_finish_kernel<<<nseg, kThreads, 0, st>>>(amax, part, off, total);
    RSG_LAUNCH_CHECK("amax_finish_kernel");
    pack16_chunks_kernel<<<nchunk, kThreads, 0, st>>>(src, dst, amax, off, cs, total);
    RSG_LAUNCH_CHECK("pack16_chunks_kernel");
    return RSG_OK;
}

int unpack16_table(const half_t *src, const unsigned *amax, float *dst, const size_t *off, const Chunk16 *cs,
                   int nchunk, size_t total, hipStream_t st) {
    if (nchunk < 1) return fail(RSG_EINVAL, "unpack16: empty segment table");
    if (((uintptr_t)src & 7) || ((uintptr_t)dst & 15)) return fail(RSG_EINVAL, "unpack16: misaligned buffers");
    unpack16_chunks_kernel<<<nchunk, kThreads, 0, st>>>(src, amax, dst, off, cs, total);
    RSG_LAUNCH_CHECK("unpack16_chunks_kernel");
    return RSG_OK;
}

int pack_interior(const float *src, void *dst, void *aux, int nx, int nz, int lpml, size_t sstride,
                  int nshot, int store, hipStream_t st) {
    if (nshot < 1 || nshot > 65535) return fail(RSG_EINVAL, "pack_interior: 1..65535 shots");
    const dim3 g((nx + kThreads - 1) / kThreads, (nz + kRows - 1) / kRows, nshot);
    unsigned *amax = (unsigned *)aux;
    if (store == kStoreS16) {
        const size_t n = (size_t)nx * nz * nshot, nblk = s16_blocks(n);
        pack_interior_s16_kernel<<<(unsigned)((nblk + kS16PerWg - 1) / kS16PerWg), kThreads, 0, st>>>(
            src, (short *)dst, (short *)aux, nx, nz, lpml, sstride, n, nblk);
        RSG_LAUNCH_CHECK("pack_interior_s16_kernel");
    } else if (store == kStoreF16) {
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
