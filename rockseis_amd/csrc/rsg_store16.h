// FP16 storage of wavefield snapshots and checkpoints with a per-snapshot
// power-of-two scale.
//
// Each stored array (one shot's snapshot, or one field / PML strip of one
// shot's checkpoint) gets its own scale 2^(15-E), where amax = m * 2^E,
// m in [0.5, 1), is the array's largest magnitude: the stored values then lie
// in (-32768, 32768), inside FP16's finite range, and the scaling itself is
// exact, so the only rounding is FP32 -> FP16 (round to nearest even,
// relative error <= 2^-11 for values above amax * 2^-29).  An all-zero array
// keeps scale 1; the exponent is capped at 126 so that arrays of tiny values
// (early PML memory variables, denormals) never get an infinite scale.  amax
// is kept as the bit pattern of a non-negative float, which orders like an
// unsigned integer, so it is reduced with atomicMax.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

namespace rsg {

using half_t = _Float16;

// Exponent s of the scale 2^s: 15 - E, capped at 126 so that both 2^s and
// 2^-s are finite normal floats (amax below 2^-111 keeps s = 126).
__device__ inline int store16_exp(unsigned amax_bits) {
    const float a = __uint_as_float(amax_bits);
    if (!(a > 0.f) || !isfinite(a)) return 0;
    int e;
    (void)frexpf(a, &e);
    const int s = 15 - e;
    return s > 126 ? 126 : s;
}

__device__ inline float store16_scale(unsigned amax_bits) { return ldexpf(1.f, store16_exp(amax_bits)); }

__device__ inline float store16_unscale(unsigned amax_bits) { return ldexpf(1.f, -store16_exp(amax_bits)); }

// Contiguous segments: segment s is src[s*seglen .. (s+1)*seglen).  amax
// must hold nseg zeroed entries (pack16_segments zeroes them first).
int pack16_segments(const float *src, half_t *dst, unsigned *amax, size_t seglen, int nseg,
                    hipStream_t st);
int unpack16_segments(const half_t *src, const unsigned *amax, float *dst, size_t seglen, int nseg,
                      hipStream_t st);

// Segments of different lengths: segment s is src[off[s] .. off[s+1]) of one
// flat buffer of total = off[nseg] values (off: nseg+1 device offsets,
// off[0] = 0), swept in chunks of kChunk16 values; cs (device) is
// store16_chunks(off) for the same offsets, nchunk its length; part is
// device scratch of nchunk values.  src/dst must be 16-byte (float) and
// 8-byte (half) aligned.
constexpr int kChunk16 = 4096;
struct Chunk16 {
    int first, last;  // segments touched by the chunk
};
std::vector<Chunk16> store16_chunks(const std::vector<size_t> &off);
int pack16_table(const float *src, half_t *dst, unsigned *amax, unsigned *part, const size_t *off,
                 const Chunk16 *cs, int nchunk, int nseg, size_t total, hipStream_t st);
int unpack16_table(const half_t *src, const unsigned *amax, float *dst, const size_t *off, const Chunk16 *cs,
                   int nchunk, size_t total, hipStream_t st);

// 16-bit block floating point (S16) storage: blocks of kS16Block
// consecutive values share one power-of-two exponent E, the block's largest
// magnitude being m * 2^E with m in [0.5, 1); each value v is kept as the
// signed 16-bit integer rint(v * 2^(15-E)) (|q| <= 32767), so its error is at
// most 2^(E-16), about 1.5e-5 of the block maximum.  A smooth wavefield varies
// little within 256 cells, so most values keep 14-15 significant bits where
// FP16 keeps 11.  An all-zero block stores exponent kS16Zero.
//
// Blocks never straddle two stored arrays: a checkpoint holds fields and PML
// memory variables whose magnitudes differ by many orders, and a block shared
// by the end of one and the start of the next would round the smaller one
// away.  The source is a flat buffer of nseg segments (segment s =
// src[off[s] .. off[s+1])); segment s owns blocks blk[s] .. blk[s+1]-1
// (s16_block_offsets), bseg[b] is block b's segment (s16_block_segments),
// and block b's 256 values sit at q[256 b ..] (the last block of a segment
// padded), its exponent at e[b].
constexpr int kS16Block = 256;
constexpr short kS16Zero = -32768;
inline size_t s16_blocks(size_t n) { return (n + kS16Block - 1) / kS16Block; }
std::vector<size_t> s16_block_offsets(const std::vector<size_t> &off);
std::vector<int> s16_block_segments(const std::vector<size_t> &blk);
// off, blk, bseg: device copies of the tables (nseg + 1, nseg + 1 and nblk
// entries); q holds nblk * kS16Block values, e nblk exponents.
int pack_s16(const float *src, short *q, short *e, const size_t *off, const size_t *blk, const int *bseg,
             size_t nblk, hipStream_t st);
int unpack_s16(const short *q, const short *e, float *dst, const size_t *off, const size_t *blk, const int *bseg,
               size_t nblk, hipStream_t st);

// Storage kinds of checkpoints (the fp16 member of the FWI settings):
// 0 FP32, 1 FP16 with one power-of-two scale per array, 2 S16.
enum { kStoreF32 = 0, kStoreF16 = 1, kStoreS16 = 2 };

// Unpadded nx x nz window of every shot's padded (lpml) field, shots sstride
// apart -> dst [nshot][nz][nx] as `store`: FP32; FP16 with one scale per
// shot in aux = unsigned amax[nshot]; or S16 over the flat snapshot with
// aux = short exponents[s16_blocks(nshot*nz*nx)].
int pack_interior(const float *src, void *dst, void *aux, int nx, int nz, int lpml, size_t sstride,
                  int nshot, int store, hipStream_t st);

// One value of a single-segment S16 buffer (q, exponents e) at flat index i.
__device__ inline float s16_value(const short *q, const short *e, size_t i) {
    const short x = e[i / kS16Block];
    return x == kS16Zero ? 0.f : ldexpf((float)q[i], x - 15);
}

}  // namespace rsg
