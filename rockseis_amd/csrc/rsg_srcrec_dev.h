// Per-trace device helpers of source injection and receiver gathers
// (WavesAcoustic2D / WavesElastic2D insertSource and recordData,
// waves.cpp:407-595, 2085-2288): bilinear weights xtri = {1-shift, shift}
// over the 2x2 cells at lpml + map.
#pragma once
#include <hip/hip_runtime.h>

namespace rsg {

// F += scale * Mod * wav * xtri * ytri (Mod null: modulus 1).  Corners of
// different traces may coincide, so the add is atomic.
// Corners at flat index >= ncell are skipped (maps of a domain model may
// put a receiver on the last row or column; the reference then writes
// past its array).
__device__ __forceinline__ void inject2d_trace(int i, float *F, const float *Mod, float scale, int nx,
                                               int lpml, const int *map, const float *shift,
                                               const float *wav, int nt, int it, size_t ncell = ~(size_t)0) {
    const int mx = map[2 * i], mz = map[2 * i + 1];
    if (mx < 0 || mz < 0) return;
    const float sx = shift[2 * i], sz = shift[2 * i + 1];
    const float xtri[2] = {1 - sx, sx}, ytri[2] = {1 - sz, sz};
    const float w = wav[(size_t)i * nt + it];
#pragma unroll
    for (int i1 = 0; i1 < 2; i1++)
#pragma unroll
        for (int i2 = 0; i2 < 2; i2++) {
            const size_t p = (size_t)(lpml + mz + i1) * nx + (lpml + mx + i2);
            if (p >= ncell) continue;
            const float m = Mod ? scale * Mod[p] : scale;
            unsafeAtomicAdd(F + p, m * w * xtri[i2] * ytri[i1]);
        }
}

// data[i*nt+it] += sum xtri*ytri*F over the 2x2 cells, in the reference's
// corner order.  With F2 set, the elastic pressure 0.5*(F+F2)/factor is
// accumulated in double as waves.cpp:2234-2237 does.
__device__ __forceinline__ void record2d_trace(int i, const float *F, const float *F2,
                                               const float *factor, int nx, int lpml, const int *map,
                                               const float *shift, float *data, int nt, int it,
                                               size_t ncell = ~(size_t)0) {
    const int mx = map[2 * i], mz = map[2 * i + 1];
    if (mx < 0 || mz < 0) return;
    const float sx = shift[2 * i], sz = shift[2 * i + 1];
    const float xtri[2] = {1 - sx, sx}, ytri[2] = {1 - sz, sz};
    float *d = data + (size_t)i * nt + it;
    float acc = *d;
    if (!F2) {
#pragma unroll
        for (int i1 = 0; i1 < 2; i1++)
#pragma unroll
            for (int i2 = 0; i2 < 2; i2++) {
                const size_t p = (size_t)(lpml + mz + i1) * nx + (lpml + mx + i2);
                if (p < ncell) acc += xtri[i2] * ytri[i1] * F[p];
            }
    } else {
        const double f = factor ? (double)factor[i] : 1.0;
#pragma unroll
        for (int i1 = 0; i1 < 2; i1++)
#pragma unroll
            for (int i2 = 0; i2 < 2; i2++) {
                const size_t p = (size_t)(lpml + mz + i1) * nx + (lpml + mx + i2);
                if (p >= ncell) continue;
                const double w = 0.5 * (double)xtri[i2] * (double)ytri[i1];
                acc = (float)((double)acc + w * (double)F[p] / f);
                acc = (float)((double)acc + w * (double)F2[p] / f);
            }
    }
    *d = acc;
}

}  // namespace rsg
