// Device helpers shared by the exact discrete adjoints (adjoint_a2d.hip,
// adjoint_e2d.hip): the transposed staggered derivative at the reduced-order
// grid edges and the reverse of one CFS-PML strip term.
#pragma once
#include "rsg_common.h"

namespace rsg {

// Transposed derivative at index j of an n-long axis, u staged around j at
// stride st: sum_i D[i,j] u_i with each row i at its own stencil order.
// FWD: D = forward difference (row i: +c at i+io+1, -c at i-io, order kfw),
// else backward (row i: +c at i+io, -c at i-io-1, order kbw).
template <bool FWD>
__device__ float dtr_edge(const float *s, int st, int j, int n, int order, float invd) {
    float acc = 0.f;
    for (int io = 0; io < order; io++) {
        if (FWD) {
            const int i1 = j - io - 1, i2 = j + io;
            if (i1 >= 0) {
                const int k = kfw(i1, n, order);
                if (io < k) acc += kCoefC[cind(k) + io] * s[-(io + 1) * st];
            }
            if (i2 < n) {
                const int k = kfw(i2, n, order);
                if (io < k) acc -= kCoefC[cind(k) + io] * s[io * st];
            }
        } else {
            const int i1 = j - io, i2 = j + io + 1;
            if (i1 >= 0) {
                const int k = kbw(i1, n, order);
                if (io < k) acc += kCoefC[cind(k) + io] * s[-io * st];
            }
            if (i2 < n) {
                const int k = kbw(i2, n, order);
                if (io < k) acc -= kCoefC[cind(k) + io] * s[(io + 1) * st];
            }
        }
    }
    return acc * invd;
}

// One strip term of a per-cell pass: adjoint of the derivative, adjoint
// strip update and gradient factor correction (psi_new: forward strip after
// the update, or computed from the old one).
__device__ __forceinline__ void strip_term(float &w, float &f, float dt, float dtM, float lf, float d, float A,
                                           float B, float C, float psi_new, float *lpsi) {
    const float lq = *lpsi;
    w = dtM * (1.f - A - C) * lf + A * lq;
    *lpsi = B * lq - dtM * B * lf;
    f -= dt * (psi_new + C * d);
}

}  // namespace rsg
