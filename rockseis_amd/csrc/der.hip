// Der<T>::dd{x,y,z}_{fw,bw} (der.cpp:155-540) as a standalone device call:
// the staggered first derivative of one array along one axis into df, with
// the reference's reduced stencil orders at the ends of the axis.  The time
// step kernels fuse these stencils; this entry point serves callers of the
// Der class API that differentiate arrays of their own.
//
// Per cell the sum is formed as der.cpp does, (f[+1]-f[0])*c[0], then
// += (f[io+1]-f[-io])*c[io] (forward) or (f[io]-f[-io-1])*c[io] (backward),
// then divided by the step; this file is compiled without contraction, so
// the result equals the reference's float arithmetic.
#include "rsg_common.h"

namespace rsg {
namespace {

// Stencil order of cell i of an n-long axis, in the reference's loop order
// (the last loop to write a cell wins): forward der.cpp:173-207, backward
// der.cpp:363-397.
__device__ inline int order_fw(int i, int n, int order) {
    if (i == n - 1) return 0;
    if (i >= n - order - 1) return n - 1 - i;
    if (i < order - 1) return i + 1;
    return order;
}
__device__ inline int order_bw(int i, int n, int order) {
    if (i >= n - order) return n - i;
    if (i == 0) return 0;
    if (i < order) return i;
    return order;
}

__global__ void der_kernel(const float *__restrict__ f, float *__restrict__ df, int nx, int ny, int nz, int axis,
                           int bw, float d, int order) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nxy = (size_t)nx * ny, ntot = nxy * nz;
    if (idx >= ntot) return;
    const int ix = (int)(idx % nx), iy = (int)((idx / nx) % ny), iz = (int)(idx / nxy);
    const int i = axis == 0 ? ix : (axis == 1 ? iy : iz);
    const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const long st = axis == 0 ? 1 : (axis == 1 ? (long)nx : (long)nxy);
    const float *s = f + idx;
    const int k = bw ? order_bw(i, n, order) : order_fw(i, n, order);
    if (k == 0) {
        df[idx] = 0.0f;
        return;
    }
    const float *c = kCoefC + cind(k);
    float acc;
    if (bw) {
        acc = (s[0] - s[-st]) * c[0];
        for (int io = 1; io < k; io++) acc += (s[io * st] - s[-(io + 1) * st]) * c[io];
    } else {
        acc = (s[st] - s[0]) * c[0];
        for (int io = 1; io < k; io++) acc += (s[(io + 1) * st] - s[-io * st]) * c[io];
    }
    df[idx] = acc / d;
}

}  // namespace
}  // namespace rsg

extern "C" int rsg_der_apply(int axis, int backward, const float *f, float *df, int nx, int ny, int nz, float d,
                             int order, void *stream) {
    if (!f || !df || nx < 1 || ny < 1 || nz < 1 || axis < 0 || axis > 2 || !(d != 0.0f))
        return rsg::fail(RSG_EINVAL, "rsg_der_apply: bad arguments");
    if (order < 1) order = 1;  // Der clamps the order (der.cpp:94-95)
    if (order > 8) order = 8;
    const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    if (n < 2 * order + 1) return rsg::fail(RSG_EINVAL, "rsg_der_apply: axis shorter than the stencil (2*order+1)");
    const size_t ntot = (size_t)nx * ny * nz;
    rsg::der_kernel<<<(unsigned)((ntot + 255) / 256), 256, 0, (hipStream_t)stream>>>(f, df, nx, ny, nz, axis,
                                                                                   backward ? 1 : 0, d, order);
    RSG_LAUNCH_CHECK("der_kernel");
    return RSG_OK;
}
