// Record resampling on the device (Interp::interp, interp.cpp:110-186):
// trace-major data from[ntr][n1] sampled at o1 + i*d1 -> to[ntr][n2] at
// o2 + j*d2, one thread per output sample.  LINEAR (2 taps), BSPLINE (4 taps)
// and SINC (6 taps with a Lanczos-3 window); output samples whose taps leave
// the input are zero.  The reference's float / double promotions are kept so
// results match its host code.
#include "rsg_common.h"

// This file is compiled with -ffp-contract=off (Makefile): every multiply and
// add is rounded separately, as in the reference's host arithmetic.

namespace rsg {

namespace {

constexpr double kPi = 3.14159265358979323846;  // utils.h:13

__device__ inline float sinc_ref(float x) {
    if (x == 0.f) return 1.f;
    const double px = __dmul_rn(kPi, (double)x);
    return (float)__ddiv_rn(sin(px), px);
}

__global__ __launch_bounds__(256) void interp_kernel(int mode, const float *__restrict__ from, int ntr, int n1,
                                                      float d1, float o1, float *__restrict__ to, int n2, float d2,
                                                      float o2) {
    const int j = blockIdx.x * 256 + threadIdx.x, i = blockIdx.y;
    if (j >= n2 || i >= ntr) return;
    const int pad = mode == 0 ? 1 : (mode == 1 ? 2 : 3);
    // explicitly rounded fp32 operations: no fused multiply-add, as the
    // reference's host build
    const float val = __fadd_rn(o2, __fmul_rn((float)j, d2));
    const long i0 = (long)__fdiv_rn(__fsub_rn(val, o1), d1) - (pad - 1);
    float *out = to + (size_t)i * n2 + j;
    if (i0 < 0 || (i0 + (2 * pad) - 1) > (n1 - 1)) {
        *out = 0.f;
        return;
    }
    const float x0 = __fadd_rn(o1, __fmul_rn((float)(i0 + (pad - 1)), d1));
    const float d = __fdiv_rn(__fsub_rn(val, x0), d1);
    const float *m = from + (size_t)i * n1 + i0;
    if (mode == 0) {
        *out = (float)(__dadd_rn(__dmul_rn(__dsub_rn(1.0, (double)d), (double)m[0]), (double)__fmul_rn(d, m[1])));
    } else if (mode == 1) {
        const double dd = d;
        const float xp1 = (float)(dd + 1.0);
        const float omx = (float)(1.0 - dd);
        const double x = xp1, o = omx;
        const float a0 = (float)__dadd_rn(__dmul_rn(__dsub_rn(__dmul_rn(__dadd_rn(__dmul_rn(-(1.0 / 6.0), x), 1.0), x), 2.0), x),
                                          (4.0 / 3.0));
        const float a1 = (float)__dadd_rn(__dmul_rn(__dmul_rn(__dsub_rn(__dmul_rn(0.5, dd), 1.0), dd), dd), (2.0 / 3.0));
        const float a2 = (float)__dadd_rn(__dmul_rn(__dmul_rn(__dsub_rn(__dmul_rn(0.5, o), 1.0), o), o), (2.0 / 3.0));
        const float a3 = (float)__dsub_rn(__dsub_rn(__dsub_rn(1.0, (double)a0), (double)a1), (double)a2);
        float acc = __fmul_rn(a0, m[0]);
        acc = __fadd_rn(acc, __fmul_rn(a1, m[1]));
        acc = __fadd_rn(acc, __fmul_rn(a2, m[2]));
        *out = __fadd_rn(acc, __fmul_rn(a3, m[3]));
    } else {
        float acc = 0.f;
        for (int k = 0; k < 6; k++) {
            const double u = __dadd_rn(2.0 - k, (double)d);
            const float s1 = sinc_ref((float)u);
            const float s2 = sinc_ref((float)__ddiv_rn(u, 3.0));
            acc = __fadd_rn(acc, __fmul_rn(__fmul_rn(m[k], s1), s2));
        }
        *out = acc;
    }
}

}  // namespace

}  // namespace rsg

extern "C" int rsg_interp(int mode, const float *from, int ntr, int n1, float d1, float o1, float *to, int n2,
                          float d2, float o2, void *stream) {
    if (mode < 0 || mode > 2) return rsg::fail(RSG_EINVAL, "rsg_interp: mode must be 0 (LINEAR), 1 (BSPLINE), 2 (SINC)");
    if (ntr < 0 || n1 < 1 || n2 < 0 || ntr > 65535 || !(d1 > 0.f) || !(d2 > 0.f) || ((!from || !to) && ntr * n2 > 0))
        return rsg::fail(RSG_EINVAL, "rsg_interp: bad sizes or null data");
    if (ntr == 0 || n2 == 0) return RSG_OK;
    rsg::interp_kernel<<<dim3((n2 + 255) / 256, ntr), 256, 0, (hipStream_t)stream>>>(mode, from, ntr, n1, d1, o1, to,
                                                                                      n2, d2, o2);
    RSG_LAUNCH_CHECK("interp_kernel");
    return RSG_OK;
}
