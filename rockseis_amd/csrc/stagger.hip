// Model*::staggerModels on the device (ModelAcoustic2D model.cpp:1350-1411,
// ModelElastic2D model.cpp:1965-2040), for callers whose model already lives
// in HBM (the autograd Functions, inversion loops): the same values as the
// host set-up rsg_stagger_a2d / rsg_stagger_e2d, bit for bit, without a
// round trip through host memory.
//
// One thread per padded cell.  Edge padding is a clamp of the unpadded
// index.  The reference staggers in place, one line at a time,
// m[i] = (m[i+1] + m[i]) / 2.0 with the last cell copying its staggered
// neighbour, so the staggered value at i is the mean of the padded inputs
// at i and i+1 (at n-1: at n-2 and n-1).  The elastic shear modulus is
// staggered along z first and then along x, so its value is the mean along
// x of two z-means.  This file is compiled without FMA contraction: every
// product and sum rounds as the host code's does.
#include "rsg_common.h"
#include "rsg_srcrec.h"

namespace rsg {
namespace {

struct PadIdx {
    const float *a;
    int ld, nx, nz, lpml;
    __device__ float operator()(int x, int z) const {
        int i = x - lpml, j = z - lpml;
        i = i < 0 ? 0 : (i >= nx ? nx - 1 : i);
        j = j < 0 ? 0 : (j >= nz ? nz - 1 : j);
        return a[(size_t)j * ld + i];
    }
};

__device__ __forceinline__ float mean2(float a, float b) {
    // (float)((double)(m[i+1] + m[i]) / 2.0), m[i+1] first as the reference adds
    return (float)((double)(b + a) / 2.0);
}

// staggered index pair along an axis of n padded cells: (i, i+1), or
// (n-2, n-1) on the last cell
__device__ __forceinline__ int lo_of(int i, int n) { return i < n - 1 ? i : n - 2; }

__global__ void stagger_a2d_kernel(PadIdx vp, PadIdx rho, int nxp, int nzp, int fs, float *L, float *Rx,
                                   float *Rz, int *bad) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, z = blockIdx.y;
    if (x >= nxp) return;
    const size_t p = (size_t)z * nxp + x;
    const float r = rho(x, z), v = vp(x, z);
    float l = r * v * v;
    const int xl = lo_of(x, nxp), zl = lo_of(z, nzp);
    const float sx = mean2(rho(xl, z), rho(xl + 1, z));
    const float sz = mean2(rho(x, zl), rho(x, zl + 1));
    if (sx == 0.0f || sz == 0.0f) atomicOr(bad, 1);
    float rx = (float)(1.0 / (double)sx);
    const float rz = (float)(1.0 / (double)sz);
    if (fs && z == vp.lpml) {
        rx *= 2.0f;
        l *= 0.0f;
    }
    L[p] = l;
    Rx[p] = rx;
    Rz[p] = rz;
}

__global__ void stagger_e2d_kernel(PadIdx vp, PadIdx vs, PadIdx rho, int nxp, int nzp, int fs, float *L,
                                   float *L2M, float *M, float *Rx, float *Rz, int *bad) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, z = blockIdx.y;
    if (x >= nxp) return;
    const int lpml = vp.lpml;
    const size_t p = (size_t)z * nxp + x;
    const float r = rho(x, z), v = vp(x, z), w = vs(x, z);
    float l2m = r * v * v;
    float l = (float)((double)(r * v * v) - 2.0 * (double)r * (double)w * (double)w);
    // unstaggered shear modulus at (x', z'), with the free-surface row's
    // value as the host code leaves it before staggering
    auto mu = [&](int xx, int zz) {
        const float rr = rho(xx, zz), ww = vs(xx, zz);
        return rr * ww * ww;
    };
    if (fs && z == lpml) {
        l = 0.0f;
        l2m = mu(x, z);
    }
    const int xl = lo_of(x, nxp), zl = lo_of(z, nzp);
    // z-mean first (at columns xl and xl+1), then the x-mean of the two
    const float mz0 = mean2(mu(xl, zl), mu(xl, zl + 1));
    const float mz1 = mean2(mu(xl + 1, zl), mu(xl + 1, zl + 1));
    const float m = mean2(mz0, mz1);
    const float sx = mean2(rho(xl, z), rho(xl + 1, z));
    const float sz = mean2(rho(x, zl), rho(x, zl + 1));
    if (sx == 0.0f || sz == 0.0f) atomicOr(bad, 1);
    float rx = (float)(1.0 / (double)sx);
    const float rz = (float)(1.0 / (double)sz);
    if (fs && z == lpml) rx *= 2.0f;
    L[p] = l;
    L2M[p] = l2m;
    M[p] = m;
    Rx[p] = rx;
    Rz[p] = rz;
}

// One zero-density flag per device, allocated at the first call on it and
// kept for the process (callers in an inversion loop restagger every
// iteration).
int *flag_for_device() {
    static int *flags[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    if (!flags[dev]) {
        int *p = nullptr;
        if (hipMalloc((void **)&p, sizeof(int)) != hipSuccess) return nullptr;
        flags[dev] = p;
    }
    return flags[dev];
}

int finish(int *flag, hipStream_t st) {
    int h = 0;
    RSG_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    RSG_HIP(hipStreamSynchronize(st));
    return h ? fail(RSG_EINVAL, "staggerModels: Zero density found.") : RSG_OK;
}

}  // namespace
}  // namespace rsg

using namespace rsg;

extern "C" int rsg_stagger_a2d_dev(const float *vp, const float *rho, int ld, int nx, int nz, int lpml, int fs,
                                   float *L, float *Rx, float *Rz, void *stream) {
    if (!vp || !rho || !L || !Rx || !Rz || nx < 2 || nz < 2 || lpml < 0 || ld < nx)
        return fail(RSG_EINVAL, "rsg_stagger_a2d_dev: bad arguments");
    const int nxp = nx + 2 * lpml, nzp = nz + 2 * lpml;
    hipStream_t st = (hipStream_t)stream;
    int *flag = flag_for_device();
    if (!flag) return fail(RSG_ENOMEM, "rsg_stagger_a2d_dev: flag allocation");
    if (int rc = launch_set_int(flag, 0, st)) return rc;
    stagger_a2d_kernel<<<dim3((nxp + 255) / 256, nzp), 256, 0, st>>>(
        PadIdx{vp, ld, nx, nz, lpml}, PadIdx{rho, ld, nx, nz, lpml}, nxp, nzp, fs, L, Rx, Rz, flag);
    RSG_LAUNCH_CHECK("stagger_a2d_kernel");
    return finish(flag, st);
}

extern "C" int rsg_stagger_e2d_dev(const float *vp, const float *vs, const float *rho, int ld, int nx, int nz,
                                   int lpml, int fs, float *L, float *L2M, float *M, float *Rx, float *Rz,
                                   void *stream) {
    if (!vp || !vs || !rho || !L || !L2M || !M || !Rx || !Rz || nx < 2 || nz < 2 || lpml < 0 || ld < nx)
        return fail(RSG_EINVAL, "rsg_stagger_e2d_dev: bad arguments");
    const int nxp = nx + 2 * lpml, nzp = nz + 2 * lpml;
    hipStream_t st = (hipStream_t)stream;
    int *flag = flag_for_device();
    if (!flag) return fail(RSG_ENOMEM, "rsg_stagger_e2d_dev: flag allocation");
    if (int rc = launch_set_int(flag, 0, st)) return rc;
    stagger_e2d_kernel<<<dim3((nxp + 255) / 256, nzp), 256, 0, st>>>(
        PadIdx{vp, ld, nx, nz, lpml}, PadIdx{vs, ld, nx, nz, lpml}, PadIdx{rho, ld, nx, nz, lpml}, nxp, nzp, fs, L,
        L2M, M, Rx, Rz, flag);
    RSG_LAUNCH_CHECK("stagger_e2d_kernel");
    return finish(flag, st);
}
