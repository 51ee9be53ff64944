// Misfit and residuals of the acoustic FWI gradients on the device
// (FwiAcoustic2D::computeMisfit / computeResiduals, fwi.cpp:574-660 and
// 772-880; FwiAcoustic3D the same with 3D maps, fwi.cpp:1468-1700):
// DIFFERENCE and CORRELATION, with optional data weights.
//
// The per-trace sums of CORRELATION are accumulated by one lane per trace in
// the reference's sample order with explicitly rounded fp32 operations (no
// fused multiply-add), so norms and residuals are those of the reference.
// The 97th-percentile clip of CORRELATION selects the clip value exactly: a
// radix select over the bit patterns of |res| (non-negative floats order as
// unsigned integers), four 8-bit passes per shot.  Misfits are reduced in
// double (the reference sums in T).
#include <vector>

#include "rsg_common.h"
#include "rsg_fwi.h"

// This file is compiled with -ffp-contract=off (Makefile): every multiply and
// add is rounded separately, as in the reference's host arithmetic.

namespace rsg {

namespace {

constexpr int kThreads = 256;

__device__ inline bool trace_valid(const int *map, int dim, size_t tr) {
    const int *m = map + (size_t)dim * tr;
    return m[0] >= 0 && m[1] >= 0 && (dim == 2 || m[2] >= 0);
}

// Norms of CORRELATION, one lane per trace: n1 = |mod|, n2 = |rec| (zero
// replaced by 1), n3 = sum(mod*rec)/(n1*n2).
__global__ __launch_bounds__(kThreads) void corr_norms_kernel(const float *__restrict__ mod,
                                                               const float *__restrict__ rec, int ntr, int nt,
                                                               float *__restrict__ norms) {
    const int tr = blockIdx.x * kThreads + threadIdx.x;
    if (tr >= ntr) return;
    const float *m = mod + (size_t)tr * nt, *r = rec + (size_t)tr * nt;
    float a = 0.f, b = 0.f, c = 0.f;
    for (int it = 0; it < nt; it++) {
        a = __fadd_rn(a, __fmul_rn(m[it], m[it]));
        b = __fadd_rn(b, __fmul_rn(r[it], r[it]));
        c = __fadd_rn(c, __fmul_rn(m[it], r[it]));
    }
    a = __fsqrt_rn(a);
    b = __fsqrt_rn(b);
    if (a == 0.f) a = 1.f;
    if (b == 0.f) b = 1.f;
    c = __fdiv_rn(c, __fmul_rn(a, b));
    norms[3 * (size_t)tr] = a;
    norms[3 * (size_t)tr + 1] = b;
    norms[3 * (size_t)tr + 2] = c;
}

// Misfit (traces with valid maps) and residuals (all traces) of every sample.
__global__ __launch_bounds__(kThreads) void misfit_res_kernel(int type, const float *__restrict__ mod,
                                                               const float *__restrict__ rec,
                                                               const float *__restrict__ wei,
                                                               const float *__restrict__ norms,
                                                               float *__restrict__ res, const int *__restrict__ map,
                                                               int dim, size_t n, int nt, int scaling,
                                                               double *misfit) {
    __shared__ double part[kThreads];
    double acc = 0.0;
    for (size_t p = blockIdx.x * (size_t)kThreads + threadIdx.x; p < n; p += (size_t)gridDim.x * kThreads) {
        const size_t tr = p / nt;
        const float m = mod[p], r = rec[p];
        const bool valid = trace_valid(map, dim, tr);
        if (type == 1) {
            const float n1 = norms[3 * tr], n2 = norms[3 * tr + 1], n3 = norms[3 * tr + 2];
            const float n12 = __fmul_rn(n1, n2);
            if (valid) {
                float v = __fdiv_rn(__fmul_rn(m, r), n12);
                v = -v;
                if (wei) v = __fmul_rn(v, wei[p]);
                acc += v;
            }
            const float d = __fsub_rn(__fdiv_rn(r, n12), __fmul_rn(__fdiv_rn(m, __fmul_rn(n1, n1)), n3));
            float q = (float)((-1.0) * scaling * (double)d);
            if (wei) q = __fmul_rn(q, __fmul_rn(wei[p], wei[p]));
            res[p] = q;
        } else {
            const float dr = __fsub_rn(m, r);
            if (valid) {
                const float v = wei ? __fmul_rn(dr, wei[p]) : dr;
                acc += 0.5 * (double)v * (double)v;
            }
            float q = __fmul_rn((float)scaling, dr);
            if (wei) q = __fmul_rn(q, __fmul_rn(wei[p], wei[p]));
            res[p] = q;
        }
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(misfit, part[0]);
}

// Histogram of the 8-bit digit at `shift` of |res| bit patterns whose higher
// bits equal prefix[shot] (under mask), blockIdx.y = shot.
__global__ __launch_bounds__(kThreads) void radix_hist_kernel(const float *__restrict__ res, size_t nper,
                                                               const unsigned *__restrict__ prefix, unsigned mask,
                                                               int shift, unsigned *hist) {
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const float *r = res + blockIdx.y * nper;
    const unsigned pf = prefix[blockIdx.y];
    for (size_t p = blockIdx.x * (size_t)kThreads + threadIdx.x; p < nper; p += (size_t)gridDim.x * kThreads) {
        const unsigned key = __float_as_uint(r[p]) & 0x7fffffffu;
        if ((key & mask) == pf) atomicAdd(&h[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[blockIdx.y * 256 + threadIdx.x], h[threadIdx.x]);
}

__global__ __launch_bounds__(kThreads) void clip_kernel(float *__restrict__ res, size_t nper,
                                                         const unsigned *__restrict__ pclip) {
    float *r = res + blockIdx.y * nper;
    const unsigned c = pclip[blockIdx.y];
    for (size_t p = blockIdx.x * (size_t)kThreads + threadIdx.x; p < nper; p += (size_t)gridDim.x * kThreads)
        if ((__float_as_uint(r[p]) & 0x7fffffffu) >= c) r[p] = 0.f;
}

__global__ void fill_u32(unsigned *p, int n) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace

int acoustic_misfit_residuals(int type, const float *mod, const float *rec, const float *wei, float *res,
                              const int *map, int dim, int ntr, int nt, int nshot, int scaling, double *misfit,
                              hipStream_t st) {
    if (type != 0 && type != 1)
        return fail(RSG_EINVAL, "fwi: misfit type must be DIFFERENCE (0) or CORRELATION (1)");
    const size_t nper = (size_t)ntr * nt, n = nper * nshot;
    if (n == 0) return RSG_OK;
    float *norms = nullptr;
    unsigned *hist = nullptr, *prefix = nullptr;
    struct Free {
        std::vector<void *> p;
        ~Free() {
            for (void *q : p) (void)hipFree(q);
        }
    } fr;
    if (type == 1) {
        if (hipMalloc((void **)&norms, sizeof(float) * 3 * (size_t)ntr * nshot) != hipSuccess ||
            hipMalloc((void **)&hist, sizeof(unsigned) * 256 * nshot) != hipSuccess ||
            hipMalloc((void **)&prefix, sizeof(unsigned) * nshot) != hipSuccess)
            return fail(RSG_ENOMEM, "fwi: misfit scratch allocation failed");
        fr.p = {norms, hist, prefix};
        const int nt_all = ntr * nshot;
        corr_norms_kernel<<<(nt_all + kThreads - 1) / kThreads, kThreads, 0, st>>>(mod, rec, nt_all, nt, norms);
        RSG_LAUNCH_CHECK("corr_norms_kernel");
    }
    misfit_res_kernel<<<1024, kThreads, 0, st>>>(type, mod, rec, wei, norms, res, map, dim, n, nt, scaling, misfit);
    RSG_LAUNCH_CHECK("misfit_res_kernel");
    if (type != 1) return RSG_OK;
    // 97th-percentile clip per shot: element (int)(97*nt*ntr/100) of the sorted |res|
    const long long kth0 = (97LL * nt * ntr) / 100;
    std::vector<unsigned long long> k(nshot, (unsigned long long)kth0);
    std::vector<unsigned> pf(nshot, 0u), h(256 * (size_t)nshot);
    unsigned blocks = (unsigned)std::min<size_t>((nper + kThreads * 8 - 1) / (kThreads * 8), 1024);
    if (blocks < 1) blocks = 1;
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 24 - 8 * pass;
        const unsigned mask = pass == 0 ? 0u : (0xffffffffu << (32 - 8 * pass));
        RSG_HIP(hipMemcpyAsync(prefix, pf.data(), sizeof(unsigned) * nshot, hipMemcpyHostToDevice, st));
        fill_u32<<<(256 * nshot + kThreads - 1) / kThreads, kThreads, 0, st>>>(hist, 256 * nshot);
        RSG_LAUNCH_CHECK("fill_u32");
        radix_hist_kernel<<<dim3(blocks, nshot), kThreads, 0, st>>>(res, nper, prefix, mask, shift, hist);
        RSG_LAUNCH_CHECK("radix_hist_kernel");
        RSG_HIP(hipMemcpyAsync(h.data(), hist, sizeof(unsigned) * 256 * nshot, hipMemcpyDeviceToHost, st));
        RSG_HIP(hipStreamSynchronize(st));
        for (int s = 0; s < nshot; s++) {
            unsigned long long cum = 0;
            int b = 0;
            for (; b < 256; b++) {
                const unsigned long long c = h[(size_t)s * 256 + b];
                if (cum + c > k[s]) break;
                cum += c;
            }
            if (b == 256) b = 255;  // unreachable while k < nper
            pf[s] |= (unsigned)b << shift;
            k[s] -= cum;
        }
    }
    RSG_HIP(hipMemcpyAsync(prefix, pf.data(), sizeof(unsigned) * nshot, hipMemcpyHostToDevice, st));
    clip_kernel<<<dim3(blocks, nshot), kThreads, 0, st>>>(res, nper, prefix);
    RSG_LAUNCH_CHECK("clip_kernel");
    RSG_HIP(hipStreamSynchronize(st));
    return RSG_OK;
}

}  // namespace rsg
