// Misfit and residuals of the acoustic FWI gradients on the device
// (FwiAcoustic2D::computeMisfit / computeResiduals, fwi.cpp:574-1027;
// FwiAcoustic3D the same with 3D maps, fwi.cpp:1468-1900): DIFFERENCE,
// CORRELATION, ADAPTIVE_GAUSS and ADAPTIVE_LINEAR, with optional data weights.
//
// The per-trace sums of CORRELATION are accumulated by one lane per trace in
// the reference's sample order with explicitly rounded fp32 operations (no
// fused multiply-add), so norms and residuals are those of the reference.
// The 97th-percentile clip of CORRELATION selects the clip value exactly: a
// radix select over the bit patterns of |res| (non-negative floats order as
// unsigned integers), four 8-bit passes per shot.  Misfits are reduced in
// double (the reference sums in T).
#include <cmath>
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
    // correctly rounded float square roots (the reference's sqrt on T):
    // through the double square root, since __fsqrt_rn is the native
    // (approximate) instruction unless OCML's rounded operations are enabled
    a = (float)sqrt((double)a);
    b = (float)sqrt((double)b);
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

// ---- ADAPTIVE misfits (fwi.cpp:152-420, 647-736, 851-1014) ----------------
// Per trace: Wiener shaping filter of the weighted record onto the weighted
// model trace (autocorrelation and crosscorrelation, 1% white noise on the
// zero lag, Levinson solve), time-weighted by H (Gaussian or linear ramp,
// tabulated on the host in the reference's double arithmetic); residual =
// the shaping filter of (H^2 - 2 misfit) shaper / norm convolved with the
// weighted record.  One lane per trace runs the reference's sequential
// float recursions in their order (Hale's pairwise convolution, Claerbout's
// Levinson recursion), so results equal the reference's bit for bit; the
// 97th-percentile clip follows as for CORRELATION.

// z = x convolved with y ([ifx, ifx+lx), [ify, ify+ly) -> [ifz, ifz+lz)),
// pairwise scheme; reads past an array end that the scheme never uses are
// not made.
__device__ void conv_dev(int lx, int ifx, const float *x, int ly, int ify, const float *y, int lz, int ifz,
                         float *z) {
    int ilx = ifx + lx - 1, ily = ify + ly - 1, ilz = ifz + lz - 1;
    int i, j, ilow, ihigh, jlow, jhigh, t;
    float sa, sb, xa, xb, ya, yb;
    if (lx > ly) {
        const float *tp;
        t = ifx; ifx = ify; ify = t;
        t = ilx; ilx = ily; ily = t;
        t = lx; lx = ly; ly = t;
        tp = x; x = y; y = tp;
    }
    x -= ifx;
    y -= ify;
    z -= ifz;
    ilow = ifz;
    ihigh = ify + ifx - 1;
    if (ihigh > ilz) ihigh = ilz;
    for (i = ilow; i <= ihigh; ++i) z[i] = 0.f;
    if (i < ify + ilx && i <= ilz) {
        jlow = ifx;
        jhigh = i - ify;
        if ((jhigh - jlow) % 2) {
            sa = 0.f;
            for (j = jlow; j <= jhigh; ++j) sa += x[j] * y[i - j];
            z[i++] = sa;
        }
    }
    ilow = i;
    ihigh = ilx + ify - 1;
    if (ihigh > ilz) ihigh = ilz;
    jlow = ifx;
    jhigh = ilow - ify;
    for (i = ilow; i < ihigh; i += 2, jhigh += 2) {
        sa = sb = 0.f;
        xb = x[jhigh + 1];
        yb = 0.f;
        for (j = jhigh; j >= jlow; j -= 2) {
            sa += xb * yb;
            ya = y[i - j];
            sb += xb * ya;
            xa = x[j];
            sa += xa * ya;
            yb = y[i + 1 - j];
            sb += xa * yb;
            if (j - 1 >= jlow) xb = x[j - 1];
        }
        z[i] = sa;
        z[i + 1] = sb;
    }
    if (i == ihigh) {
        jlow = ifx;
        jhigh = i - ify;
        sa = 0.f;
        for (j = jlow; j <= jhigh; ++j) sa += x[j] * y[i - j];
        z[i++] = sa;
    }
    ilow = i;
    ihigh = ily + ifx;
    if (ihigh > ilz) ihigh = ilz;
    jlow = ifx;
    jhigh = ilx;
    for (i = ilow; i < ihigh; i += 2) {
        sa = sb = 0.f;
        yb = y[i + 1 - jlow];
        xa = x[jlow];
        for (j = jlow; j < jhigh; j += 2) {
            sb += xa * yb;
            ya = y[i - j];
            sa += xa * ya;
            xb = x[j + 1];
            sb += xb * ya;
            yb = y[i - 1 - j];
            sa += xb * yb;
            if (j + 2 <= jhigh) xa = x[j + 2];
        }
        if ((jhigh - jlow) % 2) {
            z[i] = sa;
            z[i + 1] = sb;
        } else {
            z[i] = sa + x[jhigh] * y[i - jhigh];
            z[i + 1] = sb + x[jhigh] * y[i + 1 - jhigh];
        }
    }
    if (i == ihigh) {
        sa = 0.f;
        for (j = jlow; j <= jhigh; ++j) sa += x[j] * y[i - j];
        z[i++] = sa;
    }
    if (i <= ily + ilx && i <= ilz) {
        jlow = i - ily;
        jhigh = ilx;
        if (!((jhigh - jlow) % 2)) {
            sa = 0.f;
            for (j = jlow; j <= jhigh; ++j) sa += x[j] * y[i - j];
            z[i++] = sa;
        }
    }
    ilow = i;
    ihigh = ily + ilx;
    if (ihigh > ilz) ihigh = ilz;
    jlow = ilow - ily;
    jhigh = ilx - 2;
    for (i = ilow; i < ihigh; i += 2, jlow += 2) {
        sa = sb = 0.f;
        xa = x[jlow];
        yb = 0.f;
        for (j = jlow; j < jhigh; j += 2) {
            sb += xa * yb;
            ya = y[i - j];
            sa += xa * ya;
            xb = x[j + 1];
            sb += xb * ya;
            yb = y[i - 1 - j];
            sa += xb * yb;
            xa = x[j + 2];
        }
        sb += xa * yb;
        ya = y[i - j];
        sa += xa * ya;
        xb = x[j + 1];
        sb += xb * ya;
        yb = y[i - 1 - j];
        sa += xb * yb;
        z[i] = sa;
        z[i + 1] = sb;
    }
    if (i == ihigh) {
        jlow = i - ily;
        jhigh = ilx;
        sa = 0.f;
        for (j = jlow; j <= jhigh; ++j) sa += x[j] * y[i - j];
        z[i++] = sa;
    }
    for (; i <= ilz; ++i) z[i] = 0.f;
}

// Levinson solve of the symmetric Toeplitz system R f = g (first row r).
__device__ void stoep_dev(int n, const float *r, const float *g, float *f, float *a) {
    float v, e, c, w, bot;
    if (r[0] == 0.f) return;
    a[0] = 1.f;
    v = r[0];
    f[0] = __fdiv_rn(g[0], r[0]);
    for (int j = 1; j < n; j++) {
        int i;
        a[j] = 0.f;
        f[j] = 0.f;
        for (i = 0, e = 0.f; i < j; i++) e += a[i] * r[j - i];
        c = __fdiv_rn(e, v);
        v -= c * e;
        for (i = 0; i <= j / 2; i++) {
            bot = a[j - i] - c * a[i];
            a[i] -= c * a[j - i];
            a[j - i] = bot;
        }
        for (i = 0, w = 0.f; i < j; i++) w += f[i] * r[j - i];
        c = __fdiv_rn(w - g[j], v);
        for (i = 0; i <= j; i++) f[i] -= c * a[j - i];
    }
}

// One lane per trace (shots' traces back to back).  scratch: 7 nt floats per
// trace (weighted model, weighted record, autocorrelation, crosscorrelation,
// Levinson scratch, reversed copy, shaper).
__global__ __launch_bounds__(64) void adaptive_kernel(int type, const float *__restrict__ mod,
                                                      const float *__restrict__ rec, const float *__restrict__ wei,
                                                      float *__restrict__ res, const int *__restrict__ map, int dim,
                                                      int ntr_all, int nt, int scaling, const float *__restrict__ H,
                                                      float *__restrict__ scratch, double *misfit) {
    const int tr = blockIdx.x * blockDim.x + threadIdx.x;
    if (tr >= ntr_all) return;
    const size_t o = (size_t)tr * nt;
    const float *m = mod + o, *r = rec + o, *wt = wei ? wei + o : nullptr;
    float *w = scratch + (size_t)tr * 7 * nt;
    float *modwei = w, *recwei = w + nt, *autoc = w + 2 * nt, *cross = w + 3 * nt, *spiker = w + 4 * nt,
          *xr = w + 5 * nt, *shaper = w + 6 * nt, *rs = res + o;
    for (int it = 0; it < nt; it++) {
        modwei[it] = wt ? m[it] * wt[it] : m[it];
        recwei[it] = wt ? r[it] * wt[it] : r[it];
    }
    // autocorrelation and crosscorrelation: the record reversed, convolved
    for (int i = 0, j = nt - 1; i < nt; ++i, --j) xr[i] = recwei[j];
    conv_dev(nt, 1 - nt, xr, nt, 0, recwei, nt, 0, autoc);
    conv_dev(nt, 1 - nt, xr, nt, 0, modwei, nt, 0, cross);
    if (autoc[0] == 0.f) autoc[0] = 1.f;
    autoc[0] = (float)((double)autoc[0] * (1.0 + 0.01));
    for (int it = 0; it < nt; it++) shaper[it] = 0.f;
    stoep_dev(nt, autoc, cross, shaper, spiker);
    float norm = 0.f;
    for (int it = 0; it < nt; it++) norm += shaper[it] * shaper[it];
    if (norm == 0.f) norm = 1.f;
    if (trace_valid(map, dim, tr)) {
        double acc = 0.0;
        for (int it = 0; it < nt; it++) {
            const float v = H[it] * shaper[it];
            acc += 0.5 * (double)v * (double)v / (double)norm;
        }
        atomicAdd(misfit, type == 2 ? -acc : acc);
    }
    float mis = 0.f;
    for (int it = 0; it < nt; it++) {
        const float v = H[it] * shaper[it];
        mis = (float)((double)mis + 0.5 * (double)v * (double)v / (double)norm);
    }
    const double sgn = type == 2 ? -1.0 * scaling : (double)scaling;
    for (int it = 0; it < nt; it++)
        rs[it] = (float)(sgn * ((double)(H[it] * H[it]) - 2.0 * (double)mis) * (double)shaper[it] / (double)norm);
    stoep_dev(nt, autoc, rs, shaper, spiker);
    conv_dev(nt, 0, shaper, nt, 0, recwei, nt, 0, rs);
    if (wt)
        for (int it = 0; it < nt; it++) rs[it] *= wt[it];
}

__global__ void fill_u32(unsigned *p, int n) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace

int acoustic_misfit_residuals(int type, const float *mod, const float *rec, const float *wei, float *res,
                              const int *map, int dim, int ntr, int nt, int nshot, int scaling, double *misfit,
                              hipStream_t st) {
    if (type < 0 || type > 3)
        return fail(RSG_EINVAL,
                    "fwi: misfit type must be DIFFERENCE (0), CORRELATION (1), ADAPTIVE_GAUSS (2) or "
                    "ADAPTIVE_LINEAR (3)");
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
    if (type >= 2) {
        // time weights in the reference's arithmetic: Gaussian exp(-it^2/(2 s^2)),
        // s = (float)(0.01 nt), or the ramp min(it, 500) (fwi.cpp:397-420)
        std::vector<float> h(nt);
        const float sd = type == 2 ? (float)(0.01 * nt) : 500.f;
        for (int it = 0; it < nt; it++) {
            if (type == 2)
                h[it] = sd == 0.f ? 0.f : (float)std::exp(-1.0 * (it * it) / (2.0 * (sd * sd)));
            else
                h[it] = it < sd ? (float)it : sd;
        }
        float *hd = nullptr, *scr = nullptr;
        const size_t ntr_all = (size_t)ntr * nshot;
        if (hipMalloc((void **)&hd, sizeof(float) * nt) != hipSuccess ||
            hipMalloc((void **)&scr, sizeof(float) * 7 * nt * ntr_all) != hipSuccess) {
            if (hd) (void)hipFree(hd);
            return fail(RSG_ENOMEM, "fwi: ADAPTIVE misfit scratch allocation failed");
        }
        fr.p = {hd, scr};
        RSG_HIP(hipMemcpyAsync(hd, h.data(), sizeof(float) * nt, hipMemcpyHostToDevice, st));
        adaptive_kernel<<<(unsigned)((ntr_all + 63) / 64), 64, 0, st>>>(type, mod, rec, wei, res, map, dim,
                                                                         (int)ntr_all, nt, scaling, hd, scr, misfit);
        RSG_LAUNCH_CHECK("adaptive_kernel");
        RSG_HIP(hipStreamSynchronize(st));  // the scratch is freed on return
        if (hipMalloc((void **)&hist, sizeof(unsigned) * 256 * nshot) != hipSuccess ||
            hipMalloc((void **)&prefix, sizeof(unsigned) * nshot) != hipSuccess)
            return fail(RSG_ENOMEM, "fwi: misfit scratch allocation failed");
        fr.p.push_back(hist);
        fr.p.push_back(prefix);
    } else if (type == 1) {
        if (hipMalloc((void **)&norms, sizeof(float) * 3 * (size_t)ntr * nshot) != hipSuccess ||
            hipMalloc((void **)&hist, sizeof(unsigned) * 256 * nshot) != hipSuccess ||
            hipMalloc((void **)&prefix, sizeof(unsigned) * nshot) != hipSuccess)
            return fail(RSG_ENOMEM, "fwi: misfit scratch allocation failed");
        fr.p = {norms, hist, prefix};
        const int nt_all = ntr * nshot;
        corr_norms_kernel<<<(nt_all + kThreads - 1) / kThreads, kThreads, 0, st>>>(mod, rec, nt_all, nt, norms);
        RSG_LAUNCH_CHECK("corr_norms_kernel");
    }
    if (type <= 1) {
        misfit_res_kernel<<<1024, kThreads, 0, st>>>(type, mod, rec, wei, norms, res, map, dim, n, nt, scaling,
                                                     misfit);
        RSG_LAUNCH_CHECK("misfit_res_kernel");
    }
    if (type == 0) return RSG_OK;
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

extern "C" int rsg_misfit_residuals(int type, const float *mod, const float *rec, const float *wei, float *res,
                                    const int *map, int dim, int ntr, int nt, int nshot, int scaling, double *misfit,
                                    void *stream) {
    if (!mod || !rec || !res || !map || ntr < 0 || nt < 0 || nshot < 1 || (dim != 2 && dim != 3) ||
        (scaling != 1 && scaling != -1))
        return rsg::fail(RSG_EINVAL, "rsg_misfit_residuals: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    double *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(double)) != hipSuccess)
        return rsg::fail(RSG_ENOMEM, "rsg_misfit_residuals: allocation failed");
    int rc = rsg::launch_zero(d, sizeof(double), st);
    if (rc == RSG_OK) rc = rsg::acoustic_misfit_residuals(type, mod, rec, wei, res, map, dim, ntr, nt, nshot, scaling, d, st);
    double h = 0.0;
    if (rc == RSG_OK) {
        const hipError_t e = hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, st);
        const hipError_t e2 = hipStreamSynchronize(st);
        if (e != hipSuccess || e2 != hipSuccess) rc = rsg::check_hip(e != hipSuccess ? e : e2, "rsg_misfit_residuals");
    } else {
        (void)hipStreamSynchronize(st);
    }
    (void)hipFree(d);
    if (rc == RSG_OK && misfit) *misfit = h;
    return rc;
}
