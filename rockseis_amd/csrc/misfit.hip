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
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
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

// Shaping filter of one trace (shaper_dev): autocorrelation of the weighted
// record and its crosscorrelation with the weighted model trace (the record
// reversed, convolved), 1% white noise on the zero lag, Levinson solve.
// w: 6 nt scratch (weighted model, weighted record, autocorrelation,
// crosscorrelation, Levinson scratch, reversed copy); returns the shaper's
// squared norm (1 when zero).  shaper_weighted_dev starts from the weighted
// inputs already in w[0, nt) (model) and w[nt, 2 nt) (record).
__device__ float shaper_weighted_dev(int nt, float *w, float *shaper) {
    float *modwei = w, *recwei = w + nt, *autoc = w + 2 * nt, *cross = w + 3 * nt, *spiker = w + 4 * nt,
          *xr = w + 5 * nt;
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
    return norm;
}

__device__ float shaper_dev(const float *m, const float *r, const float *wt, int nt, float *w, float *shaper) {
    for (int it = 0; it < nt; it++) {
        w[it] = wt ? m[it] * wt[it] : m[it];
        w[nt + it] = wt ? r[it] * wt[it] : r[it];
    }
    return shaper_weighted_dev(nt, w, shaper);
}

// Residual of one trace after shaper_dev on the same scratch: the shaper's
// own misfit in float, rs = sgn (H^2 - 2 misfit) shaper / norm, solved again
// with the autocorrelation, convolved with the weighted record, weighted.
__device__ void adaptive_res_dev(int type, int nt, const float *wt, const float *H, float *w, float *shaper,
                                 float norm, int scaling, float *rs, const float *conv_rec = nullptr) {
    float mis = 0.f;
    for (int it = 0; it < nt; it++) {
        const float v = H[it] * shaper[it];
        mis = (float)((double)mis + 0.5 * (double)v * (double)v / (double)norm);
    }
    const double sgn = type == 2 ? -1.0 * scaling : (double)scaling;
    for (int it = 0; it < nt; it++)
        rs[it] = (float)(sgn * ((double)(H[it] * H[it]) - 2.0 * (double)mis) * (double)shaper[it] / (double)norm);
    stoep_dev(nt, w + 2 * nt, rs, shaper, w + 4 * nt);
    conv_dev(nt, 0, shaper, nt, 0, conv_rec ? conv_rec : w + nt, nt, 0, rs);
    if (wt)
        for (int it = 0; it < nt; it++) rs[it] *= wt[it];
}

// One lane per trace (shots' traces back to back).  scratch: 7 nt floats per
// trace (shaper_dev's six and the shaper).
__global__ __launch_bounds__(64) void adaptive_kernel(int type, const float *__restrict__ mod,
                                                      const float *__restrict__ rec, const float *__restrict__ wei,
                                                      float *__restrict__ res, const int *__restrict__ map, int dim,
                                                      int ntr_all, int nt, int scaling, const float *__restrict__ H,
                                                      float *__restrict__ scratch, double *misfit) {
    const int tr = blockIdx.x * blockDim.x + threadIdx.x;
    if (tr >= ntr_all) return;
    const size_t o = (size_t)tr * nt;
    const float *wt = wei ? wei + o : nullptr;
    float *w = scratch + (size_t)tr * 7 * nt, *shaper = w + 6 * nt;
    const float norm = shaper_dev(mod + o, rec + o, wt, nt, w, shaper);
    if (trace_valid(map, dim, tr)) {
        double acc = 0.0;
        for (int it = 0; it < nt; it++) {
            const float v = H[it] * shaper[it];
            acc += 0.5 * (double)v * (double)v / (double)norm;
        }
        atomicAdd(misfit, type == 2 ? -acc : acc);
    }
    adaptive_res_dev(type, nt, wt, H, w, shaper, norm, scaling, res + o);
}

// ---- Elastic records (FwiElastic2D::computeMisfit / computeResiduals,
// fwi.cpp:2852-3773).  blockIdx.y = shot for the sample-parallel kernels;
// the per-trace kernels find their shot in tr0 (prefix of nrec).

__device__ inline bool finite_f(float v) { return !isnan(v) && !isinf(v); }

__device__ inline bool e2d_valid(const E2DRecs &s, int tr) { return s.map[2 * tr] >= 0 && s.map[2 * tr + 1] >= 0; }

// DIFFERENCE (fwi.cpp:2915-2937, 3288-3308): misfit 0.5*(a^2+b^2+c^2) of the
// weighted differences, residuals (mod-rec)*w^2.
__global__ __launch_bounds__(kThreads) void e2d_diff_kernel(const E2DRecs *__restrict__ sd, int nt, double *misfit) {
    __shared__ double part[kThreads];
    const E2DRecs &s = sd[blockIdx.y];
    const size_t n = (size_t)s.nrec * nt;
    double acc = 0.0;
    for (size_t p = blockIdx.x * (size_t)kThreads + threadIdx.x; p < n; p += (size_t)gridDim.x * kThreads) {
        float a[3];
        for (int c = 0; c < 3; c++) {
            const float d = s.mod[c][p] - s.obs[c][p];
            a[c] = d;
            float q = d;
            if (s.wei[c]) {
                const float w = s.wei[c][p];
                a[c] = d * w;
                q = d * (w * w);
            }
            s.res[c][p] = q;
        }
        if (e2d_valid(s, (int)(p / nt))) acc += 0.5 * (double)(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(misfit, part[0]);
}

// CORRELATION norms, one lane per (trace, component): |mod|, |rec| (zero
// replaced by 1) and sum(mod*rec)/(|mod||rec|) (fwi.cpp:3310-3358).
__global__ __launch_bounds__(kThreads) void e2d_corr_norms_kernel(const E2DRecs *__restrict__ sd, int nt) {
    const E2DRecs &s = sd[blockIdx.y];
    const int tr = blockIdx.x * kThreads + threadIdx.x, c = blockIdx.z;
    if (tr >= s.nrec) return;
    const float *m = s.mod[c] + (size_t)tr * nt, *r = s.obs[c] + (size_t)tr * nt;
    float a = 0.f, b = 0.f, x = 0.f;
    for (int it = 0; it < nt; it++) {
        a += m[it] * m[it];
        b += r[it] * r[it];
        x += m[it] * r[it];
    }
    a = (float)sqrt((double)a);
    b = (float)sqrt((double)b);
    if (a == 0.f) a = 1.f;
    if (b == 0.f) b = 1.f;
    x /= (a * b);
    float *o = s.norms + 3 * ((size_t)c * s.nrec + tr);
    o[0] = a;
    o[1] = b;
    o[2] = x;
}

// CORRELATION misfit (fwi.cpp:2939-2990: -mod*rec/(n1 n2), weighted,
// non-finite terms 0, summed over the three records) and residuals
// (fwi.cpp:3359-3374: -(rec/(n1 n2) - mod/(n1 n1) n3) * w^2) before the clips.
__global__ __launch_bounds__(kThreads) void e2d_corr_kernel(const E2DRecs *__restrict__ sd, int nt, double *misfit) {
    __shared__ double part[kThreads];
    const E2DRecs &s = sd[blockIdx.y];
    const size_t n = (size_t)s.nrec * nt;
    double acc = 0.0;
    for (size_t p = blockIdx.x * (size_t)kThreads + threadIdx.x; p < n; p += (size_t)gridDim.x * kThreads) {
        const int tr = (int)(p / nt);
        const bool valid = e2d_valid(s, tr);
        float v[3];
        for (int c = 0; c < 3; c++) {
            const float *nn = s.norms + 3 * ((size_t)c * s.nrec + tr);
            const float n1 = nn[0], n2 = nn[1], n3 = nn[2];
            const float m = s.mod[c][p], r = s.obs[c][p];
            float t = -(m * r / (n1 * n2));
            if (s.wei[c]) t *= s.wei[c][p];
            v[c] = finite_f(t) ? t : 0.f;
            float q = -((r / (n1 * n2)) - (m / (n1 * n1)) * n3);
            if (s.wei[c]) q *= s.wei[c][p] * s.wei[c][p];
            s.res[c][p] = q;
        }
        if (valid) acc += (double)(v[0] + v[1] + v[2]);
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(misfit, part[0]);
}

__device__ inline int shot_of(const int *tr0, int ns, int g) {
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tr0[mid] <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// ADAPTIVE_GAUSS / ADAPTIVE_LINEAR, one lane per trace (global traces
// [j0, j0 + nj) of all shots), the Ux and Uz shapers side by side (scratch
// 14 nt): misfit -/+ 0.5*((H sx)^2/|sx|^2 + (H sz)^2/|sz|^2) summed in float
// per sample as fwi.cpp:2990-3060, then each component's residual
// (fwi.cpp:3440-3510).
__global__ __launch_bounds__(64) void e2d_adaptive_kernel(int type, const E2DRecs *__restrict__ sd,
                                                          const int *__restrict__ tr0, int ns, int j0, int nj, int nt,
                                                          const float *__restrict__ H, float *__restrict__ scratch,
                                                          double *misfit) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nj) return;
    const int g = j0 + j, k = shot_of(tr0, ns, g), tr = g - tr0[k];
    const E2DRecs &s = sd[k];
    const size_t o = (size_t)tr * nt;
    const float *wx = s.wei[1] ? s.wei[1] + o : nullptr, *wz = s.wei[2] ? s.wei[2] + o : nullptr;
    float *w = scratch + (size_t)j * 14 * nt, *shx = w + 6 * nt, *wzs = w + 7 * nt, *shz = w + 13 * nt;
    const float nx = shaper_dev(s.mod[1] + o, s.obs[1] + o, wx, nt, w, shx);
    const float nz = shaper_dev(s.mod[2] + o, s.obs[2] + o, wz, nt, wzs, shz);
    if (e2d_valid(s, tr)) {
        double acc = 0.0;
        for (int it = 0; it < nt; it++) {
            const float a = H[it] * shx[it], b = H[it] * shz[it];
            acc += 0.5 * (double)(a * a / nx + b * b / nz);
        }
        atomicAdd(misfit, type == 2 ? -acc : acc);
    }
    adaptive_res_dev(type, nt, wx, H, w, shx, nx, 1, s.res[1] + o);
    adaptive_res_dev(type, nt, wz, H, wzs, shz, nz, 1, s.res[2] + o);
}

// ADAPTIVE_SINGLE (fwi.cpp:3149-3200, 3675-3752): the weighted Ux and Uz
// traces back to back as one 2 nt sequence, one shaping filter of nt lags,
// linear time weight; no clip.  scratch 10 nt per lane: weighted model and
// record (2 nt each), reversed copy (2 nt), autocorrelation,
// crosscorrelation, Levinson scratch, shaper.
__global__ __launch_bounds__(64) void e2d_single_kernel(const E2DRecs *__restrict__ sd, const int *__restrict__ tr0,
                                                        int ns, int j0, int nj, int nt, const float *__restrict__ H,
                                                        float *__restrict__ scratch, double *misfit) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nj) return;
    const int g = j0 + j, k = shot_of(tr0, ns, g), tr = g - tr0[k];
    const E2DRecs &s = sd[k];
    const size_t o = (size_t)tr * nt;
    const float *mx = s.mod[1] + o, *rx = s.obs[1] + o, *mz = s.mod[2] + o, *rz = s.obs[2] + o;
    const float *wx = s.wei[1] ? s.wei[1] + o : nullptr, *wz = s.wei[2] ? s.wei[2] + o : nullptr;
    const int n2 = 2 * nt;
    float *w = scratch + (size_t)j * 10 * nt;
    float *modwei = w, *recwei = w + n2, *xr = w + 2 * n2, *autoc = w + 6 * nt, *cross = w + 7 * nt,
          *spiker = w + 8 * nt, *shaper = w + 9 * nt;
    for (int it = 0; it < nt; it++) {
        modwei[it] = wx ? mx[it] * wx[it] : mx[it];
        recwei[it] = wx ? rx[it] * wx[it] : rx[it];
        modwei[nt + it] = wz ? mz[it] * wz[it] : mz[it];
        recwei[nt + it] = wz ? rz[it] * wz[it] : rz[it];
    }
    for (int i = 0, q = n2 - 1; i < n2; ++i, --q) xr[i] = recwei[q];
    conv_dev(n2, 1 - n2, xr, n2, 0, recwei, nt, 0, autoc);
    conv_dev(n2, 1 - n2, xr, n2, 0, modwei, nt, 0, cross);
    if (autoc[0] == 0.f) autoc[0] = 1.f;
    autoc[0] = (float)((double)autoc[0] * (1.0 + 0.01));
    for (int it = 0; it < nt; it++) shaper[it] = 0.f;
    stoep_dev(nt, autoc, cross, shaper, spiker);
    float norm = 0.f;
    for (int it = 0; it < nt; it++) norm += shaper[it] * shaper[it];
    if (norm == 0.f) norm = 1.f;
    if (e2d_valid(s, tr)) {
        double acc = 0.0;
        for (int it = 0; it < nt; it++) {
            const float a = H[it] * shaper[it];
            acc += 0.5 * (double)(a * a / norm);
        }
        atomicAdd(misfit, acc);
    }
    float mis = 0.f;
    for (int it = 0; it < nt; it++) {
        const float v = H[it] * shaper[it];
        mis = (float)((double)mis + 0.5 * (double)v * (double)v / (double)norm);
    }
    for (int it = 0; it < nt; it++)
        cross[it] = (float)(((double)(H[it] * H[it]) - 2.0 * (double)mis) * (double)shaper[it] / (double)norm);
    stoep_dev(nt, autoc, cross, shaper, spiker);
    conv_dev(nt, 0, shaper, n2, 0, recwei, n2, 0, modwei);  // the joint residual overwrites the model traces
    float *rsx = s.res[1] + o, *rsz = s.res[2] + o;
    for (int it = 0; it < nt; it++) {
        rsx[it] = wx ? modwei[it] * wx[it] : modwei[it];
        rsz[it] = wz ? modwei[nt + it] * wz[it] : modwei[nt + it];
    }
}

// Clips of every shot: Ux and Uz zeroed together where |res_x| reaches the
// larger of their two percentiles pc[3k], pc[3k+1] (|res_z| is not tested,
// fwi.cpp:3376-3406); with `corr` non-finite values zeroed and P clipped on
// its own at pc[3k+2] (fwi.cpp:3408-3430).
__global__ __launch_bounds__(kThreads) void e2d_clip_kernel(const E2DRecs *__restrict__ sd, int nt,
                                                            const float *__restrict__ pc, int corr) {
    const int k = blockIdx.y;
    const E2DRecs &s = sd[k];
    const size_t n = (size_t)s.nrec * nt;
    float cx = pc[3 * k];
    if (pc[3 * k + 1] > cx) cx = pc[3 * k + 1];
    const float cp = pc[3 * k + 2];
    for (size_t p = blockIdx.x * (size_t)kThreads + threadIdx.x; p < n; p += (size_t)gridDim.x * kThreads) {
        float rx = s.res[1][p], rz = s.res[2][p];
        if (fabsf(rx) >= cx) {
            rx = 0.f;
            rz = 0.f;
        }
        if (corr) {
            if (!finite_f(rx)) rx = 0.f;
            if (!finite_f(rz)) rz = 0.f;
            float rp = s.res[0][p];
            if (fabsf(rp) >= cp) rp = 0.f;
            if (!finite_f(rp)) rp = 0.f;
            s.res[0][p] = rp;
        }
        s.res[1][p] = rx;
        s.res[2][p] = rz;
    }
}

// Histogram of the 8-bit digit at `shift` of |v| bit patterns whose higher
// bits equal prefix[seg] (under mask), blockIdx.y = segment of segs.
struct SelSeg {
    const float *p;
    size_t n;
};
__global__ __launch_bounds__(kThreads) void radix_hist_seg_kernel(const SelSeg *__restrict__ segs,
                                                                   const unsigned *__restrict__ prefix, unsigned mask,
                                                                   int shift, unsigned *hist) {
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const SelSeg sg = segs[blockIdx.y];
    const unsigned pf = prefix[blockIdx.y];
    for (size_t p = blockIdx.x * (size_t)kThreads + threadIdx.x; p < sg.n; p += (size_t)gridDim.x * kThreads) {
        const unsigned key = __float_as_uint(sg.p[p]) & 0x7fffffffu;
        if ((key & mask) == pf) atomicAdd(&h[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[blockIdx.y * 256 + threadIdx.x], h[threadIdx.x]);
}

__global__ void fill_u32(unsigned *p, int n) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = 0u;
}

struct FreeList {
    std::vector<void *> p;
    ~FreeList() {
        for (void *q : p) (void)hipFree(q);
    }
    template <class T>
    int alloc(T **out, size_t n, const char *what) {
        void *q = nullptr;
        if (hipMalloc(&q, sizeof(T) * (n ? n : 1)) != hipSuccess) return fail(RSG_ENOMEM, what);
        p.push_back(q);
        *out = (T *)q;
        return RSG_OK;
    }
};

// Device scratch taken from a fixed region (the no-allocation entry point):
// bump allocation in 256-byte steps, an error past the region's end.
struct Arena {
    char *base;
    size_t cap, used;
    template <class T>
    int alloc(T **out, size_t n, const char *) {
        const size_t b = (sizeof(T) * (n ? n : 1) + 255) & ~(size_t)255;
        if (used + b > cap)
            return fail(RSG_EINVAL, "rsg_misfit_residuals_noalloc: these sizes need more device scratch than the "
                                    "fixed RSG_NOALLOC_SCRATCH_BYTES");
        *out = (T *)(base + used);
        used += b;
        return RSG_OK;
    }
};

// k[s]-th smallest |v| (0-based) of every segment, exactly: a radix select
// over the bit patterns of |v| (non-negative floats order as unsigned
// integers), four 8-bit passes, all segments per launch (at most 65535 a
// launch).  out[s] = the bit pattern.  Host work arrays: out and k of nseg
// entries, h of 256 * min(nseg, 65535); device scratch from fr.
template <class Alloc>
int select_kth_impl(Alloc &fr, const SelSeg *segs, const unsigned long long *kth, int nseg, unsigned *out,
                    unsigned long long *k, unsigned *h, hipStream_t st) {
    for (int i = 0; i < nseg; i++) {
        out[i] = 0u;
        k[i] = kth[i];
    }
    if (nseg == 0) return RSG_OK;
    SelSeg *dsegs = nullptr;
    unsigned *hist = nullptr, *prefix = nullptr;
    const int cap = std::min(nseg, 65535);
    if (int rc = fr.alloc(&dsegs, nseg, "fwi: misfit scratch allocation failed")) return rc;
    if (int rc = fr.alloc(&hist, 256 * (size_t)cap, "fwi: misfit scratch allocation failed")) return rc;
    if (int rc = fr.alloc(&prefix, nseg, "fwi: misfit scratch allocation failed")) return rc;
    RSG_HIP(hipMemcpyAsync(dsegs, segs, sizeof(SelSeg) * nseg, hipMemcpyHostToDevice, st));
    size_t maxn = 1;
    for (int i = 0; i < nseg; i++) maxn = std::max(maxn, segs[i].n);
    unsigned blocks = (unsigned)std::min<size_t>((maxn + kThreads * 8 - 1) / (kThreads * 8), 1024);
    if (blocks < 1) blocks = 1;
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 24 - 8 * pass;
        const unsigned mask = pass == 0 ? 0u : (0xffffffffu << (32 - 8 * pass));
        RSG_HIP(hipMemcpyAsync(prefix, out, sizeof(unsigned) * nseg, hipMemcpyHostToDevice, st));
        for (int s0 = 0; s0 < nseg; s0 += cap) {
            const int ng = std::min(cap, nseg - s0);
            fill_u32<<<(256 * ng + kThreads - 1) / kThreads, kThreads, 0, st>>>(hist, 256 * ng);
            RSG_LAUNCH_CHECK("fill_u32");
            radix_hist_seg_kernel<<<dim3(blocks, ng), kThreads, 0, st>>>(dsegs + s0, prefix + s0, mask, shift, hist);
            RSG_LAUNCH_CHECK("radix_hist_seg_kernel");
            RSG_HIP(hipMemcpyAsync(h, hist, sizeof(unsigned) * 256 * ng, hipMemcpyDeviceToHost, st));
            RSG_HIP(hipStreamSynchronize(st));
            for (int j = 0; j < ng; j++) {
                const int sg = s0 + j;
                unsigned long long cum = 0;
                int b = 0;
                for (; b < 256; b++) {
                    const unsigned long long c = h[(size_t)j * 256 + b];
                    if (cum + c > k[sg]) break;
                    cum += c;
                }
                if (b == 256) b = 255;  // unreachable while k < n
                out[sg] |= (unsigned)b << shift;
                k[sg] -= cum;
            }
        }
    }
    return RSG_OK;
}

int select_kth(const std::vector<SelSeg> &segs, const std::vector<unsigned long long> &kth, std::vector<unsigned> &out,
               hipStream_t st) {
    const int nseg = (int)segs.size();
    out.assign(nseg, 0u);
    FreeList fr;
    std::vector<unsigned long long> k(nseg);
    std::vector<unsigned> h(256 * (size_t)std::max(1, std::min(nseg, 65535)));
    return select_kth_impl(fr, segs.data(), kth.data(), nseg, out.data(), k.data(), h.data(), st);
}

// Host table of the ADAPTIVE time weight (fwi.cpp:397-420) in the
// reference's arithmetic: type 2 Gaussian exp(-it^2/(2 s^2)), s = (float)(0.01
// nt); otherwise the ramp min(it, 500).
std::vector<float> adaptive_h(int type, int nt) {
    std::vector<float> h(nt);
    const float sd = type == 2 ? (float)(0.01 * nt) : 500.f;
    for (int it = 0; it < nt; it++) {
        if (type == 2)
            h[it] = sd == 0.f ? 0.f : (float)std::exp(-1.0 * (it * it) / (2.0 * (sd * sd)));
        else
            h[it] = it < sd ? (float)it : sd;
    }
    return h;
}

}  // namespace

namespace {

// Largest shot count of the no-allocation entry point: its per-shot host
// work arrays are fixed-size arrays on the stack.
constexpr int kNoallocShots = 64;

// The acoustic misfit and residuals with device scratch from fr; the clip's
// host work arrays in hk (kth, out, k, h as select_kth_impl takes them,
// nshot and 256 * nshot entries).
template <class Alloc>
int acoustic_misfit_residuals_impl(Alloc &fr, int type, const float *mod, const float *rec, const float *wei,
                                   float *res, const int *map, int dim, int ntr, int nt, int nshot, int scaling,
                                   double *misfit, SelSeg *segs, unsigned long long *kth, unsigned *out,
                                   unsigned long long *k, unsigned *h, hipStream_t st) {
    if (type == 4) type = 0;  // ADAPTIVE_SINGLE has no acoustic branch: the default one is DIFFERENCE
    if (type < 0 || type > 3)
        return fail(RSG_EINVAL,
                    "fwi: misfit type must be DIFFERENCE (0), CORRELATION (1), ADAPTIVE_GAUSS (2), "
                    "ADAPTIVE_LINEAR (3) or ADAPTIVE_SINGLE (4)");
    const size_t nper = (size_t)ntr * nt, n = nper * nshot;
    if (n == 0) return RSG_OK;
    float *norms = nullptr;
    if (type >= 2) {
        const std::vector<float> hh = adaptive_h(type, nt);
        float *hd = nullptr, *scr = nullptr;
        const size_t ntr_all = (size_t)ntr * nshot;
        if (int rc = fr.alloc(&hd, nt, "fwi: ADAPTIVE misfit scratch allocation failed")) return rc;
        if (int rc = fr.alloc(&scr, 7 * (size_t)nt * ntr_all, "fwi: ADAPTIVE misfit scratch allocation failed"))
            return rc;
        RSG_HIP(hipMemcpyAsync(hd, hh.data(), sizeof(float) * nt, hipMemcpyHostToDevice, st));
        adaptive_kernel<<<(unsigned)((ntr_all + 63) / 64), 64, 0, st>>>(type, mod, rec, wei, res, map, dim,
                                                                         (int)ntr_all, nt, scaling, hd, scr, misfit);
        RSG_LAUNCH_CHECK("adaptive_kernel");
    } else if (type == 1) {
        if (int rc = fr.alloc(&norms, 3 * (size_t)ntr * nshot, "fwi: misfit scratch allocation failed")) return rc;
        const int nt_all = ntr * nshot;
        corr_norms_kernel<<<(nt_all + kThreads - 1) / kThreads, kThreads, 0, st>>>(mod, rec, nt_all, nt, norms);
        RSG_LAUNCH_CHECK("corr_norms_kernel");
    }
    if (type <= 1) {
        misfit_res_kernel<<<1024, kThreads, 0, st>>>(type, mod, rec, wei, norms, res, map, dim, n, nt, scaling,
                                                     misfit);
        RSG_LAUNCH_CHECK("misfit_res_kernel");
    }
    // FwiAcoustic3D::computeResiduals (fwi.cpp:1651-1853) has no clip step
    if (type == 0 || dim == 3) return RSG_OK;
    // 97th-percentile clip per shot (FwiAcoustic2D): element (int)(97*nt*ntr/100) of the sorted |res|
    for (int q = 0; q < nshot; q++) {
        segs[q] = SelSeg{res + q * nper, nper};
        kth[q] = (unsigned long long)((97LL * nt * ntr) / 100);
    }
    if (int rc = select_kth_impl(fr, segs, kth, nshot, out, k, h, st)) return rc;
    unsigned *dpf = nullptr;
    if (int rc = fr.alloc(&dpf, nshot, "fwi: misfit scratch allocation failed")) return rc;
    RSG_HIP(hipMemcpyAsync(dpf, out, sizeof(unsigned) * nshot, hipMemcpyHostToDevice, st));
    unsigned blocks = (unsigned)std::min<size_t>((nper + kThreads * 8 - 1) / (kThreads * 8), 1024);
    if (blocks < 1) blocks = 1;
    clip_kernel<<<dim3(blocks, nshot), kThreads, 0, st>>>(res, nper, dpf);
    RSG_LAUNCH_CHECK("clip_kernel");
    RSG_HIP(hipStreamSynchronize(st));  // the scratch is released (or reused) on return
    return RSG_OK;
}

}  // namespace

int acoustic_misfit_residuals(int type, const float *mod, const float *rec, const float *wei, float *res,
                              const int *map, int dim, int ntr, int nt, int nshot, int scaling, double *misfit,
                              hipStream_t st) {
    FreeList fr;
    const size_t ns = (size_t)std::max(nshot, 1);
    std::vector<SelSeg> segs(ns);
    std::vector<unsigned long long> kth(ns), k(ns);
    std::vector<unsigned> out(ns), h(256 * std::min<size_t>(ns, 65535));
    return acoustic_misfit_residuals_impl(fr, type, mod, rec, wei, res, map, dim, ntr, nt, nshot, scaling, misfit,
                                          segs.data(), kth.data(), out.data(), k.data(), h.data(), st);
}

int elastic_misfit_residuals(int type, const std::vector<E2DRecs> &shots_in, int nt, double *misfit,
                             hipStream_t st) {
    if (type < 0 || type > 4)
        return fail(RSG_EINVAL,
                    "fwi_e2d: misfit type must be DIFFERENCE (0), CORRELATION (1), ADAPTIVE_GAUSS (2), "
                    "ADAPTIVE_LINEAR (3) or ADAPTIVE_SINGLE (4)");
    const int ns = (int)shots_in.size();
    if (ns == 0 || nt == 0) return RSG_OK;
    std::vector<E2DRecs> shots(shots_in);
    size_t maxrec = 0, ntr_all = 0;
    std::vector<int> tr0(ns + 1, 0);
    for (int k = 0; k < ns; k++) {
        maxrec = std::max(maxrec, (size_t)shots[k].nrec);
        ntr_all += shots[k].nrec;
        tr0[k + 1] = (int)ntr_all;
    }
    if (ntr_all == 0) return RSG_OK;
    FreeList fr;
    E2DRecs *dsh = nullptr;
    const char *oom = "fwi_e2d: misfit scratch allocation failed";
    float *norms = nullptr;
    if (type == 1) {
        if (int rc = fr.alloc(&norms, 9 * ntr_all, oom)) return rc;
        for (int k = 0; k < ns; k++) shots[k].norms = norms + 9 * (size_t)tr0[k];
    }
    if (int rc = fr.alloc(&dsh, ns, oom)) return rc;
    RSG_HIP(hipMemcpyAsync(dsh, shots.data(), sizeof(E2DRecs) * ns, hipMemcpyHostToDevice, st));
    const size_t maxn = maxrec * nt;
    unsigned blocks = (unsigned)std::min<size_t>((maxn + kThreads * 8 - 1) / (kThreads * 8), 1024);
    if (blocks < 1) blocks = 1;
    if (ns > 65535) return fail(RSG_EINVAL, "fwi_e2d: at most 65535 shots");
    if (type == 0) {
        e2d_diff_kernel<<<dim3(blocks, ns), kThreads, 0, st>>>(dsh, nt, misfit);
        RSG_LAUNCH_CHECK("e2d_diff_kernel");
        RSG_HIP(hipStreamSynchronize(st));  // the shot table is freed on return
        return RSG_OK;
    }
    if (type == 1) {
        e2d_corr_norms_kernel<<<dim3((unsigned)((maxrec + kThreads - 1) / kThreads), ns, 3), kThreads, 0, st>>>(dsh,
                                                                                                              nt);
        RSG_LAUNCH_CHECK("e2d_corr_norms_kernel");
        e2d_corr_kernel<<<dim3(blocks, ns), kThreads, 0, st>>>(dsh, nt, misfit);
        RSG_LAUNCH_CHECK("e2d_corr_kernel");
    } else {
        // P residuals stay zero (the reference does not write them for these types)
        for (int k = 0; k < ns; k++)
            if (int rc = launch_zero(shots[k].res[0], sizeof(float) * shots[k].nrec * (size_t)nt, st)) return rc;
        const std::vector<float> h = adaptive_h(type == 4 ? 3 : type, nt);
        float *hd = nullptr, *scr = nullptr;
        int *dtr0 = nullptr;
        const size_t per = (type == 4 ? 10 : 14) * (size_t)nt;
        // lanes per launch: scratch of at most 1 GiB (at least one wave)
        const size_t lanes = std::max<size_t>(64, std::min<size_t>(ntr_all, ((size_t)1 << 28) / per));
        if (int rc = fr.alloc(&hd, nt, oom)) return rc;
        if (int rc = fr.alloc(&dtr0, ns + 1, oom)) return rc;
        if (int rc = fr.alloc(&scr, per * lanes, oom)) return rc;
        RSG_HIP(hipMemcpyAsync(hd, h.data(), sizeof(float) * nt, hipMemcpyHostToDevice, st));
        RSG_HIP(hipMemcpyAsync(dtr0, tr0.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice, st));
        for (size_t j0 = 0; j0 < ntr_all; j0 += lanes) {
            const int nj = (int)std::min(lanes, ntr_all - j0);
            if (type == 4) {
                e2d_single_kernel<<<(nj + 63) / 64, 64, 0, st>>>(dsh, dtr0, ns, (int)j0, nj, nt, hd, scr, misfit);
                RSG_LAUNCH_CHECK("e2d_single_kernel");
            } else {
                e2d_adaptive_kernel<<<(nj + 63) / 64, 64, 0, st>>>(type, dsh, dtr0, ns, (int)j0, nj, nt, hd, scr,
                                                                   misfit);
                RSG_LAUNCH_CHECK("e2d_adaptive_kernel");
            }
        }
        if (type == 4) {
            RSG_HIP(hipStreamSynchronize(st));  // the scratch is freed on return
            return RSG_OK;
        }
    }
    // percentiles: |res_x|, |res_z| (and |res_p| for CORRELATION) of every shot
    std::vector<SelSeg> segs(3 * (size_t)ns, SelSeg{nullptr, 0});
    std::vector<unsigned long long> kth(3 * (size_t)ns, 0);
    for (int k = 0; k < ns; k++) {
        const size_t n = (size_t)shots[k].nrec * nt;
        const unsigned long long q = (unsigned long long)((97LL * nt * shots[k].nrec) / 100);
        for (int c = 0; c < 3; c++) {
            const int comp = c == 2 ? 0 : c + 1;  // order x, z, p
            if (n && (comp != 0 || type == 1)) segs[3 * k + c] = SelSeg{shots[k].res[comp], n};
            kth[3 * k + c] = q;
        }
    }
    std::vector<unsigned> bits;
    if (int rc = select_kth(segs, kth, bits, st)) return rc;
    std::vector<float> pc(3 * (size_t)ns);
    for (size_t i = 0; i < pc.size(); i++) {
        unsigned b = bits[i];
        std::memcpy(&pc[i], &b, sizeof(float));
    }
    float *dpc = nullptr;
    if (int rc = fr.alloc(&dpc, pc.size(), oom)) return rc;
    RSG_HIP(hipMemcpyAsync(dpc, pc.data(), sizeof(float) * pc.size(), hipMemcpyHostToDevice, st));
    e2d_clip_kernel<<<dim3(blocks, ns), kThreads, 0, st>>>(dsh, nt, dpc, type == 1);
    RSG_LAUNCH_CHECK("e2d_clip_kernel");
    RSG_HIP(hipStreamSynchronize(st));  // the scratch is freed on return
    return RSG_OK;
}

// ---- Elastic 3D records (FwiElastic3D::computeMisfit / computeResiduals,
// fwi.cpp:5632-6196): Ux, Uy, Uz, every trace counts.

namespace {

__device__ inline double block_sum(double acc) {
    __shared__ double part[kThreads];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    return part[0];
}

// DIFFERENCE and the default branch: misfit 0.5 (a^2 + b^2 + c^2) of the
// weighted differences; residuals (mod - rec) w^2, from sample 1 in the
// default branch.  CORRELATION (norms given): misfit -mod rec / (n1 n2)
// weighted, summed over the components; residuals from sample 1,
// -(rec / (n1 n2) - mod / (n1 n1) n3) w^2.
__global__ __launch_bounds__(kThreads) void e3d_misfit_res_kernel(int type, const E3DRecs r, int nt,
                                                                   const float *__restrict__ norms, double *misfit) {
    const size_t n = (size_t)r.nrec * nt;
    double acc = 0.0;
    for (size_t p = blockIdx.x * (size_t)kThreads + threadIdx.x; p < n; p += (size_t)gridDim.x * kThreads) {
        const int it = (int)(p % nt), tr = (int)(p / nt);
        float v[3];
        for (int c = 0; c < 3; c++) {
            const float m = r.mod[c][p], o = r.obs[c][p];
            const float w = r.wei[c] ? r.wei[c][p] : 1.f;
            if (type == 1) {
                const float *nn = norms + 3 * ((size_t)c * r.nrec + tr);
                const float n1 = nn[0], n2 = nn[1], n3 = nn[2];
                v[c] = -(m * o / (n1 * n2));
                if (r.wei[c]) v[c] *= w;
                if (it >= 1) {
                    float q = -((o / (n1 * n2)) - (m / (n1 * n1)) * n3);
                    if (r.wei[c]) q *= w * w;
                    r.res[c][p] = q;
                }
            } else {
                const float d = m - o;
                v[c] = r.wei[c] ? d * w : d;
                if (type == 0 || it >= 1) r.res[c][p] = r.wei[c] ? d * (w * w) : d;
            }
        }
        acc += type == 1 ? (double)(v[0] + v[1] + v[2]) : 0.5 * (double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) atomicAdd(misfit, t);
}

// CORRELATION norms, one lane per (trace, component): |mod|, |rec| (zero
// replaced by 1), sum(mod rec) / (|mod| |rec|).
__global__ __launch_bounds__(kThreads) void e3d_corr_norms_kernel(const E3DRecs r, int nt, float *norms) {
    const int tr = blockIdx.x * kThreads + threadIdx.x, c = blockIdx.y;
    if (tr >= r.nrec) return;
    const float *m = r.mod[c] + (size_t)tr * nt, *o = r.obs[c] + (size_t)tr * nt;
    float a = 0.f, b = 0.f, x = 0.f;
    for (int it = 0; it < nt; it++) {
        a += m[it] * m[it];
        b += o[it] * o[it];
        x += m[it] * o[it];
    }
    a = (float)sqrt((double)a);
    b = (float)sqrt((double)b);
    if (a == 0.f) a = 1.f;
    if (b == 0.f) b = 1.f;
    float *q = norms + 3 * ((size_t)c * r.nrec + tr);
    q[0] = a;
    q[1] = b;
    q[2] = x / (a * b);
}

// ADAPTIVE_GAUSS, one lane per trace (traces [j0, j0 + nj)), scratch 21 nt
// per lane (per component shaper_dev's 6 nt and the shaper).  The shaper
// inputs are the reference's (fwi.cpp:5765-5790): a weighted Ux keeps only
// the weighted model (the record input stays zero), a weighted Uy only the
// weighted record, Uz both.  Misfit +0.5 sum_t (H sx)^2/|sx|^2 + (H sy)^2/|sy|^2
// + (H sz)^2/|sz|^2; residuals as adaptive_res_dev, convolved with the
// unweighted record and not weighted (the reference's weight multiply falls
// on the next trace's sample 0, which that trace rewrites).
__global__ __launch_bounds__(64) void e3d_adaptive_kernel(const E3DRecs r, int j0, int nj, int nt,
                                                          const float *__restrict__ H, float *__restrict__ scratch,
                                                          double *misfit) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nj) return;
    const size_t o = (size_t)(j0 + j) * nt;
    float *w = scratch + (size_t)j * 21 * nt;
    float nrm[3];
    for (int c = 0; c < 3; c++) {
        float *wc = w + (size_t)c * 7 * nt;
        const float *m = r.mod[c] + o, *rc = r.obs[c] + o, *wt = r.wei[c] ? r.wei[c] + o : nullptr;
        for (int it = 0; it < nt; it++) {
            float a = m[it], b = rc[it];
            if (wt) {
                a = c == 1 ? 0.f : a * wt[it];
                b = c == 0 ? 0.f : b * wt[it];
            }
            wc[it] = a;
            wc[nt + it] = b;
        }
        nrm[c] = shaper_weighted_dev(nt, wc, wc + 6 * nt);
    }
    const float *sx = w + 6 * nt, *sy = w + 13 * nt, *sz = w + 20 * nt;
    double acc = 0.0;
    for (int it = 0; it < nt; it++) {
        const float a = H[it] * sx[it], b = H[it] * sy[it], c = H[it] * sz[it];
        acc += 0.5 * (double)(a * a / nrm[0] + b * b / nrm[1] + c * c / nrm[2]);
    }
    atomicAdd(misfit, acc);
    for (int c = 0; c < 3; c++) {
        float *wc = w + (size_t)c * 7 * nt;
        adaptive_res_dev(2, nt, nullptr, H, wc, wc + 6 * nt, nrm[c], 1, r.res[c] + o, r.obs[c] + o);
    }
}

}  // namespace

int elastic3d_misfit_residuals(int type, const E3DRecs &r, int nt, double *misfit, hipStream_t st) {
    if (type < 0 || type > 4)
        return fail(RSG_EINVAL,
                    "fwi_e3d: misfit type must be DIFFERENCE (0), CORRELATION (1), ADAPTIVE_GAUSS (2), "
                    "ADAPTIVE_LINEAR (3) or ADAPTIVE_SINGLE (4)");
    const size_t n = (size_t)r.nrec * nt;
    if (n == 0) return RSG_OK;
    FreeList fr;
    const char *oom = "fwi_e3d: misfit scratch allocation failed";
    if (type == 2) {
        const std::vector<float> h = adaptive_h(2, nt);
        const size_t per = 21 * (size_t)nt;
        const size_t lanes = std::max<size_t>(64, std::min<size_t>(r.nrec, ((size_t)1 << 28) / per));
        float *hd = nullptr, *scr = nullptr;
        if (int rc = fr.alloc(&hd, nt, oom)) return rc;
        if (int rc = fr.alloc(&scr, per * lanes, oom)) return rc;
        RSG_HIP(hipMemcpyAsync(hd, h.data(), sizeof(float) * nt, hipMemcpyHostToDevice, st));
        for (size_t j0 = 0; j0 < (size_t)r.nrec; j0 += lanes) {
            const int nj = (int)std::min(lanes, (size_t)r.nrec - j0);
            e3d_adaptive_kernel<<<(nj + 63) / 64, 64, 0, st>>>(r, (int)j0, nj, nt, hd, scr, misfit);
            RSG_LAUNCH_CHECK("e3d_adaptive_kernel");
        }
        RSG_HIP(hipStreamSynchronize(st));  // the scratch is freed on return
        return RSG_OK;
    }
    float *norms = nullptr;
    if (type == 1) {
        if (int rc = fr.alloc(&norms, 9 * (size_t)r.nrec, oom)) return rc;
        e3d_corr_norms_kernel<<<dim3((r.nrec + kThreads - 1) / kThreads, 3), kThreads, 0, st>>>(r, nt, norms);
        RSG_LAUNCH_CHECK("e3d_corr_norms_kernel");
    }
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>((n + kThreads - 1) / kThreads, 1024));
    e3d_misfit_res_kernel<<<blocks, kThreads, 0, st>>>(type, r, nt, norms, misfit);
    RSG_LAUNCH_CHECK("e3d_misfit_res_kernel");
    if (norms) RSG_HIP(hipStreamSynchronize(st));  // the scratch is freed on return
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

extern "C" int rsg_misfit_residuals_e2d(int type, const rsg_e2d_records *shots, int nshot, int nt, double *misfit,
                                        void *stream) {
    if (!shots || nshot < 1 || nshot > 65535 || nt < 0)
        return rsg::fail(RSG_EINVAL, "rsg_misfit_residuals_e2d: bad arguments");
    std::vector<rsg::E2DRecs> recs(nshot);
    for (int k = 0; k < nshot; k++) {
        const rsg_e2d_records &a = shots[k];
        if (a.nrec < 0 || (a.nrec > 0 && !a.map))
            return rsg::fail(RSG_EINVAL, "rsg_misfit_residuals_e2d: bad shot");
        rsg::E2DRecs &r = recs[k];
        r = rsg::E2DRecs{};
        r.nrec = a.nrec;
        r.map = a.map;
        for (int c = 0; c < 3; c++) {
            if (a.nrec > 0 && (!a.mod[c] || !a.obs[c] || !a.res[c]))
                return rsg::fail(RSG_EINVAL, "rsg_misfit_residuals_e2d: null record");
            r.mod[c] = a.mod[c];
            r.obs[c] = a.obs[c];
            r.wei[c] = a.weight[c];
            r.res[c] = a.res[c];
        }
    }
    hipStream_t st = (hipStream_t)stream;
    double *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(double)) != hipSuccess)
        return rsg::fail(RSG_ENOMEM, "rsg_misfit_residuals_e2d: allocation failed");
    int rc = rsg::launch_zero(d, sizeof(double), st);
    if (rc == RSG_OK) rc = rsg::elastic_misfit_residuals(type, recs, nt, d, st);
    double h = 0.0;
    if (rc == RSG_OK) {
        const hipError_t e = hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, st);
        const hipError_t e2 = hipStreamSynchronize(st);
        if (e != hipSuccess || e2 != hipSuccess)
            rc = rsg::check_hip(e != hipSuccess ? e : e2, "rsg_misfit_residuals_e2d");
    } else {
        (void)hipStreamSynchronize(st);
    }
    (void)hipFree(d);
    if (rc == RSG_OK && misfit) *misfit = h;
    return rc;
}

// ---------------------------------------------------------------------------
// rsg_misfit_residuals without device allocation.  Its scratch (misfit
// accumulator, CORRELATION norms, radix-select tables, clip values) lives in
// one file-scope device array per GPU.  In host code the array's name is a
// stand-in, not a device address: the address each device holds it at comes
// from hipGetSymbolAddress, taken once per device with that device current,
// and only that address is passed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[RSG_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
void *g_scratch_ptr[64] = {};

int scratch_of_current_device(char **out) {
    int dev = -1;
    RSG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return rsg::fail(RSG_EINVAL, "rsg_misfit_residuals_noalloc: device index past 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return rsg::check_hip(e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                         "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = p;
    }
    *out = (char *)g_scratch_ptr[dev];
    return RSG_OK;
}

}  // namespace

extern "C" int rsg_misfit_residuals_noalloc(int type, const float *mod, const float *rec, const float *wei,
                                            float *res, const int *map, int dim, int ntr, int nt, int nshot,
                                            int scaling, double *misfit, void *stream) {
    if (!mod || !rec || !res || !map || ntr < 0 || nt < 0 || nshot < 1 || (dim != 2 && dim != 3) ||
        (scaling != 1 && scaling != -1))
        return rsg::fail(RSG_EINVAL, "rsg_misfit_residuals_noalloc: bad arguments");
    if (type != 0 && type != 1 && type != 4)
        return rsg::fail(RSG_EINVAL, "rsg_misfit_residuals_noalloc: DIFFERENCE (0), CORRELATION (1) or "
                                     "ADAPTIVE_SINGLE (4, the DIFFERENCE branch) only; the ADAPTIVE types take "
                                     "rsg_misfit_residuals");
    if (nshot > rsg::kNoallocShots)
        return rsg::fail(RSG_EINVAL, "rsg_misfit_residuals_noalloc: at most 64 shots per call");
    char *base = nullptr;
    if (int rc = scratch_of_current_device(&base)) return rc;
    hipStream_t st = (hipStream_t)stream;
    rsg::Arena ar{base, (size_t)RSG_NOALLOC_SCRATCH_BYTES, 0};
    double *d = nullptr;
    if (int rc = ar.alloc(&d, 1, "")) return rc;
    // the array keeps the previous call's contents: clear the accumulator
    // with a fill kernel
    int rc = rsg::launch_zero(d, sizeof(double), st);
    rsg::SelSeg segs[rsg::kNoallocShots];
    unsigned long long kth[rsg::kNoallocShots], k[rsg::kNoallocShots];
    unsigned out[rsg::kNoallocShots], h[256 * rsg::kNoallocShots];
    if (rc == RSG_OK)
        rc = rsg::acoustic_misfit_residuals_impl(ar, type, mod, rec, wei, res, map, dim, ntr, nt, nshot, scaling, d,
                                                 segs, kth, out, k, h, st);
    double hm = 0.0;
    if (rc == RSG_OK) {
        const hipError_t e = hipMemcpyAsync(&hm, d, sizeof(double), hipMemcpyDeviceToHost, st);
        const hipError_t e2 = hipStreamSynchronize(st);
        if (e != hipSuccess || e2 != hipSuccess)
            rc = rsg::check_hip(e != hipSuccess ? e : e2, "rsg_misfit_residuals_noalloc");
    } else {
        (void)hipStreamSynchronize(st);
    }
    if (rc == RSG_OK && misfit) *misfit = hm;
    return rc;
}
