// Host-side set-up that feeds the kernels, kept bit-exact with the
// reference's host code: CFS-PML tables (pml.cpp:79-126), model staggering
// (model.cpp:93-239, 1350-1411, 1965-2040, 2720-2801) and coordinate maps
// (geometry.cpp:137-235, 303-; data.cpp:440-460).  The float/double
// promotions follow the reference's expressions so that results match its
// T=float build to the bit.
#include <cmath>
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/rsgpu.h"

namespace rsg {
int fail(int code, const std::string &msg);
}

namespace {

// One CFS profile point with Amax=0, Kmax=2: pow in double (C++11
// std::pow(float,int) promotes), exp in float, C and A quotients in double.
void cfs_point(float func, float Smax, float dt, float *A, float *B, float *C) {
    const float Amax = 0.0f, Kmax = 2.0f;
    func = (float)std::pow((double)func, 3.0);
    const float a = Amax * func;
    const float k = (float)(1.0 + (double)((Kmax - 1) * func));
    const float s = Smax * func;
    *C = (float)(1.0 - 1.0 / (double)k);
    *B = std::exp(-(a + s / k) * dt);
    *A = (float)(((double)s * (1.0 - (double)*B)) / (double)(k * (k * a + s)));
}

// Edge-replicating pad of a [nz][nx] model into [nz+2L][nx+2L].
void pad2(float *out, const float *in, int nx, int nz, int L) {
    const int nxp = nx + 2 * L, nzp = nz + 2 * L;
    for (int z = 0; z < nz; z++) std::memcpy(out + (size_t)(z + L) * nxp + L, in + (size_t)z * nx, sizeof(float) * nx);
    for (int z = 0; z < nzp; z++) {
        float *row = out + (size_t)z * nxp;
        for (int x = 0; x < L; x++) {
            row[x] = row[L];
            row[nxp - L + x] = row[nxp - L - 1];
        }
    }
    for (int z = 0; z < L; z++) {
        std::memcpy(out + (size_t)z * nxp, out + (size_t)L * nxp, sizeof(float) * nxp);
        std::memcpy(out + (size_t)(nzp - L + z) * nxp, out + (size_t)(nzp - L - 1) * nxp, sizeof(float) * nxp);
    }
}

// Arithmetic-mean staggering along a stride: m[i] <- (m[i+1]+m[i])/2.0,
// last element copies its neighbour.
void stagger_line(float *m, size_t stride, int n) {
    for (int i = 0; i < n - 1; i++) m[i * stride] = (float)((double)(m[(i + 1) * stride] + m[i * stride]) / 2.0);
    m[(n - 1) * stride] = m[(n - 2) * stride];
}

bool invert(float *m, size_t n) {
    for (size_t p = 0; p < n; p++) {
        if (m[p] == 0.0f) return false;
        m[p] = (float)(1.0 / (double)m[p]);
    }
    return true;
}

int map_axis(float coord, float o, float d, float shiftf, int n, float *frac) {
    const double dd = d, oo = o;
    const double res = ((double)coord - oo + shiftf * dd) / dd;
    int pos = (int)std::floor(res);
    const double thr = dd * 1e-2;
    if (std::fabs(res - pos - 1.0) <= thr * 1.0 && pos < 0) pos += 1;
    if (std::fabs(res - pos) <= 0.0 && pos >= n) pos -= 1;
    if (pos >= 0 && pos < n) {
        *frac = (float)((((double)coord - oo + shiftf * dd) / dd) - pos);
        return pos;
    }
    return -1;
}

}  // namespace

extern "C" {

int rsg_der_coeffs(float *out36) {
    static const float t[36] = {1.0029f, 1.188401e+0, -7.046382e-2, 1.216624e+0, -9.197724e-2,
                                1.300041e-2, 1.230862e+0, -1.034123e-1, 2.011671e-2, -3.245760e-3,
                                1.239407e+0, -1.105315e-1, 2.496329e-2, -5.804879e-3, 9.358680e-4,
                                1.245095e+0, -1.153979e-1, 2.848442e-2, -7.899473e-3, 1.898222e-3,
                                -2.935304e-4, 1.249150e+0, -1.189375e-1, 3.116091e-2, -9.623004e-3,
                                2.812069e-3, -6.655106e-4, 9.728191e-5, 1.252186e+0, -1.216289e-1,
                                3.326513e-2, -1.105777e-2, 3.647865e-3, -1.066144e-3, 2.437491e-4,
                                -3.351959e-5};
    if (!out36) return rsg::fail(RSG_EINVAL, "null out");
    std::memcpy(out36, t, sizeof(t));
    return RSG_OK;
}

int rsg_pml_tables(int lpml, float dt, float smax, int sign, float *tab) {
    if (lpml < 1 || !tab) return rsg::fail(RSG_EINVAL, "rsg_pml_tables: bad arguments");
    for (int i = 0; i < lpml; i++) {
        float *r = tab;
        const int L = lpml;
        cfs_point((float)(L - i) / L, smax, dt, &r[0 * L + i], &r[1 * L + i], &r[2 * L + i]);
        cfs_point((float)((double)(L - i) - sign * 0.5) / L, smax, dt, &r[3 * L + i], &r[4 * L + i], &r[5 * L + i]);
        cfs_point((float)((double)i + 1.0) / L, smax, dt, &r[6 * L + i], &r[7 * L + i], &r[8 * L + i]);
        cfs_point((float)((double)i + 1.0 + sign * 0.5) / L, smax, dt, &r[9 * L + i], &r[10 * L + i], &r[11 * L + i]);
    }
    return RSG_OK;
}

int rsg_stagger_a2d(const float *vp, const float *rho, int nx, int nz, int lpml, int fs, float *L,
                    float *Rx, float *Rz) {
    if (!vp || !rho || !L || !Rx || !Rz || nx < 2 || nz < 2 || lpml < 0)
        return rsg::fail(RSG_EINVAL, "rsg_stagger_a2d: bad arguments");
    const int nxp = nx + 2 * lpml, nzp = nz + 2 * lpml;
    const size_t n = (size_t)nxp * nzp;
    pad2(Rx, rho, nx, nz, lpml);
    pad2(Rz, rho, nx, nz, lpml);
    pad2(L, vp, nx, nz, lpml);
    for (size_t p = 0; p < n; p++) L[p] = Rx[p] * L[p] * L[p];
    for (int z = 0; z < nzp; z++) stagger_line(Rx + (size_t)z * nxp, 1, nxp);
    for (int x = 0; x < nxp; x++) stagger_line(Rz + x, nxp, nzp);
    if (!invert(Rx, n) || !invert(Rz, n)) return rsg::fail(RSG_EINVAL, "staggerModels: Zero density found.");
    if (fs)
        for (int x = 0; x < nxp; x++) {
            Rx[(size_t)lpml * nxp + x] *= 2.0f;
            L[(size_t)lpml * nxp + x] *= 0.0f;
        }
    return RSG_OK;
}

int rsg_stagger_e2d(const float *vp, const float *vs, const float *rho, int nx, int nz, int lpml,
                    int fs, float *L, float *L2M, float *M, float *Rx, float *Rz) {
    if (!vp || !vs || !rho || !L || !L2M || !M || !Rx || !Rz || nx < 2 || nz < 2 || lpml < 0)
        return rsg::fail(RSG_EINVAL, "rsg_stagger_e2d: bad arguments");
    const int nxp = nx + 2 * lpml, nzp = nz + 2 * lpml;
    const size_t n = (size_t)nxp * nzp;
    pad2(Rx, rho, nx, nz, lpml);
    pad2(Rz, rho, nx, nz, lpml);
    pad2(L, vp, nx, nz, lpml);
    pad2(M, vs, nx, nz, lpml);
    for (size_t p = 0; p < n; p++) {
        const float v = L[p], w = M[p], r = Rx[p];
        L2M[p] = r * v * v;
        M[p] = r * w * w;
        L[p] = (float)((double)(r * v * v) - 2.0 * (double)r * (double)w * (double)w);
    }
    if (fs)
        for (int x = 0; x < nxp; x++) {
            L[(size_t)lpml * nxp + x] = 0.0f;
            L2M[(size_t)lpml * nxp + x] = M[(size_t)lpml * nxp + x];
        }
    for (int z = 0; z < nzp; z++) stagger_line(Rx + (size_t)z * nxp, 1, nxp);
    for (int x = 0; x < nxp; x++) stagger_line(Rz + x, nxp, nzp);
    for (int x = 0; x < nxp; x++) stagger_line(M + x, nxp, nzp);
    for (int z = 0; z < nzp; z++) stagger_line(M + (size_t)z * nxp, 1, nxp);
    if (!invert(Rx, n) || !invert(Rz, n)) return rsg::fail(RSG_EINVAL, "staggerModels: Zero density found.");
    if (fs)
        for (int x = 0; x < nxp; x++) Rx[(size_t)lpml * nxp + x] *= 2.0f;
    return RSG_OK;
}

int rsg_stagger_a3d(const float *vp, const float *rho, int nx, int ny, int nz, int lpml, int fs,
                    float *L, float *Rx, float *Ry, float *Rz) {
    if (!vp || !rho || !L || !Rx || !Ry || !Rz || nx < 2 || ny < 2 || nz < 2 || lpml < 0)
        return rsg::fail(RSG_EINVAL, "rsg_stagger_a3d: bad arguments");
    const int nxp = nx + 2 * lpml, nyp = ny + 2 * lpml, nzp = nz + 2 * lpml;
    const size_t nxy = (size_t)nxp * nyp, n = nxy * nzp;
    // padmodel3d: interior, then y sides, x sides, z sides (edge replication)
    float *dst[4] = {Rx, Ry, Rz, L};
    const float *src[4] = {rho, rho, rho, vp};
    for (int a = 0; a < 4; a++) {
        float *o = dst[a];
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                std::memcpy(o + (size_t)(z + lpml) * nxy + (size_t)(y + lpml) * nxp + lpml,
                            src[a] + ((size_t)z * ny + y) * nx, sizeof(float) * nx);
        for (int z = 0; z < nzp; z++)
            for (int y = 0; y < lpml; y++)
                for (int x = 0; x < nxp; x++) {
                    o[z * nxy + (size_t)y * nxp + x] = o[z * nxy + (size_t)lpml * nxp + x];
                    o[z * nxy + (size_t)(nyp - lpml + y) * nxp + x] = o[z * nxy + (size_t)(nyp - lpml - 1) * nxp + x];
                }
        for (int z = 0; z < nzp; z++)
            for (int y = 0; y < nyp; y++) {
                float *row = o + z * nxy + (size_t)y * nxp;
                for (int x = 0; x < lpml; x++) {
                    row[x] = row[lpml];
                    row[nxp - lpml + x] = row[nxp - lpml - 1];
                }
            }
        for (int z = 0; z < lpml; z++) {
            std::memcpy(o + z * nxy, o + (size_t)lpml * nxy, sizeof(float) * nxy);
            std::memcpy(o + (size_t)(nzp - lpml + z) * nxy, o + (size_t)(nzp - lpml - 1) * nxy, sizeof(float) * nxy);
        }
    }
    for (size_t p = 0; p < n; p++) L[p] = Rx[p] * L[p] * L[p];
    for (int z = 0; z < nzp; z++)
        for (int y = 0; y < nyp; y++) stagger_line(Rx + z * nxy + (size_t)y * nxp, 1, nxp);
    for (int z = 0; z < nzp; z++)
        for (int x = 0; x < nxp; x++) stagger_line(Ry + z * nxy + x, nxp, nyp);
    for (size_t xy = 0; xy < nxy; xy++) stagger_line(Rz + xy, nxy, nzp);
    if (!invert(Rx, n) || !invert(Ry, n) || !invert(Rz, n))
        return rsg::fail(RSG_EINVAL, "staggerModels: Zero density found.");
    if (fs)
        for (size_t xy = 0; xy < nxy; xy++) {
            const size_t p = (size_t)lpml * nxy + xy;
            Rx[p] *= 2.0f;
            Ry[p] *= 2.0f;
            L[p] *= 0.0f;
        }
    return RSG_OK;
}

int rsg_stagger_e3d(const float *vp, const float *vs, const float *rho, int nx, int ny, int nz, int lpml, int fs,
                    float *L, float *L2M, float *Mxz, float *Myz, float *Mxy, float *Rx, float *Ry, float *Rz) {
    if (!vp || !vs || !rho || !L || !L2M || !Mxz || !Myz || !Mxy || !Rx || !Ry || !Rz || nx < 2 || ny < 2 || nz < 2 ||
        lpml < 0)
        return rsg::fail(RSG_EINVAL, "rsg_stagger_e3d: bad arguments");
    const int nxp = nx + 2 * lpml, nyp = ny + 2 * lpml, nzp = nz + 2 * lpml;
    const size_t nxy = (size_t)nxp * nyp, n = nxy * nzp;
    // padmodel3d: every padding cell a copy of the nearest model cell
    auto pad3 = [&](float *o, const float *m) {
        for (int z = 0; z < nzp; z++) {
            const int mz = std::min(std::max(z - lpml, 0), nz - 1);
            for (int y = 0; y < nyp; y++) {
                const int my = std::min(std::max(y - lpml, 0), ny - 1);
                const float *row = m + ((size_t)mz * ny + my) * nx;
                float *orow = o + z * nxy + (size_t)y * nxp;
                for (int x = 0; x < nxp; x++) orow[x] = row[std::min(std::max(x - lpml, 0), nx - 1)];
            }
        }
    };
    pad3(Rx, rho);
    pad3(Ry, rho);
    pad3(Rz, rho);
    pad3(L, vp);
    pad3(Mxz, vs);
    for (size_t p = 0; p < n; p++) {
        const float v = L[p], w = Mxz[p], r = Rx[p];
        L2M[p] = r * v * v;
        L[p] = (float)((double)(r * v * v) - 2.0 * (double)r * (double)w * (double)w);
        Mxz[p] = r * w * w;
        Myz[p] = r * w * w;
        Mxy[p] = r * w * w;
    }
    if (fs)
        for (size_t xy = 0; xy < nxy; xy++) {
            L[(size_t)lpml * nxy + xy] = 0.0f;
            L2M[(size_t)lpml * nxy + xy] = Mxz[(size_t)lpml * nxy + xy];
        }
    auto sx = [&](float *m) {
        for (int z = 0; z < nzp; z++)
            for (int y = 0; y < nyp; y++) stagger_line(m + z * nxy + (size_t)y * nxp, 1, nxp);
    };
    auto sy = [&](float *m) {
        for (int z = 0; z < nzp; z++)
            for (int x = 0; x < nxp; x++) stagger_line(m + z * nxy + x, nxp, nyp);
    };
    auto sz = [&](float *m) {
        for (size_t xy = 0; xy < nxy; xy++) stagger_line(m + xy, nxy, nzp);
    };
    sx(Rx);
    sy(Ry);
    sz(Rz);
    sx(Mxz);
    sz(Mxz);
    sy(Myz);
    sz(Myz);
    sx(Mxy);
    sy(Mxy);
    if (!invert(Rx, n) || !invert(Ry, n) || !invert(Rz, n))
        return rsg::fail(RSG_EINVAL, "staggerModels: Zero density found.");
    if (fs)
        for (size_t xy = 0; xy < nxy; xy++) {
            Rx[(size_t)lpml * nxy + xy] *= 2.0f;
            Ry[(size_t)lpml * nxy + xy] *= 2.0f;
        }
    return RSG_OK;
}

int rsg_makemap2d(const float *coords, int ntrace, int nx, int nz, float dx, float dz, float ox,
                  float oz, int field, int *map, float *shift) {
    if ((!coords || !map || !shift) && ntrace > 0) return rsg::fail(RSG_EINVAL, "rsg_makemap2d: null");
    const float shx = field == RSG_VX ? -0.5f : 0.0f, shz = field == RSG_VZ ? -0.5f : 0.0f;
    for (int i = 0; i < ntrace; i++) {
        float f = 0.0f;
        map[2 * i] = map_axis(coords[2 * i], ox, dx, shx, nx, &f);
        if (map[2 * i] >= 0) shift[2 * i] = f;
        map[2 * i + 1] = map_axis(coords[2 * i + 1], oz, dz, shz, nz, &f);
        if (map[2 * i + 1] >= 0) shift[2 * i + 1] = f;
    }
    return RSG_OK;
}

int rsg_makemap3d(const float *coords, int ntrace, int nx, int ny, int nz, float dx, float dy,
                  float dz, float ox, float oy, float oz, int field, int *map, float *shift) {
    if ((!coords || !map || !shift) && ntrace > 0) return rsg::fail(RSG_EINVAL, "rsg_makemap3d: null");
    const int n[3] = {nx, ny, nz};
    const float d[3] = {dx, dy, dz}, o[3] = {ox, oy, oz};
    float sh[3] = {0.f, 0.f, 0.f};
    if (field == RSG_VX) sh[0] = -0.5f;
    if (field == RSG_VY) sh[1] = -0.5f;
    if (field == RSG_VZ) sh[2] = -0.5f;
    for (int i = 0; i < ntrace; i++)
        for (int a = 0; a < 3; a++) {
            float f = 0.0f;
            map[3 * i + a] = map_axis(coords[3 * i + a], o[a], d[a], sh[a], n[a], &f);
            if (map[3 * i + a] >= 0) shift[3 * i + a] = f;
        }
    return RSG_OK;
}

// Snapshot2D/3D constructor (snap.cpp:330-335): snapnt = (nt-1)/snapinc + 1
// snapshots; enddiff = rint(((nt-1)dt - (snapnt-1) snapinc dt) / dt), in
// float as the reference computes it.
int rsg_snap_cadence(int nt, float dt, int snapinc, int *snapnt, int *enddiff) {
    if (!snapnt || !enddiff || nt < 1 || snapinc < 1 || !(dt > 0.f))
        return rsg::fail(RSG_EINVAL, "rsg_snap_cadence: bad arguments");
    const int n = (nt - 1) / snapinc + 1;
    const float snapdt = dt * (float)snapinc;
    // volatile: both products are rounded to float before the subtraction
    // (no fused multiply-add), as in the reference build
    const volatile float a = (float)(nt - 1) * dt;
    const volatile float b = (float)(n - 1) * snapdt;
    const float d = (a - b) / dt;
    *snapnt = n;
    *enddiff = (int)rintf(d);
    return RSG_OK;
}


// Model<T>::getLocalsize2d (model.cpp:271-340).  The extremes and offsets
// are doubles, the window sizes go through rintf of the double quotients
// (rintf converts its argument to float first), as in the reference.
int rsg_local_window_2d(const float *coords, int ntrace, float dx, float ox, float aperture, int map,
                        long long *start, long long *size) {
    if (!coords || ntrace < 1 || !start || !size || !(dx > 0.f))
        return rsg::fail(RSG_EINVAL, "rsg_local_window_2d: bad arguments");
    auto sx = [&](int i) { return (double)coords[4 * i]; };
    auto gx = [&](int i) { return (double)coords[4 * i + 2]; };
    size_t sz = 0;
    long long st = 0;
    if (aperture >= 0) {
        double mn = map == 0 ? sx(0) : gx(0), mx = mn, off = std::fabs(gx(0) - sx(0));
        for (int i = 1; i < ntrace; i++) {
            const double v = map == 0 ? sx(i) : gx(i);
            if (v < mn) mn = v;
            if (v > mx) mx = v;
            if (std::fabs(gx(i) - sx(i)) > off) off = std::fabs(gx(i) - sx(i));
        }
        if (aperture > 0) {
            sz = (size_t)(rintf((float)((mx - mn + aperture) / dx)) + 1);
            if (sz % 2 == 0) sz = sz + 1;
        } else {
            sz = (size_t)(2 * rintf((float)(off / dx)) + 1);
        }
        st = (long long)(rintf((float)((mn - ox) / dx)) - (float)((sz - 1) / 2));
    } else {
        double mn = sx(0), mx = sx(0);
        for (int i = 0; i < ntrace; i++) {
            if (sx(i) < mn) mn = sx(i);
            if (sx(i) > mx) mx = sx(i);
            if (gx(i) < mn) mn = gx(i);
            if (gx(i) > mx) mx = gx(i);
        }
        const double da = aperture;
        sz = (size_t)(rintf((float)((mx - mn + 2 * std::fabs(da)) / dx)) + 2);
        st = (long long)(rintf((float)((mn - ox) / dx)) - rintf((float)std::fabs(da / dx))) - 1;
    }
    *start = st;
    *size = (long long)sz;
    return RSG_OK;
}

namespace {
// Clamped source column / row of the global model for a padded domain cell
// (the lpos clamps of model.cpp:1573-1590).
inline long long clampi(long long v, long long n) { return v < 0 ? 0 : (v > n - 1 ? n - 1 : v); }
}  // namespace

int rsg_domain_model_a2d(const float *vp, const float *rho, int nx, int nz, int lpml, int fs, float dx,
                         float dz, float ox, float oz, long long start, long long size, float *Vp, float *R,
                         float *L, float *Rx, float *Rz, float *ox_out, float *oz_out) {
    if (!vp || !rho || !Vp || !R || !L || !Rx || !Rz || nx < 1 || nz < 1 || lpml < 0 || size < 1)
        return rsg::fail(RSG_EINVAL, "rsg_domain_model_a2d: bad arguments");
    const long long nxd = size + 2 * lpml, nzd = nz + 2 * (long long)lpml;
    for (long long i1 = 0; i1 < nzd; i1++) {
        const float *vrow = vp + clampi(i1 - lpml, nz) * nx;
        const float *rrow = rho + clampi(i1 - lpml, nz) * nx;
        const float *radv = rho + clampi(i1 - lpml + 1, nz) * nx;
        for (long long i2 = 0; i2 < nxd; i2++) {
            const long long lp = clampi(start + i2 - lpml, nx);
            const size_t q = (size_t)(i1 * nxd + i2);
            const float r = rrow[lp], v = vrow[lp];
            if (r <= 0.0f) return rsg::fail(RSG_EINVAL, "getDomainmodel: Zero density found.");
            Vp[q] = v;
            R[q] = r;
            L[q] = r * v * v;
            Rx[q] = lp < nx - 1 ? (float)(2.0 / (double)(r + rrow[lp + 1])) : (float)(1.0 / (double)r);
            Rz[q] = (float)(2.0 / (double)(r + radv[lp]));
        }
    }
    if (fs)
        for (long long x = 0; x < nxd; x++) {
            Rx[(size_t)(lpml * nxd + x)] *= 2.0f;
            L[(size_t)(lpml * nxd + x)] *= 0.0f;
        }
    if (ox_out) *ox_out = ox + (float)(start - lpml) * dx;
    if (oz_out) *oz_out = oz + (float)(-lpml) * dz;
    return RSG_OK;
}

int rsg_domain_model_e2d(const float *vp, const float *vs, const float *rho, int nx, int nz, int lpml, int fs,
                         float dx, float dz, float ox, float oz, long long start, long long size, float *Vp,
                         float *Vs, float *R, float *L, float *L2M, float *M, float *Rx, float *Rz,
                         float *ox_out, float *oz_out) {
    if (!vp || !vs || !rho || !Vp || !Vs || !R || !L || !L2M || !M || !Rx || !Rz || nx < 1 || nz < 1 ||
        lpml < 0 || size < 1)
        return rsg::fail(RSG_EINVAL, "rsg_domain_model_e2d: bad arguments");
    const long long nxd = size + 2 * lpml, nzd = nz + 2 * (long long)lpml;
    for (long long i1 = 0; i1 < nzd; i1++) {
        const long long z0 = clampi(i1 - lpml, nz), z1 = clampi(i1 - lpml + 1, nz);
        const float *vprow = vp + z0 * nx, *vsrow = vs + z0 * nx, *rrow = rho + z0 * nx;
        const float *vsadv = vs + z1 * nx, *radv = rho + z1 * nx;
        for (long long i2 = 0; i2 < nxd; i2++) {
            const long long lp = clampi(start + i2 - lpml, nx);
            const size_t q = (size_t)(i1 * nxd + i2);
            const float r = rrow[lp], v = vprow[lp], w = vsrow[lp];
            if (r <= 0.0f) return rsg::fail(RSG_EINVAL, "getDomainmodel: Zero density found.");
            Vp[q] = v;
            Vs[q] = w;
            R[q] = r;
            L2M[q] = r * v * v;
            L[q] = (float)((double)(r * v * v) - 2.0 * (double)r * (double)w * (double)w);
            Rx[q] = lp < nx - 1 ? (float)(2.0 / (double)(r + rrow[lp + 1])) : (float)(1.0 / (double)r);
            Rz[q] = (float)(2.0 / (double)(r + radv[lp]));
            const float m1 = r * w * w, m3 = radv[lp] * vsadv[lp] * vsadv[lp];
            const long long l2 = lp < nx - 1 ? lp + 1 : lp;
            const float m2 = rrow[l2] * vsrow[l2] * vsrow[l2], m4 = radv[l2] * vsadv[l2] * vsadv[l2];
            M[q] = (float)(0.25 * (double)(m1 + m2 + m3 + m4));
        }
    }
    if (fs)
        for (long long x = 0; x < nxd; x++) {
            const size_t q = (size_t)(lpml * nxd + x);
            Rx[q] *= 2.0f;
            L2M[q] = (float)((double)(L2M[q] - L[q]) / 2.0);
            L[q] = 0.0f;
        }
    if (ox_out) *ox_out = ox + (float)(start - lpml) * dx;
    if (oz_out) *oz_out = oz + (float)(-lpml) * dz;
    return RSG_OK;
}

int rsg_image_offset(float ol, float dl, float og, float dg) {
    const volatile float a = ol / dl, b = og / dg;
    return (int)(a - b);
}

}  // extern "C"
