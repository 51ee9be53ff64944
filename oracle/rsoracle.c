/*
 * rsoracle -- CPU restatement of RockSeis's staggered-grid FD time step,
 * model staggering, PML tables, source/receiver maps and FWI imaging
 * condition, in plain C.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, no -march, so no
 * FMA contraction, like the reference's g++ -O3 build on x86-64).
 *
 * Parity pinning: tests/test_oracle_golden.py checks every exported function
 * against fixtures written by the compiled reference (oracle/ref_harness.cpp,
 * tests/golden/make_golden.py) bit for bit, and tests/test_oracle_math.py
 * checks closed forms (coefficient consistency, PML formula, linearity).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Optimised staggered-grid coefficients, sets for stencil order 1..8
 * (der.cpp:3-53).  Entry cind(k-1)+io is coefficient io of order k. */
static const double orc_der_table[36] = {
    1.0029,
    1.188401e+0, -7.046382e-2,
    1.216624e+0, -9.197724e-2, 1.300041e-2,
    1.230862e+0, -1.034123e-1, 2.011671e-2, -3.245760e-3,
    1.239407e+0, -1.105315e-1, 2.496329e-2, -5.804879e-3, 9.358680e-4,
    1.245095e+0, -1.153979e-1, 2.848442e-2, -7.899473e-3, 1.898222e-3, -2.935304e-4,
    1.249150e+0, -1.189375e-1, 3.116091e-2, -9.623004e-3, 2.812069e-3, -6.655106e-4, 9.728191e-5,
    1.252186e+0, -1.216289e-1, 3.326513e-2, -1.105777e-2, 3.647865e-3, -1.066144e-3, 2.437491e-4, -3.351959e-5,
};
/* w11 is written as a float literal (1.0029f) in the reference. */
static const float orc_w11 = 1.0029f;

/* Stencil order used by Der at index i of an n-long axis (der.cpp loops) */
static inline int orc_kfw(int i, int n, int order)
{
    int k = order;
    if (i + 1 < k) k = i + 1;
    if (n - 1 - i < k) k = n - 1 - i;
    return k;
}
static inline int orc_kbw(int i, int n, int order)
{
    int k = order;
    if (i < k) k = i;
    if (n - i < k) k = n - i;
    return k;
}

int orc_kfw_export(int i, int n, int order) { return orc_kfw(i, n, order); }
int orc_kbw_export(int i, int n, int order) { return orc_kbw(i, n, order); }

#define REAL float
#define SUF _f32
#define REXP(x) expf(x)
#include "rsoracle_impl.h"
#undef REAL
#undef SUF
#undef REXP

#define REAL double
#define SUF _f64
#define REXP(x) exp(x)
#include "rsoracle_impl.h"
#undef REAL
#undef SUF
#undef REXP

/* ---------------------------------------------------------------------- */
/* Geometry2D/3D::makeMap (geometry.cpp:137-235, 303-)                      */
/* ---------------------------------------------------------------------- */

#define ORC_ABS(x) ((x) < 0 ? -(x) : (x))
#define ORC_CLOSE(x, y, d) ((ORC_ABS((x) - (y)) <= (d) * ORC_ABS(y)) ? 1 : 0)

/* One coordinate axis of makeMap.  Coordinates are T=float, the grid
 * origin/step are T converted to double, the shift (e.g. -0.5 for VX
 * records, data.cpp:440-460) is T.  Returns the grid index or -1 and
 * writes the fractional offset (converted to float) when in bounds. */
static int orc_map_axis(float coord, float o, float d, float shiftf, int n, int padl, int padh,
                        float *frac)
{
    double dd = (double)d, oo = (double)o;
    double res = ((double)coord - oo + shiftf * dd) / dd;
    int pos = (int)floor(res);
    if (ORC_CLOSE(res - pos, 1.0, dd * 1e-2) && (pos < padl)) pos += 1;
    if (ORC_CLOSE(res - pos, 0.0, dd * 1e-2) && (pos >= (n - padh))) pos -= 1;
    if (pos >= padl && pos < n - padh) {
        *frac = (float)((((double)coord - oo + shiftf * dd) / dd) - pos);
        return pos;
    }
    return -1;
}

/* coords: ntrace*(x,z) floats.  map: ntrace*(ix,iz) ints (initialised by the
 * caller), shift: ntrace*(fx,fz) floats (left untouched when off grid). */
void orc_makemap2d(const float *coords, int ntrace, int nx, int nz, float dx, float dz,
                   float ox, float oz, int padlx, int padlz, int padhx, int padhz,
                   float shiftx, float shiftz, int *map, float *shift)
{
    for (int i = 0; i < ntrace; i++) {
        float fr;
        int p = orc_map_axis(coords[2 * i], ox, dx, shiftx, nx, padlx, padhx, &fr);
        map[2 * i] = p;
        if (p >= 0) shift[2 * i] = fr;
        p = orc_map_axis(coords[2 * i + 1], oz, dz, shiftz, nz, padlz, padhz, &fr);
        map[2 * i + 1] = p;
        if (p >= 0) shift[2 * i + 1] = fr;
    }
}

void orc_makemap3d(const float *coords, int ntrace, const int *n, const float *d, const float *o,
                   const int *padl, const int *padh, const float *sh, int *map, float *shift)
{
    for (int i = 0; i < ntrace; i++)
        for (int a = 0; a < 3; a++) {
            float fr;
            int p = orc_map_axis(coords[3 * i + a], o[a], d[a], sh[a], n[a], padl[a], padh[a], &fr);
            map[3 * i + a] = p;
            if (p >= 0) shift[3 * i + a] = fr;
        }
}
