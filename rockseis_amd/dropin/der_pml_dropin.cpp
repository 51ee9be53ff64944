// Float members of the reference's Der class (include/der.h, der.cpp:58-540)
// and Pml<T>::computeABC (pml.cpp:79-126).
//
// The time step kernels fuse the staggered derivatives, so a Der object
// passed to a Waves object only supplies its stencil order.  The dd*
// members remain for callers that differentiate arrays of their own: the
// array goes to the device, rsg_der_apply differentiates it with the
// reference's operation order, and the result comes back in getDf().
// computeABC fills the object's CFS-PML tables with rsg_pml_tables, whose
// values equal the reference's to the bit.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "der.h"
#include "pml.h"
#include "rsgd.h"

namespace rockseis {

template <>
Der<float>::Der() {
    nx = ny = nz = 1;
    dx = dy = dz = 1.f;
    order = 1;
    coeffs = (float *)std::calloc(36, sizeof(float));
    rsgd::ok(rsg_der_coeffs(coeffs), "Der::Der");
    df = (float *)std::calloc(1, sizeof(float));
}

template <>
Der<float>::Der(const int _nx, const int _ny, const int _nz, const float _dx, const float _dy, const float _dz,
                const int _order) {
    nx = _nx;
    ny = _ny;
    nz = _nz;
    dx = _dx;
    dy = _dy;
    dz = _dz;
    order = _order;
    // the reference indexes with int arithmetic: refuse grids where it wraps
    const long lnx = nx, lny = ny, lnz = nz;
    const uint32_t ux = nx, uy = ny, uz = nz;
    const long last = (long)(int32_t)((uz - 1) * uy * ux + (uy - 1) * ux + (ux - 1));
    if (lnx * lny * lnz - 1 != last)
        rs_error("Der::Der: The model size is beyond the size this program can model.");
    if (order < 1) order = 1;
    if (order > 8) order = 8;
    coeffs = (float *)std::calloc(36, sizeof(float));
    rsgd::ok(rsg_der_coeffs(coeffs), "Der::Der");
    df = (float *)std::calloc((size_t)nx * ny * nz, sizeof(float));
    if (!coeffs || !df) rs_error("Der::Der: host allocation failed");
}

template <>
Der<float>::~Der() {
    std::free(df);
    std::free(coeffs);
}

template <>
void Der<float>::setOrder(const int _order) {
    order = _order;
}

namespace {
void der_on_device(const float *f, float *df, int nx, int ny, int nz, int axis, int bw, float d, int order) {
    const size_t n = (size_t)nx * ny * nz;
    rsgd::DevMem a, b;
    a.alloc(sizeof(float) * n);
    b.alloc(sizeof(float) * n);
    a.upload(f, sizeof(float) * n);
    rsgd::ok(rsg_der_apply(axis, bw, a.f(), b.f(), nx, ny, nz, d, order, nullptr), "Der derivative");
    b.download(df, sizeof(float) * n);
}
}  // namespace

template <>
void Der<float>::ddx_fw(float *f) {
    der_on_device(f, df, nx, ny, nz, 0, 0, dx, order);
}
template <>
void Der<float>::ddy_fw(float *f) {
    der_on_device(f, df, nx, ny, nz, 1, 0, dy, order);
}
template <>
void Der<float>::ddz_fw(float *f) {
    der_on_device(f, df, nx, ny, nz, 2, 0, dz, order);
}
template <>
void Der<float>::ddx_bw(float *f) {
    der_on_device(f, df, nx, ny, nz, 0, 1, dx, order);
}
template <>
void Der<float>::ddy_bw(float *f) {
    der_on_device(f, df, nx, ny, nz, 1, 1, dy, order);
}
template <>
void Der<float>::ddz_bw(float *f) {
    der_on_device(f, df, nx, ny, nz, 2, 1, dz, order);
}

template <>
void Pml<float>::computeABC(int sign) {
    // rsg_pml_tables is the reference's profile with Amax = 0 and Kmax = 2
    // (pml.h:14-16), the values every caller on the path keeps
    if (Amax != 0.0f || Kmax != 2.0f)
        rs_error("Pml::computeABC: ", "only Amax = 0 and Kmax = 2 are supported by the rsgpu tables");
    const int L = Lpml;
    std::vector<float> tab(12 * (size_t)L);
    rsgd::ok(rsg_pml_tables(L, dt, Smax, sign, tab.data()), "Pml::computeABC");
    float *const dst[12] = {A_ltf, B_ltf, C_ltf, A_ltf_stag, B_ltf_stag, C_ltf_stag,
                            A_rbb, B_rbb, C_rbb, A_rbb_stag, B_rbb_stag, C_rbb_stag};
    for (int r = 0; r < 12; r++) std::memcpy(dst[r], tab.data() + (size_t)r * L, sizeof(float) * L);
}

}  // namespace rockseis
