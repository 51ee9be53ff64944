// Harness around the compiled reference (barntsen/rockseis lib/*.cpp, built
// from /root/reference by oracle/Makefile into oracle/_ref/librsref.so).
//
// TEST INFRASTRUCTURE ONLY.  It drives the reference's own classes (Der,
// Pml, Model*, Data*, Waves*, Revolve, FwiAcoustic2D) on caller-provided
// arrays so tests/golden/make_golden.py can write fixtures and bench.py can
// time the reference CPU path.  Nothing in the product links it.
//
// Every entry point is extern "C" with plain pointers; all arrays are
// float32, x fastest.
#include <chrono>
#include <cstring>
#include <memory>
#include <vector>

#include "der.h"
#include "pml.h"
#include "model.h"
#include "data.h"
#include "waves.h"
#include "revolve.h"
#include "fwi.h"
#include "image.h"
#include "interp.h"

using namespace rockseis;

namespace {

template <typename P>
void fill2(std::shared_ptr<Data2D<float>> d, const float *coords, bool src) {
    Point2D<float> *c = src ? d->getGeom()->getScoords() : d->getGeom()->getGcoords();
    for (size_t i = 0; i < d->getNtrace(); i++) {
        c[i].x = coords[2 * i];
        c[i].y = coords[2 * i + 1];
    }
}

void copy_map2(std::shared_ptr<Data2D<float>> d, bool smap, int *map, float *shift) {
    Point2D<int> *m = smap ? d->getGeom()->getSmap() : d->getGeom()->getGmap();
    Point2D<float> *s = smap ? d->getGeom()->getSshift() : d->getGeom()->getGshift();
    for (size_t i = 0; i < d->getNtrace(); i++) {
        map[2 * i] = m[i].x;
        map[2 * i + 1] = m[i].y;
        shift[2 * i] = s[i].x;
        shift[2 * i + 1] = s[i].y;
    }
}

rs_field field_of(int f) {
    switch (f) {
    case 1: return VX;
    case 2: return VY;
    case 3: return VZ;
    default: return PRESSURE;
    }
}

}  // namespace

extern "C" {

// which: 0 ddx_fw 1 ddy_fw 2 ddz_fw 3 ddx_bw 4 ddy_bw 5 ddz_bw
void ref_der(const float *f, float *df, int nx, int ny, int nz, float dx, float dy, float dz,
             int order, int which) {
    Der<float> der(nx, ny, nz, dx, dy, dz, order);
    std::vector<float> buf(f, f + (size_t)nx * ny * nz);
    switch (which) {
    case 0: der.ddx_fw(buf.data()); break;
    case 1: der.ddy_fw(buf.data()); break;
    case 2: der.ddz_fw(buf.data()); break;
    case 3: der.ddx_bw(buf.data()); break;
    case 4: der.ddy_bw(buf.data()); break;
    default: der.ddz_bw(buf.data()); break;
    }
    std::memcpy(df, der.getDf(), sizeof(float) * (size_t)nx * ny * nz);
}

void ref_pml_abc(int lpml, float dt, float smax, int sign, float *tab) {
    Pml<float> p(lpml, dt);
    p.setSmax(smax);
    p.computeABC(sign);
    float *t[12] = {p.A_ltf, p.B_ltf, p.C_ltf, p.A_ltf_stag, p.B_ltf_stag, p.C_ltf_stag,
                    p.A_rbb, p.B_rbb, p.C_rbb, p.A_rbb_stag, p.B_rbb_stag, p.C_rbb_stag};
    for (int j = 0; j < 12; j++) std::memcpy(tab + (size_t)j * lpml, t[j], sizeof(float) * lpml);
}

void ref_stagger_a2d(const float *vp, const float *rho, int nx, int nz, int lpml, int fs,
                     float *L, float *Rx, float *Rz) {
    auto m = std::make_shared<ModelAcoustic2D<float>>(nx, nz, lpml, 10.f, 10.f, 0.f, 0.f, fs != 0);
    m->createModel();
    std::memcpy(m->getVp(), vp, sizeof(float) * nx * nz);
    std::memcpy(m->getR(), rho, sizeof(float) * nx * nz);
    m->staggerModels();
    size_t n = (size_t)(nx + 2 * lpml) * (nz + 2 * lpml);
    std::memcpy(L, m->getL(), sizeof(float) * n);
    std::memcpy(Rx, m->getRx(), sizeof(float) * n);
    std::memcpy(Rz, m->getRz(), sizeof(float) * n);
}

void ref_stagger_e2d(const float *vp, const float *vs, const float *rho, int nx, int nz, int lpml,
                     int fs, float *L, float *L2M, float *M, float *Rx, float *Rz) {
    auto m = std::make_shared<ModelElastic2D<float>>(nx, nz, lpml, 10.f, 10.f, 0.f, 0.f, fs != 0);
    m->createModel();
    std::memcpy(m->getVp(), vp, sizeof(float) * nx * nz);
    std::memcpy(m->getVs(), vs, sizeof(float) * nx * nz);
    std::memcpy(m->getR(), rho, sizeof(float) * nx * nz);
    m->staggerModels();
    size_t n = (size_t)(nx + 2 * lpml) * (nz + 2 * lpml);
    std::memcpy(L, m->getL(), sizeof(float) * n);
    std::memcpy(L2M, m->getL2M(), sizeof(float) * n);
    std::memcpy(M, m->getM(), sizeof(float) * n);
    std::memcpy(Rx, m->getRx(), sizeof(float) * n);
    std::memcpy(Rz, m->getRz(), sizeof(float) * n);
}

void ref_stagger_a3d(const float *vp, const float *rho, int nx, int ny, int nz, int lpml, int fs,
                     float *L, float *Rx, float *Ry, float *Rz) {
    auto m = std::make_shared<ModelAcoustic3D<float>>(nx, ny, nz, lpml, 10.f, 10.f, 10.f, 0.f, 0.f,
                                                      0.f, fs != 0);
    m->createModel();
    std::memcpy(m->getVp(), vp, sizeof(float) * nx * ny * nz);
    std::memcpy(m->getR(), rho, sizeof(float) * nx * ny * nz);
    m->staggerModels();
    size_t n = (size_t)(nx + 2 * lpml) * (ny + 2 * lpml) * (nz + 2 * lpml);
    std::memcpy(L, m->getL(), sizeof(float) * n);
    std::memcpy(Rx, m->getRx(), sizeof(float) * n);
    std::memcpy(Ry, m->getRy(), sizeof(float) * n);
    std::memcpy(Rz, m->getRz(), sizeof(float) * n);
}

// Data2D::makeMap(model geometry, SMAP) for a field (0 P, 1 VX, 3 VZ).
void ref_makemap2d(const float *coords, int ntrace, int nx, int nz, float dx, float dz, float ox,
                   float oz, int field, int *map, float *shift) {
    auto m = std::make_shared<ModelAcoustic2D<float>>(nx, nz, 0, dx, dz, ox, oz, false);
    auto d = std::make_shared<Data2D<float>>(ntrace, 1, 1.f, 0.f);
    d->setField(field_of(field));
    fill2<float>(d, coords, true);
    d->makeMap(m->getGeom(), SMAP);
    copy_map2(d, true, map, shift);
}

// Full acoustic 2D modelling loop (modelling.cpp:242-286 order):
// velocity, stress, insertSource(it), recordData(it).  Snapshots of P
// (padded grid) are copied at the steps listed in snap_its.
void ref_a2d_run(const float *vp, const float *rho, int nx, int nz, int lpml, float dx,
                 float dz, float dt, int nt, int order, int fs, int srcfield, int nsrc,
                 const float *scoords, const float *wav, int recfield, int nrec,
                 const float *gcoords, float *rec, int nsnap, const int *snap_its, float *snaps,
                 float *Pout, float *Vxout, float *Vzout) {
    auto model = std::make_shared<ModelAcoustic2D<float>>(nx, nz, lpml, dx, dz, 0.f, 0.f, fs != 0);
    model->createModel();
    std::memcpy(model->getVp(), vp, sizeof(float) * nx * nz);
    std::memcpy(model->getR(), rho, sizeof(float) * nx * nz);
    auto src = std::make_shared<Data2D<float>>(nsrc, nt, dt, 0.f);
    src->setField(field_of(srcfield));
    fill2<float>(src, scoords, true);
    std::memcpy(src->getData(), wav, sizeof(float) * nsrc * nt);
    src->makeMap(model->getGeom(), SMAP);
    auto data = std::make_shared<Data2D<float>>(nrec, nt, dt, 0.f);
    data->setField(field_of(recfield));
    fill2<float>(data, gcoords, false);
    data->makeMap(model->getGeom(), GMAP);
    model->staggerModels();
    auto waves = std::make_shared<WavesAcoustic2D<float>>(model, nt, dt, 0.f);
    auto der = std::make_shared<Der<float>>(waves->getNx_pml(), 1, waves->getNz_pml(),
                                            waves->getDx(), 1.0, waves->getDz(), order);
    waves->getPml()->setSmax(SMAX);
    waves->getPml()->computeABC();
    size_t np = (size_t)waves->getNx_pml() * waves->getNz_pml();
    int isnap = 0;
    for (int it = 0; it < nt; it++) {
        waves->forwardstepVelocity(model, der);
        waves->forwardstepStress(model, der);
        waves->insertSource(model, src, SMAP, it);
        waves->recordData(data, GMAP, it);
        if (isnap < nsnap && snap_its[isnap] == it) {
            std::memcpy(snaps + np * isnap, waves->getP(), sizeof(float) * np);
            isnap++;
        }
    }
    std::memcpy(rec, data->getData(), sizeof(float) * nrec * nt);
    std::memcpy(Pout, waves->getP(), sizeof(float) * np);
    std::memcpy(Vxout, waves->getVx(), sizeof(float) * np);
    std::memcpy(Vzout, waves->getVz(), sizeof(float) * np);
}

// Elastic 2D: nt forward steps (velocity, stress, inject, record), then nbw
// adjoint-state steps (backwardstepVelocity, backwardstepStress) with no
// source.  recfield 0 P (recip flag), 1 VX, 3 VZ.  fields: Vx,Vz,Sxx,Szz,Sxz.
void ref_e2d_run(const float *vp, const float *vs, const float *rho, int nx, int nz, int lpml,
                 float dx, float dz, float dt, int nt, int order, int fs, int srcfield, int nsrc,
                 const float *scoords, const float *wav, int recfield, int recip, int nrec,
                 const float *gcoords, float *rec, int nbw, float *fields) {
    auto model = std::make_shared<ModelElastic2D<float>>(nx, nz, lpml, dx, dz, 0.f, 0.f, fs != 0);
    model->createModel();
    std::memcpy(model->getVp(), vp, sizeof(float) * nx * nz);
    std::memcpy(model->getVs(), vs, sizeof(float) * nx * nz);
    std::memcpy(model->getR(), rho, sizeof(float) * nx * nz);
    auto src = std::make_shared<Data2D<float>>(nsrc, nt, dt, 0.f);
    src->setField(field_of(srcfield));
    fill2<float>(src, scoords, true);
    std::memcpy(src->getData(), wav, sizeof(float) * nsrc * nt);
    src->makeMap(model->getGeom(), SMAP);
    auto data = std::make_shared<Data2D<float>>(nrec, nt, dt, 0.f);
    data->setField(field_of(recfield));
    data->setReciprocity(recip != 0);
    fill2<float>(data, gcoords, false);
    data->makeMap(model->getGeom(), GMAP);
    model->staggerModels();
    auto waves = std::make_shared<WavesElastic2D<float>>(model, nt, dt, 0.f);
    auto der = std::make_shared<Der<float>>(waves->getNx_pml(), 1, waves->getNz_pml(),
                                            waves->getDx(), 1.0, waves->getDz(), order);
    waves->getPml()->setSmax(SMAX);
    waves->getPml()->computeABC();
    for (int it = 0; it < nt; it++) {
        waves->forwardstepVelocity(model, der);
        waves->forwardstepStress(model, der);
        waves->insertSource(model, src, SMAP, it);
        waves->recordData(model, data, GMAP, it);
    }
    for (int ib = 0; ib < nbw; ib++) {
        waves->backwardstepVelocity(model, der);
        waves->backwardstepStress(model, der);
    }
    size_t np = (size_t)waves->getNx_pml() * waves->getNz_pml();
    std::memcpy(rec, data->getData(), sizeof(float) * nrec * nt);
    float *f[5] = {waves->getVx(), waves->getVz(), waves->getSxx(), waves->getSzz(), waves->getSxz()};
    for (int j = 0; j < 5; j++) std::memcpy(fields + np * j, f[j], sizeof(float) * np);
}

// Timing of the reference's acoustic 2D time step (WavesAcoustic2D::
// forwardstepVelocity / forwardstepStress, insertSource, recordData of one
// PRESSURE receiver set, modelling.cpp:242-286) on a given state: P, Vx, Vz
// (each nx_pml*nz_pml), then the 8 PML strips in PmlAcoustic2D member order
// (pml.h:103-111: P_left, P_right, P_top, P_bottom, Vxx_left, Vxx_right,
// Vzz_top, Vzz_bottom).  Returns the seconds of steps it0 .. it0+n-1.
double ref_a2d_time_window(const float *vp, const float *rho, int nx, int nz, int lpml, float dx, float dz,
                           float dt, int nt, int order, int nsrc, const float *scoords, const float *wav, int nrec,
                           const float *gcoords, int it0, int n, const float *state) {
    auto model = std::make_shared<ModelAcoustic2D<float>>(nx, nz, lpml, dx, dz, 0.f, 0.f, false);
    model->createModel();
    std::memcpy(model->getVp(), vp, sizeof(float) * nx * nz);
    std::memcpy(model->getR(), rho, sizeof(float) * nx * nz);
    auto src = std::make_shared<Data2D<float>>(nsrc, nt, dt, 0.f);
    src->setField(PRESSURE);
    fill2<float>(src, scoords, true);
    std::memcpy(src->getData(), wav, sizeof(float) * nsrc * nt);
    src->makeMap(model->getGeom(), SMAP);
    auto data = std::make_shared<Data2D<float>>(nrec, nt, dt, 0.f);
    data->setField(PRESSURE);
    fill2<float>(data, gcoords, false);
    data->makeMap(model->getGeom(), GMAP);
    model->staggerModels();
    auto waves = std::make_shared<WavesAcoustic2D<float>>(model, nt, dt, 0.f);
    auto der = std::make_shared<Der<float>>(waves->getNx_pml(), 1, waves->getNz_pml(), waves->getDx(), 1.0,
                                            waves->getDz(), order);
    waves->getPml()->setSmax(SMAX);
    waves->getPml()->computeABC();
    const size_t nxp = waves->getNx_pml(), nzp = waves->getNz_pml(), np = nxp * nzp;
    float *f[3] = {waves->getP(), waves->getVx(), waves->getVz()};
    for (int j = 0; j < 3; j++) std::memcpy(f[j], state + np * j, sizeof(float) * np);
    auto pm = waves->getPml();
    float *strips[8] = {pm->P_left, pm->P_right, pm->P_top, pm->P_bottom,
                        pm->Vxx_left, pm->Vxx_right, pm->Vzz_top, pm->Vzz_bottom};
    const float *p = state + 3 * np;
    for (int j = 0; j < 8; j++) {
        const size_t len = (j % 4) < 2 ? (size_t)lpml * nzp : nxp * (size_t)lpml;
        std::memcpy(strips[j], p, sizeof(float) * len);
        p += len;
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (int it = it0; it < it0 + n; it++) {
        waves->forwardstepVelocity(model, der);
        waves->forwardstepStress(model, der);
        waves->insertSource(model, src, SMAP, it);
        waves->recordData(data, GMAP, it);
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Timing of the reference's elastic 2D time step (WavesElastic2D::
// forwardstepVelocity / forwardstepStress, insertSource, recordData of a VX
// and a VZ receiver set, as Modelling::run steps, modelling.cpp:242-286) on
// a given state: state holds the 5 fields (Vx, Vz, Sxx, Szz, Sxz, each
// nx_pml*nz_pml) then the 16 PML strips in PmlElastic2D member order
// (pml.h:159-178; x strips lpml*nz_pml, z strips nx_pml*lpml).  Steps
// it0 .. it0+n-1 run from that state; the return value is their wall time in
// seconds, set-up excluded.  Each call builds its own objects, so calls on
// separate threads run independently (the CPU baseline's shot-parallel
// figure).
double ref_e2d_time_window(const float *vp, const float *vs, const float *rho, int nx, int nz, int lpml,
                           float dx, float dz, float dt, int nt, int order, int fs, int nsrc,
                           const float *scoords, const float *wav, int nrec, const float *gcoords, int it0,
                           int n, const float *state) {
    auto model = std::make_shared<ModelElastic2D<float>>(nx, nz, lpml, dx, dz, 0.f, 0.f, fs != 0);
    model->createModel();
    std::memcpy(model->getVp(), vp, sizeof(float) * nx * nz);
    std::memcpy(model->getVs(), vs, sizeof(float) * nx * nz);
    std::memcpy(model->getR(), rho, sizeof(float) * nx * nz);
    auto src = std::make_shared<Data2D<float>>(nsrc, nt, dt, 0.f);
    src->setField(PRESSURE);
    fill2<float>(src, scoords, true);
    std::memcpy(src->getData(), wav, sizeof(float) * nsrc * nt);
    src->makeMap(model->getGeom(), SMAP);
    std::shared_ptr<Data2D<float>> rec[2];
    const rs_field rf[2] = {VX, VZ};
    for (int k = 0; k < 2; k++) {
        rec[k] = std::make_shared<Data2D<float>>(nrec, nt, dt, 0.f);
        rec[k]->setField(rf[k]);
        fill2<float>(rec[k], gcoords, false);
        rec[k]->makeMap(model->getGeom(), GMAP);
    }
    model->staggerModels();
    auto waves = std::make_shared<WavesElastic2D<float>>(model, nt, dt, 0.f);
    auto der = std::make_shared<Der<float>>(waves->getNx_pml(), 1, waves->getNz_pml(), waves->getDx(), 1.0,
                                            waves->getDz(), order);
    waves->getPml()->setSmax(SMAX);
    waves->getPml()->computeABC();
    const size_t nxp = waves->getNx_pml(), nzp = waves->getNz_pml(), np = nxp * nzp;
    float *f[5] = {waves->getVx(), waves->getVz(), waves->getSxx(), waves->getSzz(), waves->getSxz()};
    for (int j = 0; j < 5; j++) std::memcpy(f[j], state + np * j, sizeof(float) * np);
    auto pm = waves->getPml();
    float *strips[16] = {pm->Sxx_left, pm->Sxx_right, pm->Sxzx_left, pm->Sxzx_right, pm->Szz_top, pm->Szz_bottom,
                         pm->Sxzz_top, pm->Sxzz_bottom, pm->Vxx_left, pm->Vxx_right, pm->Vzx_left, pm->Vzx_right,
                         pm->Vzz_top, pm->Vzz_bottom, pm->Vxz_top, pm->Vxz_bottom};
    const float *p = state + 5 * np;
    for (int j = 0; j < 16; j++) {
        const size_t len = (j % 8) < 4 ? (size_t)lpml * nzp : nxp * (size_t)lpml;
        std::memcpy(strips[j], p, sizeof(float) * len);
        p += len;
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (int it = it0; it < it0 + n; it++) {
        waves->forwardstepVelocity(model, der);
        waves->forwardstepStress(model, der);
        waves->insertSource(model, src, SMAP, it);
        waves->recordData(model, rec[0], GMAP, it);
        waves->recordData(model, rec[1], GMAP, it);
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Acoustic 3D: coords are (x,y,z) triples.  fields: P,Vx,Vy,Vz.
void ref_a3d_run(const float *vp, const float *rho, int nx, int ny, int nz, int lpml, float dx,
                 float dy, float dz, float dt, int nt, int order, int fs, int srcfield, int nsrc,
                 const float *scoords, const float *wav, int recfield, int nrec,
                 const float *gcoords, float *rec, float *fields) {
    auto model = std::make_shared<ModelAcoustic3D<float>>(nx, ny, nz, lpml, dx, dy, dz, 0.f, 0.f,
                                                          0.f, fs != 0);
    model->createModel();
    std::memcpy(model->getVp(), vp, sizeof(float) * nx * ny * nz);
    std::memcpy(model->getR(), rho, sizeof(float) * nx * ny * nz);
    auto src = std::make_shared<Data3D<float>>(nsrc, nt, dt, 0.f);
    src->setField(field_of(srcfield));
    Point3D<float> *sc = src->getGeom()->getScoords();
    for (int i = 0; i < nsrc; i++) {
        sc[i].x = scoords[3 * i];
        sc[i].y = scoords[3 * i + 1];
        sc[i].z = scoords[3 * i + 2];
    }
    std::memcpy(src->getData(), wav, sizeof(float) * nsrc * nt);
    src->makeMap(model->getGeom(), SMAP);
    auto data = std::make_shared<Data3D<float>>(nrec, nt, dt, 0.f);
    data->setField(field_of(recfield));
    Point3D<float> *gc = data->getGeom()->getGcoords();
    for (int i = 0; i < nrec; i++) {
        gc[i].x = gcoords[3 * i];
        gc[i].y = gcoords[3 * i + 1];
        gc[i].z = gcoords[3 * i + 2];
    }
    data->makeMap(model->getGeom(), GMAP);
    model->staggerModels();
    auto waves = std::make_shared<WavesAcoustic3D<float>>(model, nt, dt, 0.f);
    auto der = std::make_shared<Der<float>>(waves->getNx_pml(), waves->getNy_pml(),
                                            waves->getNz_pml(), waves->getDx(), waves->getDy(),
                                            waves->getDz(), order);
    waves->getPml()->setSmax(SMAX);
    waves->getPml()->computeABC();
    for (int it = 0; it < nt; it++) {
        waves->forwardstepVelocity(model, der);
        waves->forwardstepStress(model, der);
        waves->insertSource(model, src, SMAP, it);
        waves->recordData(data, GMAP, it);
    }
    size_t np = (size_t)waves->getNx_pml() * waves->getNy_pml() * waves->getNz_pml();
    std::memcpy(rec, data->getData(), sizeof(float) * nrec * nt);
    float *f[4] = {waves->getP(), waves->getVx(), waves->getVy(), waves->getVz()};
    for (int j = 0; j < 4; j++) std::memcpy(fields + np * j, f[j], sizeof(float) * np);
}

// WavesElastic3D modelling (Modelling loop: forwardstepVelocity,
// forwardstepStress, insertSource, recordData) on a ModelElastic3D; fields
// out in the order Vx, Vy, Vz, Sxx, Syy, Szz, Syz, Sxz, Sxy (padded grids).
void ref_e3d_run(const float *vp, const float *vs, const float *rho, int nx, int ny, int nz, int lpml, float dx,
                 float dy, float dz, float dt, int nt, int order, int fs, int srcfield, int nsrc,
                 const float *scoords, const float *wav, int recfield, int nrec, const float *gcoords, float *rec,
                 float *fields) {
    auto model = std::make_shared<ModelElastic3D<float>>(nx, ny, nz, lpml, dx, dy, dz, 0.f, 0.f, 0.f, fs != 0);
    model->createModel();
    std::memcpy(model->getVp(), vp, sizeof(float) * nx * ny * nz);
    std::memcpy(model->getVs(), vs, sizeof(float) * nx * ny * nz);
    std::memcpy(model->getR(), rho, sizeof(float) * nx * ny * nz);
    auto src = std::make_shared<Data3D<float>>(nsrc, nt, dt, 0.f);
    src->setField(field_of(srcfield));
    Point3D<float> *sc = src->getGeom()->getScoords();
    for (int i = 0; i < nsrc; i++) {
        sc[i].x = scoords[3 * i];
        sc[i].y = scoords[3 * i + 1];
        sc[i].z = scoords[3 * i + 2];
    }
    std::memcpy(src->getData(), wav, sizeof(float) * nsrc * nt);
    src->makeMap(model->getGeom(), SMAP);
    auto data = std::make_shared<Data3D<float>>(nrec, nt, dt, 0.f);
    data->setField(field_of(recfield));
    Point3D<float> *gc = data->getGeom()->getGcoords();
    for (int i = 0; i < nrec; i++) {
        gc[i].x = gcoords[3 * i];
        gc[i].y = gcoords[3 * i + 1];
        gc[i].z = gcoords[3 * i + 2];
    }
    data->makeMap(model->getGeom(), GMAP);
    model->staggerModels();
    auto waves = std::make_shared<WavesElastic3D<float>>(model, nt, dt, 0.f);
    auto der = std::make_shared<Der<float>>(waves->getNx_pml(), waves->getNy_pml(), waves->getNz_pml(),
                                            waves->getDx(), waves->getDy(), waves->getDz(), order);
    waves->getPml()->setSmax(SMAX);
    waves->getPml()->computeABC();
    for (int it = 0; it < nt; it++) {
        waves->forwardstepVelocity(model, der);
        waves->forwardstepStress(model, der);
        waves->insertSource(model, src, SMAP, it);
        waves->recordData(model, data, GMAP, it);
    }
    size_t np = (size_t)waves->getNx_pml() * waves->getNy_pml() * waves->getNz_pml();
    std::memcpy(rec, data->getData(), sizeof(float) * nrec * nt);
    float *f[9] = {waves->getVx(),  waves->getVy(),  waves->getVz(),  waves->getSxx(), waves->getSyy(),
                   waves->getSzz(), waves->getSyz(), waves->getSxz(), waves->getSxy()};
    for (int j = 0; j < 9; j++) std::memcpy(fields + np * j, f[j], sizeof(float) * np);
}

// ModelElastic3D::staggerModels: out in the order L, L2M, M_xz, M_yz, M_xy,
// Rx, Ry, Rz (padded grids).
void ref_stagger_e3d(const float *vp, const float *vs, const float *rho, int nx, int ny, int nz, int lpml, int fs,
                     float *out) {
    auto model = std::make_shared<ModelElastic3D<float>>(nx, ny, nz, lpml, 10.f, 10.f, 10.f, 0.f, 0.f, 0.f, fs != 0);
    model->createModel();
    std::memcpy(model->getVp(), vp, sizeof(float) * nx * ny * nz);
    std::memcpy(model->getVs(), vs, sizeof(float) * nx * ny * nz);
    std::memcpy(model->getR(), rho, sizeof(float) * nx * ny * nz);
    model->staggerModels();
    const size_t np = (size_t)(nx + 2 * lpml) * (ny + 2 * lpml) * (nz + 2 * lpml);
    float *f[8] = {model->getL(), model->getL2M(), model->getM_xz(), model->getM_yz(), model->getM_xy(),
                   model->getRx(), model->getRy(),  model->getRz()};
    for (int j = 0; j < 8; j++) std::memcpy(out + np * j, f[j], sizeof(float) * np);
}

// FwiAcoustic2D::run_optimal on a staggered (non-domain) model: the
// reference's gradient of one shot (fwi.cpp:1159-1332) with ncheck in-core
// Revolve checkpoints and the DIFFERENCE misfit.  Outputs: vpgrad, rhograd
// [nz*nx], modelled and residual data [nrec*nt], wavelet gradient
// [nsrc*nt]; returns the misfit.
// FwiAcoustic2D gradient of one shot: run_optimal with ncheck in-core
// Revolve snapshots when ncheck > 0, else run() with snapshots every snapinc
// steps in the file `snapfile` (removed by run()).
static float fwi_a2d(const float *vp, const float *rho, int nx, int nz, int lpml, float dx, float dz,
                     float dt, int nt, int order, int fs, int ncheck, int snapinc, const char *snapfile,
                     int nsrc, const float *scoords, const float *wav, int nrec, const float *gcoords,
                     const float *dobs, float *vpgrad, float *rhograd, float *dmod, float *dres,
                     float *wgrad, int misfit_type = 0, const float *weight = nullptr) {
    auto model = std::make_shared<ModelAcoustic2D<float>>(nx, nz, lpml, dx, dz, 0.f, 0.f, fs != 0);
    model->createModel();
    std::memcpy(model->getVp(), vp, sizeof(float) * nx * nz);
    std::memcpy(model->getR(), rho, sizeof(float) * nx * nz);
    auto src = std::make_shared<Data2D<float>>(nsrc, nt, dt, 0.f);
    src->setField(PRESSURE);
    fill2<float>(src, scoords, true);
    std::memcpy(src->getData(), wav, sizeof(float) * nsrc * nt);
    src->makeMap(model->getGeom(), SMAP);
    auto mk = [&](const float *d) {
        auto x = std::make_shared<Data2D<float>>(nrec, nt, dt, 0.f);
        x->setField(PRESSURE);
        fill2<float>(x, gcoords, false);
        if (d) std::memcpy(x->getData(), d, sizeof(float) * nrec * nt);
        x->makeMap(model->getGeom(), GMAP);
        return x;
    };
    auto obs = mk(dobs), mod = mk(nullptr), res = mk(nullptr);
    model->staggerModels();
    auto fwi = std::make_shared<FwiAcoustic2D<float>>(model, src, obs, order, 1);
    fwi->setDatamodP(mod);
    fwi->setDataresP(res);
    auto vg = std::make_shared<Image2D<float>>("vpgrad_unused", model, 1, 1);
    auto rg = std::make_shared<Image2D<float>>("rhograd_unused", model, 1, 1);
    fwi->setVpgrad(vg);
    fwi->setRhograd(rg);
    auto wg = std::make_shared<Data2D<float>>(nsrc, nt, dt, 0.f);
    wg->setField(PRESSURE);
    fill2<float>(wg, scoords, true);
    wg->makeMap(model->getGeom(), SMAP);
    fwi->setWavgrad(wg);
    if (ncheck > 0) {
        fwi->setNcheck(ncheck);
        fwi->setIncore(true);
    } else {  // FwiAcoustic2D::run: every snapinc-th pressure to a snapshot file
        fwi->setSnapinc(snapinc);
        fwi->setSnapfile(snapfile);
    }
    fwi->setMisfit_type((rs_fwimisfit)misfit_type);  // DIFFERENCE, CORRELATION, ADAPTIVE_GAUSS, ADAPTIVE_LINEAR
    if (weight) fwi->setDataweight(mk(weight));
    if (ncheck > 0)
        fwi->run_optimal();
    else
        fwi->run();
    std::memcpy(vpgrad, vg->getImagedata(), sizeof(float) * nx * nz);
    std::memcpy(rhograd, rg->getImagedata(), sizeof(float) * nx * nz);
    std::memcpy(dmod, mod->getData(), sizeof(float) * nrec * nt);
    std::memcpy(dres, res->getData(), sizeof(float) * nrec * nt);
    std::memcpy(wgrad, wg->getData(), sizeof(float) * nsrc * nt);
    return fwi->getMisfit();
}

float ref_fwi_a2d(const float *vp, const float *rho, int nx, int nz, int lpml, float dx, float dz,
                  float dt, int nt, int order, int fs, int ncheck, int nsrc, const float *scoords,
                  const float *wav, int nrec, const float *gcoords, const float *dobs, float *vpgrad,
                  float *rhograd, float *dmod, float *dres, float *wgrad) {
    return fwi_a2d(vp, rho, nx, nz, lpml, dx, dz, dt, nt, order, fs, ncheck, 1, "", nsrc, scoords, wav,
                   nrec, gcoords, dobs, vpgrad, rhograd, dmod, dres, wgrad);
}

// run_optimal with a misfit type (0 DIFFERENCE, 1 CORRELATION) and optional
// data weights (nrec*nt).
float ref_fwi_a2d_misfit(const float *vp, const float *rho, int nx, int nz, int lpml, float dx, float dz,
                         float dt, int nt, int order, int fs, int ncheck, int nsrc, const float *scoords,
                         const float *wav, int nrec, const float *gcoords, const float *dobs, int misfit_type,
                         const float *weight, float *vpgrad, float *rhograd, float *dmod, float *dres,
                         float *wgrad) {
    return fwi_a2d(vp, rho, nx, nz, lpml, dx, dz, dt, nt, order, fs, ncheck, 1, "", nsrc, scoords, wav, nrec,
                   gcoords, dobs, vpgrad, rhograd, dmod, dres, wgrad, misfit_type, weight);
}

float ref_fwi_a2d_full(const float *vp, const float *rho, int nx, int nz, int lpml, float dx, float dz,
                       float dt, int nt, int order, int fs, int snapinc, const char *snapfile, int nsrc,
                       const float *scoords, const float *wav, int nrec, const float *gcoords,
                       const float *dobs, float *vpgrad, float *rhograd, float *dmod, float *dres,
                       float *wgrad) {
    return fwi_a2d(vp, rho, nx, nz, lpml, dx, dz, dt, nt, order, fs, 0, snapinc, snapfile, nsrc, scoords,
                   wav, nrec, gcoords, dobs, vpgrad, rhograd, dmod, dres, wgrad);
}

// FwiElastic2D::run_optimal with in-core Revolve (ncheck snapshots), P, Ux
// and Uz data (dobs = 3 records of nrec*nt: P, Ux, Uz), the source field
// srcfield (0 P, 1 VX, 3 VZ).  Outputs the scaled Vp/Vs/rho gradients, the
// modelled data and residuals (3 records each), the wavelet gradient; returns
// the misfit.  misfit_type as rs_fwimisfit; weip/weix/weiz: optional data
// weights of the P, Ux, Uz records (nullptr = not set).
float ref_fwi_e2d(const float *vp, const float *vs, const float *rho, int nx, int nz, int lpml, float dx,
                  float dz, float dt, int nt, int order, int fs, int ncheck, int srcfield, int nsrc,
                  const float *scoords, const float *wav, int nrec, const float *gcoords,
                  const float *dobs, float *vpgrad, float *vsgrad, float *rhograd, float *dmod,
                  float *dres, float *wgrad, int misfit_type, const float *weip, const float *weix,
                  const float *weiz) {
    auto model = std::make_shared<ModelElastic2D<float>>(nx, nz, lpml, dx, dz, 0.f, 0.f, fs != 0);
    model->createModel();
    std::memcpy(model->getVp(), vp, sizeof(float) * nx * nz);
    std::memcpy(model->getVs(), vs, sizeof(float) * nx * nz);
    std::memcpy(model->getR(), rho, sizeof(float) * nx * nz);
    auto src = std::make_shared<Data2D<float>>(nsrc, nt, dt, 0.f);
    src->setField(field_of(srcfield));
    fill2<float>(src, scoords, true);
    std::memcpy(src->getData(), wav, sizeof(float) * nsrc * nt);
    src->makeMap(model->getGeom(), SMAP);
    const size_t nd = (size_t)nrec * nt;
    auto mk = [&](rs_field f, const float *d) {
        auto x = std::make_shared<Data2D<float>>(nrec, nt, dt, 0.f);
        x->setField(f);
        fill2<float>(x, gcoords, false);
        if (d) std::memcpy(x->getData(), d, sizeof(float) * nd);
        x->makeMap(model->getGeom(), GMAP);
        return x;
    };
    const rs_field fields[3] = {PRESSURE, VX, VZ};
    std::shared_ptr<Data2D<float>> obs[3], mod[3], res[3];
    for (int k = 0; k < 3; k++) {
        obs[k] = mk(fields[k], dobs + nd * k);
        mod[k] = mk(fields[k], nullptr);
        res[k] = mk(fields[k], nullptr);
    }
    model->staggerModels();
    auto fwi = std::make_shared<FwiElastic2D<float>>(model, src, obs[0], obs[1], obs[2], order, 1);
    fwi->setDatamodP(mod[0]);
    fwi->setDatamodUx(mod[1]);
    fwi->setDatamodUz(mod[2]);
    fwi->setDataresP(res[0]);
    fwi->setDataresUx(res[1]);
    fwi->setDataresUz(res[2]);
    auto vg = std::make_shared<Image2D<float>>("vpgrad_unused", model, 1, 1);
    auto sg = std::make_shared<Image2D<float>>("vsgrad_unused", model, 1, 1);
    auto rg = std::make_shared<Image2D<float>>("rhograd_unused", model, 1, 1);
    fwi->setVpgrad(vg);
    fwi->setVsgrad(sg);
    fwi->setRhograd(rg);
    auto wg = std::make_shared<Data2D<float>>(nsrc, nt, dt, 0.f);
    wg->setField(field_of(srcfield));
    fill2<float>(wg, scoords, true);
    wg->makeMap(model->getGeom(), SMAP);
    fwi->setWavgrad(wg);
    fwi->setNcheck(ncheck);
    fwi->setIncore(true);
    // DIFFERENCE, CORRELATION, ADAPTIVE_GAUSS, ADAPTIVE_LINEAR, ADAPTIVE_SINGLE
    fwi->setMisfit_type((rs_fwimisfit)misfit_type);
    const float *wei[3] = {weip, weix, weiz};
    std::shared_ptr<Data2D<float>> wd[3];
    for (int k = 0; k < 3; k++) {
        if (!wei[k]) continue;
        wd[k] = mk(fields[k], wei[k]);
        if (k == 0) fwi->setDataweightp(wd[k]);
        if (k == 1) fwi->setDataweightx(wd[k]);
        if (k == 2) fwi->setDataweightz(wd[k]);
    }
    fwi->run_optimal();
    std::memcpy(vpgrad, vg->getImagedata(), sizeof(float) * nx * nz);
    std::memcpy(vsgrad, sg->getImagedata(), sizeof(float) * nx * nz);
    std::memcpy(rhograd, rg->getImagedata(), sizeof(float) * nx * nz);
    for (int k = 0; k < 3; k++) {
        std::memcpy(dmod + nd * k, mod[k]->getData(), sizeof(float) * nd);
        std::memcpy(dres + nd * k, res[k]->getData(), sizeof(float) * nd);
    }
    std::memcpy(wgrad, wg->getData(), sizeof(float) * nsrc * nt);
    return fwi->getMisfit();
}

// FwiAcoustic3D::run_optimal with in-core Revolve (ncheck snapshots), a
// PRESSURE source and PRESSURE data; coords are (x,y,z) triples; misfit_type
// as rs_fwimisfit, weight: optional data weights [nrec][nt].  Returns the
// misfit; outputs the Vp and rho gradients, modelled data, residuals and the
// wavelet gradient.
float ref_fwi_a3d(const float *vp, const float *rho, int nx, int ny, int nz, int lpml, float dx, float dy,
                  float dz, float dt, int nt, int order, int fs, int ncheck, const float *scoords,
                  const float *wav, int nrec, const float *gcoords, const float *dobs, float *vpgrad,
                  float *rhograd, float *dmod, float *dres, float *wgrad, int misfit_type, const float *weight) {
    auto model = std::make_shared<ModelAcoustic3D<float>>(nx, ny, nz, lpml, dx, dy, dz, 0.f, 0.f, 0.f,
                                                          fs != 0);
    model->createModel();
    const size_t nm = (size_t)nx * ny * nz, nd = (size_t)nrec * nt;
    std::memcpy(model->getVp(), vp, sizeof(float) * nm);
    std::memcpy(model->getR(), rho, sizeof(float) * nm);
    auto mk = [&](int ntr, const float *coords, bool src, const float *d) {
        auto x = std::make_shared<Data3D<float>>(ntr, nt, dt, 0.f);
        x->setField(PRESSURE);
        Point3D<float> *c = src ? x->getGeom()->getScoords() : x->getGeom()->getGcoords();
        for (int i = 0; i < ntr; i++) {
            c[i].x = coords[3 * i];
            c[i].y = coords[3 * i + 1];
            c[i].z = coords[3 * i + 2];
        }
        if (d) std::memcpy(x->getData(), d, sizeof(float) * ntr * nt);
        x->makeMap(model->getGeom(), src ? SMAP : GMAP);
        return x;
    };
    auto src = mk(1, scoords, true, wav);
    auto obs = mk(nrec, gcoords, false, dobs), mod = mk(nrec, gcoords, false, nullptr),
         res = mk(nrec, gcoords, false, nullptr);
    auto wg = mk(1, scoords, true, nullptr);
    model->staggerModels();
    auto fwi = std::make_shared<FwiAcoustic3D<float>>(model, src, obs, order, 1);
    fwi->setDatamodP(mod);
    fwi->setDataresP(res);
    auto vg = std::make_shared<Image3D<float>>("vpgrad_unused", model, 1, 1, 1);
    auto rg = std::make_shared<Image3D<float>>("rhograd_unused", model, 1, 1, 1);
    fwi->setVpgrad(vg);
    fwi->setRhograd(rg);
    fwi->setWavgrad(wg);
    fwi->setNcheck(ncheck);
    fwi->setIncore(true);
    fwi->setMisfit_type((rs_fwimisfit)misfit_type);
    if (weight) fwi->setDataweight(mk(nrec, gcoords, false, weight));
    fwi->run_optimal();
    std::memcpy(vpgrad, vg->getImagedata(), sizeof(float) * nm);
    std::memcpy(rhograd, rg->getImagedata(), sizeof(float) * nm);
    std::memcpy(dmod, mod->getData(), sizeof(float) * nd);
    std::memcpy(dres, res->getData(), sizeof(float) * nd);
    std::memcpy(wgrad, wg->getData(), sizeof(float) * nt);
    return fwi->getMisfit();
}

// FwiElastic3D::run_optimal with in-core Revolve (ncheck snapshots): source
// of field srcfield (0 P, 1 VX, 2 VY, 3 VZ), observed Ux, Uy, Uz records
// dobs[3][nrec][nt]; misfit_type as rs_fwimisfit; weix/weiy/weiz optional
// data weights.  Returns the misfit; outputs the Vp, Vs, rho gradients, the
// modelled data and residuals [3][nrec][nt] and the wavelet gradient.
float ref_fwi_e3d(const float *vp, const float *vs, const float *rho, int nx, int ny, int nz, int lpml, float dx,
                  float dy, float dz, float dt, int nt, int order, int fs, int ncheck, int srcfield, int nsrc,
                  const float *scoords, const float *wav, int nrec, const float *gcoords, const float *dobs,
                  float *vpgrad, float *vsgrad, float *rhograd, float *dmod, float *dres, float *wgrad,
                  int misfit_type, const float *weix, const float *weiy, const float *weiz) {
    auto model = std::make_shared<ModelElastic3D<float>>(nx, ny, nz, lpml, dx, dy, dz, 0.f, 0.f, 0.f, fs != 0);
    model->createModel();
    const size_t nm = (size_t)nx * ny * nz, nd = (size_t)nrec * nt;
    std::memcpy(model->getVp(), vp, sizeof(float) * nm);
    std::memcpy(model->getVs(), vs, sizeof(float) * nm);
    std::memcpy(model->getR(), rho, sizeof(float) * nm);
    auto mk = [&](int ntr, const float *coords, bool src, rs_field f, const float *d) {
        auto x = std::make_shared<Data3D<float>>(ntr, nt, dt, 0.f);
        x->setField(f);
        Point3D<float> *c = src ? x->getGeom()->getScoords() : x->getGeom()->getGcoords();
        for (int i = 0; i < ntr; i++) {
            c[i].x = coords[3 * i];
            c[i].y = coords[3 * i + 1];
            c[i].z = coords[3 * i + 2];
        }
        if (d) std::memcpy(x->getData(), d, sizeof(float) * ntr * nt);
        x->makeMap(model->getGeom(), src ? SMAP : GMAP);
        return x;
    };
    auto src = mk(nsrc, scoords, true, field_of(srcfield), wav);
    auto wg = mk(nsrc, scoords, true, field_of(srcfield), nullptr);
    const rs_field fields[3] = {VX, VY, VZ};
    std::shared_ptr<Data3D<float>> obs[3], mod[3], res[3], wd[3];
    const float *wei[3] = {weix, weiy, weiz};
    for (int k = 0; k < 3; k++) {
        obs[k] = mk(nrec, gcoords, false, fields[k], dobs + nd * k);
        mod[k] = mk(nrec, gcoords, false, fields[k], nullptr);
        res[k] = mk(nrec, gcoords, false, fields[k], nullptr);
        if (wei[k]) wd[k] = mk(nrec, gcoords, false, fields[k], wei[k]);
    }
    model->staggerModels();
    auto fwi = std::make_shared<FwiElastic3D<float>>(model, src, obs[0], obs[1], obs[2], order, 1);
    fwi->setDatamodUx(mod[0]);
    fwi->setDatamodUy(mod[1]);
    fwi->setDatamodUz(mod[2]);
    fwi->setDataresUx(res[0]);
    fwi->setDataresUy(res[1]);
    fwi->setDataresUz(res[2]);
    if (wd[0]) fwi->setDataweightx(wd[0]);
    if (wd[1]) fwi->setDataweighty(wd[1]);
    if (wd[2]) fwi->setDataweightz(wd[2]);
    auto vg = std::make_shared<Image3D<float>>("vpgrad_unused", model, 1, 1, 1);
    auto sg = std::make_shared<Image3D<float>>("vsgrad_unused", model, 1, 1, 1);
    auto rg = std::make_shared<Image3D<float>>("rhograd_unused", model, 1, 1, 1);
    fwi->setVpgrad(vg);
    fwi->setVsgrad(sg);
    fwi->setRhograd(rg);
    fwi->setWavgrad(wg);
    fwi->setNcheck(ncheck);
    fwi->setIncore(true);
    fwi->setMisfit_type((rs_fwimisfit)misfit_type);
    fwi->run_optimal();
    std::memcpy(vpgrad, vg->getImagedata(), sizeof(float) * nm);
    std::memcpy(vsgrad, sg->getImagedata(), sizeof(float) * nm);
    std::memcpy(rhograd, rg->getImagedata(), sizeof(float) * nm);
    for (int k = 0; k < 3; k++) {
        std::memcpy(dmod + nd * k, mod[k]->getData(), sizeof(float) * nd);
        std::memcpy(dres + nd * k, res[k]->getData(), sizeof(float) * nd);
    }
    std::memcpy(wgrad, wg->getData(), sizeof(float) * nsrc * nt);
    return fwi->getMisfit();
}

// Interp<float>(mode)::interp between two Data2D of ntr traces (mode 0
// LINEAR, 1 BSPLINE, 2 SINC).
void ref_interp(int mode, const float *from, int ntr, int n1, float d1, float o1, float *to, int n2, float d2,
                float o2) {
    auto a = std::make_shared<Data2D<float>>(ntr, n1, d1, o1);
    auto b = std::make_shared<Data2D<float>>(ntr, n2, d2, o2);
    std::memcpy(a->getData(), from, sizeof(float) * ntr * n1);
    Interp<float> ip(mode == 0 ? LINEAR : (mode == 1 ? BSPLINE : SINC));
    ip.interp(a, b);
    std::memcpy(to, b->getData(), sizeof(float) * ntr * n2);
}

// Revolve action sequence: actions[k] (advance=0 takeshot=1 restore=2
// firsturn=3 youturn=4 terminate=5 error=6) and capo after the call.
// Returns the number of actions written, or -1 if maxn was too small.
int ref_revolve(int nt, int snaps, int *actions, int *capos, int *checks, int maxn) {
    Revolve<float> r(nt, snaps, true);
    int n = 0;
    revolve_action a;
    do {
        a = r.revolve();
        if (n >= maxn) return -1;
        actions[n] = (int)a;
        capos[n] = r.getCapo();
        checks[n] = r.getCheck();
        n++;
    } while (a != terminate && a != error);
    return n;
}

int ref_revolve_numforw(int nt, int snaps) {
    Revolve<float> r(nt, snaps, true);
    return r.numforw(nt, snaps);
}

int ref_revolve_adjust(int nt) {
    Revolve<float> r(nt, 2, true);
    return r.adjust(nt);
}

// Model<T>::getLocalsize2d and ModelAcoustic2D / ModelElastic2D::
// getDomainmodel(data, aperture, map, 0, 1, 1, order) for one domain, on the
// model files vpfile / rhofile (/ vsfile) written by the caller.  coords
// [ntr][4] = sx, sz, gx, gz.  With the output arrays null only the sizes are
// returned: out_i = {start, size, nxd, nzd}, out_f = {ox, oz} of the domain.
int ref_domain_model(const char *vpfile, const char *vsfile, const char *rhofile, int lpml, int fs,
                     const float *coords, int ntr, float aperture, int map, int order, long long *out_i,
                     float *out_f, float *Vp, float *Vs, float *R, float *L, float *L2M, float *M, float *Rx,
                     float *Rz) {
    auto d = std::make_shared<Data2D<float>>(ntr, 1, 1.0f, 0.0f);
    Point2D<float> *sc = d->getGeom()->getScoords(), *gc = d->getGeom()->getGcoords();
    for (int i = 0; i < ntr; i++) {
        sc[i].x = coords[4 * i];
        sc[i].y = coords[4 * i + 1];
        gc[i].x = coords[4 * i + 2];
        gc[i].y = coords[4 * i + 3];
    }
    off_t start = 0;
    size_t size = 0;
    const bool m = map != 0;
    if (vsfile && vsfile[0]) {
        auto g = std::make_shared<ModelElastic2D<float>>(vpfile, vsfile, rhofile, lpml, fs);
        g->getLocalsize2d(d, aperture, m, &start, &size);
        auto l = g->getDomainmodel(d, aperture, m, 0, 1, 1, order);
        const size_t n = (size_t)l->getNx() * l->getNz();
        out_i[0] = start;
        out_i[1] = (long long)size;
        out_i[2] = l->getNx();
        out_i[3] = l->getNz();
        out_f[0] = l->getOx();
        out_f[1] = l->getOz();
        if (!Vp) return 0;
        std::memcpy(Vp, l->getVp(), sizeof(float) * n);
        std::memcpy(Vs, l->getVs(), sizeof(float) * n);
        std::memcpy(R, l->getR(), sizeof(float) * n);
        std::memcpy(L, l->getL(), sizeof(float) * n);
        std::memcpy(L2M, l->getL2M(), sizeof(float) * n);
        std::memcpy(M, l->getM(), sizeof(float) * n);
        std::memcpy(Rx, l->getRx(), sizeof(float) * n);
        std::memcpy(Rz, l->getRz(), sizeof(float) * n);
        return 0;
    }
    auto g = std::make_shared<ModelAcoustic2D<float>>(vpfile, rhofile, lpml, fs);
    g->getLocalsize2d(d, aperture, m, &start, &size);
    auto l = g->getDomainmodel(d, aperture, m, 0, 1, 1, order);
    const size_t n = (size_t)l->getNx() * l->getNz();
    out_i[0] = start;
    out_i[1] = (long long)size;
    out_i[2] = l->getNx();
    out_i[3] = l->getNz();
    out_f[0] = l->getOx();
    out_f[1] = l->getOz();
    if (!Vp) return 0;
    std::memcpy(Vp, l->getVp(), sizeof(float) * n);
    std::memcpy(R, l->getR(), sizeof(float) * n);
    std::memcpy(L, l->getL(), sizeof(float) * n);
    std::memcpy(Rx, l->getRx(), sizeof(float) * n);
    std::memcpy(Rz, l->getRz(), sizeof(float) * n);
    return 0;
}

// Image2D::stackImage_parallel's start column: (int)(oxl/dxl - oxg/dxg) in T.
int ref_image_offset(float ol, float dl, float og, float dg) {
    float a = ol / dl, b = og / dg;
    return (int)(a - b);
}

// Data2D<float>::write / Data3D<float>::write of ntrace traces (coords per
// trace: sx, sz, gx, gz, or sx, sy, sz, gx, gy, gz); returns 0 on success.
int ref_rss_write_data(const char *path, int dim3, const float *coords, const float *data, int ntrace, int nt,
                       float dt, float ot) {
    if (dim3) {
        auto d = std::make_shared<Data3D<float>>(ntrace, nt, dt, ot);
        d->setFile(path);
        Point3D<float> *s = d->getGeom()->getScoords(), *g = d->getGeom()->getGcoords();
        for (int i = 0; i < ntrace; i++) {
            s[i].x = coords[6 * i];
            s[i].y = coords[6 * i + 1];
            s[i].z = coords[6 * i + 2];
            g[i].x = coords[6 * i + 3];
            g[i].y = coords[6 * i + 4];
            g[i].z = coords[6 * i + 5];
        }
        std::memcpy(d->getData(), data, sizeof(float) * ntrace * nt);
        return d->write() == FILE_OK ? 0 : 1;
    }
    auto d = std::make_shared<Data2D<float>>(ntrace, nt, dt, ot);
    d->setFile(path);
    Point2D<float> *s = d->getGeom()->getScoords(), *g = d->getGeom()->getGcoords();
    for (int i = 0; i < ntrace; i++) {
        s[i].x = coords[4 * i];
        s[i].y = coords[4 * i + 1];
        g[i].x = coords[4 * i + 2];
        g[i].y = coords[4 * i + 3];
    }
    std::memcpy(d->getData(), data, sizeof(float) * ntrace * nt);
    return d->write() == FILE_OK ? 0 : 1;
}

// Data2D<float>::read / Data3D<float>::read into coords and data.
int ref_rss_read_data(const char *path, int dim3, float *coords, float *data, int ntrace, int nt) {
    if (dim3) {
        auto d = std::make_shared<Data3D<float>>(ntrace, nt, 1.0f, 0.0f);
        d->setFile(path);
        if (d->read() != FILE_OK) return 1;
        Point3D<float> *s = d->getGeom()->getScoords(), *g = d->getGeom()->getGcoords();
        for (int i = 0; i < ntrace; i++) {
            const float v[6] = {s[i].x, s[i].y, s[i].z, g[i].x, g[i].y, g[i].z};
            std::memcpy(coords + 6 * i, v, sizeof(v));
        }
        std::memcpy(data, d->getData(), sizeof(float) * ntrace * nt);
        return 0;
    }
    auto d = std::make_shared<Data2D<float>>(ntrace, nt, 1.0f, 0.0f);
    d->setFile(path);
    if (d->read() != FILE_OK) return 1;
    Point2D<float> *s = d->getGeom()->getScoords(), *g = d->getGeom()->getGcoords();
    for (int i = 0; i < ntrace; i++) {
        const float v[4] = {s[i].x, s[i].y, g[i].x, g[i].y};
        std::memcpy(coords + 4 * i, v, sizeof(v));
    }
    std::memcpy(data, d->getData(), sizeof(float) * ntrace * nt);
    return 0;
}

// The inversion driver's file stacking: Image2D::createEmpty of the global
// image at path, then Image2D::stackImage_parallel of each local image
// (images [nhz][nhx][nz][nx], concatenated in `locals`), then
// Image2D::read of the result into out.
int ref_stack_images_2d(const char *path, int nxg, int nzg, int nhx, int nhz, float dx, float dz, float oxg, float ozg,
                        int nloc, const int *nxl, const int *nzl, const float *oxl, const float *ozl,
                        const float *locals, int padlx, int padhx, int padlz, int padhz, float *out) {
    auto g = std::make_shared<Image2D<float>>(nxg, nzg, nhx, nhz, dx, dz, oxg, ozg);
    g->setImagefile(path);
    if (g->createEmpty() != FILE_OK) return 1;
    size_t off = 0;
    for (int k = 0; k < nloc; k++) {
        auto l = std::make_shared<Image2D<float>>(nxl[k], nzl[k], nhx, nhz, dx, dz, oxl[k], ozl[k]);
        l->allocateImage();
        const size_t n = (size_t)nxl[k] * nzl[k] * nhx * nhz;
        std::memcpy(l->getImagedata(), locals + off, sizeof(float) * n);
        off += n;
        if (l->stackImage_parallel(path, padlx, padhx, padlz, padhz) != FILE_OK) return 1;
    }
    auto r = std::make_shared<Image2D<float>>(nxg, nzg, nhx, nhz, dx, dz, oxg, ozg);
    r->setImagefile(path);
    if (r->read() != FILE_OK) return 1;
    std::memcpy(out, r->getImagedata(), sizeof(float) * nxg * nzg * nhx * nhz);
    return 0;
}

}  // extern "C"
