// Reference harness, part 2: RtmAcoustic2D/3D (test infrastructure).  rtm.h
// and fwi.h declare the same enum, so the RTM classes are driven from this
// translation unit of their own; extern "C" with plain pointers, arrays
// x-fastest as in ref_harness.cpp.
#include <cstring>
#include <memory>

#include "model.h"
#include "data.h"
#include "image.h"
#include "rtm.h"

using namespace rockseis;

namespace {

void set_coords(std::shared_ptr<Data2D<float>> d, const float *coords, bool src) {
    Point2D<float> *c = src ? d->getGeom()->getScoords() : d->getGeom()->getGcoords();
    for (size_t i = 0; i < d->getNtrace(); i++) {
        c[i].x = coords[2 * i];
        c[i].y = coords[2 * i + 1];
    }
}

}  // namespace

extern "C" {

// RtmAcoustic2D::run_optimal with in-core Revolve (ncheck snapshots), one
// PRESSURE source and PRESSURE data; image out [nhz][nhx][nz][nx].
void ref_rtm_a2d(const float *vp, const float *rho, int nx, int nz, int lpml, float dx, float dz, float dt, int nt,
                 int order, int fs, int ncheck, const float *scoords, const float *wav, int nrec,
                 const float *gcoords, const float *data, int nhx, int nhz, float *image) {
    auto model = std::make_shared<ModelAcoustic2D<float>>(nx, nz, lpml, dx, dz, 0.f, 0.f, fs != 0);
    model->createModel();
    std::memcpy(model->getVp(), vp, sizeof(float) * nx * nz);
    std::memcpy(model->getR(), rho, sizeof(float) * nx * nz);
    auto src = std::make_shared<Data2D<float>>(1, nt, dt, 0.f);
    src->setField(PRESSURE);
    set_coords(src, scoords, true);
    std::memcpy(src->getData(), wav, sizeof(float) * nt);
    src->makeMap(model->getGeom(), SMAP);
    auto dat = std::make_shared<Data2D<float>>(nrec, nt, dt, 0.f);
    dat->setField(PRESSURE);
    set_coords(dat, gcoords, false);
    std::memcpy(dat->getData(), data, sizeof(float) * nrec * nt);
    dat->makeMap(model->getGeom(), GMAP);
    model->staggerModels();
    auto img = std::make_shared<Image2D<float>>("rtm_image_unused", model, nhx, nhz);
    auto rtm = std::make_shared<RtmAcoustic2D<float>>(model, img, src, dat, order, 1);
    rtm->setNcheck(ncheck);
    rtm->setIncore(true);
    rtm->run_optimal();
    std::memcpy(image, img->getImagedata(), sizeof(float) * nx * nz * nhx * nhz);
}

// RtmAcoustic3D::run_optimal with in-core Revolve (ncheck snapshots), one
// PRESSURE source and PRESSURE data, (x,y,z) coordinate triples; image out
// [nhz][nhy][nhx][nz][ny][nx].
void ref_rtm_a3d(const float *vp, const float *rho, int nx, int ny, int nz, int lpml, float dx, float dy, float dz,
                 float dt, int nt, int order, int fs, int ncheck, const float *scoords, const float *wav, int nrec,
                 const float *gcoords, const float *data, int nhx, int nhy, int nhz, float *image) {
    auto model = std::make_shared<ModelAcoustic3D<float>>(nx, ny, nz, lpml, dx, dy, dz, 0.f, 0.f, 0.f, fs != 0);
    model->createModel();
    const size_t nm = (size_t)nx * ny * nz;
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
        std::memcpy(x->getData(), d, sizeof(float) * ntr * nt);
        x->makeMap(model->getGeom(), src ? SMAP : GMAP);
        return x;
    };
    auto src = mk(1, scoords, true, wav);
    auto dat = mk(nrec, gcoords, false, data);
    model->staggerModels();
    auto img = std::make_shared<Image3D<float>>("rtm_image_unused", model, nhx, nhy, nhz);
    auto rtm = std::make_shared<RtmAcoustic3D<float>>(model, img, src, dat, order, 1);
    rtm->setNcheck(ncheck);
    rtm->setIncore(true);
    rtm->run_optimal();
    std::memcpy(image, img->getImagedata(), sizeof(float) * nm * nhx * nhy * nhz);
}

}  // extern "C"
