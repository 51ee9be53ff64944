// Float members of the reference's Waves, WavesAcoustic2D, WavesAcoustic3D
// and WavesElastic2D classes (include/waves.h:30-209), replacing
// waves.cpp:17-1275 and 1294-2302 for T = float.  Each object's fields,
// PML memory strips and model arrays live in HBM (rsgd::Prop); the time
// steps, source injection and receiver gathers are the C-ABI calls of
// include/rsgpu.h.  The host arrays the header exposes (getP(), ...) are
// kept as mirrors; see rsgd.h for when they are refreshed.
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "rsgd.h"
#include "waves.h"

namespace rockseis {

// ---- Waves<float> (waves.cpp:17-84) ---------------------------------------

template <>
Waves<float>::Waves() {
    geometry = std::make_shared<Geometry<float>>();
    for (int a = 1; a <= 4; a++) {
        geometry->setN(a, 1);
        geometry->setD(a, 1.);
        geometry->setO(a, 0.);
    }
    dim = 0;
    lpml = 10;
    domdec = false;
}

template <>
Waves<float>::Waves(const int _dim, const int _nx, const int _ny, const int _nz, const int _nt, const int _lpml,
                    const float _dx, const float _dy, const float _dz, const float _dt, const float _ox,
                    const float _oy, const float _oz, const float _ot) {
    geometry = std::make_shared<Geometry<float>>();
    const int n[4] = {_nx, _ny, _nz, _nt};
    const float d[4] = {_dx, _dy, _dz, _dt}, o[4] = {_ox, _oy, _oz, _ot};
    for (int a = 0; a < 4; a++) {
        geometry->setN(a + 1, n[a]);
        geometry->setD(a + 1, d[a]);
        geometry->setO(a + 1, o[a]);
    }
    dim = _dim;
    lpml = _lpml;
    domdec = false;
}

template <>
int Waves<float>::getNx_pml() {
    return domdec ? (int)geometry->getN(1) : (int)geometry->getN(1) + 2 * lpml;
}
template <>
int Waves<float>::getNy_pml() {
    return domdec ? (int)geometry->getN(2) : (int)geometry->getN(2) + 2 * lpml;
}
template <>
int Waves<float>::getNz_pml() {
    return domdec ? (int)geometry->getN(3) : (int)geometry->getN(3) + 2 * lpml;
}

template <>
Waves<float>::~Waves() {}

}  // namespace rockseis

namespace {

using rockseis::Data2D;
using rockseis::Data3D;
using rsgd::Prop;

float *host_array(size_t n) {
    float *p = (float *)std::calloc(n ? n : 1, sizeof(float));
    if (!p) rsgd::fatal("rsgpu drop-in: host allocation failed");
    return p;
}

// Grid of a Waves object as the kernels take it: the padded sizes, the
// step sizes, and the domain flag (waves.cpp:145-182).
rsg_grid grid_of(rockseis::Waves<float> *w, int ny_pml) {
    rsg_grid g{};
    g.nx = w->getNx_pml();
    g.ny = ny_pml;
    g.nz = w->getNz_pml();
    g.lpml = w->getLpml();
    g.order = 0;
    g.fs = 0;
    g.dx = w->getDx();
    g.dy = ny_pml > 1 ? w->getDy() : 1.0f;
    g.dz = w->getDz();
    g.dt = w->getDt();
    g.domdec = w->getDomdec() ? 1 : 0;
    return g;
}

// Device fields and strips: nf fields of n cells; strips given as
// (slot, cells) pairs, all zero (the reference callocs them).
void alloc_state(Prop *p, int nf, const int *slots, const size_t *sizes, int nstrip) {
    p->fields.alloc(sizeof(float) * p->n * nf);
    size_t tot = 0;
    for (int k = 0; k < nstrip; k++) tot += sizes[k];
    p->strips.alloc(sizeof(float) * tot);
    float *q = p->strips.f();
    for (int k = 0; k < nstrip; k++) {
        p->s.psi[slots[k]] = q;
        q += sizes[k];
    }
}

// The PML of a Pml<float> object in the C-ABI's table order
// (A, B, C ltf, A, B, C ltf_stag, A, B, C rbb, A, B, C rbb_stag); every side
// must be applied (one domain).
void tables_of(Prop *p, rockseis::Pml<float> *pml) {
    for (int side = 0; side < 6; side++) {
        if (side == 2 || side == 3) {
            if (p->g.ny == 1) continue;  // front / back exist in 3D only
        }
        if (!pml->getApplypml(side))
            rsgd::fatal("rsgpu drop-in: ",
                        "a model split over several domains is not supported (one GPU holds the whole model)");
    }
    const float *const t[12] = {pml->A_ltf, pml->B_ltf, pml->C_ltf, pml->A_ltf_stag, pml->B_ltf_stag,
                                pml->C_ltf_stag, pml->A_rbb, pml->B_rbb, pml->C_rbb, pml->A_rbb_stag,
                                pml->B_rbb_stag, pml->C_rbb_stag};
    rsgd::upload_tables(p, t, pml->getLpml());
}

void set_order(Prop *p, rockseis::Der<float> *der) {
    if (!der) return;
    int o = der->getOrder();
    if (o < 1) o = 1;
    if (o > 8) o = 8;
    p->g.order = o;
    p->order_set = true;
}

void check_model_size(const Prop *p, int mnx, int mny, int mnz) {
    const int pad = p->g.domdec ? 0 : 2 * p->g.lpml;
    if (mnx + pad != p->g.nx || mnz + pad != p->g.nz || (p->g.ny > 1 && mny + pad != p->g.ny))
        rsgd::fatal("rsgpu drop-in: ", "the model does not match the grid of the Waves object");
}

template <class D>
rsgd::Traces *traces2d(Prop *p, const std::shared_ptr<D> &d, bool maptype, int nt, bool record) {
    auto *geo = d->getGeom().get();
    const void *map = maptype ? (const void *)geo->getGmap() : (const void *)geo->getSmap();
    const void *shift = maptype ? (const void *)geo->getGshift() : (const void *)geo->getSshift();
    if (!map || !shift) rsgd::fatal("rsgpu drop-in: ", "trace map not made (Data2D::makeMap)");
    return rsgd::traces(p, d.get(), d, maptype ? 1 : 0, (int)d->getField(), (int)d->getNtrace(), nt,
                        (const int *)map, (const float *)shift, d->getData(), 2, record);
}

template <class D>
rsgd::Traces *traces3d(Prop *p, const std::shared_ptr<D> &d, bool maptype, int nt, bool record) {
    auto *geo = d->getGeom().get();
    const void *map = maptype ? (const void *)geo->getGmap() : (const void *)geo->getSmap();
    const void *shift = maptype ? (const void *)geo->getGshift() : (const void *)geo->getSshift();
    if (!map || !shift) rsgd::fatal("rsgpu drop-in: ", "trace map not made (Data3D::makeMap)");
    return rsgd::traces(p, d.get(), d, maptype ? 1 : 0, (int)d->getField(), (int)d->getNtrace(), nt,
                        (const int *)map, (const float *)shift, d->getData(), 3, record);
}

// Record flush at the end of a sweep in either time direction: the last
// sample of a forward loop, sample 0 of a reverse loop.
struct SweepEnd {
    std::unordered_map<const void *, int> last;
    bool at_end(const void *key, int it, int nt) {
        auto f = last.find(key);
        const int prev = f == last.end() ? -1 : f->second;
        last[key] = it;
        if (prev < 0) return nt == 1;
        return (it == nt - 1 && it > prev) || (it == 0 && prev > 0);
    }
};
std::unordered_map<const void *, SweepEnd> &sweeps() {
    static std::unordered_map<const void *, SweepEnd> s;
    return s;
}

void after_record(Prop *p, const void *waves, rsgd::Traces *t, const void *data, int it) {
    if (rsgd::eager_mirror() || sweeps()[waves].at_end(data, it, t->t.nt)) rsgd::flush(t);
}

void check_wavelet_nt(int data_nt, int waves_nt, const char *what) {
    // insertSource indexes the wavelet with the propagator's nt (waves.cpp:421)
    if (data_nt != waves_nt) rsgd::fatal(what, ": the source has a different number of samples than the propagator");
}

void release(const void *waves) {
    sweeps().erase(waves);
    rsgd::destroy(waves);
}

}  // namespace

namespace rockseis {

// ---- WavesAcoustic2D<float> (waves.cpp:88-595) ------------------------------

namespace {
Prop *make_a2d(WavesAcoustic2D<float> *w, float *P, float *Vx, float *Vz) {
    Prop *p = rsgd::create(w, RSG_KIND_A2D);
    p->g = grid_of(w, 1);
    p->n = (size_t)p->g.nx * p->g.nz;
    const size_t xs = (size_t)p->g.lpml * p->g.nz, zs = (size_t)p->g.nx * p->g.lpml;
    const int slots[8] = {RSG_A2D_P_LEFT, RSG_A2D_P_RIGHT, RSG_A2D_VXX_LEFT, RSG_A2D_VXX_RIGHT,
                          RSG_A2D_P_TOP, RSG_A2D_P_BOTTOM, RSG_A2D_VZZ_TOP, RSG_A2D_VZZ_BOTTOM};
    const size_t sizes[8] = {xs, xs, xs, xs, zs, zs, zs, zs};
    alloc_state(p, 3, slots, sizes, 8);
    p->s.P = p->fields.f();
    p->s.Vx = p->s.P + p->n;
    p->s.Vz = p->s.P + 2 * p->n;
    rsgd::add_mirror(p, P, p->s.P);
    rsgd::add_mirror(p, Vx, p->s.Vx);
    rsgd::add_mirror(p, Vz, p->s.Vz);
    return p;
}

void prepare_a2d(Prop *p, WavesAcoustic2D<float> *w, ModelAcoustic2D<float> *m, Der<float> *der) {
    if (der && !p->order_set) set_order(p, der);
    if (p->model_ready) return;
    check_model_size(p, m->getNx(), 1, m->getNz());
    p->model.alloc(sizeof(float) * 3 * p->n);
    p->model.upload(m->getL(), sizeof(float) * p->n, 0);
    p->model.upload(m->getRx(), sizeof(float) * p->n, sizeof(float) * p->n);
    p->model.upload(m->getRz(), sizeof(float) * p->n, 2 * sizeof(float) * p->n);
    p->s.L = p->model.f();
    p->s.Rx = p->model.f() + p->n;
    p->s.Rz = p->model.f() + 2 * p->n;
    p->g.fs = m->getFs() ? 1 : 0;
    tables_of(p, w->getPml().get());
    p->model_ready = true;
}
}  // namespace

template <>
WavesAcoustic2D<float>::WavesAcoustic2D() {
    const int nx = this->getNx(), nz = this->getNz(), lpml = this->getLpml();
    this->setDim(2);
    Pml = std::make_shared<PmlAcoustic2D<float>>(nx, nz, lpml, this->getDt());
    const size_t n = (size_t)(nx + 2 * lpml) * (nz + 2 * lpml);
    P = host_array(n);
    Vx = host_array(n);
    Vz = host_array(n);
    make_a2d(this, P, Vx, Vz);
}

template <>
WavesAcoustic2D<float>::WavesAcoustic2D(const int _nx, const int _nz, const int _nt, const int _lpml,
                                        const float _dx, const float _dz, const float _dt, const float _ox,
                                        const float _oz, const float _ot)
    : Waves<float>(2, _nx, 1, _nz, _nt, _lpml, _dx, 1.0, _dz, _dt, _ox, 0.0, _oz, _ot) {
    Pml = std::make_shared<PmlAcoustic2D<float>>(_nx, _nz, _lpml, _dt);
    this->setDim(2);
    const size_t n = (size_t)(_nx + 2 * _lpml) * (_nz + 2 * _lpml);
    P = host_array(n);
    Vx = host_array(n);
    Vz = host_array(n);
    make_a2d(this, P, Vx, Vz);
}

template <>
WavesAcoustic2D<float>::WavesAcoustic2D(std::shared_ptr<ModelAcoustic2D<float>> model, int _nt, float _dt,
                                        float _ot)
    : Waves<float>() {
    const int _nx = model->getNx(), _nz = model->getNz(), _lpml = model->getLpml();
    this->setNx(_nx);
    this->setNy(model->getNy());
    this->setNz(_nz);
    this->setNt(_nt);
    this->setDx(model->getDx());
    this->setDy(model->getDy());
    this->setDz(model->getDz());
    this->setDt(_dt);
    this->setOx(model->getOx());
    this->setOy(model->getOy());
    this->setOz(model->getOz());
    this->setOt(_ot);
    this->setLpml(_lpml);
    this->setDim(model->getDim());
    size_t n;
    if (model->getDomain()->getStatus()) {
        // a domain model: the arrays already hold the padding (waves.cpp:149-175)
        this->setDomain(model->getDomain());
        auto dom = model->getDomain();
        bool low[2] = {dom->getIx0() < _lpml, dom->getIz0() < _lpml};
        bool high[2] = {dom->getIx0() + _nx >= dom->getNx_orig() - _lpml,
                        dom->getIz0() + _nz >= dom->getNz_orig() - _lpml};
        Pml = std::make_shared<PmlAcoustic2D<float>>(_nx, _nz, _lpml, _dt, &low[0], &high[0]);
        n = (size_t)_nx * _nz;
    } else {
        Pml = std::make_shared<PmlAcoustic2D<float>>(_nx, _nz, _lpml, _dt);
        this->setDim(2);
        n = (size_t)(_nx + 2 * _lpml) * (_nz + 2 * _lpml);
    }
    P = host_array(n);
    Vx = host_array(n);
    Vz = host_array(n);
    make_a2d(this, P, Vx, Vz);
}

template <>
WavesAcoustic2D<float>::~WavesAcoustic2D() {
    release(this);
    std::free(P);
    std::free(Vx);
    std::free(Vz);
}

template <>
void WavesAcoustic2D<float>::forwardstepVelocity(std::shared_ptr<ModelAcoustic2D<float>> model,
                                                 std::shared_ptr<Der<float>> der) {
    Prop *p = rsgd::prop(this);
    prepare_a2d(p, this, model.get(), der.get());
    rsgd::ok(rsg_a2d_forward_velocity(&p->g, &p->s, nullptr), "WavesAcoustic2D::forwardstepVelocity");
    if (rsgd::eager_mirror()) rsgd::mirror_fields(p);
}

template <>
void WavesAcoustic2D<float>::forwardstepStress(std::shared_ptr<ModelAcoustic2D<float>> model,
                                               std::shared_ptr<Der<float>> der) {
    Prop *p = rsgd::prop(this);
    prepare_a2d(p, this, model.get(), der.get());
    rsgd::ok(rsg_a2d_forward_stress(&p->g, &p->s, nullptr), "WavesAcoustic2D::forwardstepStress");
    if (rsgd::eager_mirror()) rsgd::mirror_fields(p);
}

template <>
void WavesAcoustic2D<float>::insertSource(std::shared_ptr<ModelAcoustic2D<float>> model,
                                          std::shared_ptr<Data2D<float>> source, bool maptype, int it) {
    Prop *p = rsgd::prop(this);
    prepare_a2d(p, this, model.get(), nullptr);
    check_wavelet_nt((int)source->getNt(), this->getNt(), "WavesAcoustic2D::insertSource");
    rsgd::Traces *t = traces2d(p, source, maptype, this->getNt(), false);
    rsgd::ok(rsg_insert_source(RSG_KIND_A2D, &p->g, &p->s, &t->t, it, nullptr), "WavesAcoustic2D::insertSource");
    if (rsgd::eager_mirror()) rsgd::mirror_fields(p);
}

template <>
void WavesAcoustic2D<float>::recordData(std::shared_ptr<Data2D<float>> data, bool maptype, int it) {
    Prop *p = rsgd::prop(this);
    rsgd::Traces *t = traces2d(p, data, maptype, (int)data->getNt(), true);
    rsgd::ok(rsg_record_data(RSG_KIND_A2D, &p->g, &p->s, &t->t, it, nullptr), "WavesAcoustic2D::recordData");
    after_record(p, this, t, data.get(), it);
}

// ---- WavesAcoustic3D<float> (waves.cpp:598-1275) ----------------------------

namespace {
Prop *make_a3d(WavesAcoustic3D<float> *w, float *P, float *Vx, float *Vy, float *Vz) {
    Prop *p = rsgd::create(w, RSG_KIND_A3D);
    p->g = grid_of(w, w->getNy_pml());
    p->n = (size_t)p->g.nx * p->g.ny * p->g.nz;
    const size_t L = p->g.lpml;
    const size_t lr = L * p->g.ny * p->g.nz, fb = (size_t)p->g.nx * L * p->g.nz, tb = (size_t)p->g.nx * p->g.ny * L;
    const int slots[12] = {RSG_A3D_P_LEFT, RSG_A3D_P_RIGHT, RSG_A3D_VXX_LEFT, RSG_A3D_VXX_RIGHT,
                           RSG_A3D_P_FRONT, RSG_A3D_P_BACK, RSG_A3D_VYY_FRONT, RSG_A3D_VYY_BACK,
                           RSG_A3D_P_TOP, RSG_A3D_P_BOTTOM, RSG_A3D_VZZ_TOP, RSG_A3D_VZZ_BOTTOM};
    const size_t sizes[12] = {lr, lr, lr, lr, fb, fb, fb, fb, tb, tb, tb, tb};
    alloc_state(p, 4, slots, sizes, 12);
    p->s.P = p->fields.f();
    p->s.Vx = p->s.P + p->n;
    p->s.Vy = p->s.P + 2 * p->n;
    p->s.Vz = p->s.P + 3 * p->n;
    rsgd::add_mirror(p, P, p->s.P);
    rsgd::add_mirror(p, Vx, p->s.Vx);
    rsgd::add_mirror(p, Vy, p->s.Vy);
    rsgd::add_mirror(p, Vz, p->s.Vz);
    return p;
}

void prepare_a3d(Prop *p, WavesAcoustic3D<float> *w, ModelAcoustic3D<float> *m, Der<float> *der) {
    if (der && !p->order_set) set_order(p, der);
    if (p->model_ready) return;
    check_model_size(p, m->getNx(), m->getNy(), m->getNz());
    p->model.alloc(sizeof(float) * 4 * p->n);
    const float *src[4] = {m->getL(), m->getRx(), m->getRy(), m->getRz()};
    for (int k = 0; k < 4; k++) p->model.upload(src[k], sizeof(float) * p->n, k * sizeof(float) * p->n);
    p->s.L = p->model.f();
    p->s.Rx = p->model.f() + p->n;
    p->s.Ry = p->model.f() + 2 * p->n;
    p->s.Rz = p->model.f() + 3 * p->n;
    p->g.fs = m->getFs() ? 1 : 0;
    tables_of(p, w->getPml().get());
    p->model_ready = true;
}
}  // namespace

template <>
WavesAcoustic3D<float>::WavesAcoustic3D() {
    // the reference's default object holds no arrays (waves.cpp:598-600)
    P = Vx = Vy = Vz = nullptr;
}

template <>
WavesAcoustic3D<float>::WavesAcoustic3D(const int _nx, const int _ny, const int _nz, const int _nt,
                                        const int _lpml, const float _dx, const float _dy, const float _dz,
                                        const float _dt, const float _ox, const float _oy, const float _oz,
                                        const float _ot)
    : Waves<float>(3, _nx, _ny, _nz, _nt, _lpml, _dx, _dy, _dz, _dt, _ox, _oy, _oz, _ot) {
    Pml = std::make_shared<PmlAcoustic3D<float>>(_nx, _ny, _nz, _lpml, _dt);
    const size_t n = (size_t)(_nx + 2 * _lpml) * (_ny + 2 * _lpml) * (_nz + 2 * _lpml);
    P = host_array(n);
    Vx = host_array(n);
    Vy = host_array(n);
    Vz = host_array(n);
    make_a3d(this, P, Vx, Vy, Vz);
}

template <>
WavesAcoustic3D<float>::WavesAcoustic3D(std::shared_ptr<ModelAcoustic3D<float>> model, int _nt, float _dt,
                                        float _ot)
    : Waves<float>() {
    const int _nx = model->getNx(), _ny = model->getNy(), _nz = model->getNz(), _lpml = model->getLpml();
    this->setNx(_nx);
    this->setNy(_ny);
    this->setNz(_nz);
    this->setNt(_nt);
    this->setDx(model->getDx());
    this->setDy(model->getDy());
    this->setDz(model->getDz());
    this->setDt(_dt);
    this->setOx(model->getOx());
    this->setOy(model->getOy());
    this->setOz(model->getOz());
    this->setOt(_ot);
    this->setLpml(_lpml);
    this->setDim(model->getDim());
    size_t n;
    if (model->getDomain()->getStatus()) {
        this->setDomain(model->getDomain());
        auto dom = model->getDomain();
        bool low[3] = {dom->getIx0() < _lpml, dom->getIy0() < _lpml, dom->getIz0() < _lpml};
        bool high[3] = {dom->getIx0() + _nx >= dom->getNx_orig() - _lpml,
                        dom->getIy0() + _ny >= dom->getNy_orig() - _lpml,
                        dom->getIz0() + _nz >= dom->getNz_orig() - _lpml};
        Pml = std::make_shared<PmlAcoustic3D<float>>(_nx, _ny, _nz, _lpml, _dt, &low[0], &high[0]);
        n = (size_t)_nx * _ny * _nz;
    } else {
        Pml = std::make_shared<PmlAcoustic3D<float>>(_nx, _ny, _nz, _lpml, _dt);
        n = (size_t)(_nx + 2 * _lpml) * (_ny + 2 * _lpml) * (_nz + 2 * _lpml);
    }
    P = host_array(n);
    Vx = host_array(n);
    Vy = host_array(n);
    Vz = host_array(n);
    make_a3d(this, P, Vx, Vy, Vz);
}

template <>
WavesAcoustic3D<float>::~WavesAcoustic3D() {
    release(this);
    std::free(P);
    std::free(Vx);
    std::free(Vy);
    std::free(Vz);
}

template <>
void WavesAcoustic3D<float>::forwardstepVelocity(std::shared_ptr<ModelAcoustic3D<float>> model,
                                                 std::shared_ptr<Der<float>> der) {
    Prop *p = rsgd::prop(this);
    prepare_a3d(p, this, model.get(), der.get());
    rsgd::ok(rsg_a3d_forward_velocity(&p->g, &p->s, nullptr), "WavesAcoustic3D::forwardstepVelocity");
    if (rsgd::eager_mirror()) rsgd::mirror_fields(p);
}

template <>
void WavesAcoustic3D<float>::forwardstepStress(std::shared_ptr<ModelAcoustic3D<float>> model,
                                               std::shared_ptr<Der<float>> der) {
    Prop *p = rsgd::prop(this);
    prepare_a3d(p, this, model.get(), der.get());
    rsgd::ok(rsg_a3d_forward_stress(&p->g, &p->s, nullptr), "WavesAcoustic3D::forwardstepStress");
    if (rsgd::eager_mirror()) rsgd::mirror_fields(p);
}

template <>
void WavesAcoustic3D<float>::insertSource(std::shared_ptr<ModelAcoustic3D<float>> model,
                                          std::shared_ptr<Data3D<float>> source, bool maptype, int it) {
    Prop *p = rsgd::prop(this);
    prepare_a3d(p, this, model.get(), nullptr);
    check_wavelet_nt((int)source->getNt(), this->getNt(), "WavesAcoustic3D::insertSource");
    rsgd::Traces *t = traces3d(p, source, maptype, this->getNt(), false);
    rsgd::ok(rsg_insert_source(RSG_KIND_A3D, &p->g, &p->s, &t->t, it, nullptr), "WavesAcoustic3D::insertSource");
    if (rsgd::eager_mirror()) rsgd::mirror_fields(p);
}

template <>
void WavesAcoustic3D<float>::recordData(std::shared_ptr<Data3D<float>> data, bool maptype, int it) {
    Prop *p = rsgd::prop(this);
    rsgd::Traces *t = traces3d(p, data, maptype, (int)data->getNt(), true);
    rsgd::ok(rsg_record_data(RSG_KIND_A3D, &p->g, &p->s, &t->t, it, nullptr), "WavesAcoustic3D::recordData");
    after_record(p, this, t, data.get(), it);
}

// ---- WavesElastic2D<float> (waves.cpp:1294-2302) ----------------------------

namespace {
Prop *make_e2d(WavesElastic2D<float> *w, float *Sxx, float *Szz, float *Sxz, float *Vx, float *Vz) {
    Prop *p = rsgd::create(w, RSG_KIND_E2D);
    p->g = grid_of(w, 1);
    p->n = (size_t)p->g.nx * p->g.nz;
    const size_t xs = (size_t)p->g.lpml * p->g.nz, zs = (size_t)p->g.nx * p->g.lpml;
    const int slots[16] = {RSG_E2D_SXX_LEFT, RSG_E2D_SXX_RIGHT, RSG_E2D_SXZX_LEFT, RSG_E2D_SXZX_RIGHT,
                           RSG_E2D_VXX_LEFT, RSG_E2D_VXX_RIGHT, RSG_E2D_VZX_LEFT, RSG_E2D_VZX_RIGHT,
                           RSG_E2D_SZZ_TOP, RSG_E2D_SZZ_BOTTOM, RSG_E2D_SXZZ_TOP, RSG_E2D_SXZZ_BOTTOM,
                           RSG_E2D_VZZ_TOP, RSG_E2D_VZZ_BOTTOM, RSG_E2D_VXZ_TOP, RSG_E2D_VXZ_BOTTOM};
    const size_t sizes[16] = {xs, xs, xs, xs, xs, xs, xs, xs, zs, zs, zs, zs, zs, zs, zs, zs};
    alloc_state(p, 5, slots, sizes, 16);
    float *f = p->fields.f();
    p->s.Sxx = f;
    p->s.Szz = f + p->n;
    p->s.Sxz = f + 2 * p->n;
    p->s.Vx = f + 3 * p->n;
    p->s.Vz = f + 4 * p->n;
    rsgd::add_mirror(p, Sxx, p->s.Sxx);
    rsgd::add_mirror(p, Szz, p->s.Szz);
    rsgd::add_mirror(p, Sxz, p->s.Sxz);
    rsgd::add_mirror(p, Vx, p->s.Vx);
    rsgd::add_mirror(p, Vz, p->s.Vz);
    return p;
}

void prepare_e2d(Prop *p, WavesElastic2D<float> *w, ModelElastic2D<float> *m, Der<float> *der) {
    if (der && !p->order_set) set_order(p, der);
    if (p->model_ready) return;
    check_model_size(p, m->getNx(), 1, m->getNz());
    p->model.alloc(sizeof(float) * 5 * p->n);
    const float *src[5] = {m->getL(), m->getL2M(), m->getM(), m->getRx(), m->getRz()};
    for (int k = 0; k < 5; k++) p->model.upload(src[k], sizeof(float) * p->n, k * sizeof(float) * p->n);
    p->s.L = p->model.f();
    p->s.L2M = p->model.f() + p->n;
    p->s.M = p->model.f() + 2 * p->n;
    p->s.Rx = p->model.f() + 3 * p->n;
    p->s.Rz = p->model.f() + 4 * p->n;
    p->g.fs = m->getFs() ? 1 : 0;
    tables_of(p, w->getPml().get());
    p->model_ready = true;
}

// Elastic pressure records: divisor rho*(Vp^2 - Vs^2) at each receiver's
// cell under reciprocity, else 1 (waves.cpp:2219-2228), in the reference's
// float arithmetic.
void elastic_factor(rsgd::Traces *t, ModelElastic2D<float> *m, bool reciprocity) {
    if (t->t.field != RSG_PRESSURE || t->factor.p()) return;
    const int ntr = t->t.ntrace, mnx = m->getNx();
    std::vector<int> map(2 * (size_t)ntr);
    t->map.download(map.data(), sizeof(int) * map.size());
    std::vector<float> f(ntr, 1.0f);
    const float *R = m->getR(), *Vp = m->getVp(), *Vs = m->getVs();
    if (reciprocity)
        for (int i = 0; i < ntr; i++) {
            const int x = map[2 * i], z = map[2 * i + 1];
            if (x < 0 || z < 0) continue;
            const size_t q = (size_t)z * mnx + x;
            const float a = R[q] * Vp[q] * Vp[q], b = R[q] * Vs[q] * Vs[q];
            f[i] = a - b;
        }
    t->factor.alloc(sizeof(float) * ntr);
    t->factor.upload(f.data(), sizeof(float) * ntr);
    t->t.factor = t->factor.f();
}
}  // namespace

template <>
WavesElastic2D<float>::WavesElastic2D() {
    // the reference's default object holds no arrays (waves.cpp:1294-1297)
    Sxx = Szz = Sxz = Vx = Vz = wrk = nullptr;
    adjoint = false;
}

template <>
WavesElastic2D<float>::WavesElastic2D(const int _nx, const int _nz, const int _nt, const int _lpml,
                                      const float _dx, const float _dz, const float _dt, const float _ox,
                                      const float _oz, const float _ot)
    : Waves<float>(2, _nx, 1, _nz, _nt, _lpml, _dx, 1.0, _dz, _dt, _ox, 0.0, _oz, _ot) {
    Pml = std::make_shared<PmlElastic2D<float>>(_nx, _nz, _lpml, _dt);
    const size_t n = (size_t)(_nx + 2 * _lpml) * (_nz + 2 * _lpml);
    Sxx = host_array(n);
    Szz = host_array(n);
    Sxz = host_array(n);
    Vx = host_array(n);
    Vz = host_array(n);
    wrk = nullptr;
    adjoint = false;
    make_e2d(this, Sxx, Szz, Sxz, Vx, Vz);
}

template <>
WavesElastic2D<float>::WavesElastic2D(std::shared_ptr<ModelElastic2D<float>> model, int _nt, float _dt,
                                      float _ot)
    : Waves<float>() {
    const int _nx = model->getNx(), _nz = model->getNz(), _lpml = model->getLpml();
    this->setNx(_nx);
    this->setNy(model->getNy());
    this->setNz(_nz);
    this->setNt(_nt);
    this->setDx(model->getDx());
    this->setDy(model->getDy());
    this->setDz(model->getDz());
    this->setDt(_dt);
    this->setOx(model->getOx());
    this->setOy(model->getOy());
    this->setOz(model->getOz());
    this->setOt(_ot);
    this->setLpml(_lpml);
    this->setDim(model->getDim());
    adjoint = false;
    wrk = nullptr;
    size_t n;
    if (model->getDomain()->getStatus()) {
        this->setDomain(model->getDomain());
        auto dom = model->getDomain();
        bool low[2] = {dom->getIx0() < _lpml, dom->getIz0() < _lpml};
        bool high[2] = {dom->getIx0() + _nx >= dom->getNx_orig() - _lpml,
                        dom->getIz0() + _nz >= dom->getNz_orig() - _lpml};
        Pml = std::make_shared<PmlElastic2D<float>>(_nx, _nz, _lpml, _dt, &low[0], &high[0]);
        n = (size_t)_nx * _nz;
    } else {
        Pml = std::make_shared<PmlElastic2D<float>>(_nx, _nz, _lpml, _dt);
        this->setDim(2);
        n = (size_t)(_nx + 2 * _lpml) * (_nz + 2 * _lpml);
    }
    Sxx = host_array(n);
    Szz = host_array(n);
    Sxz = host_array(n);
    Vx = host_array(n);
    Vz = host_array(n);
    make_e2d(this, Sxx, Szz, Sxz, Vx, Vz);
}

template <>
WavesElastic2D<float>::~WavesElastic2D() {
    if (Sxx) release(this);
    std::free(Sxx);
    std::free(Szz);
    std::free(Sxz);
    std::free(Vx);
    std::free(Vz);
    std::free(wrk);
}

template <>
void WavesElastic2D<float>::setAdjoint() {
    // the adjoint-state kernels form the L2M*Sxx+L*Szz style products while
    // loading their tiles, so no work array is needed (waves.cpp:1406-1413)
    if (adjoint) rs_error("WavesElastic2D<T>::setAdjoint:Adjoint already set.");
    adjoint = true;
}

template <>
void WavesElastic2D<float>::forwardstepVelocity(std::shared_ptr<ModelElastic2D<float>> model,
                                                std::shared_ptr<Der<float>> der) {
    Prop *p = rsgd::prop(this);
    prepare_e2d(p, this, model.get(), der.get());
    rsgd::ok(rsg_e2d_forward_velocity(&p->g, &p->s, nullptr), "WavesElastic2D::forwardstepVelocity");
    if (rsgd::eager_mirror()) rsgd::mirror_fields(p);
}

template <>
void WavesElastic2D<float>::forwardstepStress(std::shared_ptr<ModelElastic2D<float>> model,
                                              std::shared_ptr<Der<float>> der) {
    Prop *p = rsgd::prop(this);
    prepare_e2d(p, this, model.get(), der.get());
    rsgd::ok(rsg_e2d_forward_stress(&p->g, &p->s, nullptr), "WavesElastic2D::forwardstepStress");
    if (rsgd::eager_mirror()) rsgd::mirror_fields(p);
}

template <>
void WavesElastic2D<float>::backwardstepVelocity(std::shared_ptr<ModelElastic2D<float>> model,
                                                 std::shared_ptr<Der<float>> der) {
    Prop *p = rsgd::prop(this);
    if (!adjoint) rs_error("WavesElastic2D::backwardstepVelocity: ", "setAdjoint() not called");
    prepare_e2d(p, this, model.get(), der.get());
    rsgd::ok(rsg_e2d_backward_velocity(&p->g, &p->s, nullptr), "WavesElastic2D::backwardstepVelocity");
    if (rsgd::eager_mirror()) rsgd::mirror_fields(p);
}

template <>
void WavesElastic2D<float>::backwardstepStress(std::shared_ptr<ModelElastic2D<float>> model,
                                               std::shared_ptr<Der<float>> der) {
    Prop *p = rsgd::prop(this);
    if (!adjoint) rs_error("WavesElastic2D::backwardstepStress: ", "setAdjoint() not called");
    prepare_e2d(p, this, model.get(), der.get());
    rsgd::ok(rsg_e2d_backward_stress(&p->g, &p->s, nullptr), "WavesElastic2D::backwardstepStress");
    if (rsgd::eager_mirror()) rsgd::mirror_fields(p);
}

template <>
void WavesElastic2D<float>::insertSource(std::shared_ptr<ModelElastic2D<float>> model,
                                         std::shared_ptr<Data2D<float>> source, bool maptype, int it) {
    Prop *p = rsgd::prop(this);
    prepare_e2d(p, this, model.get(), nullptr);
    check_wavelet_nt((int)source->getNt(), this->getNt(), "WavesElastic2D::insertSource");
    rsgd::Traces *t = traces2d(p, source, maptype, this->getNt(), false);
    rsgd::ok(rsg_insert_source(RSG_KIND_E2D, &p->g, &p->s, &t->t, it, nullptr), "WavesElastic2D::insertSource");
    if (rsgd::eager_mirror()) rsgd::mirror_fields(p);
}

template <>
void WavesElastic2D<float>::recordData(std::shared_ptr<ModelElastic2D<float>> model,
                                       std::shared_ptr<Data2D<float>> data, bool maptype, int it) {
    Prop *p = rsgd::prop(this);
    rsgd::Traces *t = traces2d(p, data, maptype, (int)data->getNt(), true);
    elastic_factor(t, model.get(), data->getReciprocity());
    rsgd::ok(rsg_record_data(RSG_KIND_E2D, &p->g, &p->s, &t->t, it, nullptr), "WavesElastic2D::recordData");
    after_record(p, this, t, data.get(), it);
}

}  // namespace rockseis
