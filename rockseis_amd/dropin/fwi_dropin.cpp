// Float members of FwiAcoustic2D (include/fwi.h:106-170; fwi.cpp:511-1332)
// over the device gradient of the C-ABI.
//
// run_optimal() and run() hand the whole shot to rsg_fwi_a2d_gradient:
// forward steps with the source, the modelled record, misfit and residuals,
// the adjoint wavefield driven by the residuals and the zero-lag imaging
// condition at every reverse step, with Revolve checkpoints (run_optimal)
// or the snapshot cadence of run() kept in HBM.  The result is what the
// reference's member functions leave in the objects the caller set: the
// modelled and residual data, the wavelet gradient, the Vp / rho gradient
// images (and the source illumination when set) and the misfit.  A model
// built by getDomainmodel (domain decomposition mode, one domain) keeps the
// reference's imaging window: the whole padded domain (fwi.cpp:535-544).
//
// crossCorr, computeMisfit and computeResiduals keep their public
// signatures for callers that use them on host arrays; they run on the
// device as well.
#include <cstring>
#include <vector>

#include "fwi.h"
#include "rsgd.h"

namespace {

using rockseis::Data2D;
using rockseis::Image2D;
using rockseis::ModelAcoustic2D;

struct DevTraces {
    rsgd::DevMem map, shift, data;
    rsg_traces t{};
    void from(const std::shared_ptr<Data2D<float>> &d, bool gmap, const char *what) {
        auto *geo = d->getGeom().get();
        const void *m = gmap ? (const void *)geo->getGmap() : (const void *)geo->getSmap();
        const void *s = gmap ? (const void *)geo->getGshift() : (const void *)geo->getSshift();
        if (!m || !s) rsgd::fatal(what, ": trace map not made (Data2D::makeMap)");
        const size_t ntr = d->getNtrace(), nt = d->getNt();
        map.alloc(sizeof(int) * 2 * ntr);
        shift.alloc(sizeof(float) * 2 * ntr);
        data.alloc(sizeof(float) * ntr * nt);
        map.upload(m, sizeof(int) * 2 * ntr);
        shift.upload(s, sizeof(float) * 2 * ntr);
        data.upload(d->getData(), sizeof(float) * ntr * nt);
        t.field = (int)d->getField();
        t.ntrace = (int)ntr;
        t.nt = (int)nt;
        t.map = map.i();
        t.shift = shift.f();
        t.data = data.f();
        t.factor = nullptr;
    }
};

// Everything a gradient of one shot needs from an FwiAcoustic2D object.
struct A2DJob {
    ModelAcoustic2D<float> *model;
    std::shared_ptr<Data2D<float>> source, dataP, datamodP, dataresP, dataweight, wavgrad;
    std::shared_ptr<Image2D<float>> vpgrad, rhograd, srcilum;
    int order, snapinc, ncheck, misfit_type;
    bool filter;
};

void need_pressure(const std::shared_ptr<Data2D<float>> &d, const char *what) {
    if (d && d->getField() != rockseis::PRESSURE)
        rsgd::fatal(what, ": the device gradient takes PRESSURE sources and receivers");
}

void same_map(const std::shared_ptr<Data2D<float>> &a, const std::shared_ptr<Data2D<float>> &b) {
    const size_t n = a->getNtrace();
    if ((size_t)b->getNtrace() != n || b->getNt() != a->getNt())
        rsgd::fatal("FwiAcoustic2D: ", "the wavelet gradient must have the source's traces and samples");
    auto *ga = a->getGeom().get();
    auto *gb = b->getGeom().get();
    if (!gb->getSmap() || std::memcmp(ga->getSmap(), gb->getSmap(), sizeof(int) * 2 * n) != 0 ||
        std::memcmp(ga->getSshift(), gb->getSshift(), sizeof(float) * 2 * n) != 0)
        rsgd::fatal("FwiAcoustic2D: ", "the wavelet gradient must be mapped like the source (SMAP)");
}

// Image of the gradient on the device: the host image (allocated by the
// caller's objects as the reference allocates them) is uploaded first, so
// the device adds to it as crossCorr adds to the host image.
struct DevImage {
    rsgd::DevMem m;
    float *host = nullptr;
    size_t n = 0;
    float *bind(const std::shared_ptr<Image2D<float>> &img, size_t cells, bool zero_first) {
        n = cells;
        m.alloc(sizeof(float) * n);
        if (!img) return m.f();
        if (!img->getAllocated()) img->allocateImage();
        if ((size_t)img->getNx() * img->getNz() * img->getNhx() * img->getNhz() != cells)
            rsgd::fatal("FwiAcoustic2D: ", "gradient image size does not match the model");
        host = img->getImagedata();
        if (zero_first) std::memset(host, 0, sizeof(float) * n);
        m.upload(host, sizeof(float) * n);
        return m.f();
    }
    void back() {
        if (host) m.download(host, sizeof(float) * n);
    }
};

double run_gradient(const A2DJob &j, int mode) {
    ModelAcoustic2D<float> *m = j.model;
    need_pressure(j.source, "FwiAcoustic2D");
    need_pressure(j.dataP, "FwiAcoustic2D");
    if (j.filter && j.misfit_type >= rockseis::ADAPTIVE_GAUSS && j.misfit_type <= rockseis::ADAPTIVE_LINEAR)
        rsgd::fatal("FwiAcoustic2D: ", "the residual band-pass filter of the adaptive misfits is not supported");
    const bool domdec = m->getDomain()->getStatus();
    const int lpml = m->getLpml(), nx = m->getNx(), nz = m->getNz();
    rsg_grid g{};
    g.nx = domdec ? nx : nx + 2 * lpml;
    g.ny = 1;
    g.nz = domdec ? nz : nz + 2 * lpml;
    g.lpml = lpml;
    g.order = j.order < 1 ? 1 : (j.order > 8 ? 8 : j.order);
    g.fs = m->getFs() ? 1 : 0;
    g.dx = m->getDx();
    g.dy = 1.0f;
    g.dz = m->getDz();
    g.dt = j.source->getDt();
    g.domdec = domdec ? 1 : 0;
    const size_t npad = (size_t)g.nx * g.nz, nimg = (size_t)nx * nz;

    rsgd::DevMem model, tab;
    model.alloc(sizeof(float) * (3 * npad + 2 * nimg));
    model.upload(m->getL(), sizeof(float) * npad, 0);
    model.upload(m->getRx(), sizeof(float) * npad, sizeof(float) * npad);
    model.upload(m->getRz(), sizeof(float) * npad, 2 * sizeof(float) * npad);
    model.upload(m->getVp(), sizeof(float) * nimg, 3 * sizeof(float) * npad);
    model.upload(m->getR(), sizeof(float) * nimg, sizeof(float) * (3 * npad + nimg));
    // run / run_optimal: setSmax(SMAX) and computeABC() (fwi.cpp:1050-1051, 1183-1186)
    std::vector<float> t(12 * (size_t)lpml);
    rsgd::ok(rsg_pml_tables(lpml, g.dt, SMAX, 1, t.data()), "FwiAcoustic2D: PML tables");
    tab.alloc(sizeof(float) * t.size());
    tab.upload(t.data(), sizeof(float) * t.size());

    DevTraces src, obs;
    src.from(j.source, false, "FwiAcoustic2D source");
    obs.from(j.dataP, true, "FwiAcoustic2D data");
    const size_t ndat = (size_t)obs.t.ntrace * obs.t.nt;
    rsgd::DevMem dmod, dres, wgrad, wei;
    dmod.alloc(sizeof(float) * ndat);
    dres.alloc(sizeof(float) * ndat);
    if (j.wavgrad) {
        same_map(j.source, j.wavgrad);
        wgrad.alloc(sizeof(float) * src.t.ntrace * src.t.nt);
    }
    if (j.dataweight) {
        if ((size_t)j.dataweight->getNtrace() * j.dataweight->getNt() != ndat)
            rsgd::fatal("FwiAcoustic2D: ", "data weights must match the recorded data");
        wei.alloc(sizeof(float) * ndat);
        wei.upload(j.dataweight->getData(), sizeof(float) * ndat);
    }
    DevImage vg, rg, si;
    rsg_fwi_a2d f{};
    f.L = model.f();
    f.Rx = model.f() + npad;
    f.Rz = model.f() + 2 * npad;
    f.tab = tab.f();
    f.Vp = model.f() + 3 * npad;
    f.Rho = model.f() + 3 * npad + nimg;
    f.nx = nx;
    f.nz = nz;
    f.src = src.t;
    f.obs = obs.t;
    f.dmod = dmod.f();
    f.dres = dres.f();
    f.wavgrad = j.wavgrad ? wgrad.f() : nullptr;
    f.vpgrad = vg.bind(j.vpgrad, nimg, true);  // run_optimal reallocates vpgrad (fwi.cpp:1190)
    f.rhograd = rg.bind(j.rhograd, nimg, false);
    f.srcilum = j.srcilum ? si.bind(j.srcilum, nimg, false) : nullptr;
    double misfit = 0.0;
    f.misfit = &misfit;
    f.ncheck = j.ncheck;
    f.mode = mode;
    f.mode_used = nullptr;
    f.snapinc = j.snapinc;
    f.fp16 = 0;
    f.misfit_type = j.misfit_type;
    f.weight = j.dataweight ? wei.f() : nullptr;
    f.adjoint = RSG_ADJOINT_REFERENCE;
    rsgd::ok(rsg_fwi_a2d_gradient(&g, &f, nullptr), "FwiAcoustic2D gradient");

    if (j.datamodP) dmod.download(j.datamodP->getData(), sizeof(float) * ndat);
    if (j.dataresP) dres.download(j.dataresP->getData(), sizeof(float) * ndat);
    if (j.wavgrad) wgrad.download(j.wavgrad->getData(), sizeof(float) * src.t.ntrace * src.t.nt);
    vg.back();
    rg.back();
    si.back();
    return misfit;
}

// computeMisfit / computeResiduals on the device (misfit.hip): map, weights
// and records as the reference reads them (fwi.cpp:574-1027).
double misfit_residuals(int type, const std::shared_ptr<Data2D<float>> &mod, const std::shared_ptr<Data2D<float>> &rec,
                        const std::shared_ptr<Data2D<float>> &wei, float *res_host, int scaling) {
    const size_t ntr = mod->getNtrace(), nt = mod->getNt();
    if ((size_t)rec->getNtrace() != ntr)
        rockseis::rs_error("Mismatch between number of traces in the modelled and recorded data.");
    if ((size_t)rec->getNt() != nt)
        rockseis::rs_error("Mismatch between number of time samples in the modelled and recorded data.");
    const void *map = mod->getGeom()->getGmap();
    if (!map) rsgd::fatal("FwiAcoustic2D: ", "modelled data not mapped (Data2D::makeMap)");
    const size_t n = ntr * nt;
    rsgd::DevMem dm, dr, dw, ds, dmap;
    dm.alloc(sizeof(float) * n);
    dr.alloc(sizeof(float) * n);
    ds.alloc(sizeof(float) * n);
    dmap.alloc(sizeof(int) * 2 * ntr);
    dm.upload(mod->getData(), sizeof(float) * n);
    dr.upload(rec->getData(), sizeof(float) * n);
    dmap.upload(map, sizeof(int) * 2 * ntr);
    if (wei) {
        dw.alloc(sizeof(float) * n);
        dw.upload(wei->getData(), sizeof(float) * n);
    }
    double misfit = 0.0;
    rsgd::ok(rsg_misfit_residuals(type, dm.f(), dr.f(), wei ? dw.f() : nullptr, ds.f(), dmap.i(), 2, (int)ntr,
                                  (int)nt, 1, scaling, &misfit, nullptr),
             "FwiAcoustic2D misfit");
    if (res_host) ds.download(res_host, sizeof(float) * n);
    return misfit;
}

}  // namespace

namespace rockseis {

template <>
int FwiAcoustic2D<float>::run_optimal() {
    if (!vpgradset && !rhogradset) rs_error("FwiAcoustic2D:crossCorr: No gradient set in fwi class");
    this->createLog(this->getLogfile());
    this->writeLog("Running 2D Acoustic full-waveform inversion gradient with optimal checkpointing (rsgpu).");
    A2DJob j{model.get(), source, dataP, datamodPset ? datamodP : nullptr, dataresPset ? dataresP : nullptr,
             dataweightset ? dataweight : nullptr, wavgradset ? wavgrad : nullptr, vpgradset ? vpgrad : nullptr,
             rhogradset ? rhograd : nullptr, srcilumset ? srcilum : nullptr, this->getOrder(), this->getSnapinc(),
             this->getNcheck(), (int)this->getMisfit_type(), this->getFilter()};
    this->setMisfit((float)run_gradient(j, RSG_FWI_REVOLVE));
    this->writeLog("\nGradient computation completed.");
    return FWI_OK;
}

template <>
int FwiAcoustic2D<float>::run() {
    if (!datamodPset || !dataresPset)
        rs_error("FwiAcoustic2D::run: datamodP and dataresP must be set before running the simulation.");
    if (!this->checkStability())
        rs_error("FwiAcoustic2D::run: Wavelet sampling interval (dt) does not match the stability criteria.");
    if (!vpgradset && !rhogradset) rs_error("FwiAcoustic2D:crossCorr: No gradient set in fwi class");
    this->createLog(this->getLogfile());
    this->writeLog("Running 2D Acoustic full-waveform inversion gradient with full checkpointing (rsgpu).");
    A2DJob j{model.get(), source, dataP, datamodP, dataresP, dataweightset ? dataweight : nullptr,
             wavgradset ? wavgrad : nullptr, vpgradset ? vpgrad : nullptr, rhogradset ? rhograd : nullptr,
             srcilumset ? srcilum : nullptr, this->getOrder(), this->getSnapinc(), this->getNcheck(),
             (int)this->getMisfit_type(), this->getFilter()};
    this->setMisfit((float)run_gradient(j, RSG_FWI_FULL));
    this->writeLog("\nGradient computation completed.");
    return FWI_OK;
}

template <>
int FwiAcoustic2D<float>::run_edge() {
    rs_error("FwiAcoustic2D::run_edge: ", "boundary-saving reconstruction is not provided by the rsgpu drop-in; use run_optimal");
    return FWI_ERR;
}

template <>
void FwiAcoustic2D<float>::crossCorr(float *wsp, int pads, float *wrp, float *wrx, float *wrz, int padr,
                                     std::shared_ptr<ModelAcoustic2D<float>> mdl, bool domdec) {
    (void)wrp;
    if (!vpgradset && !rhogradset) rs_error("FwiAcoustic2D:crossCorr: No gradient set in fwi class");
    const int nx = vpgrad->getNx(), nz = vpgrad->getNz();
    if (domdec) pads = padr = 0;
    const size_t ns = (size_t)(nx + 2 * pads) * (nz + 2 * pads), nr = (size_t)(nx + 2 * padr) * (nz + 2 * padr);
    const size_t ni = (size_t)nx * nz;
    rsgd::DevMem a;
    a.alloc(sizeof(float) * (ns + 4 * nr + 2 * ni));
    float *dwsp = a.f(), *dwrx = dwsp + ns, *dwrz = dwrx + nr, *dRx = dwrz + nr, *dRz = dRx + nr, *dVp = dRz + nr,
          *dRho = dVp + ni;
    a.upload(wsp, sizeof(float) * ns, 0);
    a.upload(wrx, sizeof(float) * nr, sizeof(float) * ns);
    a.upload(wrz, sizeof(float) * nr, sizeof(float) * (ns + nr));
    a.upload(mdl->getRx(), sizeof(float) * nr, sizeof(float) * (ns + 2 * nr));
    a.upload(mdl->getRz(), sizeof(float) * nr, sizeof(float) * (ns + 3 * nr));
    a.upload(mdl->getVp(), sizeof(float) * ni, sizeof(float) * (ns + 4 * nr));
    a.upload(mdl->getR(), sizeof(float) * ni, sizeof(float) * (ns + 4 * nr + ni));
    DevImage vg, rg, si;
    float *gv = vg.bind(vpgrad, ni, false), *gr = rg.bind(rhograd, ni, false);
    float *gs = srcilumset ? si.bind(srcilum, ni, false) : nullptr;
    rsgd::ok(rsg_crosscorr_a2d(dwsp, pads, dwrx, dwrz, padr, dVp, dRho, dRx, dRz, nx, nz, vpgrad->getDx(),
                               vpgrad->getDz(), domdec ? 1 : 0, gv, gr, gs, nullptr),
             "FwiAcoustic2D::crossCorr");
    vg.back();
    rg.back();
    si.back();
}

template <>
void FwiAcoustic2D<float>::computeMisfit() {
    this->setMisfit((float)misfit_residuals((int)this->getMisfit_type(), datamodP, dataP,
                                            dataweightset ? dataweight : nullptr, nullptr, 1));
}

template <>
void FwiAcoustic2D<float>::computeResiduals() {
    const int scaling = (dataresP->getField() == VX || dataresP->getField() == VZ) ? -1 : 1;
    misfit_residuals((int)this->getMisfit_type(), datamodP, dataP, dataweightset ? dataweight : nullptr,
                     dataresP->getData(), scaling);
}

}  // namespace rockseis

// ---- FwiElastic2D<float> (fwi.h:243-312; fwi.cpp:2591-3773, 3937-4143) -------
//
// run_optimal() hands the shot to rsg_fwi_e2d_gradient_batch as a batch of
// one window at offset 0: forward steps, the P / Ux / Uz records, misfit and
// residuals, the adjoint-state sweep with Revolve checkpoints in HBM, the
// lambda / mu / rho imaging condition, and scaleGrad into Vp / Vs / rho
// gradients, as run_optimal ends (fwi.cpp:4132).  Domain models keep the
// reference's imaging window (fwi.cpp:2666-2675).

namespace {

using rockseis::ModelElastic2D;

struct E2DJob {
    ModelElastic2D<float> *model;
    std::shared_ptr<Data2D<float>> source, obs[3], mod[3], res[3], wei[3], wavgrad;
    std::shared_ptr<Image2D<float>> grad[3];
    int order, ncheck, misfit_type;
};

double run_e2d(const E2DJob &j) {
    ModelElastic2D<float> *m = j.model;
    const int want[3] = {rockseis::PRESSURE, rockseis::VX, rockseis::VZ};
    for (int c = 0; c < 3; c++) {
        if (!j.obs[c] || j.obs[c]->getField() != want[c])
            rsgd::fatal("FwiElastic2D: ", "the device gradient takes the P, Ux (VX) and Uz (VZ) records together");
        if (j.obs[c]->getReciprocity())
            rsgd::fatal("FwiElastic2D: ", "reciprocity-scaled pressure records are not supported");
        if (!j.grad[c]) rsgd::fatal("FwiElastic2D: ", "the Vp, Vs and rho gradient images must all be set");
    }
    const int sf = (int)j.source->getField();
    if (sf != rockseis::PRESSURE && sf != rockseis::VX && sf != rockseis::VZ)
        rsgd::fatal("FwiElastic2D: ", "source field must be PRESSURE, VX or VZ");
    const bool domdec = m->getDomain()->getStatus();
    const int lpml = m->getLpml(), nx = m->getNx(), nz = m->getNz();
    rsg_fwi_e2d_shot s{};
    s.g.nx = domdec ? nx : nx + 2 * lpml;
    s.g.ny = 1;
    s.g.nz = domdec ? nz : nz + 2 * lpml;
    s.g.lpml = lpml;
    s.g.order = j.order < 1 ? 1 : (j.order > 8 ? 8 : j.order);
    s.g.fs = m->getFs() ? 1 : 0;
    s.g.dx = m->getDx();
    s.g.dy = 1.0f;
    s.g.dz = m->getDz();
    s.g.dt = j.source->getDt();
    s.g.domdec = domdec ? 1 : 0;
    const size_t npad = (size_t)s.g.nx * s.g.nz, nimg = (size_t)nx * nz;
    rsgd::DevMem model, tab;
    model.alloc(sizeof(float) * (5 * npad + 3 * nimg));
    const float *padded[5] = {m->getL(), m->getL2M(), m->getM(), m->getRx(), m->getRz()};
    for (int k = 0; k < 5; k++) model.upload(padded[k], sizeof(float) * npad, k * sizeof(float) * npad);
    const float *img[3] = {m->getVp(), m->getVs(), m->getR()};
    for (int k = 0; k < 3; k++) model.upload(img[k], sizeof(float) * nimg, sizeof(float) * (5 * npad + k * nimg));
    float *b = model.f();
    s.L = b;
    s.L2M = b + npad;
    s.M = b + 2 * npad;
    s.Rx = b + 3 * npad;
    s.Rz = b + 4 * npad;
    s.Vp = b + 5 * npad;
    s.Vs = b + 5 * npad + nimg;
    s.Rho = b + 5 * npad + 2 * nimg;
    s.ox = 0;
    std::vector<float> t(12 * (size_t)lpml);
    rsgd::ok(rsg_pml_tables(lpml, s.g.dt, SMAX, 1, t.data()), "FwiElastic2D: PML tables");
    tab.alloc(sizeof(float) * t.size());
    tab.upload(t.data(), sizeof(float) * t.size());

    DevTraces src, obs[3];
    src.from(j.source, false, "FwiElastic2D source");
    s.src = src.t;
    rsgd::DevMem dmod[3], dres[3], wei[3], wgrad, grads;
    size_t ndat = 0;
    for (int c = 0; c < 3; c++) {
        obs[c].from(j.obs[c], true, "FwiElastic2D data");
        s.obs[c] = obs[c].t;
        ndat = (size_t)obs[c].t.ntrace * obs[c].t.nt;
        dmod[c].alloc(sizeof(float) * ndat);
        dres[c].alloc(sizeof(float) * ndat);
        s.dmod[c] = dmod[c].f();
        s.dres[c] = dres[c].f();
        if (j.wei[c]) {
            if ((size_t)j.wei[c]->getNtrace() * j.wei[c]->getNt() != ndat)
                rsgd::fatal("FwiElastic2D: ", "data weights must match the recorded data");
            wei[c].alloc(sizeof(float) * ndat);
            wei[c].upload(j.wei[c]->getData(), sizeof(float) * ndat);
            s.weight[c] = wei[c].f();
        }
    }
    if (j.wavgrad) {
        same_map(j.source, j.wavgrad);
        wgrad.alloc(sizeof(float) * src.t.ntrace * src.t.nt);
        s.wavgrad = wgrad.f();
    }
    // run_optimal allocates fresh (zero) images (fwi.cpp:3973-3975)
    grads.alloc(sizeof(float) * 3 * nimg);
    for (int c = 0; c < 3; c++) {
        if (!j.grad[c]->getAllocated()) j.grad[c]->allocateImage();
        if ((size_t)j.grad[c]->getNx() * j.grad[c]->getNz() * j.grad[c]->getNhx() * j.grad[c]->getNhz() != nimg)
            rsgd::fatal("FwiElastic2D: ", "gradient image size does not match the model");
    }
    rsg_fwi_e2d f{};
    f.tab = tab.f();
    f.ncheck = j.ncheck;
    f.fp16 = 0;
    f.scale = 1;
    f.nx = nx;
    f.nz = nz;
    for (int c = 0; c < 3; c++) f.grad[c] = grads.f() + c * nimg;
    double misfit = 0.0;
    f.misfit = &misfit;
    f.misfit_type = j.misfit_type;
    rsgd::ok(rsg_fwi_e2d_gradient_batch(&f, &s, 1, nullptr), "FwiElastic2D gradient");
    for (int c = 0; c < 3; c++) {
        grads.download(j.grad[c]->getImagedata(), sizeof(float) * nimg, sizeof(float) * c * nimg);
        if (j.mod[c]) dmod[c].download(j.mod[c]->getData(), sizeof(float) * ndat);
        if (j.res[c]) dres[c].download(j.res[c]->getData(), sizeof(float) * ndat);
    }
    if (j.wavgrad) wgrad.download(j.wavgrad->getData(), sizeof(float) * src.t.ntrace * src.t.nt);
    return misfit;
}

// FwiElastic2D::computeMisfit / computeResiduals on the device.
double e2d_misfit(int type, const std::shared_ptr<Data2D<float>> mod[3], const std::shared_ptr<Data2D<float>> rec[3],
                  const std::shared_ptr<Data2D<float>> wei[3], const std::shared_ptr<Data2D<float>> res[3]) {
    const size_t ntr = mod[1]->getNtrace(), nt = mod[1]->getNt(), n = ntr * nt;
    const void *map = mod[1]->getGeom()->getGmap();
    if (!map) rsgd::fatal("FwiElastic2D: ", "modelled data not mapped (Data2D::makeMap)");
    rsgd::DevMem buf, dmap;
    buf.alloc(sizeof(float) * 12 * n);
    dmap.alloc(sizeof(int) * 2 * ntr);
    dmap.upload(map, sizeof(int) * 2 * ntr);
    rsg_e2d_records r{};
    r.nrec = (int)ntr;
    r.map = dmap.i();
    for (int c = 0; c < 3; c++) {
        if ((size_t)mod[c]->getNtrace() != ntr || (size_t)rec[c]->getNtrace() != ntr || (size_t)rec[c]->getNt() != nt)
            rockseis::rs_error("Mismatch between number of traces or samples in the modelled and recorded data.");
        float *pm = buf.f() + (4 * c) * n, *pr = pm + n, *pw = pr + n, *ps = pw + n;
        buf.upload(mod[c]->getData(), sizeof(float) * n, sizeof(float) * (4 * c) * n);
        buf.upload(rec[c]->getData(), sizeof(float) * n, sizeof(float) * (4 * c + 1) * n);
        r.mod[c] = pm;
        r.obs[c] = pr;
        if (wei[c]) {
            buf.upload(wei[c]->getData(), sizeof(float) * n, sizeof(float) * (4 * c + 2) * n);
            r.weight[c] = pw;
        }
        r.res[c] = ps;
    }
    double misfit = 0.0;
    rsgd::ok(rsg_misfit_residuals_e2d(type, &r, 1, (int)nt, &misfit, nullptr), "FwiElastic2D misfit");
    for (int c = 0; c < 3; c++)
        if (res[c]) buf.download(res[c]->getData(), sizeof(float) * n, sizeof(float) * (4 * c + 3) * n);
    return misfit;
}

}  // namespace

namespace rockseis {

template <>
int FwiElastic2D<float>::run_optimal() {
    if (!vpgradset && !vsgradset && !wavgradset) {
        rs_warning("FwiElastic2D::run: No image set");
        return FWI_ERR;
    }
    this->createLog(this->getLogfile());
    this->writeLog("Running 2D Elastic full-waveform inversion gradient with optimal checkpointing (rsgpu).");
    E2DJob j{};
    j.model = model.get();
    j.source = source;
    j.obs[0] = dataPset ? dataP : nullptr;
    j.obs[1] = dataUxset ? dataUx : nullptr;
    j.obs[2] = dataUzset ? dataUz : nullptr;
    j.mod[0] = datamodPset ? datamodP : nullptr;
    j.mod[1] = datamodUxset ? datamodUx : nullptr;
    j.mod[2] = datamodUzset ? datamodUz : nullptr;
    j.res[0] = dataresPset ? dataresP : nullptr;
    j.res[1] = dataresUxset ? dataresUx : nullptr;
    j.res[2] = dataresUzset ? dataresUz : nullptr;
    j.wei[0] = dataweightpset ? dataweightp : nullptr;
    j.wei[1] = dataweightxset ? dataweightx : nullptr;
    j.wei[2] = dataweightzset ? dataweightz : nullptr;
    j.wavgrad = wavgradset ? wavgrad : nullptr;
    j.grad[0] = vpgradset ? vpgrad : nullptr;
    j.grad[1] = vsgradset ? vsgrad : nullptr;
    j.grad[2] = rhogradset ? rhograd : nullptr;
    j.order = this->getOrder();
    j.ncheck = this->getNcheck();
    j.misfit_type = (int)this->getMisfit_type();
    this->setMisfit((float)run_e2d(j));
    this->writeLog("\nGradient computation completed.");
    return FWI_OK;
}

template <>
int FwiElastic2D<float>::run() {
    rs_error("FwiElastic2D::run: ", "full snapshot saving is not provided by the rsgpu drop-in; use run_optimal "
                                    "(snapmethod OPTIMAL)");
    return FWI_ERR;
}

template <>
void FwiElastic2D<float>::computeMisfit() {
    const std::shared_ptr<Data2D<float>> mod[3] = {datamodP, datamodUx, datamodUz};
    const std::shared_ptr<Data2D<float>> rec[3] = {dataP, dataUx, dataUz};
    const std::shared_ptr<Data2D<float>> wei[3] = {dataweightpset ? dataweightp : nullptr,
                                                   dataweightxset ? dataweightx : nullptr,
                                                   dataweightzset ? dataweightz : nullptr};
    const std::shared_ptr<Data2D<float>> none[3] = {nullptr, nullptr, nullptr};
    this->setMisfit((float)e2d_misfit((int)this->getMisfit_type(), mod, rec, wei, none));
}

template <>
void FwiElastic2D<float>::computeResiduals() {
    const std::shared_ptr<Data2D<float>> mod[3] = {datamodP, datamodUx, datamodUz};
    const std::shared_ptr<Data2D<float>> rec[3] = {dataP, dataUx, dataUz};
    const std::shared_ptr<Data2D<float>> wei[3] = {dataweightpset ? dataweightp : nullptr,
                                                   dataweightxset ? dataweightx : nullptr,
                                                   dataweightzset ? dataweightz : nullptr};
    const std::shared_ptr<Data2D<float>> res[3] = {dataresP, dataresUx, dataresUz};
    e2d_misfit((int)this->getMisfit_type(), mod, rec, wei, res);
}

}  // namespace rockseis
