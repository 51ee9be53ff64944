// Float members of Snapshot2D/3D::writeSnap (snap.cpp:638-669, 1138-1172)
// and of the Revolve class (revolve.cpp:8-305, 373-459, 622-1271) for the
// propagators the drop-in holds in HBM.
//
// writeSnap: the fields a snapshot was pointed at (setData(waves->getP(),
// 0), ...) are the host mirrors of device fields; the interior sum of the
// snapshot is formed on the device (rsg_snap_interior) and written to the
// snapshot file in one block, the same bytes the reference writes value
// by value.  Fields that are not device mirrors are summed on the host.
//
// Revolve: the schedule is the C-ABI's state machine (rsg_revolve_init /
// rsg_revolve_next, bit-exact with revolve.cpp); checkpoints of the
// acoustic 2D/3D and elastic 2D propagators (fields and PML strips) are
// device-to-device copies into HBM, for any `incore` flag, so no
// checkpoint file is written.  Other propagator kinds keep the reference's
// own checkpoint members.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "revolve.h"
#include "rsgd.h"
#include "snap.h"

namespace {

// Interior sum val = d1 (+ d2) (+ d3) in the reference's loop order, on the
// device when every given field is a device mirror.
void snapshot_block(float *const data[3], int nx, int ny, int nz, int lpml, bool is3d, std::vector<float> &out) {
    const size_t nxp = nx + 2 * lpml, nyp = is3d ? ny + 2 * lpml : 1, nzp = nz + 2 * lpml;
    const size_t n = (size_t)nx * (is3d ? ny : 1) * nz;
    out.assign(n, 0.0f);
    const float *dev[3] = {nullptr, nullptr, nullptr};
    bool on_device = data[0] != nullptr;  // a missing first field keeps the host path's running value
    for (int k = 0; k < 3; k++) {
        if (!data[k]) continue;
        dev[k] = rsgd::device_of(data[k]);
        if (!dev[k]) on_device = false;
    }
    if (on_device) {
        rsg_grid g{};
        g.nx = (int)nxp;
        g.ny = (int)nyp;
        g.nz = (int)nzp;
        g.lpml = lpml;
        rsgd::DevMem buf;
        buf.alloc(sizeof(float) * n);
        rsgd::ok(rsg_snap_interior(dev[0], dev[1], dev[2], &g, lpml, buf.f(), nullptr), "Snapshot::writeSnap");
        buf.download(out.data(), sizeof(float) * n);
        return;
    }
    float val = 0;
    size_t q = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < (is3d ? ny : 1); j++)
            for (int i = 0; i < nx; i++) {
                const size_t p = ((size_t)(k + lpml) * nyp + (size_t)(is3d ? j + lpml : 0)) * nxp + (i + lpml);
                if (data[0]) val = data[0][p];
                if (data[1]) val += data[1][p];
                if (data[2]) val += data[2][p];
                out[q++] = val;
            }
}

}  // namespace

namespace rockseis {

template <>
void Snapshot2D<float>::writeSnap(const int it) {
    std::shared_ptr<File> Fp = this->getFp();
    if (this->getOpen() && (it % this->getSnapinc()) == 0) {
        this->setSnapit(this->getSnapit() + 1);
        float *const d[3] = {this->getData(0), this->getData(1), this->getData(2)};
        std::vector<float> blk;
        snapshot_block(d, this->getNx(), 1, this->getNz(), this->getLpml(), false, blk);
        Fp->write(blk.data(), blk.size());
    }
    if (Fp->getFail()) rs_error("Snapshot2D::writeSnap: Error writting to file.");
}

template <>
void Snapshot3D<float>::writeSnap(const int it) {
    std::shared_ptr<File> Fp = this->getFp();
    if (this->getOpen() && (it % this->getSnapinc()) == 0) {
        this->setSnapit(this->getSnapit() + 1);
        float *const d[3] = {this->getData(0), this->getData(1), this->getData(2)};
        std::vector<float> blk;
        snapshot_block(d, this->getNx(), this->getNy(), this->getNz(), this->getLpml(), true, blk);
        Fp->write(blk.data(), blk.size());
    }
    if (Fp->getFail()) rs_error("Snapshot3D::writeSnap: Error writting to file.");
}

}  // namespace rockseis

// ---- Revolve<float> ---------------------------------------------------------

namespace {

struct RevolveDev {
    rsg_revolve_state st{};
    rsgd::DevMem store;  // snaps checkpoints of `slot` bytes
    size_t fields = 0, strips = 0, slot = 0;
};

std::unordered_map<const void *, std::unique_ptr<RevolveDev>> &revolves() {
    static std::unordered_map<const void *, std::unique_ptr<RevolveDev>> r;
    return r;
}

RevolveDev *rev_of(const void *self) {
    auto it = revolves().find(self);
    if (it == revolves().end()) rsgd::fatal("rsgpu drop-in: ", "unknown Revolve object");
    return it->second.get();
}

// Checkpoint store for the propagator's fields and strips (created on 'w').
void open_store(const void *self, const void *waves, char flag, int snaps) {
    RevolveDev *r = rev_of(self);
    rsgd::Prop *p = rsgd::prop(waves);
    const size_t f = p->fields.bytes(), s = p->strips.bytes();
    if (flag == 'w' || r->slot == 0) {
        r->fields = f;
        r->strips = s;
        r->slot = f + s;
        r->store.alloc(r->slot * (size_t)snaps);
    } else if (r->fields != f || r->strips != s) {
        rsgd::fatal("Revolve::openCheck: ", "checkpoints of a different propagator size");
    }
}

void write_slot(const void *self, const void *waves, int check) {
    RevolveDev *r = rev_of(self);
    rsgd::Prop *p = rsgd::prop(waves);
    if (!r->slot || check < 0) rsgd::fatal("Revolve::writeCheck: ", "checkpoints not opened");
    char *dst = (char *)r->store.p() + r->slot * (size_t)check;
    rsgd::ok(rsg_memcpy_d2d(dst, p->fields.p(), r->fields, nullptr), "Revolve::writeCheck");
    rsgd::ok(rsg_memcpy_d2d(dst + r->fields, p->strips.p(), r->strips, nullptr), "Revolve::writeCheck");
}

void read_slot(const void *self, const void *waves, int check) {
    RevolveDev *r = rev_of(self);
    rsgd::Prop *p = rsgd::prop(waves);
    if (!r->slot || check < 0) rsgd::fatal("Revolve::readCheck: ", "checkpoints not opened");
    const char *src = (const char *)r->store.p() + r->slot * (size_t)check;
    rsgd::ok(rsg_memcpy_d2d(p->fields.p(), src, r->fields, nullptr), "Revolve::readCheck");
    rsgd::ok(rsg_memcpy_d2d(p->strips.p(), src + r->fields, r->strips, nullptr), "Revolve::readCheck");
    if (rsgd::eager_mirror()) rsgd::mirror_fields(p);
}

}  // namespace

namespace rockseis {

template <>
Revolve<float>::Revolve(unsigned int _nt, unsigned int _snaps, bool _incore) {
    auto d = std::make_unique<RevolveDev>();
    int adjusted = 0;
    rsgd::ok(rsg_revolve_init(&d->st, (int)_nt, (int)_snaps, &adjusted), "Revolve::Revolve");
    if (adjusted) rs_warning("snaps invalid - setting to default optimal value.");
    steps = d->st.steps;
    snaps = d->st.snaps;
    Fc = std::make_shared<File>();
    capo = d->st.capo;
    fine = d->st.fine;
    check = d->st.check;
    info = 0;
    incore = _incore;
    open = false;
    allocated = false;
    checkpoints = nullptr;
    revolves()[this] = std::move(d);
}

template <>
Revolve<float>::~Revolve() {
    if (allocated) std::free(checkpoints);
    revolves().erase(this);
}

template <>
int Revolve<float>::adjust(int nsteps) {
    return rsg_revolve_adjust(nsteps);
}

template <>
int Revolve<float>::numforw(int nsteps, int nsnaps) {
    return rsg_revolve_numforw(nsteps, nsnaps);
}

template <>
double Revolve<float>::expense(int nsteps, int nsnaps) {
    // revolve.cpp:65-80: forward steps per step of the plain sweep
    if (nsnaps < 1) {
        std::fprintf(stderr, " error occurs in expense: snaps < 0\n");
        return -1;
    }
    if (nsteps < 1) {
        std::fprintf(stderr, " error occurs in expense: steps < 0\n");
        return -1;
    }
    return (double)numforw(nsteps, nsnaps) / (double)nsteps;
}

template <>
revolve_action Revolve<float>::revolve() {
    RevolveDev *d = rev_of(this);
    const int a = rsg_revolve_next(&d->st);
    capo = d->st.capo;
    fine = d->st.fine;
    check = d->st.check;
    switch (a) {
    case 0: return advance;
    case 1: return takeshot;
    case 2: return restore;
    case 3: return firsturn;
    case 4: return youturn;
    case 5: return terminate;
    default: return error;
    }
}

template <>
void Revolve<float>::openCheck(std::string _filename, std::shared_ptr<WavesAcoustic2D<float>> waves, char flag) {
    filename = _filename;
    checksize = (size_t)3 * waves->getNx_pml() * waves->getNz_pml() +
                (size_t)4 * waves->getLpml() * (waves->getNx_pml() + waves->getNz_pml());
    open_store(this, waves.get(), flag, snaps);
}

template <>
void Revolve<float>::openCheck(std::string _filename, std::shared_ptr<WavesAcoustic3D<float>> waves, char flag) {
    filename = _filename;
    const size_t nx = waves->getNx_pml(), ny = waves->getNy_pml(), nz = waves->getNz_pml(), L = waves->getLpml();
    checksize = 4 * nx * ny * nz + 4 * L * (ny * nz + nx * nz + nx * ny);
    open_store(this, waves.get(), flag, snaps);
}

template <>
void Revolve<float>::openCheck(std::string _filename, std::shared_ptr<WavesElastic2D<float>> waves, char flag) {
    filename = _filename;
    checksize = (size_t)5 * waves->getNx_pml() * waves->getNz_pml() +
                (size_t)8 * waves->getLpml() * (waves->getNx_pml() + waves->getNz_pml());
    open_store(this, waves.get(), flag, snaps);
}

template <>
void Revolve<float>::writeCheck(std::shared_ptr<WavesAcoustic2D<float>> waves) {
    write_slot(this, waves.get(), check);
}
template <>
void Revolve<float>::readCheck(std::shared_ptr<WavesAcoustic2D<float>> waves) {
    read_slot(this, waves.get(), check);
}
template <>
void Revolve<float>::writeCheck(std::shared_ptr<WavesAcoustic3D<float>> waves) {
    write_slot(this, waves.get(), check);
}
template <>
void Revolve<float>::readCheck(std::shared_ptr<WavesAcoustic3D<float>> waves) {
    read_slot(this, waves.get(), check);
}
template <>
void Revolve<float>::writeCheck(std::shared_ptr<WavesElastic2D<float>> waves) {
    write_slot(this, waves.get(), check);
}
template <>
void Revolve<float>::readCheck(std::shared_ptr<WavesElastic2D<float>> waves) {
    read_slot(this, waves.get(), check);
}

template <>
void Revolve<float>::closeCheck() {
    // HBM checkpoints stay in place between the forward and reverse sweeps;
    // a checkpoint file of another propagator kind is closed as the
    // reference does (revolve.cpp:622-629)
    if (!incore && open) {
        Fc->close();
        open = false;
    }
}

template <>
void Revolve<float>::removeCheck() {
    RevolveDev *d = rev_of(this);
    const bool hbm = d->slot != 0;
    d->store.alloc(0);
    d->slot = d->fields = d->strips = 0;
    if (hbm || incore) return;
    // a checkpoint file of another propagator kind (revolve.cpp:631-645)
    if (open) {
        Fc->close();
        open = false;
    }
    if (!filename.empty() && std::remove(filename.c_str()) != 0)
        rs_error("Snapshot::removeSnap: Error deleting file: ", filename);
}

}  // namespace rockseis
