// Internal state of the class-API drop-in (librsdropin.a): device buffers
// kept per reference object, over the C-ABI of include/rsgpu.h only.
//
// The drop-in defines the float members of the reference's Waves*, Der,
// Pml::computeABC, Revolve, Fwi* and Snapshot*::writeSnap classes (declared
// by the reference's own headers, which it is compiled against).  Those
// classes' layouts are fixed by the headers, so the device side of each
// object lives in a registry keyed by the object's address: created by the
// constructor, released by the destructor.
//
// Host mirrors.  The reference classes expose their fields as raw host
// pointers (getP(), getVx(), ...).  The drop-in keeps the fields in HBM and
// refreshes a host mirror only where a caller of the path reads it:
// Snapshot*::writeSnap (through the registry), and records, which are
// flushed into the caller's Data2D/3D arrays at the first and last time
// sample and when the Waves object is destroyed.  Setting RSGPU_HOST_MIRROR=1
// copies every field back after every half step as well, for other callers
// that read getP() between steps.  One caller thread per object, as in the
// reference; all work is ordered on the default stream.
#pragma once

#include <cstddef>
#include <memory>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/rsgpu.h"

namespace rsgd {

// A non-zero C-ABI return ends the program through the reference's fatal
// path (rs_error: "Fatal error: ..." and exit(1), utils.cpp:51-62).
void ok(int rc, const char *what);
[[noreturn]] void fatal(const char *what, const char *detail = "");

// One zeroed device allocation, freed with its owner.
class DevMem {
public:
    DevMem() = default;
    ~DevMem();
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    void alloc(size_t bytes);
    void upload(const void *host, size_t bytes, size_t offset = 0);
    void download(void *host, size_t bytes, size_t offset = 0) const;
    float *f() const { return (float *)p_; }
    int *i() const { return (int *)p_; }
    void *p() const { return p_; }
    size_t bytes() const { return n_; }

private:
    void *p_ = nullptr;
    size_t n_ = 0;
};

// Device copy of one trace set (a Data2D/3D object with one of its maps).
struct Traces {
    DevMem map, shift, data, factor;
    rsg_traces t{};
    float *host = nullptr;     // the caller's array (records are flushed into it)
    size_t count = 0;          // floats in data
    bool record = false;
    std::shared_ptr<void> keep;  // the Data object stays alive until the flush
};

// Device side of one Waves* object.
struct Prop {
    int kind = 0;  // RSG_KIND_A2D / _E2D / _A3D
    rsg_grid g{};
    rsg_state s{};
    size_t n = 0;  // cells of the grid
    DevMem fields, strips, model, tab;
    bool model_ready = false, order_set = false, adjoint = false;
    const void *model_key = nullptr;  // the model object whose arrays were uploaded
    std::vector<std::pair<float *, float *>> mirrors;  // (host mirror, device field)
    std::unordered_map<const void *, std::unique_ptr<Traces>> sets[2];  // [SMAP / GMAP]
};

Prop *create(const void *waves, int kind);
Prop *prop(const void *waves);  // the object's state (fatal when unknown)
void destroy(const void *waves);
void add_mirror(Prop *p, float *host, float *dev);

// Device field behind a host mirror pointer, or nullptr.
const float *device_of(const float *host);

// Trace set of `data` with map type `map` (0 SMAP, 1 GMAP) for propagator p:
// maps, shifts and samples are uploaded on first use.  dim 2 or 3.
Traces *traces(Prop *p, const void *data, std::shared_ptr<void> keep, int map, int field, int ntrace, int nt,
               const int *map_host, const float *shift_host, float *samples, int dim, bool record);
void flush(Traces *t);
void flush_records(Prop *p);
void mirror_fields(Prop *p);  // every device field into its host mirror
bool eager_mirror();

// PML tables of a Pml<float> object in the C-ABI's packed order.
void upload_tables(Prop *p, const float *const tables[12], int lpml);

}  // namespace rsgd
