// Registry and device-buffer helpers of the class-API drop-in.
#include "rsgd.h"

#include <cstdlib>
#include <cstring>
#include <string>

#include "utils.h"

namespace rsgd {

namespace {
std::unordered_map<const void *, std::unique_ptr<Prop>> &registry() {
    static std::unordered_map<const void *, std::unique_ptr<Prop>> r;
    return r;
}
// host mirror -> device field of every registered propagator
std::unordered_map<const float *, const float *> &mirror_index() {
    static std::unordered_map<const float *, const float *> m;
    return m;
}
}  // namespace

void fatal(const char *what, const char *detail) {
    rockseis::rs_error(std::string(what), std::string(detail ? detail : ""));
    std::exit(1);  // rs_error exits; kept for [[noreturn]]
}

void ok(int rc, const char *what) {
    if (rc != RSG_OK) {
        const char *m = rsg_last_error();
        fatal((std::string(what) + ": ").c_str(), m ? m : "");
    }
}

DevMem::~DevMem() {
    if (p_) (void)rsg_free(p_);
}

void DevMem::alloc(size_t bytes) {
    if (p_) (void)rsg_free(p_);
    p_ = nullptr;
    n_ = bytes;
    ok(rsg_malloc(&p_, bytes ? bytes : 4), "rsgpu drop-in: device allocation");
    ok(rsg_zero(p_, bytes ? bytes : 4, nullptr), "rsgpu drop-in: zero");
}

void DevMem::upload(const void *host, size_t bytes, size_t offset) {
    if (bytes) ok(rsg_memcpy_h2d((char *)p_ + offset, host, bytes, nullptr), "rsgpu drop-in: upload");
}

void DevMem::download(void *host, size_t bytes, size_t offset) const {
    if (!bytes) return;
    ok(rsg_memcpy_d2h(host, (const char *)p_ + offset, bytes, nullptr), "rsgpu drop-in: download");
    ok(rsg_stream_sync(nullptr), "rsgpu drop-in: synchronize");
}

Prop *create(const void *waves, int kind) {
    auto &r = registry();
    auto it = r.find(waves);
    if (it != r.end()) destroy(waves);  // an object rebuilt at the same address
    auto p = std::make_unique<Prop>();
    p->kind = kind;
    Prop *raw = p.get();
    r[waves] = std::move(p);
    return raw;
}

Prop *prop(const void *waves) {
    auto &r = registry();
    auto it = r.find(waves);
    if (it == r.end()) fatal("rsgpu drop-in: unknown Waves object");
    return it->second.get();
}

void destroy(const void *waves) {
    auto &r = registry();
    auto it = r.find(waves);
    if (it == r.end()) return;
    flush_records(it->second.get());
    for (auto &m : it->second->mirrors) mirror_index().erase(m.first);
    r.erase(it);
}

void add_mirror(Prop *p, float *host, float *dev) {
    p->mirrors.emplace_back(host, dev);
    mirror_index()[host] = dev;
}

const float *device_of(const float *host) {
    auto &m = mirror_index();
    auto it = m.find(host);
    return it == m.end() ? nullptr : it->second;
}

Traces *traces(Prop *p, const void *data, std::shared_ptr<void> keep, int map, int field, int ntrace, int nt,
               const int *map_host, const float *shift_host, float *samples, int dim, bool record) {
    auto &set = p->sets[map ? 1 : 0];
    auto it = set.find(data);
    if (it != set.end()) {
        Traces *t = it->second.get();
        if (t->t.ntrace == ntrace && t->t.nt == nt && t->host == samples && t->t.field == field) {
            t->record = t->record || record;
            return t;
        }
        if (t->record) flush(t);
        set.erase(it);
    }
    auto t = std::make_unique<Traces>();
    const size_t nm = (size_t)ntrace * dim;
    t->map.alloc(sizeof(int) * nm);
    t->shift.alloc(sizeof(float) * nm);
    t->map.upload(map_host, sizeof(int) * nm);
    t->shift.upload(shift_host, sizeof(float) * nm);
    t->count = (size_t)ntrace * nt;
    t->data.alloc(sizeof(float) * t->count);
    t->data.upload(samples, sizeof(float) * t->count);  // records accumulate (waves.cpp:545)
    t->host = samples;
    t->record = record;
    t->keep = std::move(keep);
    t->t.field = field;
    t->t.ntrace = ntrace;
    t->t.nt = nt;
    t->t.map = t->map.i();
    t->t.shift = t->shift.f();
    t->t.data = t->data.f();
    t->t.factor = nullptr;
    Traces *raw = t.get();
    set[data] = std::move(t);
    return raw;
}

void flush(Traces *t) {
    if (t->record && t->host) t->data.download(t->host, sizeof(float) * t->count);
}

void flush_records(Prop *p) {
    for (auto &set : p->sets)
        for (auto &kv : set)
            if (kv.second->record) flush(kv.second.get());
}

void mirror_fields(Prop *p) {
    if (p->mirrors.empty()) return;
    for (auto &m : p->mirrors)
        ok(rsg_memcpy_d2h(m.first, m.second, sizeof(float) * p->n, nullptr), "rsgpu drop-in: field download");
    ok(rsg_stream_sync(nullptr), "rsgpu drop-in: synchronize");
}

bool eager_mirror() {
    static const int v = [] {
        const char *e = std::getenv("RSGPU_HOST_MIRROR");
        return (e && e[0] == '1') ? 1 : 0;
    }();
    return v != 0;
}

void upload_tables(Prop *p, const float *const tables[12], int lpml) {
    std::vector<float> tab(12 * (size_t)lpml);
    for (int r = 0; r < 12; r++) std::memcpy(tab.data() + (size_t)r * lpml, tables[r], sizeof(float) * lpml);
    p->tab.alloc(sizeof(float) * tab.size());
    p->tab.upload(tab.data(), sizeof(float) * tab.size());
    p->s.tab = p->tab.f();
}

}  // namespace rsgd
