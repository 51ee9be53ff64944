// Source injection and receiver gathers (WavesAcoustic2D/3D and
// WavesElastic2D insertSource / recordData, waves.cpp:407-595, 998-1275,
// 2085-2288): one lane per trace, bilinear (2D) or trilinear (3D) weights
// xtri = {1-shift, shift} over the 2x2(x2) cells at lpml + map.
//
// The time index is either a constant or read from device memory plus an
// offset, so that a captured hipGraph of several steps can be replayed.
#include "rsg_common.h"
#include "rsg_srcrec.h"
#include "rsg_srcrec_dev.h"

namespace rsg {

__device__ __forceinline__ int time_index(const int *it_dev, int it) {
    return it_dev ? *it_dev + it : it;
}

// F2 (optional) receives the same term: the elastic explosive source drives
// Sxx and Szz together.
// blockIdx.y is the shot of a batch: its field sits at fstride elements
// from shot 0's and its traces follow shot 0's (maps, wavelets).
__global__ void inject2d_kernel(float *F, float *F2, const float *Mod, float scale, int nx, int lpml,
                                int ntrace, const int *map, const float *shift, const float *wav,
                                int nt, const int *it_dev, int it0, size_t fstride, size_t ncell) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntrace) return;
    const int it = time_index(it_dev, it0);
    const size_t k = blockIdx.y;
    const int *m = map + k * 2 * ntrace;
    const float *sh = shift + k * 2 * ntrace;
    const float *w = wav + k * (size_t)ntrace * nt;
    inject2d_trace(i, F + k * fstride, Mod, scale, nx, lpml, m, sh, w, nt, it, ncell);
    if (F2) inject2d_trace(i, F2 + k * fstride, Mod, scale, nx, lpml, m, sh, w, nt, it, ncell);
}

__device__ __forceinline__ void inject3d_trace(int i, float *F, const float *Mod, float scale, int nx,
                                               int ny, int lpml, const int *map, const float *shift,
                                               const float *wav, int nt, int it, size_t ncell) {
    const int mx = map[3 * i], my = map[3 * i + 1], mz = map[3 * i + 2];
    if (mx < 0 || my < 0 || mz < 0) return;
    const float sx = shift[3 * i], sy = shift[3 * i + 1], sz = shift[3 * i + 2];
    const float xtri[2] = {1 - sx, sx}, ytri[2] = {1 - sy, sy}, ztri[2] = {1 - sz, sz};
    const float w = wav[(size_t)i * nt + it];
    const size_t nxy = (size_t)nx * ny;
#pragma unroll
    for (int i1 = 0; i1 < 2; i1++)
#pragma unroll
        for (int i2 = 0; i2 < 2; i2++)
#pragma unroll
            for (int i3 = 0; i3 < 2; i3++) {
                const size_t p = (size_t)(lpml + mz + i1) * nxy + (size_t)(lpml + my + i2) * nx +
                                 (lpml + mx + i3);
                if (p >= ncell) continue;
                const float m = Mod ? scale * Mod[p] : scale;
                unsafeAtomicAdd(F + p, m * w * xtri[i3] * ytri[i2] * ztri[i1]);
            }
}

__global__ void inject3d_kernel(float *F, const float *Mod, float scale, int nx, int ny, int lpml,
                                int ntrace, const int *map, const float *shift, const float *wav,
                                int nt, const int *it_dev, int it0, size_t ncell) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntrace) return;
    inject3d_trace(i, F, Mod, scale, nx, ny, lpml, map, shift, wav, nt, time_index(it_dev, it0), ncell);
}

__global__ void record2d_kernel(const float *F, const float *F2, const float *factor, int nx,
                                int lpml, int ntrace, const int *map, const float *shift,
                                float *data, int nt, const int *it_dev, int it0, size_t fstride, size_t ncell) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntrace) return;
    const size_t k = blockIdx.y;
    record2d_trace(i, F + k * fstride, F2 ? F2 + k * fstride : nullptr, factor ? factor + k * ntrace : nullptr,
                   nx, lpml, map + k * 2 * ntrace, shift + k * 2 * ntrace, data + k * (size_t)ntrace * nt, nt,
                   time_index(it_dev, it0), ncell);
}

__device__ __forceinline__ void record3d_trace(int i, const float *F, int nx, int ny, int lpml,
                                               const int *map, const float *shift, float *data,
                                               int nt, int it, size_t ncell) {
    const int mx = map[3 * i], my = map[3 * i + 1], mz = map[3 * i + 2];
    if (mx < 0 || my < 0 || mz < 0) return;
    const float sx = shift[3 * i], sy = shift[3 * i + 1], sz = shift[3 * i + 2];
    const float xtri[2] = {1 - sx, sx}, ytri[2] = {1 - sy, sy}, ztri[2] = {1 - sz, sz};
    const size_t nxy = (size_t)nx * ny;
    float *d = data + (size_t)i * nt + it;
    float acc = *d;
#pragma unroll
    for (int i1 = 0; i1 < 2; i1++)
#pragma unroll
        for (int i2 = 0; i2 < 2; i2++)
#pragma unroll
            for (int i3 = 0; i3 < 2; i3++) {
                const size_t p = (size_t)(lpml + mz + i1) * nxy + (size_t)(lpml + my + i2) * nx + (lpml + mx + i3);
                if (p < ncell) acc += xtri[i3] * ytri[i2] * ztri[i1] * F[p];
            }
    *d = acc;
}

__global__ void record3d_kernel(const float *F, int nx, int ny, int lpml, int ntrace, const int *map,
                                const float *shift, float *data, int nt, const int *it_dev, int it0, size_t ncell) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntrace) return;
    record3d_trace(i, F, nx, ny, lpml, map, shift, data, nt, time_index(it_dev, it0), ncell);
}

// All receiver sets of a step in one launch: lane g of the grid gathers
// trace g of the concatenated sets (set boundaries in `first`).
struct RecDesc {
    const float *F, *F2, *factor;
    int ntrace, nt;
    const int *map;
    const float *shift;
    float *data;
};
constexpr int kMaxFusedRec = 4;
struct RecSets {
    int dim, nx, ny, lpml, nrec;
    size_t ncell;
    int first[kMaxFusedRec + 1];
    RecDesc rec[kMaxFusedRec];
};

__global__ void record_sets_kernel(const RecSets rs, const int *it_dev, int it0) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= rs.first[rs.nrec]) return;
    int r = 0;
#pragma unroll
    for (int k = 1; k < kMaxFusedRec; k++)
        if (k < rs.nrec && g >= rs.first[k]) r = k;
    const RecDesc &rd = rs.rec[r];
    const int i = g - rs.first[r];
    const int it = time_index(it_dev, it0);
    if (rs.dim == 3)
        record3d_trace(i, rd.F, rs.nx, rs.ny, rs.lpml, rd.map, rd.shift, rd.data, rd.nt, it, rs.ncell);
    else
        record2d_trace(i, rd.F, rd.F2, rd.factor, rs.nx, rs.lpml, rd.map, rd.shift, rd.data, rd.nt, it, rs.ncell);
}

__global__ void add_int_kernel(int *p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p += v;
}
__global__ void set_int_kernel(int *p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// Field and modulus a source of `field` drives, per propagator kind
// (the switch statements of the insertSource members).
static int source_target(int kind, const rsg_state *s, int field, float dt, float **F, float **F2,
                         const float **Mod, float *scale) {
    *F2 = nullptr;
    *Mod = nullptr;
    *scale = dt;
    if (kind == RSG_KIND_E2D) {
        switch (field) {
        case RSG_PRESSURE:  // explosive: Sxx, Szz -= dt*wav*w (no modulus)
            *F = s->Sxx;
            *F2 = s->Szz;
            *scale = -dt;
            return RSG_OK;
        case RSG_VX: *F = s->Vx; *Mod = s->Rx; return RSG_OK;
        case RSG_VZ: *F = s->Vz; *Mod = s->Rz; return RSG_OK;
        default: return fail(RSG_EINVAL, "elastic 2D source field must be P, VX or VZ");
        }
    }
    switch (field) {
    case RSG_PRESSURE: *F = s->P; *Mod = s->L; return RSG_OK;
    case RSG_VX: *F = s->Vx; *Mod = s->Rx; return RSG_OK;
    case RSG_VY:
        if (kind != RSG_KIND_A3D) break;
        *F = s->Vy;
        *Mod = s->Ry;
        return RSG_OK;
    case RSG_VZ: *F = s->Vz; *Mod = s->Rz; return RSG_OK;
    default: break;
    }
    return fail(RSG_EINVAL, "unsupported source field for this propagator");
}

static int record_target(int kind, const rsg_state *s, int field, const float **F, const float **F2) {
    *F = nullptr;
    *F2 = nullptr;
    switch (field) {
    case RSG_PRESSURE:
        if (kind == RSG_KIND_E2D) {
            *F = s->Sxx;
            *F2 = s->Szz;
        } else {
            *F = s->P;
        }
        break;
    case RSG_VX: *F = s->Vx; break;
    case RSG_VY: *F = kind == RSG_KIND_A3D ? s->Vy : nullptr; break;
    case RSG_VZ: *F = s->Vz; break;
    default: break;
    }
    if (!*F) return fail(RSG_EINVAL, "recordData: field not available for this propagator");
    return RSG_OK;
}

// Tail of one modelling step: insertSource(it), then every receiver set
// gathered by one launch (sets beyond kMaxFusedRec get their own launches).
int step_io(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src, int nrec,
            const rsg_traces *recs, const int *it_dev, int it, hipStream_t st) {
    if (int rc = insert_source(kind, g, s, src, it_dev, it, st)) return rc;
    RecSets rs{};
    rs.dim = kind == RSG_KIND_A3D ? 3 : 2;
    rs.nx = g->nx;
    rs.ny = g->ny;
    rs.lpml = map_pad(g);
    rs.ncell = (size_t)g->nx * g->ny * g->nz;
    rs.nrec = 0;
    rs.first[0] = 0;
    for (int r = 0; r < nrec; r++) {
        if (recs[r].ntrace <= 0) continue;
        if (rs.nrec == kMaxFusedRec) {
            if (int rc = record_data(kind, g, s, &recs[r], it_dev, it, st)) return rc;
            continue;
        }
        if (!it_dev && (it < 0 || it >= recs[r].nt)) return fail(RSG_EINVAL, "recordData: it out of range");
        RecDesc &rd = rs.rec[rs.nrec];
        if (int rc = record_target(kind, s, recs[r].field, &rd.F, &rd.F2)) return rc;
        rd.factor = recs[r].factor;
        rd.ntrace = recs[r].ntrace;
        rd.nt = recs[r].nt;
        rd.map = recs[r].map;
        rd.shift = recs[r].shift;
        rd.data = recs[r].data;
        rs.first[rs.nrec + 1] = rs.first[rs.nrec] + recs[r].ntrace;
        rs.nrec++;
    }
    if (rs.nrec == 0) return RSG_OK;
    const int total = rs.first[rs.nrec], threads = 128;
    record_sets_kernel<<<(total + threads - 1) / threads, threads, 0, st>>>(rs, it_dev, it);
    RSG_LAUNCH_CHECK("record_sets_kernel");
    return RSG_OK;
}

int insert_source(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src,
                  const int *it_dev, int it, hipStream_t st, int nshot, size_t fstride) {
    if (!src || src->ntrace <= 0) return RSG_OK;
    if (!it_dev && (it < 0 || it >= src->nt)) return fail(RSG_EINVAL, "insertSource: it out of range");
    float *F, *F2;
    const float *Mod;
    float scale;
    if (int rc = source_target(kind, s, src->field, g->dt, &F, &F2, &Mod, &scale)) return rc;
    if (!F) return fail(RSG_EINVAL, "insertSource: target field not allocated");
    const int threads = 64, blocks = (src->ntrace + threads - 1) / threads;
    if (kind == RSG_KIND_A3D) {
        if (nshot != 1) return fail(RSG_EINVAL, "3D batches are not supported");
        inject3d_kernel<<<blocks, threads, 0, st>>>(F, Mod, scale, g->nx, g->ny, map_pad(g), src->ntrace,
                                                    src->map, src->shift, src->data, src->nt, it_dev, it,
                                                    (size_t)g->nx * g->ny * g->nz);
    } else {
        inject2d_kernel<<<dim3(blocks, nshot), threads, 0, st>>>(F, F2, Mod, scale, g->nx, map_pad(g),
                                                                 src->ntrace, src->map, src->shift, src->data,
                                                                 src->nt, it_dev, it, fstride,
                                                                 (size_t)g->nx * g->nz);
    }
    RSG_LAUNCH_CHECK("inject kernel");
    return RSG_OK;
}

int record_data(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *rec,
                const int *it_dev, int it, hipStream_t st, int nshot, size_t fstride) {
    if (!rec || rec->ntrace <= 0) return RSG_OK;
    if (!it_dev && (it < 0 || it >= rec->nt)) return fail(RSG_EINVAL, "recordData: it out of range");
    const float *F, *F2;
    if (int rc = record_target(kind, s, rec->field, &F, &F2)) return rc;
    const int threads = 64, blocks = (rec->ntrace + threads - 1) / threads;
    if (kind == RSG_KIND_A3D && nshot != 1) return fail(RSG_EINVAL, "3D batches are not supported");
    if (kind == RSG_KIND_A3D)
        record3d_kernel<<<blocks, threads, 0, st>>>(F, g->nx, g->ny, map_pad(g), rec->ntrace, rec->map,
                                                    rec->shift, rec->data, rec->nt, it_dev, it,
                                                    (size_t)g->nx * g->ny * g->nz);
    else
        record2d_kernel<<<dim3(blocks, nshot), threads, 0, st>>>(F, F2, rec->factor, g->nx, map_pad(g),
                                                                 rec->ntrace, rec->map, rec->shift, rec->data,
                                                                 rec->nt, it_dev, it, fstride,
                                                                 (size_t)g->nx * g->nz);
    RSG_LAUNCH_CHECK("record kernel");
    return RSG_OK;
}

int launch_add_int(int *p, int v, hipStream_t st) {
    add_int_kernel<<<1, 64, 0, st>>>(p, v);
    RSG_LAUNCH_CHECK("add_int_kernel");
    return RSG_OK;
}

int launch_set_int(int *p, int v, hipStream_t st) {
    set_int_kernel<<<1, 64, 0, st>>>(p, v);
    RSG_LAUNCH_CHECK("set_int_kernel");
    return RSG_OK;
}

}  // namespace rsg

extern "C" int rsg_insert_source(int kind, const rsg_grid *g, const rsg_state *s,
                                 const rsg_traces *src, int it, void *stream) {
    if (!g || !s) return rsg::fail(RSG_EINVAL, "null grid/state");
    return rsg::insert_source(kind, g, s, src, nullptr, it, (hipStream_t)stream);
}

extern "C" int rsg_record_data(int kind, const rsg_grid *g, const rsg_state *s,
                               const rsg_traces *rec, int it, void *stream) {
    if (!g || !s) return rsg::fail(RSG_EINVAL, "null grid/state");
    return rsg::record_data(kind, g, s, rec, nullptr, it, (hipStream_t)stream);
}
