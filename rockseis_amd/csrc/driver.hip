// Modelling time loop on the device (ModellingAcoustic2D/Elastic2D/
// Acoustic3D::run, modelling.cpp:242-286, 363-473, 545-643): per step the
// velocity half step, the stress half step, insertSource(it) and
// recordData(it) for every receiver set, all queued on one stream.
//
// Small grids are launch-bound (a few microseconds of HBM time per step), so
// the loop can be captured as a hipGraph of `graph_chunk` steps and replayed;
// inside the graph the time index comes from a device counter that the last
// node of each chunk advances.
#include <vector>

#include "rsg_common.h"
#include "rsg_fd2d.h"
#include "rsg_srcrec.h"

namespace rsg {

int stress_fused(int kind, const rsg_grid *g, const rsg_state *s, const FusedIO &io, int rz, hipStream_t st);
int tile_rows();

int step_velocity(int kind, const rsg_grid *g, const rsg_state *s, hipStream_t st) {
    switch (kind) {
    case RSG_KIND_A2D: return rsg_a2d_forward_velocity(g, s, st);
    case RSG_KIND_E2D: return rsg_e2d_forward_velocity(g, s, st);
    case RSG_KIND_A3D: return rsg_a3d_forward_velocity(g, s, st);
    default: return fail(RSG_EINVAL, "unknown propagator kind");
    }
}

int step_stress(int kind, const rsg_grid *g, const rsg_state *s, hipStream_t st) {
    switch (kind) {
    case RSG_KIND_A2D: return rsg_a2d_forward_stress(g, s, st);
    case RSG_KIND_E2D: return rsg_e2d_forward_stress(g, s, st);
    case RSG_KIND_A3D: return rsg_a3d_forward_stress(g, s, st);
    default: return fail(RSG_EINVAL, "unknown propagator kind");
    }
}

// What a 2D modelling run fuses into its stress launches (FusedIO): up to
// kFusedSrc PRESSURE source traces and, for elastic runs, up to kFusedRec VX /
// VZ receiver sets, with per-tile receiver tables for the run's tile height.
// The rest goes through the I/O kernels of step_io.
struct IoPlan {
    bool active = false;
    FusedIO io{};
    int rz = 2;
    bool src_fused = false;
    std::vector<rsg_traces> rest;
    std::vector<void *> dev;
    IoPlan() = default;
    IoPlan(const IoPlan &) = delete;
    IoPlan &operator=(const IoPlan &) = delete;
    ~IoPlan() {
        for (void *p : dev) (void)hipFree(p);
    }
};

static int make_io_plan(int kind, const rsg_grid *g, const rsg_traces *src, int nrec, const rsg_traces *recs,
                        hipStream_t st, IoPlan &p) {
    p.rest.assign(recs, recs + nrec);
    if ((kind != RSG_KIND_A2D && kind != RSG_KIND_E2D) || !tuning().fuse_io) return RSG_OK;
    FusedIO &io = p.io;
    p.rz = tile_rows();
    const int tz = BZ * p.rz, ntx = (g->nx + BX - 1) / BX, ntiles = ntx * ((g->nz + tz - 1) / tz);
    RSG_HIP(hipStreamSynchronize(st));  // maps may still be in flight on st
    if (src && src->ntrace > 0 && src->ntrace <= kFusedSrc && src->field == RSG_PRESSURE) {
        std::vector<int> m(2 * src->ntrace);
        std::vector<float> sh(2 * src->ntrace);
        RSG_HIP(hipMemcpy(m.data(), src->map, sizeof(int) * m.size(), hipMemcpyDeviceToHost));
        RSG_HIP(hipMemcpy(sh.data(), src->shift, sizeof(float) * sh.size(), hipMemcpyDeviceToHost));
        io.nsrc = src->ntrace;
        for (int i = 0; i < src->ntrace; i++) {
            const bool ok = m[2 * i] >= 0 && m[2 * i + 1] >= 0 && g->lpml + m[2 * i] + 1 < g->nx &&
                            g->lpml + m[2 * i + 1] + 1 < g->nz;
            io.sx[i] = ok ? g->lpml + m[2 * i] : -1;
            io.sz[i] = ok ? g->lpml + m[2 * i + 1] : -1;
            io.wx[i][0] = 1 - sh[2 * i];
            io.wx[i][1] = sh[2 * i];
            io.wz[i][0] = 1 - sh[2 * i + 1];
            io.wz[i][1] = sh[2 * i + 1];
        }
        io.wav = src->data;
        io.src_nt = src->nt;
        io.scale = kind == RSG_KIND_E2D ? -g->dt : g->dt;
        p.src_fused = true;
    }
    if (kind == RSG_KIND_E2D) {
        std::vector<rsg_traces> rest;
        for (int r = 0; r < nrec; r++) {
            const rsg_traces &rc = recs[r];
            if (io.nrec == kFusedRec || rc.ntrace <= 0 || rc.factor ||
                (rc.field != RSG_VX && rc.field != RSG_VZ)) {
                rest.push_back(rc);
                continue;
            }
            std::vector<int> m(2 * rc.ntrace);
            RSG_HIP(hipMemcpy(m.data(), rc.map, sizeof(int) * m.size(), hipMemcpyDeviceToHost));
            // counting sort of the valid receivers by the tile of their base cell
            std::vector<int> start(ntiles + 1, 0), list;
            bool inside = true;
            for (int i = 0; i < rc.ntrace; i++) {
                if (m[2 * i] < 0 || m[2 * i + 1] < 0) continue;
                const int px = g->lpml + m[2 * i], pz = g->lpml + m[2 * i + 1];
                if (px + 1 >= g->nx || pz + 1 >= g->nz) inside = false;
                else start[(pz / tz) * ntx + px / BX + 1]++;
            }
            if (!inside) {
                rest.push_back(rc);
                continue;
            }
            for (int t = 0; t < ntiles; t++) start[t + 1] += start[t];
            list.resize(start[ntiles]);
            std::vector<int> fill(start.begin(), start.end() - 1);
            for (int i = 0; i < rc.ntrace; i++) {
                if (m[2 * i] < 0 || m[2 * i + 1] < 0) continue;
                const int px = g->lpml + m[2 * i], pz = g->lpml + m[2 * i + 1];
                list[fill[(pz / tz) * ntx + px / BX]++] = i;
            }
            int *d = nullptr;
            if (hipMalloc((void **)&d, sizeof(int) * (start.size() + list.size() + 1)) != hipSuccess)
                return fail(RSG_ENOMEM, "rsg_modelling_run: receiver table allocation");
            p.dev.push_back(d);
            RSG_HIP(hipMemcpy(d, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice));
            if (!list.empty())
                RSG_HIP(hipMemcpy(d + start.size(), list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice));
            const int k = io.nrec++;
            io.rvz[k] = rc.field == RSG_VZ;
            io.rmap[k] = rc.map;
            io.rshift[k] = rc.shift;
            io.rdata[k] = rc.data;
            io.rnt[k] = rc.nt;
            io.rtile[k] = d;
            io.rlist[k] = d + start.size();
        }
        p.rest = rest;
    }
    p.active = io.nsrc > 0 || io.nrec > 0;
    return RSG_OK;
}

// One modelling step; it_dev null means `it` is the absolute time index.
static int one_step(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src, const IoPlan &p,
                    const int *it_dev, int it, hipStream_t st) {
    if (int rc = step_velocity(kind, g, s, st)) return rc;
    if (p.active) {
        FusedIO io = p.io;
        io.it_dev = it_dev;
        io.it = it;
        if (int rc = stress_fused(kind, g, s, io, p.rz, st)) return rc;
    } else if (int rc = step_stress(kind, g, s, st)) {
        return rc;
    }
    return step_io(kind, g, s, p.src_fused ? nullptr : src, (int)p.rest.size(), p.rest.data(), it_dev, it, st);
}

}  // namespace rsg

using namespace rsg;

extern "C" int rsg_modelling_run(int kind, const rsg_grid *g, const rsg_state *s,
                                 const rsg_traces *src, int nrec, const rsg_traces *recs, int it0,
                                 int nsteps, int graph_chunk, void *stream) {
    if (!g || !s || nsteps < 0 || it0 < 0 || nrec < 0 || (nrec > 0 && !recs))
        return fail(RSG_EINVAL, "rsg_modelling_run: bad arguments");
    if (src && it0 + nsteps > src->nt) return fail(RSG_EINVAL, "rsg_modelling_run: past wavelet end");
    for (int r = 0; r < nrec; r++)
        if (it0 + nsteps > recs[r].nt) return fail(RSG_EINVAL, "rsg_modelling_run: past record end");
    hipStream_t st = (hipStream_t)stream;
    // fusing pays for its set-up (map downloads, receiver tables) over a few
    // tens of steps; short calls keep the separate I/O kernels
    IoPlan plan;
    if (nsteps >= 32)
        if (int rc = make_io_plan(kind, g, src, nrec, recs, st, plan)) return rc;
    if (!plan.active) plan.rest.assign(recs, recs + nrec);
    if (graph_chunk <= 1 || nsteps < 2 * graph_chunk) {
        int rc = RSG_OK;
        for (int it = it0; it < it0 + nsteps && rc == RSG_OK; it++) rc = one_step(kind, g, s, src, plan, nullptr, it, st);
        // the receiver tables are freed with the plan: drain first
        if (plan.active) {
            const hipError_t e = hipStreamSynchronize(st);
            if (rc == RSG_OK && e != hipSuccess) rc = check_hip(e, "hipStreamSynchronize");
        }
        return rc;
    }
    // Graph path: capture graph_chunk steps once, replay nsteps/graph_chunk times.
    bool own = false;
    if (!st) {
        // A private stream does not order against the null stream: drain it.
        RSG_HIP(hipDeviceSynchronize());
        RSG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        own = true;
    }
    int rc = RSG_OK;
    int *it_dev = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const int nchunk = nsteps / graph_chunk;
    do {
        if (hipMalloc(&it_dev, sizeof(int)) != hipSuccess) {
            rc = fail(RSG_ENOMEM, "rsg_modelling_run: counter allocation");
            break;
        }
        if ((rc = launch_set_int(it_dev, it0, st))) break;
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = fail(RSG_EHIP, "hipStreamBeginCapture failed");
            break;
        }
        for (int k = 0; k < graph_chunk && rc == RSG_OK; k++)
            rc = one_step(kind, g, s, src, plan, it_dev, k, st);
        if (rc == RSG_OK) rc = launch_add_int(it_dev, graph_chunk, st);
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc) break;
        if (e != hipSuccess) {
            rc = check_hip(e, "hipStreamEndCapture");
            break;
        }
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            rc = check_hip(e, "hipGraphInstantiate");
            break;
        }
        for (int c = 0; c < nchunk && rc == RSG_OK; c++) {
            e = hipGraphLaunch(exec, st);
            if (e != hipSuccess) rc = check_hip(e, "hipGraphLaunch");
        }
        if (rc) break;
        for (int it = it0 + nchunk * graph_chunk; it < it0 + nsteps && rc == RSG_OK; it++)
            rc = one_step(kind, g, s, src, plan, nullptr, it, st);
        if (rc) break;
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = check_hip(e, "hipStreamSynchronize");
    } while (0);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (it_dev) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(it_dev);
    }
    if (own) (void)hipStreamDestroy(st);
    return rc;
}

extern "C" int rsg_repeat_step(int kind, int which, const rsg_grid *g, const rsg_state *s, int n,
                               void *stream) {
    hipStream_t st = (hipStream_t)stream;
    for (int i = 0; i < n; i++) {
        int rc;
        switch (which) {
        case 0: rc = step_velocity(kind, g, s, st); break;
        case 1: rc = step_stress(kind, g, s, st); break;
        case 2:
            rc = kind == RSG_KIND_E2D ? rsg_e2d_backward_velocity(g, s, st)
                                      : fail(RSG_EINVAL, "adjoint-state steps exist for elastic 2D only");
            break;
        case 3:
            rc = kind == RSG_KIND_E2D ? rsg_e2d_backward_stress(g, s, st)
                                      : fail(RSG_EINVAL, "adjoint-state steps exist for elastic 2D only");
            break;
        default: rc = fail(RSG_EINVAL, "rsg_repeat_step: which must be 0..3");
        }
        if (rc) return rc;
    }
    return RSG_OK;
}
