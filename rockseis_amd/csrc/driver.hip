// Modelling time loop on the device (ModellingAcoustic2D/Elastic2D/
// Acoustic3D::run, modelling.cpp:242-286, 363-473, 545-643): per step the
// velocity half step, the stress half step, insertSource(it) and
// recordData(it) for every receiver set, all queued on one stream.
//
// Small grids are launch-bound (a few microseconds of HBM time per step), so
// the loop can be captured as a hipGraph of `graph_chunk` steps and replayed;
// inside the graph the time index comes from a device counter that the last
// node of each chunk advances.
#include <algorithm>
#include <new>
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
    // the fused gathers index tiles from lpml-offset maps: domain models
    // (maps without the offset) take the separate kernels
    if ((kind != RSG_KIND_A2D && kind != RSG_KIND_E2D) || !tuning().fuse_io || g->domdec) return RSG_OK;
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

// Snapshot2D/3D::writeSnap (snap.cpp:638-669, 1138-1172) on the device:
// out = f0 (+ f1) (+ f2), added in that order as the reference adds
// data1 + data2 + data3, over the interior cells (i+pad, j+pad(3D), k+pad).
// With snapinc > 0 the kernel writes only at steps t % snapinc == 0, into
// slot t / snapinc of out (t = *it_dev + it0, or it0).
__global__ void snap_interior_kernel(const float *__restrict__ f0, const float *__restrict__ f1,
                                     const float *__restrict__ f2, int nx, int ny, int pad, int ypad, long nxi,
                                     long nyi, long ntot, float *__restrict__ out, const int *it_dev, int it0,
                                     int snapinc) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ntot) return;
    long slot = 0;
    if (snapinc > 0) {
        const int t = it_dev ? *it_dev + it0 : it0;
        if (t % snapinc) return;
        slot = t / snapinc;
    }
    const long i = idx % nxi, r = idx / nxi, j = r % nyi, k = r / nyi;
    const size_t p = ((size_t)(k + pad) * ny + (size_t)(j + ypad)) * nx + (size_t)(i + pad);
    float v = f0 ? f0[p] : 0.f;
    if (f1) v += f1[p];
    if (f2) v += f2[p];
    out[slot * ntot + idx] = v;
}

static int launch_snap(const rsg_snapset &sn, const rsg_grid *g, const int *it_dev, int it, hipStream_t st) {
    const int pad = g->domdec ? 0 : g->lpml;
    const int ypad = g->ny > 1 ? pad : 0;
    const long nxi = g->nx - 2 * pad, nyi = g->ny - 2 * ypad, nzi = g->nz - 2 * pad;
    const long ntot = nxi * nyi * nzi;
    if (ntot <= 0) return RSG_OK;
    snap_interior_kernel<<<(unsigned)((ntot + 255) / 256), 256, 0, st>>>(
        sn.f[0], sn.f[1], sn.f[2], g->nx, g->ny, pad, ypad, nxi, nyi, ntot, sn.out, it_dev, it, sn.snapinc);
    RSG_LAUNCH_CHECK("snap_interior_kernel");
    return RSG_OK;
}

// One modelling step; it_dev null means `it` is the absolute time index.
static int one_step(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src, const IoPlan &p,
                    int nsnap, const rsg_snapset *snaps, const int *it_dev, int it, hipStream_t st) {
    if (int rc = step_velocity(kind, g, s, st)) return rc;
    if (p.active) {
        FusedIO io = p.io;
        io.it_dev = it_dev;
        io.it = it;
        if (int rc = stress_fused(kind, g, s, io, p.rz, st)) return rc;
    } else if (int rc = step_stress(kind, g, s, st)) {
        return rc;
    }
    if (int rc = step_io(kind, g, s, p.src_fused ? nullptr : src, (int)p.rest.size(), p.rest.data(), it_dev, it, st))
        return rc;
    for (int k = 0; k < nsnap; k++)
        if (int rc = launch_snap(snaps[k], g, it_dev, it, st)) return rc;
    return RSG_OK;
}

static int modelling_run(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src, int nrec,
                         const rsg_traces *recs, int nsnap, const rsg_snapset *snaps, int it0, int nsteps,
                         int graph_chunk, hipStream_t st);

}  // namespace rsg

using namespace rsg;

extern "C" int rsg_snap_interior(const float *f0, const float *f1, const float *f2, const rsg_grid *g, int pad,
                                 float *out, void *stream) {
    if (!g || !out || g->nx < 1 || g->ny < 1 || g->nz < 1 || pad < 0)
        return fail(RSG_EINVAL, "rsg_snap_interior: bad arguments");
    if (g->nx <= 2 * pad || g->nz <= 2 * pad || (g->ny > 1 && g->ny <= 2 * pad))
        return fail(RSG_EINVAL, "rsg_snap_interior: grid smaller than its padding");
    rsg_grid gg = *g;
    gg.lpml = pad;
    gg.domdec = 0;
    rsg_snapset sn{{f0, f1, f2}, out, 0};
    return launch_snap(sn, &gg, nullptr, 0, (hipStream_t)stream);
}

extern "C" int rsg_modelling_run(int kind, const rsg_grid *g, const rsg_state *s,
                                 const rsg_traces *src, int nrec, const rsg_traces *recs, int it0,
                                 int nsteps, int graph_chunk, void *stream) {
    return modelling_run(kind, g, s, src, nrec, recs, 0, nullptr, it0, nsteps, graph_chunk, (hipStream_t)stream);
}

extern "C" int rsg_modelling_run_snap(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src,
                                      int nrec, const rsg_traces *recs, int nsnap, const rsg_snapset *snaps,
                                      int it0, int nsteps, int graph_chunk, void *stream) {
    if (nsnap < 0 || (nsnap > 0 && !snaps)) return fail(RSG_EINVAL, "rsg_modelling_run_snap: bad snapshot sets");
    for (int k = 0; k < nsnap; k++)
        if (!snaps[k].out || snaps[k].snapinc < 1 || (!snaps[k].f[0] && !snaps[k].f[1] && !snaps[k].f[2]))
            return fail(RSG_EINVAL, "rsg_modelling_run_snap: a snapshot set needs a field, out and snapinc >= 1");
    return modelling_run(kind, g, s, src, nrec, recs, nsnap, snaps, it0, nsteps, graph_chunk, (hipStream_t)stream);
}

namespace rsg {

static int modelling_run(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src, int nrec,
                         const rsg_traces *recs, int nsnap, const rsg_snapset *snaps, int it0, int nsteps,
                         int graph_chunk, hipStream_t st) {
    if (!g || !s || nsteps < 0 || it0 < 0 || nrec < 0 || (nrec > 0 && !recs))
        return fail(RSG_EINVAL, "rsg_modelling_run: bad arguments");
    if (src && it0 + nsteps > src->nt) return fail(RSG_EINVAL, "rsg_modelling_run: past wavelet end");
    for (int r = 0; r < nrec; r++)
        if (it0 + nsteps > recs[r].nt) return fail(RSG_EINVAL, "rsg_modelling_run: past record end");
    // fusing pays for its set-up (map downloads, receiver tables) over a few
    // tens of steps; short calls keep the separate I/O kernels unless the
    // caller asked for a captured graph
    IoPlan plan;
    if (nsteps >= 32 || graph_chunk > 1)
        if (int rc = make_io_plan(kind, g, src, nrec, recs, st, plan)) return rc;
    if (!plan.active) plan.rest.assign(recs, recs + nrec);
    if (graph_chunk <= 1 || nsteps < 2 * graph_chunk) {
        int rc = RSG_OK;
        for (int it = it0; it < it0 + nsteps && rc == RSG_OK; it++) rc = one_step(kind, g, s, src, plan, nsnap, snaps, nullptr, it, st);
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
            rc = one_step(kind, g, s, src, plan, nsnap, snaps, it_dev, k, st);
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
            rc = one_step(kind, g, s, src, plan, nsnap, snaps, nullptr, it, st);
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

}  // namespace rsg

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

// ---------------------------------------------------------------------------
// Modelling plan: the time loop of one modelling run captured once as
// hipGraphs of 1, 2, 4, ..., 2^(nlev-1) steps and replayed for any step
// range.  A range of n steps replays the largest graph while it fits, then
// the binary digits of the rest, so no step of a run is launched eagerly and
// no capture happens after set-up (a caller timing short ranges times the
// replay, not the capture).  The time index comes from a device counter that
// run() sets before the first replay and that every graph's last node
// advances by its own step count.
struct rsg_modelling_plan {
    int kind = 0;
    rsg_grid g{};
    rsg_state s{};
    rsg_traces src{};
    bool has_src = false;
    std::vector<rsg_traces> recs;
    std::vector<rsg_snapset> snaps;
    int nt_max = 0;  // steps it0 + nsteps may not pass (wavelet and record lengths)
    rsg::IoPlan io;
    int *it_dev = nullptr;
    std::vector<hipGraphExec_t> exec;  // exec[k] runs 2^k steps
    ~rsg_modelling_plan() {
        // replays may still run on the caller's stream: drain before the
        // graphs, the counter and the receiver tables (io) are released
        (void)hipDeviceSynchronize();
        for (hipGraphExec_t e : exec) (void)hipGraphExecDestroy(e);
        if (it_dev) (void)hipFree(it_dev);
    }
};

extern "C" int rsg_modelling_plan_create(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src,
                                         int nrec, const rsg_traces *recs, int nsnap, const rsg_snapset *snaps,
                                         int max_chunk, void *stream, rsg_modelling_plan **out) {
    if (!out) return fail(RSG_EINVAL, "rsg_modelling_plan_create: out is NULL");
    *out = nullptr;
    if (!g || !s || nrec < 0 || (nrec > 0 && !recs) || nsnap < 0 || (nsnap > 0 && !snaps) || max_chunk < 1 ||
        max_chunk > (1 << 12))
        return fail(RSG_EINVAL, "rsg_modelling_plan_create: bad arguments (max_chunk 1..4096)");
    if (kind != RSG_KIND_A2D && kind != RSG_KIND_E2D && kind != RSG_KIND_A3D)
        return fail(RSG_EINVAL, "rsg_modelling_plan_create: unknown propagator kind");
    for (int k = 0; k < nsnap; k++)
        if (!snaps[k].out || snaps[k].snapinc < 1 || (!snaps[k].f[0] && !snaps[k].f[1] && !snaps[k].f[2]))
            return fail(RSG_EINVAL, "rsg_modelling_plan_create: a snapshot set needs a field, out and snapinc >= 1");
    rsg_modelling_plan *p = new (std::nothrow) rsg_modelling_plan;
    if (!p) return fail(RSG_ENOMEM, "rsg_modelling_plan_create: host allocation");
    p->kind = kind;
    p->g = *g;
    p->s = *s;
    p->has_src = src != nullptr;
    if (src) p->src = *src;
    p->recs.assign(recs, recs + nrec);
    p->snaps.assign(snaps, snaps + nsnap);
    p->nt_max = src ? src->nt : 0x7fffffff;
    for (int r = 0; r < nrec; r++) p->nt_max = std::min(p->nt_max, recs[r].nt);
    hipStream_t cst = (hipStream_t)stream;
    hipStream_t cap = nullptr;
    int rc = RSG_OK;
    do {
        if ((rc = make_io_plan(kind, &p->g, p->has_src ? &p->src : nullptr, nrec, p->recs.data(), cst, p->io)))
            break;
        if (!p->io.active) p->io.rest = p->recs;
        if (hipMalloc((void **)&p->it_dev, sizeof(int)) != hipSuccess) {
            rc = fail(RSG_ENOMEM, "rsg_modelling_plan_create: counter allocation");
            break;
        }
        // captured work does not run: a private stream serves every capture
        if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) {
            rc = fail(RSG_EHIP, "rsg_modelling_plan_create: capture stream");
            break;
        }
        for (int n = 1; n <= max_chunk && rc == RSG_OK; n *= 2) {
            hipGraph_t graph = nullptr;
            if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                rc = fail(RSG_EHIP, "hipStreamBeginCapture failed");
                break;
            }
            for (int k = 0; k < n && rc == RSG_OK; k++)
                rc = one_step(kind, &p->g, &p->s, p->has_src ? &p->src : nullptr, p->io, nsnap, p->snaps.data(),
                              p->it_dev, k, cap);
            if (rc == RSG_OK) rc = launch_add_int(p->it_dev, n, cap);
            const hipError_t e = hipStreamEndCapture(cap, &graph);
            if (rc == RSG_OK && e != hipSuccess) rc = check_hip(e, "hipStreamEndCapture");
            if (rc == RSG_OK) {
                hipGraphExec_t ex = nullptr;
                const hipError_t ei = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
                if (ei != hipSuccess) {
                    rc = check_hip(ei, "hipGraphInstantiate");
                } else {
                    p->exec.push_back(ex);
                    // upload now: a graph's first launch otherwise pays for
                    // it, inside the caller's first run of that size
                    const hipError_t eu = hipGraphUpload(ex, cap);
                    if (eu != hipSuccess) rc = check_hip(eu, "hipGraphUpload");
                }
            }
            if (graph) (void)hipGraphDestroy(graph);
        }
    } while (0);
    if (cap) {
        const hipError_t e = hipStreamSynchronize(cap);  // uploads done before any replay
        if (rc == RSG_OK && e != hipSuccess) rc = check_hip(e, "hipStreamSynchronize");
        (void)hipStreamDestroy(cap);
    }
    if (rc) {
        delete p;
        return rc;
    }
    *out = p;
    return RSG_OK;
}

extern "C" int rsg_modelling_plan_run(rsg_modelling_plan *p, int it0, int nsteps, void *stream) {
    if (!p || it0 < 0 || nsteps < 0) return fail(RSG_EINVAL, "rsg_modelling_plan_run: bad arguments");
    if ((long)it0 + nsteps > p->nt_max) return fail(RSG_EINVAL, "rsg_modelling_plan_run: past wavelet or record end");
    if (nsteps == 0) return RSG_OK;
    hipStream_t st = (hipStream_t)stream;
    if (int rc = launch_set_int(p->it_dev, it0, st)) return rc;
    for (int k = (int)p->exec.size() - 1; k >= 0; k--) {
        const int n = 1 << k;
        for (; nsteps >= n; nsteps -= n) {
            const hipError_t e = hipGraphLaunch(p->exec[k], st);
            if (e != hipSuccess) return check_hip(e, "hipGraphLaunch");
        }
    }
    return RSG_OK;
}

extern "C" int rsg_modelling_plan_destroy(rsg_modelling_plan *p) {
    delete p;
    return RSG_OK;
}
