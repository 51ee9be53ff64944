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
#include "rsg_srcrec.h"

namespace rsg {

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

// One modelling step; it_dev null means `it` is the absolute time index.
static int one_step(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src, int nrec,
                    const rsg_traces *recs, const int *it_dev, int it, hipStream_t st) {
    if (int rc = step_velocity(kind, g, s, st)) return rc;
    if (int rc = step_stress(kind, g, s, st)) return rc;
    return step_io(kind, g, s, src, nrec, recs, it_dev, it, st);
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
    if (graph_chunk <= 1 || nsteps < 2 * graph_chunk) {
        for (int it = it0; it < it0 + nsteps; it++)
            if (int rc = one_step(kind, g, s, src, nrec, recs, nullptr, it, st)) return rc;
        return RSG_OK;
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
            rc = one_step(kind, g, s, src, nrec, recs, it_dev, k, st);
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
            rc = one_step(kind, g, s, src, nrec, recs, nullptr, it, st);
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
