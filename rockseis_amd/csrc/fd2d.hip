// Uniform-batch launchers of the 2D staggered-grid kernels (one grid, one
// or many shots; blockIdx.y = shot) and their C-ABI entry points.  The
// kernels themselves are in rsg_fd2d.h.
#include "rsg_fd2d.h"

using namespace rsg;

#define GO(KERNEL, O, RZ)                                                                         \
    do {                                                                                          \
        K2D a = make_k2d(g, s, BZ * RZ);                                                          \
        a.sf = (long)b->field_stride;                                                             \
        a.sxs = (long)b->xstrip_stride;                                                           \
        a.szs = (long)b->zstrip_stride;                                                           \
        KERNEL<<<dim3(nblocks(a), b->nshot), dim3(BX, BZ), 0, (hipStream_t)stream>>>(a);          \
    } while (0)

static const rsg_batch kOneShot = {1, 0, 0, 0};

static int check_batch(const rsg_batch *b) {
    if (!b || b->nshot < 1 || b->nshot > 65535) return fail(RSG_EINVAL, "batch: nshot must be 1..65535");
    return RSG_OK;
}

extern "C" int rsg_a2d_forward_velocity_batch(const rsg_grid *g, const rsg_state *s, const rsg_batch *b,
                                              void *stream) {
    if (int rc = check_grid2d(g, s)) return rc;
    if (int rc = check_batch(b)) return rc;
    if (!s->P || !s->Vx || !s->Vz || !s->Rx || !s->Rz) return fail(RSG_EINVAL, "a2d: null array");
#define L_(O, RZ) GO((a2d_velocity_kernel<O, RZ>), O, RZ)
    RSG_DISPATCH(L_);
#undef L_
    RSG_LAUNCH_CHECK("a2d_velocity_kernel");
    return RSG_OK;
}

extern "C" int rsg_a2d_forward_stress_batch(const rsg_grid *g, const rsg_state *s, const rsg_batch *b,
                                            void *stream) {
    if (int rc = check_grid2d(g, s)) return rc;
    if (int rc = check_batch(b)) return rc;
    if (!s->P || !s->Vx || !s->Vz || !s->L) return fail(RSG_EINVAL, "a2d: null array");
#define L_(O, RZ) GO((a2d_stress_kernel<O, RZ>), O, RZ)
    RSG_DISPATCH(L_);
#undef L_
    RSG_LAUNCH_CHECK("a2d_stress_kernel");
    return RSG_OK;
}

namespace rsg {

// One forward acoustic step of a batch (velocity, then stress) out of place:
// the previous state is read from `in` (a checkpoint buffer, left as it is)
// and the new one written to `out`, bit for bit what the in-place pair of
// launches gives on a copy of `in`.
int a2d_forward_step_from(const rsg_grid *g, const rsg_state *in, const rsg_state *out, const rsg_batch *b,
                          hipStream_t stream) {
    if (int rc = check_grid2d(g, out)) return rc;
    if (int rc = check_batch(b)) return rc;
    if (!in || !in->P || !in->Vx || !in->Vz || !out->P || !out->Vx || !out->Vz || !out->Rx || !out->Rz || !out->L)
        return fail(RSG_EINVAL, "a2d: null array");
    for (int i = 0; i < RSG_A2D_NSTRIP; i++)
        if (!in->psi[i] || !out->psi[i]) return fail(RSG_EINVAL, "a2d: null PML strip");
    const rsg_state *s = out;
#define GOI(KERNEL, O, RZ)                                                                        \
    do {                                                                                          \
        K2D a = make_k2d(g, s, BZ * RZ);                                                          \
        a.sf = (long)b->field_stride;                                                             \
        a.sxs = (long)b->xstrip_stride;                                                           \
        a.szs = (long)b->zstrip_stride;                                                           \
        a.iP = in->P;                                                                             \
        a.iVx = in->Vx;                                                                           \
        a.iVz = in->Vz;                                                                           \
        for (int i = 0; i < RSG_A2D_NSTRIP; i++) a.ipsi[i] = in->psi[i];                          \
        KERNEL<<<dim3(nblocks(a), b->nshot), dim3(BX, BZ), 0, stream>>>(a);                       \
    } while (0)
#define L_(O, RZ) GOI((a2d_velocity_kernel<O, RZ>), O, RZ)
    RSG_DISPATCH(L_);
#undef L_
    RSG_LAUNCH_CHECK("a2d_velocity_kernel (out of place)");
#define L_(O, RZ) GOI((a2d_stress_kernel<O, RZ>), O, RZ)
    RSG_DISPATCH(L_);
#undef L_
#undef GOI
    RSG_LAUNCH_CHECK("a2d_stress_kernel (out of place)");
    return RSG_OK;
}

}  // namespace rsg

static int e2d_check(const rsg_grid *g, const rsg_state *s) {
    if (int rc = check_grid2d(g, s)) return rc;
    if (!s->Vx || !s->Vz || !s->Sxx || !s->Szz || !s->Sxz || !s->L || !s->L2M || !s->M || !s->Rx ||
        !s->Rz)
        return fail(RSG_EINVAL, "e2d: null array");
    return RSG_OK;
}

extern "C" int rsg_e2d_velocity_batch(const rsg_grid *g, const rsg_state *s, const rsg_batch *b,
                                      int adjoint, void *stream) {
    if (int rc = e2d_check(g, s)) return rc;
    if (int rc = check_batch(b)) return rc;
    if (adjoint) {
#define L_(O, RZ) GO((e2d_velocity_kernel<O, RZ, true>), O, RZ)
        RSG_DISPATCH(L_);
#undef L_
    } else {
#define L_(O, RZ) GO((e2d_velocity_kernel<O, RZ, false>), O, RZ)
        RSG_DISPATCH(L_);
#undef L_
    }
    RSG_LAUNCH_CHECK("e2d_velocity_kernel");
    return RSG_OK;
}

extern "C" int rsg_e2d_stress_batch(const rsg_grid *g, const rsg_state *s, const rsg_batch *b,
                                    int adjoint, void *stream) {
    if (int rc = e2d_check(g, s)) return rc;
    if (int rc = check_batch(b)) return rc;
    if (adjoint) {
#define L_(O, RZ) GO((e2d_stress_kernel<O, RZ, true>), O, RZ)
        RSG_DISPATCH(L_);
#undef L_
    } else {
#define L_(O, RZ) GO((e2d_stress_kernel<O, RZ, false>), O, RZ)
        RSG_DISPATCH(L_);
#undef L_
    }
    RSG_LAUNCH_CHECK("e2d_stress_kernel");
    return RSG_OK;
}

extern "C" int rsg_a2d_forward_velocity(const rsg_grid *g, const rsg_state *s, void *stream) {
    return rsg_a2d_forward_velocity_batch(g, s, &kOneShot, stream);
}
extern "C" int rsg_a2d_forward_stress(const rsg_grid *g, const rsg_state *s, void *stream) {
    return rsg_a2d_forward_stress_batch(g, s, &kOneShot, stream);
}
extern "C" int rsg_e2d_forward_velocity(const rsg_grid *g, const rsg_state *s, void *stream) {
    return rsg_e2d_velocity_batch(g, s, &kOneShot, 0, stream);
}
extern "C" int rsg_e2d_forward_stress(const rsg_grid *g, const rsg_state *s, void *stream) {
    return rsg_e2d_stress_batch(g, s, &kOneShot, 0, stream);
}
extern "C" int rsg_e2d_backward_velocity(const rsg_grid *g, const rsg_state *s, void *stream) {
    return rsg_e2d_velocity_batch(g, s, &kOneShot, 1, stream);
}
extern "C" int rsg_e2d_backward_stress(const rsg_grid *g, const rsg_state *s, void *stream) {
    return rsg_e2d_stress_batch(g, s, &kOneShot, 1, stream);
}

namespace rsg {

// Stress half step of a single-shot modelling step with the step's source
// injection and velocity gathers fused in (io), at `rz` rows per thread (the
// tile height the receiver tables were built for).
int stress_fused(int kind, const rsg_grid *g, const rsg_state *s, const FusedIO &io, int rz, hipStream_t st) {
    if (int rc = check_grid2d(g, s)) return rc;
    if (kind != RSG_KIND_A2D && kind != RSG_KIND_E2D) return fail(RSG_EINVAL, "fused stress: 2D kinds only");
    if (kind == RSG_KIND_A2D && io.nrec) return fail(RSG_EINVAL, "fused stress: gathers are elastic only");
#define FG(KERNEL, RZ)                                                                                 \
    do {                                                                                               \
        K2D a = make_k2d(g, s, BZ * RZ);                                                               \
        a.io = io;                                                                                     \
        KERNEL<<<dim3(nblocks(a), 1), dim3(BX, BZ), 0, st>>>(a);                                       \
    } while (0)
#define A_(O, RZ) FG((a2d_stress_kernel<O, RZ>), RZ)
#define E_(O, RZ) FG((e2d_stress_kernel<O, RZ, false>), RZ)
#define BY_RZ(L)                    \
    switch (rz) {                   \
    case 1: RSG_ORDER_CASES(L, 1) break; \
    case 2: RSG_ORDER_CASES(L, 2) break; \
    case 4: RSG_ORDER_CASES(L, 4) break; \
    default: RSG_ORDER_CASES(L, 8) break; \
    }
    if (kind == RSG_KIND_A2D) {
        BY_RZ(A_)
    } else {
        BY_RZ(E_)
    }
#undef BY_RZ
#undef E_
#undef A_
#undef FG
    RSG_LAUNCH_CHECK("fused stress kernel");
    return RSG_OK;
}

int tile_rows() { return rz2d(); }

}  // namespace rsg
