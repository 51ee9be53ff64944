// Ragged batches of 2D shots: many shots, each on a grid of its own size (an
// aperture window of a larger model), advanced by one launch per half step.
// A plan holds the per-shot kernel arguments and the tile prefix sums on the
// device; the kernels (rsg_fd2d.h) find their shot by binary search over
// the prefix.  Results per shot equal the single-shot entry points.
#include <vector>

#include "rsg_fd2d.h"
#include "rsg_ragged.h"

namespace rsg {

int ragged2d_create(int kind, int nshot, const rsg_grid *grids, const rsg_state *states, Ragged2D *out) {
    if (!grids || !states || !out) return fail(RSG_EINVAL, "ragged2d: null argument");
    if (kind != RSG_KIND_A2D && kind != RSG_KIND_E2D) return fail(RSG_EINVAL, "ragged2d: kind must be A2D or E2D");
    if (nshot < 1) return fail(RSG_EINVAL, "ragged2d: nshot must be >= 1");
    std::vector<K2D> desc(nshot);
    std::vector<int> tile0(4 * (size_t)(nshot + 1), 0);
    for (int k = 0; k < nshot; k++) {
        const rsg_grid *g = grids + k;
        const rsg_state *s = states + k;
        if (int rc = check_grid2d(g, s)) return rc;
        if (g->order != grids[0].order) return fail(RSG_EINVAL, "ragged2d: every shot must use one stencil order");
        if (kind == RSG_KIND_A2D && (!s->P || !s->Vx || !s->Vz || !s->Rx || !s->Rz || !s->L))
            return fail(RSG_EINVAL, "ragged2d: a2d shot with a null array");
        if (kind == RSG_KIND_E2D && (!s->Vx || !s->Vz || !s->Sxx || !s->Szz || !s->Sxz || !s->L || !s->L2M ||
                                     !s->M || !s->Rx || !s->Rz))
            return fail(RSG_EINVAL, "ragged2d: e2d shot with a null array");
        desc[k] = make_k2d(g, s, BZ * 2);
        for (int r = 0; r < 4; r++) {
            const int tz = BZ * (1 << r);
            const long nt = (long)desc[k].ntx * ((g->nz + tz - 1) / tz);
            const long next = tile0[r * (nshot + 1) + k] + nt;
            if (next > (1L << 30)) return fail(RSG_EINVAL, "ragged2d: too many tiles");
            tile0[r * (nshot + 1) + k + 1] = (int)next;
        }
    }
    Ragged2D p;
    p.kind = kind;
    p.nshot = nshot;
    p.order = grids[0].order;
    for (int r = 0; r < 4; r++) p.total[r] = tile0[r * (nshot + 1) + nshot];
    if (hipMalloc((void **)&p.d_desc, sizeof(K2D) * nshot) != hipSuccess)
        return fail(RSG_ENOMEM, "ragged2d: device allocation failed");
    if (hipMalloc((void **)&p.d_tile0, sizeof(int) * tile0.size()) != hipSuccess) {
        (void)hipFree(p.d_desc);
        return fail(RSG_ENOMEM, "ragged2d: device allocation failed");
    }
    hipError_t e = hipMemcpy(p.d_desc, desc.data(), sizeof(K2D) * nshot, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p.d_tile0, tile0.data(), sizeof(int) * tile0.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p.d_desc);
        (void)hipFree(p.d_tile0);
        return check_hip(e, "ragged2d: upload");
    }
    *out = p;
    return RSG_OK;
}

void ragged2d_destroy(Ragged2D *p) {
    if (!p) return;
    if (p->d_desc) (void)hipFree(p->d_desc);
    if (p->d_tile0) (void)hipFree(p->d_tile0);
    p->d_desc = nullptr;
    p->d_tile0 = nullptr;
}

int ragged2d_step(const Ragged2D *p, int which, hipStream_t stream) {
    if (!p || !p->d_desc) return fail(RSG_EINVAL, "ragged2d: no plan");
    if (which < 0 || which > 3 || (p->kind == RSG_KIND_A2D && which > 1))
        return fail(RSG_EINVAL, "ragged2d: which must be 0..1 (A2D) or 0..3 (E2D)");
    const int rz = rz2d_ragged();
    const int r = rz == 1 ? 0 : (rz == 2 ? 1 : (rz == 4 ? 2 : 3));
    const int *t0 = p->d_tile0 + r * (p->nshot + 1);
    const unsigned nb = 8u * (unsigned)((p->total[r] + 7) / 8);
    const K2D *d = p->d_desc;
    const int ns = p->nshot;
    const int order = p->order;
    const rsg_grid gg{1, 1, 1, 1, order, 0, 1.f, 1.f, 1.f, 1.f};
    const rsg_grid *g = &gg;  // RSG_ORDER_CASES switches on g->order
#define RG(KERNEL) KERNEL<<<dim3(nb), dim3(BX, BZ), 0, stream>>>(d, t0, ns)
    if (p->kind == RSG_KIND_A2D) {
        if (which == 0) {
#define L_(O, RZ) RG((a2d_velocity_ragged_kernel<O, RZ>))
            RSG_DISPATCH_RZ(L_, rz);
#undef L_
        } else {
#define L_(O, RZ) RG((a2d_stress_ragged_kernel<O, RZ>))
            RSG_DISPATCH_RZ(L_, rz);
#undef L_
        }
    } else if (which == 0) {
#define L_(O, RZ) RG((e2d_velocity_ragged_kernel<O, RZ, false>))
        RSG_DISPATCH_RZ(L_, rz);
#undef L_
    } else if (which == 1) {
#define L_(O, RZ) RG((e2d_stress_ragged_kernel<O, RZ, false>))
        RSG_DISPATCH_RZ(L_, rz);
#undef L_
    } else if (which == 2) {
#define L_(O, RZ) RG((e2d_velocity_ragged_kernel<O, RZ, true>))
        RSG_DISPATCH_RZ(L_, rz);
#undef L_
    } else {
#define L_(O, RZ) RG((e2d_stress_ragged_kernel<O, RZ, true>))
        RSG_DISPATCH_RZ(L_, rz);
#undef L_
    }
#undef RG
    RSG_LAUNCH_CHECK("ragged 2D step kernel");
    return RSG_OK;
}

}  // namespace rsg

struct rsg_ragged {
    rsg::Ragged2D p;
};

extern "C" int rsg_ragged2d_create(int kind, int nshot, const rsg_grid *grids, const rsg_state *states,
                                   rsg_ragged **plan) {
    if (!plan) return rsg::fail(RSG_EINVAL, "rsg_ragged2d_create: null plan");
    *plan = nullptr;
    rsg::Ragged2D p;
    if (int rc = rsg::ragged2d_create(kind, nshot, grids, states, &p)) return rc;
    *plan = new rsg_ragged{p};
    return RSG_OK;
}

extern "C" int rsg_ragged2d_step(const rsg_ragged *plan, int which, void *stream) {
    if (!plan) return rsg::fail(RSG_EINVAL, "rsg_ragged2d_step: null plan");
    return rsg::ragged2d_step(&plan->p, which, (hipStream_t)stream);
}

extern "C" int rsg_ragged2d_destroy(rsg_ragged *plan) {
    if (!plan) return RSG_OK;
    rsg::ragged2d_destroy(&plan->p);
    delete plan;
    return RSG_OK;
}
