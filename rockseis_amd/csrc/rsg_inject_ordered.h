// Trace-ordered injection of a 2D receiver set (FWI residuals, RTM data),
// shared by the acoustic 2D FWI and RTM drivers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>
#include <vector>

#include "rsg_common.h"

namespace rsg {
namespace {

// Receiver-side injection in trace order (insertSource of dataresP,
// fwi.cpp:1284-1295, and of dataP in RtmAcoustic2D, rtm.cpp:488, 513):
// off-grid receivers of one line put non-zero weights on shared corner cells,
// and per-trace atomic adds would leave those sums in scheduling order.
// At set-up the host lists, per shot, the touched cells of the padded grid,
// each with its (trace, corner) contributions in trace order; one thread per
// cell adds them in that order, as the reference does.
struct OrderedInject2D {
    const int *cells = nullptr, *starts = nullptr, *ents = nullptr, *coff = nullptr;
    int maxcells = 0;
    template <class Mem>
    int build(Mem &db, const int *map_dev, int ntr, int ns, int nxp, int lpml) {
        std::vector<int> mp(2 * (size_t)ntr * ns), hc, hs, he, ho(ns + 1, 0);
        if (!mp.empty()) RSG_HIP(hipMemcpy(mp.data(), map_dev, sizeof(int) * mp.size(), hipMemcpyDeviceToHost));
        for (int k = 0; k < ns; k++) {
            std::vector<std::pair<int, int>> e;  // (cell, trace*4 + corner)
            e.reserve(4 * (size_t)ntr);
            for (int i = 0; i < ntr; i++) {
                const int mx = mp[2 * ((size_t)k * ntr + i)], mz = mp[2 * ((size_t)k * ntr + i) + 1];
                if (mx < 0 || mz < 0) continue;
                for (int i1 = 0; i1 < 2; i1++)
                    for (int i2 = 0; i2 < 2; i2++)
                        e.emplace_back((lpml + mz + i1) * nxp + (lpml + mx + i2), i * 4 + i1 * 2 + i2);
            }
            std::stable_sort(e.begin(), e.end(), [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
                return x.first < y.first;
            });
            for (size_t q = 0; q < e.size(); q++) {
                if (q == 0 || e[q].first != e[q - 1].first) {
                    hc.push_back(e[q].first);
                    hs.push_back((int)he.size());
                }
                he.push_back(e[q].second);
            }
            ho[k + 1] = (int)hc.size();
            maxcells = std::max(maxcells, ho[k + 1] - ho[k]);
        }
        hs.push_back((int)he.size());
        // one device block: cells, starts, entries, per-shot cell offsets
        std::vector<int> all(hc);
        all.insert(all.end(), hs.begin(), hs.end());
        all.insert(all.end(), he.begin(), he.end());
        all.insert(all.end(), ho.begin(), ho.end());
        int *d = nullptr;
        if (int rc = db.alloc(&d, all.size())) return rc;
        RSG_HIP(hipMemcpy(d, all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice));
        cells = d;
        starts = cells + hc.size();
        ents = starts + hs.size();
        coff = ents + he.size();
        return RSG_OK;
    }
};

// P[k] += dt * L * res * xtri * ytri, cell by cell in trace order (blockIdx.y = shot;
// shot k's field at k*fstride, its traces after shot k-1's).
__global__ void inject2d_ordered_kernel(float *F, const float *Mod, float scale, const int *__restrict__ cells,
                                        const int *__restrict__ starts, const int *__restrict__ ents,
                                        const int *__restrict__ coff, const float *__restrict__ shift,
                                        const float *__restrict__ wav, int ntr, int nt, int it, size_t fstride) {
    const int k = blockIdx.y, j = coff[k] + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= coff[k + 1]) return;
    const size_t p = (size_t)cells[j];
    const float m = Mod ? scale * Mod[p] : scale;
    float *f = F + k * fstride + p;
    float a = *f;
    for (int q = starts[j]; q < starts[j + 1]; q++) {
        const int e = ents[q], i = e >> 2, i1 = (e >> 1) & 1, i2 = e & 1;
        const size_t tr = (size_t)k * ntr + i;
        const float sx = shift[2 * tr], sz = shift[2 * tr + 1];
        const float xt = i2 ? sx : 1 - sx, yt = i1 ? sz : 1 - sz;
        a += m * wav[tr * nt + it] * xt * yt;
    }
    *f = a;
}

}  // namespace
}  // namespace rsg
