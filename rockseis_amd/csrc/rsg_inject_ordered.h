// Trace-ordered injection of a 2D or 3D receiver set (FWI residuals, RTM
// data), shared by the acoustic FWI and RTM drivers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
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
    // map_dev is read on stream st (the caller may have written it there);
    // nxp x nzp is the padded grid, pad the map offset (lpml, or 0 on a
    // domain model).  A receiver whose 2x2 cells leave the grid is an error.
    template <class Mem>
    int build(Mem &db, const int *map_dev, int ntr, int ns, int nxp, int nzp, int pad, hipStream_t st) {
        std::vector<int> mp(2 * (size_t)ntr * ns), hc, hs, he, ho(ns + 1, 0);
        if (!mp.empty()) {
            RSG_HIP(hipMemcpyAsync(mp.data(), map_dev, sizeof(int) * mp.size(), hipMemcpyDeviceToHost, st));
            RSG_HIP(hipStreamSynchronize(st));
        }
        // a receiver on the last column of a domain model puts its x+1
        // corners on the next row's first cell, as the reference's flat
        // index does; only cells past the array are an error
        for (size_t q = 0; q < mp.size(); q += 2) {
            const int mx = mp[q], mz = mp[q + 1];
            if (mx < 0 || mz < 0) continue;
            if (((long)pad + mz + 1) * nxp + pad + mx + 1 >= (long)nxp * nzp)
                return fail(RSG_EINVAL, "2D injection: receiver map outside the padded grid");
        }
        const int lpml = pad;
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
        RSG_HIP(hipMemcpyAsync(d, all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice, st));
        // the host vector goes out of scope with this call
        RSG_HIP(hipStreamSynchronize(st));
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


// Receiver-side injection in trace order in 3D (insertSource of dataresP,
// fwi.cpp:2093 and 2132, and of dataP in RtmAcoustic3D; the trace loop of
// WavesAcoustic3D::insertSource, waves.cpp:1033-1057): neighbouring off-grid receivers put non-zero weights
// on shared corner cells, and per-trace atomic adds would leave those sums in
// scheduling order.  At set-up the host lists the touched cells of the padded
// grid, each with its (trace, corner) contributions in trace order (entry =
// trace*8 + z*4 + y*2 + x corner bit); one thread per cell adds them in that
// order.
struct OrderedInject3D {
    const size_t *cells = nullptr;
    const int *starts = nullptr, *ents = nullptr;
    int ncells = 0;
    template <class Mem>
    int build(Mem &dm, const int *map_dev, int ntr, int nx, int ny, int nz, int lpml, hipStream_t st) {
        if (ntr > (INT_MAX >> 3)) return fail(RSG_EINVAL, "3D injection: too many receiver traces");
        std::vector<int> mp(3 * (size_t)ntr);
        RSG_HIP(hipMemcpyAsync(mp.data(), map_dev, sizeof(int) * mp.size(), hipMemcpyDeviceToHost, st));
        RSG_HIP(hipStreamSynchronize(st));
        const size_t nxy = (size_t)nx * ny;
        std::vector<std::pair<size_t, int>> e;  // (cell, trace*8 + corner)
        e.reserve(8 * (size_t)ntr);
        for (int i = 0; i < ntr; i++) {
            const int mx = mp[3 * (size_t)i], my = mp[3 * (size_t)i + 1], mz = mp[3 * (size_t)i + 2];
            if (mx < 0 || my < 0 || mz < 0) continue;
            if ((size_t)(lpml + mz + 1) * nxy + (size_t)(lpml + my + 1) * nx + (lpml + mx + 1) >= nxy * nz)
                return fail(RSG_EINVAL, "3D injection: receiver map outside the padded grid");
            for (int i1 = 0; i1 < 2; i1++)
                for (int i2 = 0; i2 < 2; i2++)
                    for (int i3 = 0; i3 < 2; i3++)
                        e.emplace_back((size_t)(lpml + mz + i1) * nxy + (size_t)(lpml + my + i2) * nx + (lpml + mx + i3),
                                       i * 8 + i1 * 4 + i2 * 2 + i3);
        }
        std::stable_sort(e.begin(), e.end(), [](const std::pair<size_t, int> &x, const std::pair<size_t, int> &y) {
            return x.first < y.first;
        });
        std::vector<size_t> hc;
        std::vector<int> hs, he;
        for (size_t q = 0; q < e.size(); q++) {
            if (q == 0 || e[q].first != e[q - 1].first) {
                hc.push_back(e[q].first);
                hs.push_back((int)he.size());
            }
            he.push_back(e[q].second);
        }
        hs.push_back((int)he.size());
        ncells = (int)hc.size();
        if (ncells == 0) return RSG_OK;
        size_t *dc = nullptr;
        int *di = nullptr;
        if (int rc = dm.alloc(&dc, hc.size())) return rc;
        if (int rc = dm.alloc(&di, hs.size() + he.size())) return rc;
        RSG_HIP(hipMemcpyAsync(dc, hc.data(), sizeof(size_t) * hc.size(), hipMemcpyHostToDevice, st));
        RSG_HIP(hipMemcpyAsync(di, hs.data(), sizeof(int) * hs.size(), hipMemcpyHostToDevice, st));
        RSG_HIP(hipMemcpyAsync(di + hs.size(), he.data(), sizeof(int) * he.size(), hipMemcpyHostToDevice, st));
        // the host vectors go out of scope with this call
        RSG_HIP(hipStreamSynchronize(st));
        cells = dc;
        starts = di;
        ents = di + hs.size();
        return RSG_OK;
    }
};

// P += dt * L * res * xtri * ytri * ztri (L null: dt * res * ...), cell by
// cell in trace order.
__global__ void inject3d_ordered_kernel(float *P, const float *__restrict__ L, float dt,
                                        const size_t *__restrict__ cells, const int *__restrict__ starts,
                                        const int *__restrict__ ents, int ncells, const float *__restrict__ shift,
                                        const float *__restrict__ wav, int nt, int it) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ncells) return;
    const size_t p = cells[j];
    const float m = L ? dt * L[p] : dt;
    float a = P[p];
    for (int q = starts[j]; q < starts[j + 1]; q++) {
        const int e = ents[q], i = e >> 3, i1 = (e >> 2) & 1, i2 = (e >> 1) & 1, i3 = e & 1;
        const float sx = shift[3 * (size_t)i], sy = shift[3 * (size_t)i + 1], sz = shift[3 * (size_t)i + 2];
        const float xt = i3 ? sx : 1 - sx, yt = i2 ? sy : 1 - sy, zt = i1 ? sz : 1 - sz;
        a += m * wav[(size_t)i * nt + it] * xt * yt * zt;
    }
    P[p] = a;
}

}  // namespace
}  // namespace rsg
