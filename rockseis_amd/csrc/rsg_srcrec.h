// Internal launch helpers shared by the step, source/receiver and driver code.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/rsgpu.h"

namespace rsg {
// nshot/fstride: a batch of shots laid out as in rsg_batch (traces of shot k
// follow those of shot k-1)
int insert_source(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src,
                  const int *it_dev, int it, hipStream_t st, int nshot = 1, size_t fstride = 0);
int record_data(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *rec,
                const int *it_dev, int it, hipStream_t st, int nshot = 1, size_t fstride = 0);
int step_io(int kind, const rsg_grid *g, const rsg_state *s, const rsg_traces *src, int nrec,
            const rsg_traces *recs, const int *it_dev, int it, hipStream_t st);
int launch_add_int(int *p, int v, hipStream_t st);
int launch_set_int(int *p, int v, hipStream_t st);
}  // namespace rsg
