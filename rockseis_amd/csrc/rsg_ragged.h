// Internal interface of the ragged 2D batches (fd2d_ragged.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/rsgpu.h"

namespace rsg {

struct K2D;

// Device copy of every shot's kernel arguments and of the tile prefix sums
// for the four tile heights (rows per thread 1, 2, 4, 8).
struct Ragged2D {
    int kind = 0, nshot = 0, order = 0;
    K2D *d_desc = nullptr;
    int *d_tile0 = nullptr;  // 4 x (nshot + 1)
    int total[4] = {0, 0, 0, 0};
};

int ragged2d_create(int kind, int nshot, const rsg_grid *grids, const rsg_state *states, Ragged2D *out);
void ragged2d_destroy(Ragged2D *p);
// which: 0 velocity, 1 stress, 2 adjoint-state velocity, 3 adjoint-state stress (E2D)
int ragged2d_step(const Ragged2D *p, int which, hipStream_t stream);

}  // namespace rsg
