// Internal declarations of the FWI driver.
#pragma once
#include <vector>

#include "../../include/rsgpu.h"

namespace rsg {

// Binomial checkpointing schedule (revolve.cpp:8-305).
struct Revolve {
    enum Action { ADVANCE = 0, TAKESHOT, RESTORE, FIRSTURN, YOUTURN, TERMINATE, ERROR };
    static constexpr int kCheckUp = 512;  // revolve.h:17
    static constexpr int kRepsUp = 64;    // revolve.h:18
    int steps, snaps, capo, fine, check, turn, oldfine;
    std::vector<int> ch;
    Revolve(int steps, int snaps);
    Action next();
    static int maxrange(int ss, int tt);
    static int adjust(int steps);
    static int numforw(int steps, int snaps);
};

}  // namespace rsg
