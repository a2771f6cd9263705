// Debug bounds checks.  A library built with GFC_BOUNDS_CHECK defined
// checks the indices that come from data or from the work queue (a
// reaction's specie and efficiency-set indices, the cell of a work-queue
// ticket, the recorded step and checkpoint slot of the adjoint) before they
// address memory.  A failed check sets its bit in the
// evaluator's device word (DeviceMech::bounds) and replaces the index by 0,
// so the kernel goes on without the out-of-range access; the host turns a
// nonzero word into GFC_ERR_BOUNDS after the launch.  In the default build
// every check folds to the index itself.
#pragma once

#include <hip/hip_runtime.h>

#include "constants.hpp"

namespace gfc {

enum BoundCheck : unsigned {
    kBoundSpecies = 1u, // a reaction's specie index outside [0, nS)
    kBoundSet = 2u,     // an efficiency-set index outside [0, nSets)
    kBoundCell = 4u,    // a cell index outside [0, nCells)
    kBoundStep = 8u,    // a recorded step outside [0, maxSteps)
    kBoundSlot = 16u,   // a checkpoint slot outside [0, nck)
};

#if defined(GFC_BOUNDS_CHECK)
constexpr bool kBoundsChecked = true;
#else
constexpr bool kBoundsChecked = false;
#endif

// i when 0 <= i < n; otherwise (checked build only) records `code` in
// *word (when given) and returns 0.
GFC_HD __forceinline__ int bound(int i, int n, unsigned* word, unsigned code) {
    if (kBoundsChecked && (i < 0 || i >= n)) {
        if (word) {
#if defined(__HIP_DEVICE_COMPILE__)
            atomicOr(word, code);
#else
            *word |= code;
#endif
        }
        return 0;
    }
    return i;
}

} // namespace gfc
