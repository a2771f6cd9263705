// Launch arguments and helpers shared by the chemistry kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gpufoam_amd.h"
#include "device_mech.hpp"

namespace gfc {

// Rosenbrock tableaux: gpuRosenbrock34.H:13-46 (Shampine),
// gpuRosenbrock23.H:13-37, gpuRosenbrock12.H:14-31.
struct RosTableau {
    int stages; // 2, 3 or 4
    double gamma;
    double a21, a31, a32;
    double c21, c31, c32, c41, c42, c43;
    double b1, b2, b3, b4;
    double e1, e2, e3, e4;
    double d1, d2, d3, d4;
};

inline RosTableau tableau(int solver) {
    RosTableau t{};
    if (solver == 34) {
        t.stages = 4;
        t.gamma = 1.0 / 2.0;
        t.a21 = 2.0; t.a31 = 48.0 / 25.0; t.a32 = 6.0 / 25.0;
        t.c21 = -8.0; t.c31 = 372.0 / 25.0; t.c32 = 12.0 / 5.0;
        t.c41 = -112.0 / 125.0; t.c42 = -54.0 / 125.0; t.c43 = -2.0 / 5.0;
        t.b1 = 19.0 / 9.0; t.b2 = 1.0 / 2.0; t.b3 = 25.0 / 108.0;
        t.b4 = 125.0 / 108.0;
        t.e1 = 34.0 / 108.0; t.e2 = 7.0 / 36.0; t.e3 = 0.0; t.e4 = 125.0 / 108.0;
        t.d1 = 1.0 / 2.0; t.d2 = -3.0 / 2.0; t.d3 = 605.0 / 250.0;
        t.d4 = 29.0 / 250.0;
    } else if (solver == 23) {
        t.stages = 3;
        t.gamma = 0.43586652150845899941601945119356;
        t.a21 = 1; t.a31 = 1; t.a32 = 0;
        t.c21 = -1.0156171083877702091975600115545;
        t.c31 = 4.0759956452537699824805835358067;
        t.c32 = 9.2076794298330791242156818474003;
        t.b1 = 1; t.b2 = 6.1697947043828245592553615689730;
        t.b3 = -0.4277225654321857332623837380651;
        t.e1 = 0.5; t.e2 = -2.9079558716805469821718236208017;
        t.e3 = 0.2235406989781156962736090927619;
        t.d1 = 0.43586652150845899941601945119356;
        t.d2 = 0.24291996454816804366592249683314;
        t.d3 = 2.1851380027664058511513169485832;
    } else {
        t.stages = 2;
        const double g = 1.0 + 1.0 / 1.414213562373095;
        t.gamma = g;
        t.a21 = 1.0 / g;
        t.c21 = -2.0 / g;
        t.b1 = (3.0 / 2.0) / g;
        t.b2 = (1.0 / 2.0) / g;
        t.e1 = t.b1 - 1.0 / g;
        t.e2 = t.b2;
        t.d1 = g;
        t.d2 = -g;
    }
    return t;
}

// Step-size controller inputs (gpuAdaptiveSolver.H:10-50,
// gpuODESolverBase.H:12-57) with the per-launch constants.
struct SolveArgs {
    DeviceMech m;
    RosTableau tab;
    double absTol, relTol;
    int maxSteps;
    double safeScale, alphaInc, alphaDec, minScale, maxScale;
    double incThreshold; // pow(maxScale/safeScale, -1/alphaInc)
    double Treact;
    double deltaT;
    double deltaTChemMax;
    int nCells;
    int ldY;
    const double* rho;
    const double* Y0;
    double* Ynew;
    double* RR;
    double* dtChem;
    unsigned long long* dtMinBits; // atomicMin target (positive doubles)
    int* stats;                    // [3*ldY] or null
    // work queue of the persistent wavefront kernel: ticket t < qBulk/qChunk
    // is the chunk of cells [t*qChunk, (t+1)*qChunk), integrated one after
    // the other by the wavefront that drew it; later tickets are single
    // cells from qBulk on (qChunk = 1, qBulk = 0: one cell per ticket)
    int qChunk;
    int qBulk;
};

// Largest qChunk: 16 doubles are one 128-byte line of a state row.
constexpr int kQueueChunkMax = 16;

// Device-level evaluation of one kernel family's physics (the functions its
// solve uses), one cell per wavefront / 16-lane team / workgroup:
// mode 0 the derivative, 1 the Jacobian with the derivative computed
// alongside it, 2 the factorisation of a given matrix and one solve.
struct EvalArgs {
    DeviceMech m;
    int mode;
    int nCells;
    long ldY;
    const double* Y;  // [nS+2][ldY]: species, T, p (modes 0, 1)
    double* dYdt;     // [n][ldY] (modes 0, 1)
    double* J;        // [nCells][n][n] row-major (mode 1)
    const double* M;  // [nCells][n][n] row-major (mode 2)
    const double* b;  // [n][ldY] (mode 2)
    double* x;        // [n][ldY] (mode 2)
    int lu;           // workgroup kernel's factorisation variant (mode 2)
};

// Capacity of the workgroup-per-cell kernel (block_kernel.hip): n =
// nSpecie+1 in (64, kBlockMaxN], at most kBlockMaxR reactions (net rates in
// LDS) and kBlockMaxSets distinct third-body efficiency sets.
constexpr int kBlockMaxN = 208;
constexpr int kBlockMaxR = 1280;
constexpr int kBlockMaxSets = 512;

// Doubles of the kernel's shared LDS region (diagonal factor blocks / row
// sums / per-reaction derivatives) for n = nSpecie+1: 16 x NA rows and
// NC 32-wide column blocks with NA = 8, NC = 4 up to n = 128 and
// NA = 13, NC = 7 above.
constexpr long block_lds_doubles(int n) {
    return n <= 128 ? (4 * 32 * 33 > 32 * 128 ? 4 * 32 * 33 : 32 * 128)
                    : (7 * 32 * 33 > 32 * 208 ? 7 * 32 * 33 : 32 * 208);
}

__device__ __forceinline__ void atomic_min_positive(unsigned long long* p,
                                                    double v) {
    atomicMin(p, static_cast<unsigned long long>(__double_as_longlong(v)));
}

} // namespace gfc
