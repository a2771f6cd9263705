// Plain per-cell kernel: one lane integrates one cell over the CFD time
// step, the mapping of gpuFoam's for_each_index(singleCellSolver)
// (gpu_utils/common/for_each_index.H:7-48,
// gpu_chemistry/src/gpuKernelEvaluator/gpuKernelEvaluator/singleCellSolver.H:27-51).
//
// Per-cell scratch (Jacobian, LU, stage vectors) lives in HBM in a
// structure-of-arrays layout: element e of a cell's array sits at
// base[e*stride + slot], so the 64 lanes of a wavefront touch 64
// consecutive doubles for every element (fully coalesced), unlike the
// reference's cell-major buffers (memoryResource.H:61-92).  The effective
// ODE dimension is n = nSpecie+1: the pressure row and column of the
// reference's (nSpecie+2)^2 Jacobian are identically zero, so pressure is
// carried as a per-cell constant.
//
// The arithmetic follows the reference operation by operation; this
// translation unit disables fused multiply-add contraction so that the only
// differences from the reference built for the host are the last-ulp
// differences of exp/log/pow.
#include <hip/hip_runtime.h>

#include "lane_physics.hpp"

namespace gfc {
using namespace lane;

// singleCellSolver + computeRR epilogue for one cell in one lane.
__global__ void __launch_bounds__(64, 4)
lane_solve_kernel(SolveArgs A, double* scratch, long per) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= A.nCells) return;
    const int nS = A.m.nS, n = nS + 1;
    const LaneBuf b{wave_slot(scratch, cell, per), kSlotStride, nS + 1};
    const long ld = A.ldY;
    for (int k = 0; k < n; ++k) b.y()[k] = A.Y0[k * ld + cell];
    const double p = A.Y0[(long)(nS + 1) * ld + cell];
    double dtChem = A.dtChem[cell];
    int status = 0, attempts = 0, accepted = 0;
    NoRecord rec;
    if (b.y()[nS] > A.Treact) {
        double timeLeft = A.deltaT;
        while (timeLeft > kSmall) {
            const double dt = timeLeft;
            status |= ode_solve(A, p, dt, dtChem, b, attempts, accepted, rec);
            for (int i = 0; i < nS; i++) b.y()[i] = fmax(0.0, b.y()[i]);
            timeLeft -= dt;
            if (status & GFC_CELL_STEP_UNDERFLOW) break;
        }
    } else {
        dtChem = A.deltaT;
        status = GFC_CELL_SKIPPED;
    }
    if (A.RR) {
        for (int i = 0; i < nS; ++i)
            A.RR[i * ld + cell] = A.rho[cell] * (b.y()[i] - A.Y0[i * ld + cell]) / A.deltaT;
    }
    if (A.Ynew) {
        for (int k = 0; k < n; ++k) A.Ynew[k * ld + cell] = b.y()[k];
        A.Ynew[(long)(nS + 1) * ld + cell] = p;
    }
    if (A.dtMinBits) atomic_min_positive(A.dtMinBits, dtChem);
    A.dtChem[cell] = fmin(dtChem, A.deltaTChemMax);
    if (A.stats) {
        A.stats[cell] = status;
        A.stats[ld + cell] = attempts;
        A.stats[2 * ld + cell] = accepted;
    }
}

long lane_scratch_per_cell(int nS) { return lane_scratch_doubles(nS); }

} // namespace gfc
