// Wavefront-per-cell chemistry kernel.
//
// One 64-lane wavefront integrates one cell.  The work of gpuFoam's
// per-cell solver (singleCellSolver -> odeSolve -> adaptiveSolve ->
// Rosenbrock step -> jacobian / LU / derivatives; see
// gpu_chemistry/src/gpuOde/*.H) is spread over the lanes:
//   * species-parallel: lane i owns component i of every state and stage
//     vector (y, dydx0, dfdx, k1..k4), in registers; mixture sums are DPP
//     reductions;
//   * reaction-parallel: lane l evaluates reactions l, l+64, ... (stored
//     grouped by rate type so a pass of 64 runs one code path).  Rate
//     constants are evaluated as exp(ln A + beta ln T - Ta/T) and the
//     reverse rate as exp(... - ln Kc) with ln T, 1/T and ln(pStd/RT)
//     computed once per cell; third-body concentrations are computed once
//     per distinct efficiency vector of the mechanism.  Contributions to
//     dN/dt and to the Jacobian are accumulated in LDS with FP64 LDS
//     atomics;
//   * row-parallel LU: lane i holds row i of -J + I/(gamma*dx) in N
//     registers.  Crout with implicit scaling (ludecompose.H:9-92) becomes
//     right-looking elimination with the same per-entry operation sequence;
//     the pivot (largest scaled value, ties to the last row position) is a
//     DPP argmax, the pivot row is broadcast with v_readlane;
//   * the Jacobian stays in LDS across rejected attempts (it depends only on
//     y0), so a retry only re-factorises.
// N is the compile-time row capacity (n = nSpecie+1 <= N <= 64); rows
// n..N-1 are identity rows.
#include <hip/hip_runtime.h>

#include "team16_common.hpp"
#include "team_device.hpp"

namespace gfc {

using namespace wave;

// ---------------------------------------------------------------- kernel --

// The persistent (queue) or drop-in (queue == nullptr: cell = block index)
// solve of one wavefront.  With REC it also records the accepted steps for
// the discrete adjoint (the per-cell arrays of t16::T16Record, wave-blocked):
// the initial state, every accepted step's size, every K-th accepted step's
// initial state, the final state before the non-negativity clamp and the
// number of accepted steps.
template <int N, bool REC>
__device__ __forceinline__ void team_run(const SolveArgs& A, int* queue, const t16::T16Record& R,
                                         TeamLDS<N>& L) {
    const int lane = threadIdx.x;
    const DeviceMech& m = A.m;
    const int nS = m.nS, n = nS + 1;
    const long ld = A.ldY;
    const RosTableau& tb = A.tab;

    // the cells of one chunk share the cache lines of their state rows, so
    // one wavefront (one XCD's L2) reads each line once
    // The results of a chunk's cells are held back and written when the
    // chunk ends: lane k keeps row k of every cell (the pressure in lane n),
    // lane s the time step and counters of cell s, so each row of the chunk
    // is written as one whole line.
    int cell = queue ? 0 : blockIdx.x;
    int cBase = cell, cEnd = cell + 1;
    double yb[kQueueChunkMax];
#pragma unroll
    for (int i = 0; i < kQueueChunkMax; ++i) yb[i] = 0.0;
    double dtb = 0.0;
    int sb0 = 0, sb1 = 0, sb2 = 0;
    if (queue) cEnd = 0;
    for (;;) {
        if (queue && ++cell >= cEnd) {
            int t = 0;
            if (lane == 0) t = atomicAdd(queue, 1);
            t = __builtin_amdgcn_readfirstlane(t);
            const int nChunks = A.qBulk / A.qChunk;
            if (t < nChunks) {
                cell = t * A.qChunk;
                cEnd = cell + A.qChunk;
            } else {
                cell = A.qBulk + (t - nChunks);
                cEnd = cell + 1;
            }
            cBase = cell;
        }
        if (cell >= A.nCells) break;
        cell = bound(cell, A.nCells, m.bounds, kBoundCell);

        double y = (lane < n) ? A.Y0[lane * ld + cell] : 0.0;
        const double y0cell = y;
        const double p = A.Y0[(long)(nS + 1) * ld + cell];
        double dtChem = A.dtChem[cell];
        int status = 0, attempts = 0, accepted = 0;
        int count = 0; // recorded accepted steps
        if constexpr (REC) {
            double* y0c = t16::rec_slot(R.y0c, cell, n + 1);
            if (lane < n) y0c[lane * t16::kRecStride] = y;
            if (lane == 0) y0c[(long)n * t16::kRecStride] = p;
        }
        if (readlane(y, nS) > A.Treact) {
            double timeLeft = A.deltaT;
            while (timeLeft > kSmall) {
                const double xEnd = timeLeft;
                // odeSolve (gpuODESolverBase.H:67-111)
                double stepDx = dtChem;
                bool last = false, reached = false;
                double x = 0.0;
                for (int nStep = 0; nStep < A.maxSteps; nStep++) {
                    const double dxTry0 = stepDx;
                    if ((x + stepDx - xEnd) * (x + stepDx - 0.0) > 0) {
                        last = true;
                        stepDx = xEnd - x;
                    }
                    // adaptiveSolve (gpuAdaptiveSolver.H:83-134)
                    double dx = stepDx;
                    const double T = readlane(y, nS);
                    // the Jacobian call evaluates the right-hand side at y0
                    // alongside (its dfdx); it equals derivatives(y0) unless
                    // T lies outside the reaction temperature limits, where
                    // omega() clips T and the Jacobian does not
                    const double dfdx = team_jacobian<N>(m, p, T, y, L, lane);
                    const double dydx0 = (T >= m.Tlow && T <= m.Thigh)
                                             ? dfdx
                                             : team_derivatives<N>(m, p, T, y, L, lane);
                    double err, yNew;
                    for (;;) {
                        // -J + I/(gamma dx), factorised in registers
                        double row[N];
                        const double tdiag = 1.0 / (tb.gamma * dx);
#pragma unroll
                        for (int k = 0; k < N; ++k) {
                            double v = 0.0;
                            if (k < n && lane < n) v = -L.J[lane * TeamLDS<N>::P + k];
                            if (k == lane) v += (lane < n) ? tdiag : 1.0;
                            row[k] = v;
                        }
                        int ps, plv;
                        double rd;
                        team_lu<N>(row, n, lane, ps, rd, plv);
                        // Rosenbrock stages (gpuRosenbrock{12,23,34}.H::solve)
                        double k1 = 0, k2 = 0, k3 = 0, k4 = 0, dydx = 0;
#pragma unroll 1
                        for (int st = 0; st < tb.stages; ++st) {
                            if (st == 1 || (st == 2 && tb.stages == 4)) {
                                const double yt = (st == 1) ? y + tb.a21 * k1
                                                            : y + tb.a31 * k1 + tb.a32 * k2;
                                dydx = team_derivatives<N>(m, p, readlane(yt, nS), yt, L, lane);
                            }
                            double rhs;
                            if (st == 0) rhs = dydx0 + dx * tb.d1 * dfdx;
                            else if (st == 1) rhs = dydx + dx * tb.d2 * dfdx + tb.c21 * k1 / dx;
                            else if (st == 2) rhs = dydx + dx * tb.d3 * dfdx + (tb.c31 * k1 + tb.c32 * k2) / dx;
                            else rhs = dydx + dx * tb.d4 * dfdx + (tb.c41 * k1 + tb.c42 * k2 + tb.c43 * k3) / dx;
                            const double kx = team_solve<N>(row, lane, ps, rd, plv, rhs);
                            if (st == 0) k1 = kx;
                            else if (st == 1) k2 = kx;
                            else if (st == 2) k3 = kx;
                            else k4 = kx;
                        }
                        double e;
                        if (tb.stages == 2) {
                            yNew = y + tb.b1 * k1 + tb.b2 * k2;
                            e = tb.e1 * k1 + tb.e2 * k2;
                        } else if (tb.stages == 3) {
                            yNew = y + tb.b1 * k1 + tb.b2 * k2 + tb.b3 * k3;
                            e = tb.e1 * k1 + tb.e2 * k2 + tb.e3 * k3;
                        } else {
                            yNew = y + tb.b1 * k1 + tb.b2 * k2 + tb.b3 * k3 + tb.b4 * k4;
                            e = tb.e1 * k1 + tb.e2 * k2 + tb.e4 * k4;
                        }
                        double ne = 0.0;
                        if (lane < n) {
                            const double tol = A.absTol + A.relTol * fmax(fabs(y), fabs(yNew));
                            ne = fabs(e) / tol;
                        }
                        err = wave_max(ne);
                        ++attempts;
                        if (err > 1) {
                            const double scale = fmax(A.safeScale * pow(err, -A.alphaDec), A.minScale);
                            dx *= scale;
                            if (dx < kVSmall) {
                                status |= GFC_CELL_STEP_UNDERFLOW;
                                break;
                            }
                        } else {
                            break;
                        }
                    }
                    if (status & GFC_CELL_STEP_UNDERFLOW) break;
                    ++accepted;
                    if constexpr (REC) {
                        // the state at the start of the step and its size
                        const int ct = bound(count, R.maxSteps, m.bounds, kBoundStep);
                        if (lane == 0)
                            t16::rec_slot(R.hs, cell, R.maxSteps)[(long)ct * t16::kRecStride] = dx;
                        if (count % R.K == 0 && lane < n)
                            t16::rec_slot(R.ck, cell, R.nck * n)
                                [((long)bound(count / R.K, R.nck, m.bounds, kBoundSlot) * n + lane) *
                                 t16::kRecStride] = y;
                        ++count;
                    }
                    x += dx;
                    y = yNew;
                    if (err > A.incThreshold) {
                        stepDx = fmin(fmax(A.safeScale * pow(err, -A.alphaInc), A.minScale),
                                      A.maxScale) * dx;
                    } else {
                        stepDx = A.safeScale * A.maxScale * dx;
                    }
                    if ((x - xEnd) * (xEnd - 0.0) >= 0) {
                        if (nStep > 0 && last) stepDx = dxTry0;
                        dtChem = stepDx;
                        reached = true;
                        break;
                    }
                }
                if (!reached && !(status & GFC_CELL_STEP_UNDERFLOW))
                    status |= GFC_CELL_MAX_STEPS;
                if constexpr (REC) {
                    if (lane < n) t16::rec_slot(R.yEnd, cell, n)[lane * t16::kRecStride] = y;
                }
                if (lane < nS) y = fmax(0.0, y);
                timeLeft -= xEnd;
                if (status & GFC_CELL_STEP_UNDERFLOW) break;
            }
        } else {
            dtChem = A.deltaT;
            status = GFC_CELL_SKIPPED;
        }
        if (A.RR && lane < nS)
            A.RR[lane * ld + cell] = A.rho[cell] * (y - y0cell) / A.deltaT;
        if constexpr (REC) {
            if (lane == 0) R.nAcc[cell] = count;
        }
        if (lane == 0 && A.dtMinBits) atomic_min_positive(A.dtMinBits, dtChem);
        const int slot = cell - cBase;
        const double yOut = (lane < n) ? y : p;
#pragma unroll
        for (int i = 0; i < kQueueChunkMax; ++i) yb[i] = (i == slot) ? yOut : yb[i];
        if (lane == slot) {
            dtb = fmin(dtChem, A.deltaTChemMax);
            sb0 = status;
            sb1 = attempts;
            sb2 = accepted;
        }
        if (cell + 1 >= cEnd || cell + 1 >= A.nCells) {
            const int cnt = min(cEnd, A.nCells) - cBase;
            if (A.Ynew && lane <= n) {
                double* out = A.Ynew + (long)lane * ld + cBase;
#pragma unroll
                for (int i = 0; i < kQueueChunkMax; ++i)
                    if (i < cnt) out[i] = yb[i];
            }
            if (lane < cnt) {
                A.dtChem[cBase + lane] = dtb;
                if (A.stats) {
                    A.stats[cBase + lane] = sb0;
                    A.stats[ld + cBase + lane] = sb1;
                    A.stats[2 * ld + cBase + lane] = sb2;
                }
            }
        }
        if (!queue) break;
        __syncthreads();
    }
}

// Small mechanisms (N <= 16) fit in 256 registers: two wavefronts per SIMD
// hide each other's latency.  Larger ones need the AGPRs as well and run
// one wavefront per SIMD.
template <int N>
__global__ void __launch_bounds__(64, (N <= 16) ? 2 : 1)
team_solve_kernel(SolveArgs A, int* queue) {
    __shared__ TeamLDS<N> L;
    team_run<N, false>(A, queue, t16::T16Record{}, L);
}

// The same solve recording its accepted steps for the discrete adjoint.
template <int N>
__global__ void __launch_bounds__(64, (N <= 16) ? 2 : 1)
team_record_kernel(SolveArgs A, int* queue, t16::T16Record R) {
    __shared__ TeamLDS<N> L;
    team_run<N, true>(A, queue, R, L);
}

// Device-level evaluations with this kernel's own physics and LU (one cell
// per wavefront; grid = nCells).
template <int N>
__global__ void __launch_bounds__(64) team_eval_kernel(EvalArgs E) {
    __shared__ TeamLDS<N> L;
    const int lane = threadIdx.x;
    const int cell = blockIdx.x;
    if (cell >= E.nCells) return;
    const DeviceMech& m = E.m;
    const int nS = m.nS, n = nS + 1;
    const long ld = E.ldY;
    if (E.mode == 2) {
        const double* M = E.M + (long)cell * n * n;
        double row[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double v = (k == lane) ? 1.0 : 0.0;  // identity rows beyond n
            if (lane < n) v = (k < n) ? M[lane * n + k] : 0.0;
            row[k] = v;
        }
        int ps, plv;
        double rd;
        team_lu<N>(row, n, lane, ps, rd, plv);
        const double bl = (lane < n) ? E.b[lane * ld + cell] : 0.0;
        const double xl = team_solve<N>(row, lane, ps, rd, plv, bl);
        if (lane < n) E.x[lane * ld + cell] = xl;
        return;
    }
    const double y = (lane < n) ? E.Y[lane * ld + cell] : 0.0;
    const double p = E.Y[(long)(nS + 1) * ld + cell];
    const double T = readlane(y, nS);
    if (E.mode == 0) {
        const double d = team_derivatives<N>(m, p, T, y, L, lane);
        if (lane < n) E.dYdt[lane * ld + cell] = d;
    } else {
        const double d = team_jacobian<N>(m, p, T, y, L, lane);
        if (lane < n) E.dYdt[lane * ld + cell] = d;
        double* Jc = E.J + (long)cell * n * n;
        for (int e = lane; e < n * n; e += kLanes) Jc[e] = L.J[(e / n) * TeamLDS<N>::P + e % n];
    }
}

template __global__ void team_eval_kernel<11>(EvalArgs);
template __global__ void team_eval_kernel<32>(EvalArgs);
template __global__ void team_eval_kernel<54>(EvalArgs);
template __global__ void team_eval_kernel<64>(EvalArgs);
template __global__ void team_solve_kernel<11>(SolveArgs, int*);
template __global__ void team_solve_kernel<32>(SolveArgs, int*);
template __global__ void team_solve_kernel<54>(SolveArgs, int*);
template __global__ void team_solve_kernel<64>(SolveArgs, int*);
template __global__ void team_record_kernel<11>(SolveArgs, int*, t16::T16Record);
template __global__ void team_record_kernel<32>(SolveArgs, int*, t16::T16Record);
template __global__ void team_record_kernel<54>(SolveArgs, int*, t16::T16Record);
template __global__ void team_record_kernel<64>(SolveArgs, int*, t16::T16Record);

} // namespace gfc
