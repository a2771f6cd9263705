// Four cells per wavefront for small mechanisms (nSpecie+1 <= 16).
//
// The wavefront-per-cell kernel (team_kernel.hip) leaves most lanes idle
// when a mechanism has a dozen species: H2 uses 11 of 64 lanes for the
// specie-parallel work and 27 for its reactions.  Here a wavefront holds
// four teams of 16 lanes (one DPP row each), and each team integrates its
// own cell with the same per-cell algorithm:
//   * member i owns component i of the state and stage vectors;
//   * reactions run 16 per pass over the type-sorted records;
//   * sums and maxima over a team are row_shr trees inside the DPP row,
//     broadcast from the row's last lane, so every member holds the same
//     value;
//   * the LU keeps row i in member i's registers; pivot values and rows
//     are broadcast inside the team with ds_bpermute;
//   * each team has its own LDS block and its own step control, so the
//     four cells of a wavefront take independent step sequences; the
//     wavefront runs until its last team has no cell left.
// gpuFoam's steps (gpu_chemistry/src/gpuOde/*.H) are followed as in
// team_kernel.hip; the reduction trees differ, so results agree with the
// oracle to the parity tolerances rather than bit for bit with the
// 64-lane kernel.
#include <hip/hip_runtime.h>

#include "team16_common.hpp"

namespace gfc {

using namespace t16;

// ---------------------------------------------------------------- kernel --

// The persistent solve: each team takes cells from the queue until it is
// empty (queue == nullptr: one cell per team, fixed by the block index).  The kernels below run two wavefronts per SIMD: the register cap
// (256) spills ~350 B per lane, but the second wavefront hides the LDS and
// DPP latency that bounds a lone one; on 65536 H2 cells this runs 1.7x
// faster than at one per SIMD.
template <int N, bool REC>
__device__ __forceinline__ void t16_run(const SolveArgs& A, int* queue, const T16Record& R,
                                        Team16LDS<N>* Ls) {
    const int lane = threadIdx.x;
    const int team = lane >> 4, mem = lane & 15;
    Team16LDS<N>& L = Ls[team];
    const DeviceMech& m = A.m;
    const int nS = m.nS, n = nS + 1;
    const long ld = A.ldY;
    const RosTableau& tb = A.tab;

    // without a queue (the drop-in launch, grid = cells/4) team t of
    // block b integrates cell 4b+t once
    bool taken = false;
    for (;;) {
        int cell;
        if (queue) {
            int c = 0;
            if (mem == 0) c = atomicAdd(queue, 1);
            cell = team_bcast_i(c, lane, 0);
        } else {
            if (taken) break;
            taken = true;
            cell = (int)blockIdx.x * (kLanes / kTW) + team;
        }
        if (cell >= A.nCells) break;
        cell = bound(cell, A.nCells, A.m.bounds, kBoundCell);

        double y = (mem < n) ? A.Y0[mem * ld + cell] : 0.0;
        const double y0cell = y;
        const double p = A.Y0[(long)(nS + 1) * ld + cell];
        double dtChem = A.dtChem[cell];
        int status = 0, attempts = 0, accepted = 0;
        int count = 0; // recorded accepted steps
        if constexpr (REC) {
            double* y0c = rec_slot(R.y0c, cell, n + 1);
            if (mem < n) y0c[mem * kRecStride] = y;
            if (mem == 0) y0c[(long)n * kRecStride] = p;
        }
        if (team_bcast(y, lane, nS) > A.Treact) {
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
                    const double T = team_bcast(y, lane, nS);
                    const double dfdx = t16_jacobian<N>(m, p, T, y, L, lane, mem);
                    const double dydx0 = (T >= m.Tlow && T <= m.Thigh)
                                             ? dfdx
                                             : t16_derivatives<N>(m, p, T, y, L, lane, mem);
                    double err, yNew;
                    for (;;) {
                        double row[N];
                        const double tdiag = 1.0 / (tb.gamma * dx);
#pragma unroll
                        for (int k = 0; k < N; ++k) {
                            double v = 0.0;
                            if (k < n && mem < n) v = -L.J[mem * Team16LDS<N>::P + k];
                            if (k == mem) v += (mem < n) ? tdiag : 1.0;
                            row[k] = v;
                        }
                        int ps, plv;
                        double rd;
                        t16_lu<N>(row, n, lane, mem, ps, rd, plv);
                        t16_pivot_order<N>(row, rd, lane, plv);
                        // Rosenbrock stages (gpuRosenbrock{12,23,34}.H::solve)
                        double k1 = 0, k2 = 0, k3 = 0, k4 = 0, dydx = 0;
#pragma unroll 1
                        for (int st = 0; st < tb.stages; ++st) {
                            if (st == 1 || (st == 2 && tb.stages == 4)) {
                                const double yt = (st == 1) ? y + tb.a21 * k1
                                                            : y + tb.a31 * k1 + tb.a32 * k2;
                                dydx = t16_derivatives<N>(m, p, team_bcast(yt, lane, nS), yt,
                                                          L, lane, mem);
                            }
                            double rhs;
                            if (st == 0) rhs = dydx0 + dx * tb.d1 * dfdx;
                            else if (st == 1) rhs = dydx + dx * tb.d2 * dfdx + tb.c21 * k1 / dx;
                            else if (st == 2) rhs = dydx + dx * tb.d3 * dfdx + (tb.c31 * k1 + tb.c32 * k2) / dx;
                            else rhs = dydx + dx * tb.d4 * dfdx + (tb.c41 * k1 + tb.c42 * k2 + tb.c43 * k3) / dx;
                            const double kx = t16_solve<N>(row, lane, mem, rd, plv, rhs);
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
                        if (mem < n) {
                            const double tol = A.absTol + A.relTol * fmax(fabs(y), fabs(yNew));
                            ne = fabs(e) / tol;
                        }
                        err = team_max(ne, lane);
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
                        if (mem == 0) rec_slot(R.hs, cell, R.maxSteps)[(long)count * kRecStride] = dx;
                        if (count % R.K == 0 && mem < n)
                            rec_slot(R.ck, cell, R.nck * n)[((long)(count / R.K) * n + mem) *
                                                            kRecStride] = y;
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
                    if (mem < n) rec_slot(R.yEnd, cell, n)[mem * kRecStride] = y;
                }
                if (mem < nS) y = fmax(0.0, y);
                timeLeft -= xEnd;
                if (status & GFC_CELL_STEP_UNDERFLOW) break;
            }
        } else {
            dtChem = A.deltaT;
            status = GFC_CELL_SKIPPED;
        }
        if (A.RR && mem < nS)
            A.RR[mem * ld + cell] = A.rho[cell] * (y - y0cell) / A.deltaT;
        if (A.Ynew) {
            if (mem < n) A.Ynew[mem * ld + cell] = y;
            if (mem == 0) A.Ynew[(long)n * ld + cell] = p;
        }
        if (mem == 0) {
            if constexpr (REC) R.nAcc[cell] = count;
            if (A.dtMinBits) atomic_min_positive(A.dtMinBits, dtChem);
            A.dtChem[cell] = fmin(dtChem, A.deltaTChemMax);
            if (A.stats) {
                A.stats[cell] = status;
                A.stats[ld + cell] = attempts;
                A.stats[2 * ld + cell] = accepted;
            }
        }
        wave_sync();
    }
}

template <int N>
__global__ void __launch_bounds__(64, 2)
team16_solve_kernel(SolveArgs A, int* queue) {
    __shared__ Team16LDS<N> Ls[kLanes / kTW];
    t16_run<N, false>(A, queue, T16Record{}, Ls);
}

// The same solve recording its accepted steps for the discrete adjoint.
template <int N>
__global__ void __launch_bounds__(64, 2)
team16_record_kernel(SolveArgs A, int* queue, T16Record R) {
    __shared__ Team16LDS<N> Ls[kLanes / kTW];
    t16_run<N, true>(A, queue, R, Ls);
}

// Device-level evaluations with the 16-lane teams' own physics and LU: team
// t of block b takes cell 4b + t (grid = ceil(nCells / 4)).
template <int N>
__global__ void __launch_bounds__(64) team16_eval_kernel(EvalArgs E) {
    __shared__ Team16LDS<N> Ls[kLanes / kTW];
    const int lane = threadIdx.x;
    const int team = lane >> 4, mem = lane & 15;
    const int cell = (int)blockIdx.x * (kLanes / kTW) + team;
    if (cell >= E.nCells) return; // the teams are independent
    Team16LDS<N>& L = Ls[team];
    const DeviceMech& m = E.m;
    const int nS = m.nS, n = nS + 1;
    const long ld = E.ldY;
    if (E.mode == 2) {
        const double* M = E.M + (long)cell * n * n;
        double row[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double v = (k == mem) ? 1.0 : 0.0;  // identity rows beyond n
            if (mem < n) v = (k < n) ? M[mem * n + k] : 0.0;
            row[k] = v;
        }
        int ps, plv;
        double rd;
        t16_lu<N>(row, n, lane, mem, ps, rd, plv);
        t16_pivot_order<N>(row, rd, lane, plv);
        const double bl = (mem < n) ? E.b[mem * ld + cell] : 0.0;
        const double xl = t16_solve<N>(row, lane, mem, rd, plv, bl);
        if (mem < n) E.x[mem * ld + cell] = xl;
        return;
    }
    const double y = (mem < n) ? E.Y[mem * ld + cell] : 0.0;
    const double p = E.Y[(long)(nS + 1) * ld + cell];
    const double T = team_bcast(y, lane, nS);
    if (E.mode == 0) {
        const double d = t16_derivatives<N>(m, p, T, y, L, lane, mem);
        if (mem < n) E.dYdt[mem * ld + cell] = d;
    } else {
        const double d = t16_jacobian<N>(m, p, T, y, L, lane, mem);
        if (mem < n) E.dYdt[mem * ld + cell] = d;
        wave_sync();
        double* Jc = E.J + (long)cell * n * n;
        for (int e = mem; e < n * n; e += kTW) Jc[e] = L.J[(e / n) * Team16LDS<N>::P + e % n];
    }
}

template __global__ void team16_eval_kernel<11>(EvalArgs);
template __global__ void team16_eval_kernel<16>(EvalArgs);
template __global__ void team16_solve_kernel<11>(SolveArgs, int*);
template __global__ void team16_solve_kernel<16>(SolveArgs, int*);
template __global__ void team16_record_kernel<11>(SolveArgs, int*, T16Record);
template __global__ void team16_record_kernel<16>(SolveArgs, int*, T16Record);

} // namespace gfc
