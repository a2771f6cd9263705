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

#include "team_common.hpp"

namespace gfc {

namespace {

using namespace wave;

template <int N>
struct TeamLDS {
    static constexpr int P = (N % 2 == 0) ? N + 1 : N; // odd row pitch
    double J[N * P];
    double c[kLanes];
    double dN[kLanes];
    double Ha[kLanes];
    double colA[kLanes];            // rhoM / W_j
    double colB[kLanes];            // rhoM * v_j
    double M[kTeamMaxSets];         // third-body concentration per set
    double gL[kLanes], gH[kLanes];  // W_i Gstd_i with the low / high polynomial
    double hL[kLanes], hH[kLanes];  // W_i Ha_i with the low / high polynomial
    double Wd[kMaxSets * kLanes];   // dense Jacobian weights per set, specie
};

// Third-body concentrations of every efficiency set, M[s] = sum_i eff_si c_i
// (gpuThirdBodyEfficiencies.H:21-27, summed in specie order): lane s sums
// set s, so the sets are computed side by side.  L.c must hold c.
template <int N>
__device__ __forceinline__ void third_body(const DeviceMech& m, TeamLDS<N>& L, int lane) {
    __syncthreads();
    const int nS = m.nS;
    for (int s = lane; s < m.nSets; s += kLanes) {
        const double* e = m.effSets + s * kLanes;
        double acc = 0.0;
#pragma unroll 4
        for (int i = 0; i < nS; ++i) acc += e[i] * L.c[i];
        L.M[s] = acc;
    }
}

// Specie Gibbs energies (and enthalpies) with both polynomial sets, times W:
// a reaction's Y*Gstd of its composite thermo (gpuThermo.H:354-470,
// makeGpuReactions.H:18-59) is the stoichiometric sum of these with the set
// chosen by the composite's Tcommon.
template <int N, bool ENTHALPY>
__device__ __forceinline__ void specie_gibbs(const DeviceMech& m, TeamLDS<N>& L,
                                             const TFun& f, int lane) {
    if (lane < m.nS) {
        const double W = m.W[lane];
        const double* lo = m.low + 7 * lane;
        const double* hi = m.high + 7 * lane;
        L.gL[lane] = W * poly_Gstd(lo, f.T, f.logT);
        L.gH[lane] = W * poly_Gstd(hi, f.T, f.logT);
        if (ENTHALPY) {
            L.hL[lane] = W * poly_Ha(lo, f.T);
            L.hH[lane] = W * poly_Ha(hi, f.T);
        }
    }
}

// ---------------------------------------------------------- derivatives --
// gpuODESystem::derivatives (gpuODESystem.H:164-196).  yl: component of
// this lane; returns the dydx component of this lane.
template <int N>
__device__ __forceinline__ double team_derivatives(const DeviceMech& m, double p,
                                                   double T, double yl,
                                                   TeamLDS<N>& L, int lane) {
    const int nS = m.nS;
    double rhoPart = 0.0, cpPart = 0.0, Ha = 0.0, cu = 0.0;
    if (lane < nS) {
        const double yi = fmax(yl, 0.0);
        const double rhoi = p / (m.R[lane] * T);
        const double* a = coeffs_of(m.high + 7 * lane, m.low + 7 * lane, m.Tcommon[lane], T);
        rhoPart = yi / rhoi;
        cpPart = yi * poly_Cp(a, T);
        Ha = poly_Ha(a, T);
        cu = yi / m.W[lane];
    }
    const double rhoM = 1.0 / wave_sum(rhoPart);
    const double CpM = wave_sum(cpPart);
    const double cl = cu * rhoM;
    L.c[lane] = cl;
    L.dN[lane] = 0.0;
    third_body<N>(m, L, lane);
    // omega() evaluates the rates at T clipped to [Tlow, Thigh]
    const TFun f = tfun(fmin(fmax(T, m.Tlow), m.Thigh));
    specie_gibbs<N, false>(m, L, f, lane);
    __syncthreads();
#pragma unroll 1
    for (int pass = 0; pass < m.nPass; ++pass) {
        const int r = pass * kLanes + lane;
        if (r < m.nR) {
            const Rxn x = load_rxn(m, pass, lane);
            double kArr, ddT, pw[kTeamWidth];
            const double kf = forward_rate(x, f, L.M, kArr, ddT);
            const double kr = x.irr ? 0.0 : kf * exp(-ln_kc(x, f, L, nullptr));
            const double Cf = side_powers(x.li, x.lei, x.le, x.nl, L.c, pw);
            const double Cr = side_powers(x.ri, x.rei, x.re, x.nr, L.c, pw);
            const double w = kf * Cf - kr * Cr;
#pragma unroll
            for (int q = 0; q < kTeamWidth; ++q) {
                if (q < x.nl) atomicAdd(&L.dN[x.li[q]], -(x.ls[q] * w));
                if (q < x.nr) atomicAdd(&L.dN[x.ri[q]], x.rs[q] * w);
            }
        }
    }
    __syncthreads();
    double d = 0.0;
    if (lane < nS) d = L.dN[lane] * (m.W[lane] / rhoM);
    const double dT = -wave_sum(d * Ha) / CpM;
    if (lane == nS) d = dT;
    return d;
}

// ------------------------------------------------------------- jacobian --
// gpuODESystem::jacobian (gpuODESystem.H:198-281).  Leaves J (n x n, the
// pressure row/column dropped) in L.J and returns this lane's component of
// the derivative computed alongside (the reference's "dfdx").
template <int N>
__device__ __forceinline__ double team_jacobian(const DeviceMech& m, double p, double T,
                                             double yl, TeamLDS<N>& L, int lane) {
    constexpr int P = TeamLDS<N>::P;
    const int nS = m.nS, n = nS + 1;
    double rhoPart = 0, cpPart = 0, dcpPart = 0, avPart = 0, Ha = 0, Cp = 0, cu = 0;
    if (lane < nS) {
        const double yi = fmax(yl, 0.0);
        const double rhoi = p / (m.R[lane] * T);
        const double vi = 1.0 / rhoi;
        const double* a = coeffs_of(m.high + 7 * lane, m.low + 7 * lane, m.Tcommon[lane], T);
        Cp = poly_Cp(a, T);
        Ha = poly_Ha(a, T);
        rhoPart = yi / rhoi;
        cpPart = yi * Cp;
        dcpPart = yi * poly_dCpdT(a, T);
        avPart = yi * vi * (1.0 / T);
        cu = yi / m.W[lane];
    }
    const double rhoM = 1.0 / wave_sum(rhoPart);
    const double CpM = wave_sum(cpPart);
    const double dCpMdT = wave_sum(dcpPart);
    const double alphavM = wave_sum(avPart) * rhoM;
    const double cl = cu * rhoM;
    L.c[lane] = cl;
    L.dN[lane] = 0.0;
    L.Ha[lane] = Ha;
    if (lane < nS) {
        L.colA[lane] = rhoM / m.W[lane];
        L.colB[lane] = rhoM * (1.0 / (p / (m.R[lane] * T)));
    }
    third_body<N>(m, L, lane);
    const TFun f = tfun(T);
    specie_gibbs<N, true>(m, L, f, lane);
    for (int e = lane; e < n * P; e += kLanes) L.J[e] = 0.0;
    const int nChunks = m.nSets > 0 ? (m.nSets + kMaxSets - 1) / kMaxSets : 1;
    const bool fusedDense = m.nSets <= kSetsT;
    bool poisoned = false; // a fall-off rate with Cf - Cr/Kc not finite

#pragma unroll 1
    for (int chunk = 0; chunk < nChunks; ++chunk) {
        for (int e = lane; e < kMaxSets * kLanes; e += kLanes) L.Wd[e] = 0.0;
        __syncthreads();
        const int s0 = chunk * kMaxSets;
#pragma unroll 1
        for (int pass = 0; pass < m.nPass; ++pass) {
            const int r = pass * kLanes + lane;
            if (r >= m.nR) continue;
            const Rxn x = load_rxn(m, pass, lane);
            const bool denseHere =
                x.type == GFC_RATE_THIRD_BODY && x.set >= s0 && x.set < s0 + kMaxSets;
            // sparse part in the first chunk; dense part with its set's chunk
            if (chunk > 0 && !denseHere) continue;
            double kArr, ddT, pwl[kTeamWidth], pwr[kTeamWidth];
            const double kf = forward_rate(x, f, L.M, kArr, ddT);
            const double Cf = side_powers(x.li, x.lei, x.le, x.nl, L.c, pwl);
            double dwdT = Cf * ddT;
            double w = Cf * kf;
            double Cr = 0.0, Kc = 1.0, kr = 0.0;
            if (!x.irr) {
                double dKc;
                const double lnKc = ln_kc(x, f, L, &dKc);
                Kc = exp(lnKc);
                const double rKc = exp(-lnKc);
                kr = kf * rKc;
                Cr = side_powers(x.ri, x.rei, x.re, x.nr, L.c, pwr);
                const double dkrdT = ddT * rKc - kr * dKc;
                dwdT -= Cr * dkrdT;
                w -= Cr * kr;
            }
            const double temp = (Cf - Cr / Kc);
            if (denseHere) {
                // ddc_j = eff_j k (Cf - Cr/Kc): row weights of the set
                const double g = kArr * temp;
                double* Wd = &L.Wd[(x.set - s0) * kLanes];
#pragma unroll
                for (int i = 0; i < kTeamWidth; ++i) {
                    if (i < x.nl) atomicAdd(&Wd[x.li[i]], -(x.ls[i] * g));
                    if (i < x.nr) atomicAdd(&Wd[x.ri[i]], x.rs[i] * g);
                }
            }
            if (chunk > 0) continue;
            if (x.type > GFC_RATE_THIRD_BODY && !isfinite(temp)) poisoned = true;
            // dN/dt (the reference's dfdx) and dCdT (gpuReaction.H:595-612)
#pragma unroll
            for (int i = 0; i < kTeamWidth; ++i) {
                if (i < x.nl) {
                    atomicAdd(&L.dN[x.li[i]], -(x.ls[i] * w));
                    atomicAdd(&L.J[x.li[i] * P + nS], -(x.ls[i] * dwdT));
                }
                if (i < x.nr) {
                    atomicAdd(&L.dN[x.ri[i]], x.rs[i] * w);
                    atomicAdd(&L.J[x.ri[i] * P + nS], x.rs[i] * dwdT);
                }
            }
            // jac_dCfdcj_contribution (gpuReaction.H:547-569)
#pragma unroll
            for (int j = 0; j < kTeamWidth; ++j) {
                if (j >= x.nl) break;
                const double dCf = side_dpow(x.li, x.lei, x.le, x.nl, j, L.c, pwl) * kf;
                const int sj = x.li[j];
#pragma unroll
                for (int i = 0; i < kTeamWidth; ++i) {
                    if (i < x.nl) atomicAdd(&L.J[x.li[i] * P + sj], -(x.ls[i] * dCf));
                    if (i < x.nr) atomicAdd(&L.J[x.ri[i] * P + sj], x.rs[i] * dCf);
                }
            }
            // jac_dCrdcj_contribution (:571-593)
            if (!x.irr) {
#pragma unroll
                for (int j = 0; j < kTeamWidth; ++j) {
                    if (j >= x.nr) break;
                    const double dCr = side_dpow(x.ri, x.rei, x.re, x.nr, j, L.c, pwr) * kr;
                    const int sj = x.ri[j];
#pragma unroll
                    for (int i = 0; i < kTeamWidth; ++i) {
                        if (i < x.nl) atomicAdd(&L.J[x.li[i] * P + sj], x.ls[i] * dCr);
                        if (i < x.nr) atomicAdd(&L.J[x.ri[i] * P + sj], -(x.rs[i] * dCr));
                    }
                }
            }
        }
        __syncthreads();
        // dense third-body part, J(i, j) += sum_s Wd[s][i] eff_s[j]: with at
        // most kSetsT sets it is added row-parallel inside the ddY_dcdY
        // transform below; otherwise column-parallel here, one chunk of
        // kMaxSets sets at a time
        if (!fusedDense && lane < nS) {
            const int ns = min(kMaxSets, m.nSets - s0);
            double e[kMaxSets];
#pragma unroll
            for (int s = 0; s < kMaxSets; ++s)
                e[s] = (s < ns) ? m.effSets[(s0 + s) * kLanes + lane] : 0.0;
#pragma unroll 2
            for (int i = 0; i < nS; ++i) {
                double acc = 0.0;
#pragma unroll
                for (int s = 0; s < kMaxSets; ++s) acc += L.Wd[s * kLanes + i] * e[s];
                L.J[i * P + lane] += acc;
            }
        }
        __syncthreads();
    }
    // a fall-off reaction with a non-finite Cf - Cr/Kc makes the reference's
    // dense (zero) concentration derivative NaN for its species rows
    if (__any(poisoned)) {
        const double qnan = __longlong_as_double(0x7ff8000000000000LL);
#pragma unroll 1
        for (int pass = 0; pass < m.nPass; ++pass) {
            const int r = pass * kLanes + lane;
            if (r < m.nR) {
                const Rxn x = load_rxn(m, pass, lane);
                if (x.type > GFC_RATE_THIRD_BODY) {
                    double pw[kTeamWidth];
                    const double Cf = side_powers(x.li, x.lei, x.le, x.nl, L.c, pw);
                    const double Cr = side_powers(x.ri, x.rei, x.re, x.nr, L.c, pw);
                    const double temp = Cf - Cr / exp(ln_kc(x, f, L, nullptr));
                    if (!isfinite(temp)) {
                        for (int i = 0; i < x.nl; ++i)
                            for (int j = 0; j < nS; ++j) L.J[x.li[i] * P + j] = qnan;
                        for (int i = 0; i < x.nr; ++i)
                            for (int j = 0; j < nS; ++j) L.J[x.ri[i] * P + j] = qnan;
                    }
                }
            }
        }
        __syncthreads();
    }
    // derivative: dY/dt and dT/dt
    double d = 0.0;
    if (lane < nS) d = L.dN[lane] * (m.W[lane] / rhoM);
    const double dfdxT = -wave_sum(d * Ha) / CpM;
    // ddY_dcdY (gpuODESystem.H:65-103): lane a transforms row a
    if (lane < nS) {
        double* row = &L.J[lane * P];
        const double WiByrhoM = m.W[lane] / rhoM;
        double ddNdT = row[nS];
        if (fusedDense) {
            // this row's weight of every set; the sets' efficiencies of
            // column j are uniform (scalar loads of the transposed table)
            double w[kSetsT];
#pragma unroll
            for (int q = 0; q < kSetsT; ++q) w[q] = L.Wd[q * kLanes + lane];
#pragma unroll 2
            for (int j = 0; j < nS; ++j) {
                const double* e = m.effSetsT + j * kSetsT;
                double acc = 0.0;
#pragma unroll
                for (int q = 0; q < kSetsT; ++q) acc += w[q] * e[q];
                const double tj = row[j] + acc;
                ddNdT -= tj * L.c[j] * alphavM;
                row[j] = WiByrhoM * tj * L.colA[j] + L.colB[j] * d;
            }
        } else {
#pragma unroll 4
            for (int j = 0; j < nS; ++j) {
                const double tj = row[j];
                ddNdT -= tj * L.c[j] * alphavM;
                row[j] = WiByrhoM * tj * L.colA[j] + L.colB[j] * d;
            }
        }
        row[nS] = WiByrhoM * ddNdT + alphavM * d;
    }
    __syncthreads();
    // ddT_dtdY (:108-126): lane i sums column i
    if (lane < nS) {
        double v = 0.0;
#pragma unroll 4
        for (int j = 0; j < nS; ++j) v -= L.J[j * P + lane] * L.Ha[j];
        v -= Cp * dfdxT;
        L.J[nS * P + lane] = v / CpM;
    }
    // ddT_dtdT (:129-146)
    const double term = (lane < nS) ? d * Cp + L.J[lane * P + nS] * Ha : 0.0;
    const double s = wave_sum(term);
    if (lane == 0) L.J[nS * P + nS] = (-s - dfdxT * dCpMdT) / CpM;
    __syncthreads();
    if (lane == nS) d = dfdxT;
    return d;
}

// --------------------------------------------------------------- LU ----
// Row-per-lane factorisation.  On return `ps` is the step at which this
// lane's row was the pivot, `rd` the reciprocal of its pivot and lane j's
// `plv` the pivot lane of step j.
template <int N>
__device__ __forceinline__ void team_lu(double (&row)[N], int n, int lane, int& ps,
                                        double& rd, int& plv) {
    double amax = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) amax = fmax(amax, fabs(row[k]));
    const double vv = 1.0 / amax;
    int pos = lane;
    bool done = lane >= N;
    ps = lane;
    rd = 1.0;
    plv = lane;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        // candidates: rows not yet pivots (padding rows only once the real
        // columns are done); NaN keys never win (the reference's >=)
        double key = (done || (lane >= n && j < n)) ? -1.0 : vv * fabs(row[j]);
        if (!(key >= 0.0)) key = -1.0;
        const double kmax = wave_max(key);
        // one candidate (the rule case) needs no tie-break reduction; ties go
        // to the largest row position
        const unsigned long long cand = __ballot(key == kmax && !done);
        int pv;
        if (__builtin_popcountll(cand) == 1) {
            pv = __builtin_amdgcn_readfirstlane((int)__builtin_ctzll(cand));
        } else {
            const int Pt = wave_max_i((key == kmax && !done) ? pos : -1);
            pv = __builtin_amdgcn_readfirstlane(
                (int)__builtin_ctzll(__ballot(pos == Pt && !done)));
        }
        const int P = __builtin_amdgcn_readlane(pos, pv);
        if (lane == j) plv = pv;
        if (pos == j) pos = P;
        if (lane == pv) {
            pos = j;
            done = true;
            ps = j;
        }
        {
            // settle this step's per-lane bookkeeping here: left as select
            // chains, every step's lane masks stay live to the end of the
            // factorisation (SGPRs spilled into VGPR lanes)
            int dn = done ? 1 : 0;
            asm volatile("" : "+v"(ps), "+v"(rd), "+v"(plv), "+v"(pos), "+v"(dn));
            done = dn != 0;
        }
        double dj = readlane(row[j], pv);
        if (dj == 0.0) {
            dj = kSmall;
            if (lane == pv) row[j] = kSmall;
        }
        const double rDiag = 1.0 / dj;
        if (lane == pv) rd = rDiag;
        const double l = done ? 0.0 : row[j] * rDiag;
        if (!done) row[j] = l;
        // update in groups of 8 columns; the scheduling barrier keeps the
        // broadcast values of one group in flight at a time
#pragma unroll
        for (int k = j + 1; k < N; ++k) {
            row[k] = fma(-l, readlane(row[k], pv), row[k]);
            if ((k - j) % 8 == 0) __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// Solves (LU) x = b for b held species-parallel; returns x species-parallel.
template <int N>
__device__ __forceinline__ double team_solve(const double (&row)[N], int lane, int ps,
                                             double rd, int plv, double b) {
    // keep the per-step pivot lanes from being hoisted (and held in scalar
    // registers) across the calls of one attempt
    asm volatile("" : "+v"(plv));
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int pv = __builtin_amdgcn_readlane(plv, j);
        const double yj = readlane(b, pv);
        if (ps > j) b = fma(-row[j], yj, b);
    }
    double x = 0.0;
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
        const int pv = __builtin_amdgcn_readlane(plv, j);
        const double xj = readlane(b * rd, pv);
        if (ps < j) b = fma(-row[j], xj, b);
        if (lane == j) x = xj;
    }
    return x;
}

} // namespace

// ---------------------------------------------------------------- kernel --

// Small mechanisms (N <= 16) fit in 256 registers: two wavefronts per SIMD
// hide each other's latency.  Larger ones need the AGPRs as well and run
// one wavefront per SIMD.
template <int N>
__global__ void __launch_bounds__(64, (N <= 16) ? 2 : 1)
team_solve_kernel(SolveArgs A, int* queue) {
    __shared__ TeamLDS<N> L;
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

        double y = (lane < n) ? A.Y0[lane * ld + cell] : 0.0;
        const double y0cell = y;
        const double p = A.Y0[(long)(nS + 1) * ld + cell];
        double dtChem = A.dtChem[cell];
        int status = 0, attempts = 0, accepted = 0;
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

} // namespace gfc
