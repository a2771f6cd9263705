// Wavefront-per-cell chemistry kernel.
//
// One 64-lane wavefront integrates one cell.  The work of gpuFoam's
// per-cell solver (singleCellSolver -> odeSolve -> adaptiveSolve ->
// Rosenbrock step -> jacobian / LU / derivatives; see
// gpu_chemistry/src/gpuOde/*.H) is spread over the lanes:
//   * species-parallel: lane i owns component i of every state and stage
//     vector (y, dydx0, dfdx, k1..k4), in registers;
//   * reaction-parallel: lane l evaluates reactions l, l+64, ...; their
//     contributions to dN/dt and to the Jacobian are accumulated in LDS
//     with FP64 LDS atomics (the reactions are stored grouped by rate type
//     so that a pass of 64 reactions runs one code path);
//   * row-parallel LU: lane i holds row i of -J + I/(gamma*dx) in N
//     registers; Crout with implicit scaling becomes right-looking
//     elimination with the same per-entry operation sequence, the pivot
//     row is broadcast with v_readlane, and the pivot choice (largest
//     scaled value, ties to the last row position, ludecompose.H:9-92)
//     uses a wavefront argmax over (value, position);
//   * the Jacobian stays in LDS across rejected attempts (it only depends
//     on y0), so a retry only re-factorises.
// N is the compile-time row capacity (n = nSpecie+1 <= N <= 64); unused
// rows are identity rows and are skipped with uniform branches.
#include <hip/hip_runtime.h>

#include "kernel_common.hpp"

namespace gfc {

namespace {

constexpr int kLanes = 64;

__device__ __forceinline__ double readlane(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) v = fmax(v, __shfl_xor(v, s));
    return v;
}

__device__ __forceinline__ double uniform(double v) {
    return readlane(v, 0);
}

template <int N>
struct TeamLDS {
    static constexpr int P = (N % 2 == 0) ? N + 1 : N; // odd row pitch
    double J[N * P];
    double c[kLanes];
    double dN[kLanes];
    double Ha[kLanes];
    double colA[kLanes]; // rhoM / W_j
    double colB[kLanes]; // rhoM * v_j
    int pl[kLanes];      // pivot lane of LU step j
};

// Product of concentration powers of one reaction side, in written order
// (lhsPowers + calcCf, gpuReaction.H:348-386).
__device__ __forceinline__ double side_pow(const int* idx, const int* ei,
                                           const double* er, int n,
                                           const double* c) {
    double C = 1.0;
#pragma unroll 1
    for (int q = 0; q < n; ++q) C *= specie_pow(c[idx[q]], ei[q], er[q]);
    return C;
}

// d(product)/dc_j, the product of the other factors times the derivative of
// factor j (calcdCfdcj, gpuReaction.H:388-429)
__device__ __forceinline__ double side_dpow(const int* idx, const int* ei,
                                            const double* er, int n, int j,
                                            const double* c) {
    double v = 1;
#pragma unroll 1
    for (int i = 0; i < n; ++i) {
        const double ci = c[idx[i]];
        if (i == j) v *= er[i] * specie_pow_m1(ci, ei[i], er[i]);
        else v *= specie_pow(ci, ei[i], er[i]);
    }
    return v;
}

// M = sum_i eff_i c_i for the reaction at (pass, lane)
__device__ __forceinline__ double third_body_M(const DeviceMech& m, int pass,
                                               int lane, const double* c) {
    const double* e = m.effP + (long)pass * m.nS * kLanes + lane;
    double M = 0;
    for (int i = 0; i < m.nS; ++i) M += e[(long)i * kLanes] * c[i];
    return M;
}

// ---------------------------------------------------------- derivatives --
// gpuODESystem::derivatives (gpuODESystem.H:164-196).  yl: component of
// this lane; returns dydx component of this lane.
template <int N>
__device__ __forceinline__ double team_derivatives(const DeviceMech& m, double p, double T,
                                   double yl, TeamLDS<N>& L, int lane) {
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
    L.c[lane] = cu * rhoM;
    L.dN[lane] = 0.0;
    __syncthreads();
    const double Tc = fmin(fmax(T, m.Tlow), m.Thigh);
    for (int pass = 0; pass < m.nPass; ++pass) {
        const int r = pass * kLanes + lane;
        if (r >= m.nR) continue;
        const int t = m.type[r];
        const double* k = m.k + 3 * r;
        double kf;
        if (t == GFC_RATE_ARRHENIUS) {
            kf = arrhenius(k, Tc);
        } else {
            const double M = third_body_M(m, pass, lane, L.c);
            if (t == GFC_RATE_THIRD_BODY) {
                kf = M * arrhenius(k, Tc);
            } else {
                const double k0 = arrhenius(k, Tc);
                const double kInf = arrhenius(m.kInf + 3 * r, Tc);
                const double Pr = k0 / kInf * M;
                const double F = (t == GFC_RATE_TROE) ? troe_F(m.troe + 4 * r, Tc, Pr) : 1.0;
                kf = kInf * (Pr / (1.0 + Pr)) * F;
            }
        }
        const double kr = m.irreversible[r] ? 0.0 : kf / reaction_Kc(m, r, Tc);
        const long o = (long)r * kWidth;
        const int nl = m.nl[r], nr = m.nr[r];
        const double Cf = side_pow(m.li + o, m.lei + o, m.le + o, nl, L.c);
        const double Cr = side_pow(m.ri + o, m.rei + o, m.re + o, nr, L.c);
        const double w = kf * Cf - kr * Cr;
#pragma unroll 1
        for (int q = 0; q < nl; ++q) atomicAdd(&L.dN[m.li[o + q]], -(m.ls[o + q] * w));
#pragma unroll 1
        for (int q = 0; q < nr; ++q) atomicAdd(&L.dN[m.ri[o + q]], m.rs[o + q] * w);
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
    L.c[lane] = cu * rhoM;
    L.dN[lane] = 0.0;
    L.Ha[lane] = Ha;
    if (lane < nS) {
        L.colA[lane] = rhoM / m.W[lane];
        L.colB[lane] = rhoM * (1.0 / (p / (m.R[lane] * T)));
    }
    for (int e = lane; e < n * P; e += kLanes) L.J[e] = 0.0;
    __syncthreads();

    for (int pass = 0; pass < m.nPass; ++pass) {
        const int r = pass * kLanes + lane;
        if (r >= m.nR) continue;
        const int t = m.type[r];
        const double* k = m.k + 3 * r;
        const long o = (long)r * kWidth;
        const int nl = m.nl[r], nr = m.nr[r];
        double kf, ddT, ddcScale = 0.0;
        if (t == GFC_RATE_ARRHENIUS) {
            const double Ak = arrhenius(k, T);
            kf = Ak;
            ddT = Ak * (k[1] + k[2] / T) / T;
        } else {
            const double M = third_body_M(m, pass, lane, L.c);
            if (t == GFC_RATE_THIRD_BODY) {
                const double Ak = arrhenius(k, T);
                const double dT = Ak * (k[1] + k[2] / T) / T;
                kf = M * Ak;
                ddT = M * dT;
                ddcScale = Ak;
            } else {
                const double k0 = arrhenius(k, T);
                const double* ki = m.kInf + 3 * r;
                const double kInf = arrhenius(ki, T);
                const double Pr = k0 / kInf * M;
                const double F = (t == GFC_RATE_TROE) ? troe_F(m.troe + 4 * r, T, Pr) : 1.0;
                const double temp = (Pr / (1.0 + Pr)) * F;
                kf = kInf * temp;
                ddT = (kInf * (ki[1] + ki[2] / T) / T) * temp;
            }
        }
        const double Cf = side_pow(m.li + o, m.lei + o, m.le + o, nl, L.c);
        double dwdT = Cf * ddT;
        double w = Cf * kf;
        double Cr = 0.0, Kc = 1.0, kr = 0.0;
        const bool rev = !m.irreversible[r];
        if (rev) {
            double dKc;
            reaction_Kc_dKc(m, r, T, Kc, dKc);
            kr = kf / Kc;
            Cr = side_pow(m.ri + o, m.rei + o, m.re + o, nr, L.c);
            const double dkrdT = ddT / Kc - kr * dKc;
            dwdT -= Cr * dkrdT;
            w -= Cr * kr;
        }
        // dN/dt (the reference's dfdx)
#pragma unroll 1
        for (int i = 0; i < nl; ++i) atomicAdd(&L.dN[m.li[o + i]], -(m.ls[o + i] * w));
#pragma unroll 1
        for (int i = 0; i < nr; ++i) atomicAdd(&L.dN[m.ri[o + i]], m.rs[o + i] * w);
        // jac_dCfdcj_contribution (gpuReaction.H:547-569)
#pragma unroll 1
        for (int j = 0; j < nl; ++j) {
            const double dCf = side_dpow(m.li + o, m.lei + o, m.le + o, nl, j, L.c) * kf;
            const int sj = m.li[o + j];
#pragma unroll 1
            for (int i = 0; i < nl; ++i)
                atomicAdd(&L.J[m.li[o + i] * P + sj], -(m.ls[o + i] * dCf));
#pragma unroll 1
            for (int i = 0; i < nr; ++i)
                atomicAdd(&L.J[m.ri[o + i] * P + sj], m.rs[o + i] * dCf);
        }
        // jac_dCrdcj_contribution (:571-593)
        if (rev) {
#pragma unroll 1
            for (int j = 0; j < nr; ++j) {
                const double dCr = side_dpow(m.ri + o, m.rei + o, m.re + o, nr, j, L.c) * kr;
                const int sj = m.ri[o + j];
#pragma unroll 1
                for (int i = 0; i < nl; ++i)
                    atomicAdd(&L.J[m.li[o + i] * P + sj], m.ls[o + i] * dCr);
#pragma unroll 1
                for (int i = 0; i < nr; ++i)
                    atomicAdd(&L.J[m.ri[o + i] * P + sj], -(m.rs[o + i] * dCr));
            }
        }
        // jac_dCdT_contribution (:595-612)
#pragma unroll 1
        for (int i = 0; i < nl; ++i) atomicAdd(&L.J[m.li[o + i] * P + nS], -(m.ls[o + i] * dwdT));
#pragma unroll 1
        for (int i = 0; i < nr; ++i) atomicAdd(&L.J[m.ri[o + i] * P + nS], m.rs[o + i] * dwdT);
        // dense concentration derivative of third-body rates
        if (t != GFC_RATE_ARRHENIUS) {
            const double temp = (Cf - Cr / Kc);
            const double* e = m.effP + (long)pass * nS * kLanes + lane;
            if (t == GFC_RATE_THIRD_BODY || !isfinite(temp)) {
                const double scale = (t == GFC_RATE_THIRD_BODY) ? ddcScale : 0.0;
#pragma unroll 1
                for (int j = 0; j < nS; ++j) {
                    const double ddc = (t == GFC_RATE_THIRD_BODY)
                                           ? (e[(long)j * kLanes] * scale) * temp
                                           : 0.0 * temp;
#pragma unroll 1
                    for (int i = 0; i < nl; ++i)
                        atomicAdd(&L.J[m.li[o + i] * P + j], -(m.ls[o + i] * ddc));
#pragma unroll 1
                    for (int i = 0; i < nr; ++i)
                        atomicAdd(&L.J[m.ri[o + i] * P + j], m.rs[o + i] * ddc);
                }
            }
        }
    }
    __syncthreads();
    // derivative: dY/dt and dT/dt
    double d = 0.0;
    if (lane < nS) d = L.dN[lane] * (m.W[lane] / rhoM);
    const double dfdxT = -wave_sum(d * Ha) / CpM;
    // ddY_dcdY (gpuODESystem.H:65-103): lane a transforms row a
    if (lane < nS) {
        double* row = &L.J[lane * P];
        const double WiByrhoM = m.W[lane] / rhoM;
        double ddNdT = row[nS];
        for (int j = 0; j < nS; ++j) {
            const double tj = row[j];
            ddNdT -= tj * L.c[j] * alphavM;
            row[j] = WiByrhoM * tj * L.colA[j] + L.colB[j] * d;
        }
        row[nS] = WiByrhoM * ddNdT + alphavM * d;
    }
    __syncthreads();
    // ddT_dtdY (:108-126): lane i sums column i
    if (lane < nS) {
        double v = 0.0;
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
// Factorises the rows held in registers (row-per-lane), writing the pivot
// lane of each step to L.pl.  `ps` receives the step at which this lane
// was the pivot and `rd` the reciprocal of its pivot.
template <int N>
__device__ __forceinline__ void team_lu(double (&row)[N], int n, int lane,
                                        TeamLDS<N>& L, int& ps, double& rd) {
    double amax = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) amax = fmax(amax, fabs(row[k]));
    // the reference keeps the first strict maximum; the maximum value is
    // the same either way
    const double vv = 1.0 / amax;
    int pos = lane;
    bool done = lane >= n;
    ps = lane;
    rd = 1.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (j >= n) continue;
        double key = done ? -1.0 : vv * fabs(row[j]);
        if (!(key >= 0.0)) key = -1.0; // NaN never wins (the reference's >=)
        int tag = (pos << 8) | lane;
#pragma unroll
        for (int s = 32; s >= 1; s >>= 1) {
            const double ok = __shfl_xor(key, s);
            const int ot = __shfl_xor(tag, s);
            if (ok > key || (ok == key && ot > tag)) {
                key = ok;
                tag = ot;
            }
        }
        const int t = __builtin_amdgcn_readfirstlane(tag);
        const int pv = t & 0xff;      // pivot lane
        const int P = t >> 8;         // its position (the reference's iMax)
        if (pos == j) pos = P;
        if (lane == pv) {
            pos = j;
            done = true;
            ps = j;
        }
        if (lane == 0) L.pl[j] = pv;
        double dj = readlane(row[j], pv);
        if (dj == 0.0) {
            dj = kSmall;
            if (lane == pv) row[j] = kSmall;
        }
        const double rDiag = 1.0 / dj;
        if (lane == pv) rd = rDiag;
        if (!done) {
            const double l = row[j] * rDiag;
            row[j] = l;
#pragma unroll
            for (int k = j + 1; k < N; ++k) {
                if (k < n) row[k] -= l * readlane(row[k], pv);
            }
        }
    }
}

// Solves (LU) x = b for the right-hand side held species-parallel (lane i:
// component i); returns x species-parallel.
template <int N>
__device__ __forceinline__ double team_solve(const double (&row)[N], int n,
                                             int lane, const TeamLDS<N>& L,
                                             int ps, double rd, double b) {
    // forward: lane p_j's b is final at step j
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (j >= n) continue;
        const int pv = __builtin_amdgcn_readfirstlane(L.pl[j]);
        const double yj = readlane(b, pv);
        if (ps > j) b -= row[j] * yj;
    }
    double x = 0.0;
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
        if (j >= n) continue;
        const int pv = __builtin_amdgcn_readfirstlane(L.pl[j]);
        const double xj = readlane(b * rd, pv);
        if (ps < j) b -= row[j] * xj;
        if (lane == j) x = xj;
    }
    return x;
}

} // namespace

// ---------------------------------------------------------------- kernel --

template <int N>
__global__ void __launch_bounds__(64)
team_solve_kernel(SolveArgs A, int* queue) {
    __shared__ TeamLDS<N> L;
    const int lane = threadIdx.x;
    const DeviceMech& m = A.m;
    const int nS = m.nS, n = nS + 1;
    const long ld = A.ldY;
    const RosTableau& tb = A.tab;

    int cell = queue ? 0 : blockIdx.x;
    for (;;) {
        if (queue) {
            int c = 0;
            if (lane == 0) c = atomicAdd(queue, 1);
            cell = __builtin_amdgcn_readfirstlane(c);
        }
        if (cell >= A.nCells) break;

        double y = (lane < n) ? A.Y0[lane * ld + cell] : 0.0;
        const double y0cell = y;
        const double p = A.Y0[(long)(nS + 1) * ld + cell];
        double dtChem = A.dtChem[cell];
        int status = 0, attempts = 0, accepted = 0;
        double T = uniform(readlane(y, nS));
        if (T > A.Treact) {
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
                    T = readlane(y, nS);
                    const double dydx0 = team_derivatives<N>(m, p, T, y, L, lane);
                    const double dfdx = team_jacobian<N>(m, p, T, y, L, lane);
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
                        int ps;
                        double rd;
                        team_lu<N>(row, n, lane, L, ps, rd);
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
                            const double kx = team_solve<N>(row, n, lane, L, ps, rd, rhs);
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
                        err = uniform(wave_max(ne));
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
        if (A.Ynew) {
            if (lane < n) A.Ynew[lane * ld + cell] = y;
            if (lane == n) A.Ynew[(long)n * ld + cell] = p;
        }
        if (lane == 0) {
            if (A.dtMinBits) atomic_min_positive(A.dtMinBits, dtChem);
            A.dtChem[cell] = fmin(dtChem, A.deltaTChemMax);
            if (A.stats) {
                A.stats[cell] = status;
                A.stats[ld + cell] = attempts;
                A.stats[2 * ld + cell] = accepted;
            }
        }
        if (!queue) break;
        __syncthreads();
    }
}

template __global__ void team_solve_kernel<16>(SolveArgs, int*);
template __global__ void team_solve_kernel<56>(SolveArgs, int*);
template __global__ void team_solve_kernel<64>(SolveArgs, int*);

} // namespace gfc
