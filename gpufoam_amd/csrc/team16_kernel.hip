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

#include "team_common.hpp"

namespace gfc {

namespace {

using namespace wave;

constexpr int kTW = 16;        // lanes per team
constexpr int kSets16 = 16;    // efficiency sets held per team

template <int N>
struct Team16LDS {
    static constexpr int P = (N % 2 == 0) ? N + 1 : N; // odd row pitch
    double J[N * P];
    double c[kTW], dN[kTW], Ha[kTW], colA[kTW], colB[kTW], M[kSets16];
    double gL[kTW], gH[kTW], hL[kTW], hH[kTW];
    double Wd[kSets16 * kTW];  // dense Jacobian weights per set, specie
};

// Makes the wavefront's LDS writes visible to its own later reads (one
// wavefront per workgroup: no hardware barrier is needed).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double team_bcast(double v, int lane, int src) {
    const int l = (lane & ~15) | src;
    const int lo = __shfl(__double2loint(v), l);
    const int hi = __shfl(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ int team_bcast_i(int v, int lane, int src) {
    return __shfl(v, (lane & ~15) | src);
}

// Sum over the team in a fixed tree (row_shr 1, 2, 4, 8); every member gets
// the value of the row's last lane.
__device__ __forceinline__ double team_sum(double v, int lane) {
    v += dpp<0x111>(0.0, v);
    v += dpp<0x112>(0.0, v);
    v += dpp<0x114>(0.0, v);
    v += dpp<0x118>(0.0, v);
    return team_bcast(v, lane, 15);
}

__device__ __forceinline__ double team_max(double v, int lane) {
    const double lo = -HUGE_VAL;
    v = fmax(v, dpp<0x111>(lo, v));
    v = fmax(v, dpp<0x112>(lo, v));
    v = fmax(v, dpp<0x114>(lo, v));
    v = fmax(v, dpp<0x118>(lo, v));
    return team_bcast(v, lane, 15);
}

__device__ __forceinline__ int team_max_i(int v, int lane) {
    const int lo = -2147483647 - 1;
    v = max(v, dppi<0x111>(lo, v));
    v = max(v, dppi<0x112>(lo, v));
    v = max(v, dppi<0x114>(lo, v));
    v = max(v, dppi<0x118>(lo, v));
    return team_bcast_i(v, lane, 15);
}

template <int N>
__device__ __forceinline__ void third_body16(const DeviceMech& m, Team16LDS<N>& L, int mem) {
    wave_sync();
    const int nS = m.nS;
    for (int s = mem; s < m.nSets; s += kTW) {
        const double* e = m.effSets + s * kLanes;
        double acc = 0.0;
#pragma unroll 4
        for (int i = 0; i < nS; ++i) acc += e[i] * L.c[i];
        L.M[s] = acc;
    }
}

template <int N, bool ENTHALPY>
__device__ __forceinline__ void gibbs16(const DeviceMech& m, Team16LDS<N>& L, const TFun& f,
                                        int mem) {
    if (mem < m.nS) {
        const double W = m.W[mem];
        const double* lo = m.low + 7 * mem;
        const double* hi = m.high + 7 * mem;
        L.gL[mem] = W * poly_Gstd(lo, f.T, f.logT);
        L.gH[mem] = W * poly_Gstd(hi, f.T, f.logT);
        if (ENTHALPY) {
            L.hL[mem] = W * poly_Ha(lo, f.T);
            L.hH[mem] = W * poly_Ha(hi, f.T);
        }
    }
}

// gpuODESystem::derivatives (gpuODESystem.H:164-196) for the team's cell.
template <int N>
__device__ __forceinline__ double t16_derivatives(const DeviceMech& m, double p, double T,
                                                  double yl, Team16LDS<N>& L, int lane,
                                                  int mem) {
    const int nS = m.nS;
    double rhoPart = 0.0, cpPart = 0.0, Ha = 0.0, cu = 0.0;
    if (mem < nS) {
        const double yi = fmax(yl, 0.0);
        const double rhoi = p / (m.R[mem] * T);
        const double* a = coeffs_of(m.high + 7 * mem, m.low + 7 * mem, m.Tcommon[mem], T);
        rhoPart = yi / rhoi;
        cpPart = yi * poly_Cp(a, T);
        Ha = poly_Ha(a, T);
        cu = yi / m.W[mem];
    }
    const double rhoM = 1.0 / team_sum(rhoPart, lane);
    const double CpM = team_sum(cpPart, lane);
    L.c[mem] = cu * rhoM;
    L.dN[mem] = 0.0;
    third_body16<N>(m, L, mem);
    const TFun f = tfun(fmin(fmax(T, m.Tlow), m.Thigh));
    gibbs16<N, false>(m, L, f, mem);
    wave_sync();
    const int nPass = (m.nR + kTW - 1) / kTW;
#pragma unroll 1
    for (int pass = 0; pass < nPass; ++pass) {
        const int r = pass * kTW + mem;
        if (r < m.nR) {
            const Rxn x = load_rxn(m, r >> 6, r & 63);
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
    wave_sync();
    double d = 0.0;
    if (mem < nS) d = L.dN[mem] * (m.W[mem] / rhoM);
    const double dT = -team_sum(d * Ha, lane) / CpM;
    if (mem == nS) d = dT;
    return d;
}

// gpuODESystem::jacobian (gpuODESystem.H:198-281) for the team's cell:
// leaves J in L.J and returns this member's component of the derivative
// computed alongside (dfdx).  The mechanism's efficiency sets (at most 16)
// are added row-parallel in the transform, as in the 64-lane kernel.
template <int N>
__device__ __forceinline__ double t16_jacobian(const DeviceMech& m, double p, double T,
                                               double yl, Team16LDS<N>& L, int lane, int mem) {
    constexpr int P = Team16LDS<N>::P;
    const int nS = m.nS, n = nS + 1;
    double rhoPart = 0, cpPart = 0, dcpPart = 0, avPart = 0, Ha = 0, Cp = 0, cu = 0;
    if (mem < nS) {
        const double yi = fmax(yl, 0.0);
        const double rhoi = p / (m.R[mem] * T);
        const double vi = 1.0 / rhoi;
        const double* a = coeffs_of(m.high + 7 * mem, m.low + 7 * mem, m.Tcommon[mem], T);
        Cp = poly_Cp(a, T);
        Ha = poly_Ha(a, T);
        rhoPart = yi / rhoi;
        cpPart = yi * Cp;
        dcpPart = yi * poly_dCpdT(a, T);
        avPart = yi * vi * (1.0 / T);
        cu = yi / m.W[mem];
    }
    const double rhoM = 1.0 / team_sum(rhoPart, lane);
    const double CpM = team_sum(cpPart, lane);
    const double dCpMdT = team_sum(dcpPart, lane);
    const double alphavM = team_sum(avPart, lane) * rhoM;
    L.c[mem] = cu * rhoM;
    L.dN[mem] = 0.0;
    L.Ha[mem] = Ha;
    if (mem < nS) {
        L.colA[mem] = rhoM / m.W[mem];
        L.colB[mem] = rhoM * (1.0 / (p / (m.R[mem] * T)));
    }
    third_body16<N>(m, L, mem);
    const TFun f = tfun(T);
    gibbs16<N, true>(m, L, f, mem);
    for (int e = mem; e < n * P; e += kTW) L.J[e] = 0.0;
    for (int e = mem; e < kSets16 * kTW; e += kTW) L.Wd[e] = 0.0;
    wave_sync();
    bool poisoned = false;
    const int nPass = (m.nR + kTW - 1) / kTW;
#pragma unroll 1
    for (int pass = 0; pass < nPass; ++pass) {
        const int r = pass * kTW + mem;
        if (r >= m.nR) continue;
        const Rxn x = load_rxn(m, r >> 6, r & 63);
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
        if (x.type == GFC_RATE_THIRD_BODY) {
            const double g = kArr * temp;
            double* Wd = &L.Wd[x.set * kTW];
#pragma unroll
            for (int i = 0; i < kTeamWidth; ++i) {
                if (i < x.nl) atomicAdd(&Wd[x.li[i]], -(x.ls[i] * g));
                if (i < x.nr) atomicAdd(&Wd[x.ri[i]], x.rs[i] * g);
            }
        }
        if (x.type > GFC_RATE_THIRD_BODY && !isfinite(temp)) poisoned = true;
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
    wave_sync();
    // a fall-off reaction with a non-finite Cf - Cr/Kc makes the reference's
    // dense (zero) concentration derivative NaN for its species rows
    const bool anyPoison = team_max_i(poisoned ? 1 : 0, lane) != 0;
    if (anyPoison) {
        const double qnan = __longlong_as_double(0x7ff8000000000000LL);
#pragma unroll 1
        for (int pass = 0; pass < nPass; ++pass) {
            const int r = pass * kTW + mem;
            if (r < m.nR) {
                const Rxn x = load_rxn(m, r >> 6, r & 63);
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
        wave_sync();
    }
    double d = 0.0;
    if (mem < nS) d = L.dN[mem] * (m.W[mem] / rhoM);
    const double dfdxT = -team_sum(d * Ha, lane) / CpM;
    // ddY_dcdY (gpuODESystem.H:65-103): member a transforms row a, the
    // dense third-body part added from the sets' weights
    if (mem < nS) {
        double* row = &L.J[mem * P];
        const double WiByrhoM = m.W[mem] / rhoM;
        double ddNdT = row[nS];
        double w[kSetsT];
#pragma unroll
        for (int q = 0; q < kSetsT; ++q) w[q] = L.Wd[q * kTW + mem];
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
        row[nS] = WiByrhoM * ddNdT + alphavM * d;
    }
    wave_sync();
    // ddT_dtdY (:108-126): member i sums column i
    if (mem < nS) {
        double v = 0.0;
#pragma unroll 4
        for (int j = 0; j < nS; ++j) v -= L.J[j * P + mem] * L.Ha[j];
        v -= Cp * dfdxT;
        L.J[nS * P + mem] = v / CpM;
    }
    // ddT_dtdT (:129-146)
    const double term = (mem < nS) ? d * Cp + L.J[mem * P + nS] * Ha : 0.0;
    const double s = team_sum(term, lane);
    if (mem == 0) L.J[nS * P + nS] = (-s - dfdxT * dCpMdT) / CpM;
    wave_sync();
    if (mem == nS) d = dfdxT;
    return d;
}

// Row-per-member factorisation (as team_lu in team_kernel.hip, with the
// pivot information broadcast inside the team).
template <int N>
__device__ __forceinline__ void t16_lu(double (&row)[N], int n, int lane, int mem, int& ps,
                                       double& rd, int& plv) {
    double amax = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) amax = fmax(amax, fabs(row[k]));
    const double vv = 1.0 / amax;
    int pos = mem;
    bool done = mem >= N;
    ps = mem;
    rd = 1.0;
    plv = mem;
    const int tshift = lane & ~15;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double key = (done || (mem >= n && j < n)) ? -1.0 : vv * fabs(row[j]);
        if (!(key >= 0.0)) key = -1.0;
        const double kmax = team_max(key, lane);
        const unsigned cand =
            (unsigned)((__ballot(key == kmax && !done) >> tshift) & 0xffffull);
        int pv;
        if (__builtin_popcount(cand) == 1) {
            pv = __builtin_ctz(cand);
        } else {
            const int Pt = team_max_i((key == kmax && !done) ? pos : -1, lane);
            const unsigned at =
                (unsigned)((__ballot(pos == Pt && !done) >> tshift) & 0xffffull);
            pv = __builtin_ctz(at);
        }
        const int P = team_bcast_i(pos, lane, pv);
        if (mem == j) plv = pv;
        if (pos == j) pos = P;
        if (mem == pv) {
            pos = j;
            done = true;
            ps = j;
        }
        double dj = team_bcast(row[j], lane, pv);
        if (dj == 0.0) {
            dj = kSmall;
            if (mem == pv) row[j] = kSmall;
        }
        const double rDiag = 1.0 / dj;
        if (mem == pv) rd = rDiag;
        const double l = done ? 0.0 : row[j] * rDiag;
        if (!done) row[j] = l;
#pragma unroll
        for (int k = j + 1; k < N; ++k) row[k] = fma(-l, team_bcast(row[k], lane, pv), row[k]);
    }
}

template <int N>
__device__ __forceinline__ double t16_solve(const double (&row)[N], int lane, int mem, int ps,
                                            double rd, int plv, double b) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int pv = team_bcast_i(plv, lane, j);
        const double yj = team_bcast(b, lane, pv);
        if (ps > j) b = fma(-row[j], yj, b);
    }
    double x = 0.0;
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
        const int pv = team_bcast_i(plv, lane, j);
        const double xj = team_bcast(b * rd, lane, pv);
        if (ps < j) b = fma(-row[j], xj, b);
        if (mem == j) x = xj;
    }
    return x;
}

} // namespace

// ---------------------------------------------------------------- kernel --

template <int N>
__global__ void __launch_bounds__(64)
team16_solve_kernel(SolveArgs A, int* queue) {
    __shared__ Team16LDS<N> Ls[kLanes / kTW];
    const int lane = threadIdx.x;
    const int team = lane >> 4, mem = lane & 15;
    Team16LDS<N>& L = Ls[team];
    const DeviceMech& m = A.m;
    const int nS = m.nS, n = nS + 1;
    const long ld = A.ldY;
    const RosTableau& tb = A.tab;

    for (;;) {
        int c = 0;
        if (mem == 0) c = atomicAdd(queue, 1);
        const int cell = team_bcast_i(c, lane, 0);
        if (cell >= A.nCells) break;

        double y = (mem < n) ? A.Y0[mem * ld + cell] : 0.0;
        const double y0cell = y;
        const double p = A.Y0[(long)(nS + 1) * ld + cell];
        double dtChem = A.dtChem[cell];
        int status = 0, attempts = 0, accepted = 0;
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
                            const double kx = t16_solve<N>(row, lane, mem, ps, rd, plv, rhs);
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

template __global__ void team16_solve_kernel<11>(SolveArgs, int*);
template __global__ void team16_solve_kernel<16>(SolveArgs, int*);

} // namespace gfc
