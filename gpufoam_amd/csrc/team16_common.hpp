// Building blocks of the 16-lane team kernels: a wavefront holds four
// teams of 16 lanes (one DPP row each), and each team integrates one cell.
// Used by the four-cells-per-wavefront solve (team16_kernel.hip) and its
// discrete adjoint (team16_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "team_common.hpp"

namespace gfc {
namespace t16 {

using namespace wave;

// Accepted-step record of the discrete adjoint.  Every per-cell array is
// wave-blocked as in the one-lane-per-cell kernels: the 64 cells of a block
// share one contiguous slab, element e of a cell at [e*64] from its slot.
constexpr long kRecStride = 64;
__device__ __forceinline__ double* rec_slot(double* base, long cell, long per) {
    return base + (cell / kRecStride) * (per * kRecStride) + (cell % kRecStride);
}
struct T16Record {
    double* ck;     // every K-th accepted step's initial state, (slot*n + e)
    double* hs;     // every accepted step's size
    int* nAcc;      // accepted steps per cell
    double* yEnd;   // final state before the non-negativity clamp
    double* y0c;    // initial state (Y, T, p)
    long nck;       // checkpoint slots per cell
    long maxSteps;  // step-size slots per cell
    int K;          // checkpoint interval
};

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

// Member S's value to its whole team by DPP row_newbcast (a VALU move, no
// LDS round trip).  Every lane of the team must be active.
template <int S>
__device__ __forceinline__ int team_bcast_ci(int v) {
    static_assert(S >= 0 && S < 16, "a member of the team");
    return __builtin_amdgcn_update_dpp(0, v, 0x150 + S, 0xF, 0xF, false);
}
template <int S>
__device__ __forceinline__ double team_bcast_c(double v) {
    const int lo = team_bcast_ci<S>(__double2loint(v));
    const int hi = team_bcast_ci<S>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}
// The same for a member index known at compile time once the caller's loop
// is unrolled (the switch folds to one case).
__device__ __forceinline__ int team_bcast_ki(int v, int j) {
    switch (j) {
    case 0: return team_bcast_ci<0>(v);
    case 1: return team_bcast_ci<1>(v);
    case 2: return team_bcast_ci<2>(v);
    case 3: return team_bcast_ci<3>(v);
    case 4: return team_bcast_ci<4>(v);
    case 5: return team_bcast_ci<5>(v);
    case 6: return team_bcast_ci<6>(v);
    case 7: return team_bcast_ci<7>(v);
    case 8: return team_bcast_ci<8>(v);
    case 9: return team_bcast_ci<9>(v);
    case 10: return team_bcast_ci<10>(v);
    case 11: return team_bcast_ci<11>(v);
    case 12: return team_bcast_ci<12>(v);
    case 13: return team_bcast_ci<13>(v);
    case 14: return team_bcast_ci<14>(v);
    default: return team_bcast_ci<15>(v);
    }
}
__device__ __forceinline__ double team_bcast_k(double v, int j) {
    const int lo = team_bcast_ki(__double2loint(v), j);
    const int hi = team_bcast_ki(__double2hiint(v), j);
    return __hiloint2double(hi, lo);
}

// Sum over the team in a fixed tree (row_shr 1, 2, 4, 8); every member gets
// the value of the row's last lane.
__device__ __forceinline__ double team_sum(double v, int lane) {
    v += dpp<0x111>(0.0, v);
    v += dpp<0x112>(0.0, v);
    v += dpp<0x114>(0.0, v);
    v += dpp<0x118>(0.0, v);
    return team_bcast_c<15>(v);
}

__device__ __forceinline__ double team_max(double v, int lane) {
    const double lo = -HUGE_VAL;
    v = fmax(v, dpp<0x111>(lo, v));
    v = fmax(v, dpp<0x112>(lo, v));
    v = fmax(v, dpp<0x114>(lo, v));
    v = fmax(v, dpp<0x118>(lo, v));
    return team_bcast_c<15>(v);
}

__device__ __forceinline__ int team_max_i(int v, int lane) {
    const int lo = -2147483647 - 1;
    v = max(v, dppi<0x111>(lo, v));
    v = max(v, dppi<0x112>(lo, v));
    v = max(v, dppi<0x114>(lo, v));
    v = max(v, dppi<0x118>(lo, v));
    return team_bcast_ci<15>(v);
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
// are added row-parallel in the transform, as in the 64-lane kernel.  With
// RAW, two quantities of each species row before the transform (dN_i/dc_j
// with the dense third-body part, and dN_i/dT) are also stored:
// raw[i] = sum_j (dN_i/dc_j) c_j, summed in column order, and
// raw[16 + i] = dN_i/dT.
template <int N, bool RAW = false>
__device__ __forceinline__ double t16_jacobian(const DeviceMech& m, double p, double T,
                                               double yl, Team16LDS<N>& L, int lane, int mem,
                                               double* raw = nullptr) {
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
        double Cr = 0.0, lnKc = 0.0, kr = 0.0;
        if (!x.irr) {
            double dKc;
            lnKc = ln_kc(x, f, L, &dKc);
            const double rKc = exp(-lnKc);
            kr = kf * rKc;
            Cr = side_powers(x.ri, x.rei, x.re, x.nr, L.c, pwr);
            const double dkrdT = ddT * rKc - kr * dKc;
            dwdT -= Cr * dkrdT;
            w -= Cr * kr;
        }
        // Cf - Cr/Kc: used by the third-body and fall-off rates only, so the
        // Arrhenius passes skip its exponential
        double temp = Cf;
        if (x.type != GFC_RATE_ARRHENIUS) temp = Cf - Cr / exp(lnKc);
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
        if constexpr (RAW) raw[kTW + mem] = ddNdT;
        double sc = 0.0;
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
            if constexpr (RAW) sc += tj * L.c[j];
            ddNdT -= tj * L.c[j] * alphavM;
            row[j] = WiByrhoM * tj * L.colA[j] + L.colB[j] * d;
        }
        row[nS] = WiByrhoM * ddNdT + alphavM * d;
        if constexpr (RAW) raw[mem] = sc;
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

// Factors into pivot order: member m takes the row (and the reciprocal
// pivot) at pivot position m from the member holding it (plv), so that the
// solves broadcast from fixed members (DPP) instead of data-dependent ones.
template <int N>
__device__ __forceinline__ void t16_pivot_order(double (&row)[N], double& rd, int lane, int plv) {
#pragma unroll
    for (int k = 0; k < N; ++k) row[k] = team_bcast(row[k], lane, plv);
    rd = team_bcast(rd, lane, plv);
}

// Solves (LU) x = b with the factors in pivot order (t16_pivot_order); b
// and x species-parallel.  The same operations as the solve over the rows
// in place, member m now standing for the row at pivot position m.
template <int N>
__device__ __forceinline__ double t16_solve(const double (&row)[N], int lane, int mem, double rd,
                                            int plv, double b) {
    double bp = team_bcast(b, lane, plv); // the right-hand side of position mem
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const double yj = team_bcast_k(bp, j);
        if (mem > j) bp = fma(-row[j], yj, bp);
    }
    double x = 0.0;
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
        const double xj = team_bcast_k(bp * rd, j);
        if (mem < j) bp = fma(-row[j], xj, bp);
        if (mem == j) x = xj;
    }
    return x;
}

} // namespace t16
} // namespace gfc
