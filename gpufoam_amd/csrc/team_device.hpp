// Device functions of the wavefront-per-cell kernels (team_kernel.hip: the
// solve; wave_adjoint.hip: its recording form and the discrete adjoint):
// one 64-lane wavefront per cell, lane i owning component i, reactions
// evaluated 64 per pass, the Jacobian assembled in LDS and factorised in
// registers (one row per lane).
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

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
// the derivative computed alongside (the reference's "dfdx").  With RAW,
// two quantities of each species row before the transform (dN_i/dc_j with
// the dense third-body part, and dN_i/dT) are also stored:
// raw[i] = sum_j (dN_i/dc_j) c_j, summed in column order, and
// raw[64 + i] = dN_i/dT.
template <int N, bool RAW = false>
__device__ __forceinline__ double team_jacobian(const DeviceMech& m, double p, double T,
                                             double yl, TeamLDS<N>& L, int lane,
                                             double* raw = nullptr) {
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
        double sumc = 0.0;
        if constexpr (RAW) raw[kLanes + lane] = ddNdT;
        if (fusedDense) {
            // this row's weight of every set; the sets' efficiencies of
            // column j are uniform.  Column j+1's efficiencies are in flight
            // while column j is summed (two partial sums, even and odd sets)
            double w[kSetsT];
#pragma unroll
            for (int q = 0; q < kSetsT; ++q) w[q] = L.Wd[q * kLanes + lane];
            double en[kSetsT];
            dense_row_load(m.effSetsT, en);
#pragma unroll 1
            for (int j = 0; j < nS; ++j) {
                double e[kSetsT];
#pragma unroll
                for (int q = 0; q < kSetsT; ++q) e[q] = en[q];
                dense_row_load(m.effSetsT + min(j + 1, nS - 1) * kSetsT, en);
                const double acc = dense_dot(w, e);
                const double tj = row[j] + acc;
                if constexpr (RAW) sumc += tj * L.c[j];
                ddNdT -= tj * L.c[j] * alphavM;
                row[j] = WiByrhoM * tj * L.colA[j] + L.colB[j] * d;
            }
        } else {
#pragma unroll 4
            for (int j = 0; j < nS; ++j) {
                const double tj = row[j];
                if constexpr (RAW) sumc += tj * L.c[j];
                ddNdT -= tj * L.c[j] * alphavM;
                row[j] = WiByrhoM * tj * L.colA[j] + L.colB[j] * d;
            }
        }
        row[nS] = WiByrhoM * ddNdT + alphavM * d;
        if constexpr (RAW) raw[lane] = sumc;
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
//
// Look-ahead: step j updates column j+1 first and runs the DPP reduction
// of column j+1's pivot search (one stage per group of eight columns)
// between the groups of its remaining column updates, so the search's
// latency overlaps the trailing update.  The arithmetic of every entry and
// every comparison is that of a plain column-by-column elimination.

// Candidate key of a row for the pivot of column j: scaled magnitude, -1
// for rows already pivots (and padding rows while real columns remain);
// NaN keys never win (the reference's >=).
__device__ __forceinline__ double lu_key(double a, double vv, bool done, bool pad) {
    double key = (done || pad) ? -1.0 : vv * fabs(a);
    if (!(key >= 0.0)) key = -1.0;
    return key;
}

// Stage s of wave_max (row_shr 1, 2, 4, 8, then row_bcast 15 and 31).
template <int S>
__device__ __forceinline__ double wave_max_stage(double v) {
    const double lo = -HUGE_VAL;
    if (S == 0) return fmax(v, dpp<0x111>(lo, v));
    if (S == 1) return fmax(v, dpp<0x112>(lo, v));
    if (S == 2) return fmax(v, dpp<0x114>(lo, v));
    if (S == 3) return fmax(v, dpp<0x118>(lo, v));
    if (S == 4) return fmax(v, dpp<0x142, 0xa>(lo, v));
    return fmax(v, dpp<0x143, 0xc>(lo, v));
}

struct TeamLU {
    double vv;
    int pos, ps, plv;
    bool done;
    double rd;
    double key, kmax; // column j's keys and their maximum (found by step j-1)
};

// Step J: pivot of column J, multipliers, and the update of the columns
// right of J with column J+1's search interleaved.  One instantiation per
// step keeps every register index constant.
template <int N, int J>
__device__ __forceinline__ void team_lu_step(double (&row)[N], int n, int lane, TeamLU& S) {
    // pivot of column J from its reduced key: one candidate (the rule
    // case) needs no tie-break reduction; ties go to the largest row
    // position
    const unsigned long long cand = __ballot(S.key == S.kmax && !S.done);
    int pv;
    if (__builtin_popcountll(cand) == 1) {
        pv = __builtin_amdgcn_readfirstlane((int)__builtin_ctzll(cand));
    } else {
        const int Pt = wave_max_i((S.key == S.kmax && !S.done) ? S.pos : -1);
        pv = __builtin_amdgcn_readfirstlane(
            (int)__builtin_ctzll(__ballot(S.pos == Pt && !S.done)));
    }
    const int P = __builtin_amdgcn_readlane(S.pos, pv);
    if (lane == J) S.plv = pv;
    if (S.pos == J) S.pos = P;
    if (lane == pv) {
        S.pos = J;
        S.done = true;
        S.ps = J;
    }
    {
        // settle this step's per-lane bookkeeping here: left as select
        // chains, every step's lane masks stay live to the end of the
        // factorisation (SGPRs spilled into VGPR lanes)
        int dn = S.done ? 1 : 0;
        asm volatile("" : "+v"(S.ps), "+v"(S.rd), "+v"(S.plv), "+v"(S.pos), "+v"(dn));
        S.done = dn != 0;
    }
    double dj = readlane(row[J], pv);
    if (dj == 0.0) {
        dj = kSmall;
        if (lane == pv) row[J] = kSmall;
    }
    const double rDiag = 1.0 / dj;
    if (lane == pv) S.rd = rDiag;
    const double l = S.done ? 0.0 : row[J] * rDiag;
    if (!S.done) row[J] = l;
    if (J + 1 < N) {
        constexpr int J1 = J + 1 < N ? J + 1 : J;
        // column J+1 first, then its pivot search between the groups of
        // eight of the remaining updates
        row[J1] = fma(-l, readlane(row[J1], pv), row[J1]);
        S.key = lu_key(row[J1], S.vv, S.done, lane >= n && J1 < n);
        double v = S.key;
        constexpr int rest = N - J - 2;              // columns J+2 .. N-1
        constexpr int covered = rest / 8 < 6 ? rest / 8 : 6;
#pragma unroll
        for (int k = J + 2; k < N; ++k) {
            row[k] = fma(-l, readlane(row[k], pv), row[k]);
            if ((k - J - 1) % 8 == 0) {
                const int g = (k - J - 1) / 8 - 1;
                if (g == 0) v = wave_max_stage<0>(v);
                else if (g == 1) v = wave_max_stage<1>(v);
                else if (g == 2) v = wave_max_stage<2>(v);
                else if (g == 3) v = wave_max_stage<3>(v);
                else if (g == 4) v = wave_max_stage<4>(v);
                else if (g == 5) v = wave_max_stage<5>(v);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // the stages the update did not cover
        if (covered <= 0) v = wave_max_stage<0>(v);
        if (covered <= 1) v = wave_max_stage<1>(v);
        if (covered <= 2) v = wave_max_stage<2>(v);
        if (covered <= 3) v = wave_max_stage<3>(v);
        if (covered <= 4) v = wave_max_stage<4>(v);
        if (covered <= 5) v = wave_max_stage<5>(v);
        S.kmax = readlane(v, 63);
    }
}

template <int N, int... Js>
__device__ __forceinline__ void team_lu_steps(double (&row)[N], int n, int lane, TeamLU& S,
                                              std::integer_sequence<int, Js...>) {
    (team_lu_step<N, Js>(row, n, lane, S), ...);
}

template <int N>
__device__ __forceinline__ void team_lu(double (&row)[N], int n, int lane, int& ps,
                                        double& rd, int& plv) {
    double amax = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) amax = fmax(amax, fabs(row[k]));
    TeamLU S;
    S.vv = 1.0 / amax;
    S.pos = lane;
    S.done = lane >= N;
    S.ps = lane;
    S.rd = 1.0;
    S.plv = lane;
    // the search of column 0
    S.key = lu_key(row[0], S.vv, S.done, lane >= n && 0 < n);
    S.kmax = wave_max(S.key);
    team_lu_steps<N>(row, n, lane, S, std::make_integer_sequence<int, N>{});
    ps = S.ps;
    rd = S.rd;
    plv = S.plv;
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

} // namespace gfc
