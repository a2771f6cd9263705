// Two-wavefront-per-cell chemistry kernel (nSpecie+1 in 33..64).
//
// The work of gpuFoam's per-cell solver (singleCellSolver -> odeSolve ->
// adaptiveSolve -> Rosenbrock step -> jacobian / LU / derivatives; see
// gpu_chemistry/src/gpuOde/*.H) is the wavefront kernel's
// (team_kernel.hip), spread over a workgroup of two wavefronts that share
// one cell:
//   * the reaction passes of the Jacobian and of each derivative are split
//     between the wavefronts (passes 2p and 2p+1);
//   * the Jacobian row transforms and the temperature row are split by
//     columns;
//   * the matrix -J + I/(gamma dx) is held column-cyclically: lane i of
//     wavefront w holds columns 2q+w of row i (N/2 registers instead of N).
//     The wavefront owning column j searches its pivot and hands the pivot
//     lane, reciprocal pivot and multipliers to the other through LDS; both
//     then update their columns with the pivot row broadcast by v_readlane;
//   * the factors are written back over the Jacobian tile and wavefront 0
//     runs the triangular solves from LDS; the stage vector goes back to
//     wavefront 1 through LDS;
//   * species-parallel state (y, stage vectors, error norm, step control)
//     is held by both wavefronts, computed from the same data, so they take
//     identical control paths.
// The LDS of one cell (Jacobian tile, concentrations, thermo tables) is
// that of the wavefront kernel, now shared by two wavefronts; with half
// the row registers a compute unit holds four cells with two wavefronts
// each, against four cells of one wavefront.
//
// Determinism: every LDS accumulation has a fixed order.  Derivative sums
// go to one array per wavefront, added in wavefront order; the Jacobian's
// atomics of one pass pair are applied by wavefront 0, then, after a
// barrier, by wavefront 1.
// On a rejected attempt the Jacobian is rebuilt (the tile holds the
// factors by then); it depends only on y0, so the rebuilt one is the same.
#include <hip/hip_runtime.h>

#include <utility>

#include "team_common.hpp"

namespace gfc {

namespace {

using namespace wave;

template <int N>
struct PairLDS {
    static constexpr int P = N + 1;  // odd row pitch (N is even)
    double J[N * P];                 // Jacobian, then its factors
    double c[kLanes];
    double dN[2][kLanes];            // per-wavefront dN/dt sums
    double Ha[kLanes];
    double Cp[kLanes];
    double colA[kLanes];             // rhoM / W_j
    double colB[kLanes];             // rhoM * v_j
    double M[kTeamMaxSets];          // third-body concentration per set
    double gL[kLanes], gH[kLanes];   // W_i Gstd_i with the low / high polynomial
    double hL[kLanes], hH[kLanes];   // W_i Ha_i with the low / high polynomial
    double Wd[kSetsT * kLanes];      // dense third-body Jacobian weights per set
    double xl[2][kLanes];            // LU multipliers of the last two steps
    double xv[kLanes];               // vector handed from one wavefront to the other
    double rdv[2];                   // reciprocal pivots of the last two steps
    int pv[2];                       // pivot lanes of the last two steps
    int ticket;
};

static_assert(sizeof(PairLDS<54>) <= 40960, "four GRI cells per compute unit");

// Specie Gibbs energies (and enthalpies) times W for Kc: wavefront 0 fills
// the low-polynomial tables, wavefront 1 the high ones.
template <int N, bool ENTHALPY>
__device__ __forceinline__ void pair_gibbs(const DeviceMech& m, PairLDS<N>& L,
                                           const TFun& f, int lane, int w) {
    if (lane < m.nS) {
        const double W = m.W[lane];
        const double* a = (w == 0 ? m.low : m.high) + 7 * lane;
        const double g = W * poly_Gstd(a, f.T, f.logT);
        if (w == 0) L.gL[lane] = g;
        else L.gH[lane] = g;
        if (ENTHALPY) {
            const double h = W * poly_Ha(a, f.T);
            if (w == 0) L.hL[lane] = h;
            else L.hH[lane] = h;
        }
    }
}

// Third-body concentrations of every efficiency set (lane s + 64w sums set
// s + 64w in specie order).  L.c must hold c (a barrier precedes).
template <int N>
__device__ __forceinline__ void pair_third_body(const DeviceMech& m, PairLDS<N>& L,
                                                int lane, int w) {
    const int s = lane + kLanes * w;
    if (s < m.nSets) {
        const double* e = m.effSets + s * kLanes;
        double acc = 0.0;
#pragma unroll 4
        for (int i = 0; i < m.nS; ++i) acc += e[i] * L.c[i];
        L.M[s] = acc;
    }
}

// ---------------------------------------------------------- derivatives --
// gpuODESystem::derivatives (gpuODESystem.H:164-196); yl: this lane's
// component; returns this lane's dydx component (both wavefronts).
template <int N>
__device__ __forceinline__ double pair_derivatives(const DeviceMech& m, double p, double T,
                                                   double yl, PairLDS<N>& L, int lane,
                                                   int w) {
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
    if (w == 0) L.c[lane] = cu * rhoM;
    L.dN[w][lane] = 0.0;
    // omega() evaluates the rates at T clipped to [Tlow, Thigh]
    const TFun f = tfun(fmin(fmax(T, m.Tlow), m.Thigh));
    pair_gibbs<N, false>(m, L, f, lane, w);
    __syncthreads();
    pair_third_body<N>(m, L, lane, w);
    __syncthreads();
    double* dN = L.dN[w];
#pragma unroll 1
    for (int pass = w; pass < m.nPass; pass += 2) {
        const int r = pass * kLanes + lane;
        if (r < m.nR) {
            const Rxn x = load_rxn(m, pass, lane);
            double kArr, ddT, pw[kTeamWidth];
            const double kf = forward_rate(x, f, L.M, kArr, ddT);
            const double kr = x.irr ? 0.0 : kf * exp(-ln_kc(x, f, L, nullptr));
            const double Cf = side_powers(x.li, x.lei, x.le, x.nl, L.c, pw);
            const double Cr = side_powers(x.ri, x.rei, x.re, x.nr, L.c, pw);
            const double wr = kf * Cf - kr * Cr;
#pragma unroll
            for (int q = 0; q < kTeamWidth; ++q) {
                if (q < x.nl) atomicAdd(&dN[x.li[q]], -(x.ls[q] * wr));
                if (q < x.nr) atomicAdd(&dN[x.ri[q]], x.rs[q] * wr);
            }
        }
    }
    __syncthreads();
    double d = 0.0;
    if (lane < nS) d = (L.dN[0][lane] + L.dN[1][lane]) * (m.W[lane] / rhoM);
    const double dT = -wave_sum(d * Ha) / CpM;
    if (lane == nS) d = dT;
    return d;
}

// ------------------------------------------------------------- jacobian --
// gpuODESystem::jacobian (gpuODESystem.H:198-281).  Leaves J (n x n, the
// pressure row/column dropped) in L.J and returns this lane's component of
// the derivative computed alongside (the reference's "dfdx").  The
// mechanism has at most kSetsT efficiency sets (the dense third-body part
// is added in the row transform).
template <int N>
__device__ __forceinline__ double pair_jacobian(const DeviceMech& m, double p, double T,
                                                double yl, PairLDS<N>& L, int lane, int w) {
    constexpr int P = PairLDS<N>::P;
    const int nS = m.nS;
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
    if (w == 0) {
        L.c[lane] = cu * rhoM;
        L.Ha[lane] = Ha;
        L.Cp[lane] = Cp;
        if (lane < nS) {
            L.colA[lane] = rhoM / m.W[lane];
            L.colB[lane] = rhoM * (1.0 / (p / (m.R[lane] * T)));
        }
    }
    L.dN[w][lane] = 0.0;
    const TFun f = tfun(T);
    pair_gibbs<N, true>(m, L, f, lane, w);
    for (int e = threadIdx.x; e < N * P; e += 2 * kLanes) L.J[e] = 0.0;
    for (int e = threadIdx.x; e < kSetsT * kLanes; e += 2 * kLanes) L.Wd[e] = 0.0;
    __syncthreads();
    pair_third_body<N>(m, L, lane, w);
    __syncthreads();
    bool poisoned = false; // a fall-off rate with Cf - Cr/Kc not finite
    double* dN = L.dN[w];
    const int nPairs = (m.nPass + 1) / 2;
#pragma unroll 1
    for (int pp = 0; pp < nPairs; ++pp) {
        const int pass = 2 * pp + w;
        const int r = pass * kLanes + lane;
        const bool act = pass < m.nPass && r < m.nR;
        Rxn x;
        double kf = 0, kr = 0, w_ = 0, dwdT = 0, g = 0;
        double pwl[kTeamWidth], pwr[kTeamWidth];
        if (act) {
            x = load_rxn(m, pass, lane);
            double kArr, ddT;
            kf = forward_rate(x, f, L.M, kArr, ddT);
            const double Cf = side_powers(x.li, x.lei, x.le, x.nl, L.c, pwl);
            dwdT = Cf * ddT;
            w_ = Cf * kf;
            double Cr = 0.0, lnKc = 0.0;
            if (!x.irr) {
                double dKc;
                lnKc = ln_kc(x, f, L, &dKc);
                const double rKc = exp(-lnKc);
                kr = kf * rKc;
                Cr = side_powers(x.ri, x.rei, x.re, x.nr, L.c, pwr);
                const double dkrdT = ddT * rKc - kr * dKc;
                dwdT -= Cr * dkrdT;
                w_ -= Cr * kr;
            }
            // Cf - Cr/Kc: used by the third-body and fall-off rates only, so the
            // Arrhenius passes skip its exponential
            double temp = Cf;
            if (x.type != GFC_RATE_ARRHENIUS) temp = Cf - Cr / exp(lnKc);
            g = (x.type == GFC_RATE_THIRD_BODY) ? kArr * temp : 0.0;
            if (x.type > GFC_RATE_THIRD_BODY && !isfinite(temp)) poisoned = true;
            // dN/dt (the reference's dfdx): per-wavefront sums
#pragma unroll
            for (int i = 0; i < kTeamWidth; ++i) {
                if (i < x.nl) atomicAdd(&dN[x.li[i]], -(x.ls[i] * w_));
                if (i < x.nr) atomicAdd(&dN[x.ri[i]], x.rs[i] * w_);
            }
        }
        // the Jacobian's atomics: wavefront 0's pass, then wavefront 1's
#pragma unroll 1
        for (int turn = 0; turn < 2; ++turn) {
            if (turn == w && act) {
                if (x.type == GFC_RATE_THIRD_BODY) {
                    // ddc_j = eff_j k (Cf - Cr/Kc): row weights of the set
                    double* Wd = &L.Wd[x.set * kLanes];
#pragma unroll
                    for (int i = 0; i < kTeamWidth; ++i) {
                        if (i < x.nl) atomicAdd(&Wd[x.li[i]], -(x.ls[i] * g));
                        if (i < x.nr) atomicAdd(&Wd[x.ri[i]], x.rs[i] * g);
                    }
                }
                // dCdT (gpuReaction.H:595-612)
#pragma unroll
                for (int i = 0; i < kTeamWidth; ++i) {
                    if (i < x.nl) atomicAdd(&L.J[x.li[i] * P + nS], -(x.ls[i] * dwdT));
                    if (i < x.nr) atomicAdd(&L.J[x.ri[i] * P + nS], x.rs[i] * dwdT);
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
        }
    }
    // a fall-off reaction with a non-finite Cf - Cr/Kc makes the reference's
    // dense (zero) concentration derivative NaN for its species rows
    if (__syncthreads_or(poisoned ? 1 : 0)) {
        if (w == 0) {
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
        }
        __syncthreads();
    }
    // derivative: dY/dt and dT/dt
    double d = 0.0;
    if (lane < nS) d = (L.dN[0][lane] + L.dN[1][lane]) * (m.W[lane] / rhoM);
    const double dfdxT = -wave_sum(d * Ha) / CpM;
    // ddY_dcdY (gpuODESystem.H:65-103): lane a transforms row a; wavefront
    // 0 takes columns [0, h), wavefront 1 [h, nS); the row sums of the two
    // halves are added in that order
    const int h = (nS + 1) / 2;
    double part = 0.0;
    if (lane < nS) {
        double* row = &L.J[lane * P];
        const double WiByrhoM = m.W[lane] / rhoM;
        double wq[kSetsT];
#pragma unroll
        for (int q = 0; q < kSetsT; ++q) wq[q] = L.Wd[q * kLanes + lane];
        const int j0 = w == 0 ? 0 : h, j1 = w == 0 ? h : nS;
        double en[kSetsT];
        dense_row_load(m.effSetsT + j0 * kSetsT, en);
#pragma unroll 1
        for (int j = j0; j < j1; ++j) {
            double e[kSetsT];
#pragma unroll
            for (int q = 0; q < kSetsT; ++q) e[q] = en[q];
            dense_row_load(m.effSetsT + min(j + 1, j1 - 1) * kSetsT, en);
            const double acc = dense_dot(wq, e);
            const double tj = row[j] + acc;
            part -= tj * L.c[j] * alphavM;
            row[j] = WiByrhoM * tj * L.colA[j] + L.colB[j] * d;
        }
        if (w == 1) L.xv[lane] = part;
    }
    __syncthreads();
    if (w == 0 && lane < nS) {
        double* row = &L.J[lane * P];
        const double ddNdT = (row[nS] + part) + L.xv[lane];
        row[nS] = (m.W[lane] / rhoM) * ddNdT + alphavM * d;
    }
    __syncthreads();
    // ddT_dtdY (:108-126): column i summed in row order, by lane i of
    // wavefront 0 (i < 32) or lane i-32 of wavefront 1
    {
        const int i = lane + 32 * w;
        if (lane < 32 && i < nS) {
            double v = 0.0;
#pragma unroll 4
            for (int j = 0; j < nS; ++j) v -= L.J[j * P + i] * L.Ha[j];
            v -= L.Cp[i] * dfdxT;
            L.J[nS * P + i] = v / CpM;
        }
    }
    // ddT_dtdT (:129-146)
    const double term = (lane < nS) ? d * Cp + L.J[lane * P + nS] * Ha : 0.0;
    const double s = wave_sum(term);
    if (w == 0 && lane == 0) L.J[nS * P + nS] = (-s - dfdxT * dCpMdT) / CpM;
    __syncthreads();
    if (lane == nS) d = dfdxT;
    return d;
}

// --------------------------------------------------------------- LU ----
// Column-cyclic factorisation of the row-per-lane matrix: this wavefront
// (W) holds columns 2q+W in rh[q].  Same arithmetic per entry as the
// wavefront kernel's team_lu (ludecompose.H:9-92: implicit scaling, the
// largest scaled value, ties to the last row position, a zero pivot
// replaced by DBL_EPSILON).  On return `ps` is the step at which this
// lane's row was the pivot, `rd` its reciprocal pivot and lane j's `plv`
// the pivot lane of step j.  Each elimination step is its own
// instantiation, so every register index is a constant.
struct LUState {
    double vv;
    int pos, ps, plv;
    bool done;
    double rd;
    // the next step's pivot, found ahead by the wavefront that owns it
    int pvN;
    double rdN, lN;
};

// Pivot search of column J in this wavefront's register rh[J/2] (J owned
// by this wavefront), the reciprocal pivot and the multipliers, published
// to the other wavefront; `done` is the state before step J.
template <int N, int J>
__device__ __forceinline__ void pair_lu_search(double (&rh)[N / 2], int n, int lane,
                                               PairLDS<N>& L, LUState& S) {
    constexpr int qj = J / 2;
    double key = (S.done || (lane >= n && J < n)) ? -1.0 : S.vv * fabs(rh[qj]);
    if (!(key >= 0.0)) key = -1.0;
    const double kmax = wave_max(key);
    const unsigned long long cand = __ballot(key == kmax && !S.done);
    int pv;
    if (__builtin_popcountll(cand) == 1) {
        pv = __builtin_amdgcn_readfirstlane((int)__builtin_ctzll(cand));
    } else {
        const int Pt = wave_max_i((key == kmax && !S.done) ? S.pos : -1);
        pv = __builtin_amdgcn_readfirstlane(
            (int)__builtin_ctzll(__ballot(S.pos == Pt && !S.done)));
    }
    double dj = readlane(rh[qj], pv);
    if (dj == 0.0) {
        dj = kSmall;
        if (lane == pv) rh[qj] = kSmall;
    }
    const double rDiag = 1.0 / dj;
    const bool doneAfter = S.done || lane == pv;
    const double l = doneAfter ? 0.0 : rh[qj] * rDiag;
    if (!doneAfter) rh[qj] = l;
    L.xl[J & 1][lane] = l;
    if (lane == 0) {
        L.pv[J & 1] = pv;
        L.rdv[J & 1] = rDiag;
    }
    S.pvN = pv;
    S.rdN = rDiag;
    S.lN = l;
}

// Step J.  The owner of column J found its pivot during step J-1 (or
// before step 0) and published it; the owner of column J+1 updates that
// column first and searches its pivot before finishing step J, so the
// search overlaps the other wavefront's update.
template <int N, int W, int J>
__device__ __forceinline__ void pair_lu_step(double (&rh)[N / 2], int n, int lane,
                                             PairLDS<N>& L, LUState& S) {
    constexpr int H = N / 2;
    constexpr bool own = (J % 2) == W;
    constexpr bool ownNext = !own && J + 1 < N;
    __syncthreads();
    int pv;
    double rDiag, l;
    if (own) {
        pv = S.pvN;
        rDiag = S.rdN;
        l = S.lN;
    } else {
        pv = __builtin_amdgcn_readfirstlane(L.pv[J & 1]);
        rDiag = L.rdv[J & 1];
        l = L.xl[J & 1][lane];
    }
    // bookkeeping, identical in both wavefronts
    const int Pp = __builtin_amdgcn_readlane(S.pos, pv);
    if (lane == J) S.plv = pv;
    if (S.pos == J) S.pos = Pp;
    if (lane == pv) {
        S.pos = J;
        S.done = true;
        S.ps = J;
        S.rd = rDiag;
    }
    {
        int dn = S.done ? 1 : 0;
        asm volatile("" : "+v"(S.ps), "+v"(S.rd), "+v"(S.plv), "+v"(S.pos), "+v"(dn));
        S.done = dn != 0;
    }
    if (ownNext) {
        constexpr int qn = (J + 1) / 2;
        rh[qn] = fma(-l, readlane(rh[qn], pv), rh[qn]);
        pair_lu_search<N, J + 1>(rh, n, lane, L, S);
    }
    // this wavefront's other columns right of J, in groups of 8
#pragma unroll
    for (int q = 0; q < H; ++q) {
        if (2 * q + W > J && !(ownNext && 2 * q + W == J + 1)) {
            rh[q] = fma(-l, readlane(rh[q], pv), rh[q]);
            if ((q - J / 2) % 8 == 0) __builtin_amdgcn_sched_barrier(0);
        }
    }
}

template <int N, int W, int... Js>
__device__ __forceinline__ void pair_lu_steps(double (&rh)[N / 2], int n, int lane,
                                              PairLDS<N>& L, LUState& S,
                                              std::integer_sequence<int, Js...>) {
    (pair_lu_step<N, W, Js>(rh, n, lane, L, S), ...);
}

template <int N, int W>
__device__ __forceinline__ void pair_lu(double (&rh)[N / 2], int n, int lane, PairLDS<N>& L,
                                        int& ps, double& rd, int& plv) {
    constexpr int H = N / 2;
    double am = 0.0;
#pragma unroll
    for (int q = 0; q < H; ++q) am = fmax(am, fabs(rh[q]));
    // row maxima: the halves meet in LDS (xl[1] and xv are free here)
    if (W == 0) L.xl[1][lane] = am;
    else L.xv[lane] = am;
    __syncthreads();
    LUState S;
    S.vv = 1.0 / fmax(L.xl[1][lane], L.xv[lane]);
    S.pos = lane;
    S.done = lane >= N;
    S.ps = lane;
    S.rd = 1.0;
    S.plv = lane;
    S.pvN = 0;
    S.rdN = 1.0;
    S.lN = 0.0;
    if (W == 0) pair_lu_search<N, 0>(rh, n, lane, L, S);
    pair_lu_steps<N, W>(rh, n, lane, L, S, std::make_integer_sequence<int, N>{});
    ps = S.ps;
    rd = S.rd;
    plv = S.plv;
}

// Solves (LU) x = b from the factors in L.J (rows in lane order, pivot
// bookkeeping of pair_lu); b and x species-parallel.  One wavefront.
template <int N>
__device__ __forceinline__ double pair_solve(const PairLDS<N>& L, int lane, int ps, double rd,
                                             int plv, double b) {
    // the row is loaded into registers before the substitutions (the
    // factorisation's registers are free by now): the chains then wait on
    // v_readlane only
    const double* Rp = &L.J[min(lane, N - 1) * PairLDS<N>::P];
    double R[N];
#pragma unroll
    for (int j = 0; j < N; ++j) R[j] = Rp[j];
    asm volatile("" : "+v"(plv));
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int pv = __builtin_amdgcn_readlane(plv, j);
        const double yj = readlane(b, pv);
        if (ps > j) b = fma(-R[j], yj, b);
    }
    double x = 0.0;
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
        const int pv = __builtin_amdgcn_readlane(plv, j);
        const double xj = readlane(b * rd, pv);
        if (ps < j) b = fma(-R[j], xj, b);
        if (lane == j) x = xj;
    }
    return x;
}

// Factorises -J + I/(gamma dx) (J in L.J; diag = 1/(gamma dx), or 0 to
// factorise L.J itself) and leaves the factors in L.J.  Both wavefronts.
template <int N>
__device__ __forceinline__ void pair_factorise(PairLDS<N>& L, int n, int lane, int w,
                                               double diag, bool negate, int& ps, double& rd,
                                               int& plv) {
    constexpr int H = N / 2, P = PairLDS<N>::P;
    double rh[H];
    const double dg = (lane < n) ? diag : 1.0;
    const double* Jr = &L.J[min(lane, N - 1) * P];
#pragma unroll
    for (int q = 0; q < H; ++q) {
        const int k = 2 * q + w;
        const double a = Jr[k];
        double v = (lane < N) ? (negate ? 0.0 - a : a) : 0.0;
        if (k == lane) v += dg;
        rh[q] = v;
    }
    if (w == 0) pair_lu<N, 0>(rh, n, lane, L, ps, rd, plv);
    else pair_lu<N, 1>(rh, n, lane, L, ps, rd, plv);
    // every wavefront has read its columns of J before the first step's barrier
    if (lane < N) {
        double* Jw = &L.J[lane * P];
#pragma unroll
        for (int q = 0; q < H; ++q) Jw[2 * q + w] = rh[q];
    }
    __syncthreads();
}

} // namespace

// ---------------------------------------------------------------- kernel --

// EXACT: the mechanism has n = N (no padding rows), which the compiler
// folds into the factorisation's candidate masks.
template <int N, bool EXACT>
__global__ void __launch_bounds__(128, 2)
pair_solve_kernel(SolveArgs A, int* queue) {
    __shared__ PairLDS<N> L;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const DeviceMech& m = A.m;
    const int n = EXACT ? N : m.nS + 1;
    const int nS = n - 1;
    const long ld = A.ldY;
    const RosTableau& tb = A.tab;

    // work queue as in the wavefront kernel: 16-cell tickets, results of a
    // chunk held back (wavefront 0: row k in lane k) and written as lines
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
            if (threadIdx.x == 0) L.ticket = atomicAdd(queue, 1);
            __syncthreads();
            const int t = __builtin_amdgcn_readfirstlane(L.ticket);
            __syncthreads();
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
        cell = bound(cell, A.nCells, A.m.bounds, kBoundCell);

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
                    const double dfdx = pair_jacobian<N>(m, p, T, y, L, lane, w);
                    const double dydx0 = (T >= m.Tlow && T <= m.Thigh)
                                             ? dfdx
                                             : pair_derivatives<N>(m, p, T, y, L, lane, w);
                    double err, yNew;
                    for (int attempt = 0;; ++attempt) {
                        // the tile holds the previous attempt's factors
                        if (attempt > 0) (void)pair_jacobian<N>(m, p, T, y, L, lane, w);
                        int ps = 0, plv = 0;
                        double rd = 1.0;
                        pair_factorise<N>(L, n, lane, w, 1.0 / (tb.gamma * dx), true, ps, rd,
                                          plv);
                        // Rosenbrock stages (gpuRosenbrock{12,23,34}.H::solve)
                        double k1 = 0, k2 = 0, k3 = 0, k4 = 0, dydx = 0;
#pragma unroll 1
                        for (int st = 0; st < tb.stages; ++st) {
                            if (st == 1 || (st == 2 && tb.stages == 4)) {
                                const double yt = (st == 1) ? y + tb.a21 * k1
                                                            : y + tb.a31 * k1 + tb.a32 * k2;
                                dydx = pair_derivatives<N>(m, p, readlane(yt, nS), yt, L,
                                                           lane, w);
                            }
                            double rhs;
                            if (st == 0) rhs = dydx0 + dx * tb.d1 * dfdx;
                            else if (st == 1) rhs = dydx + dx * tb.d2 * dfdx + tb.c21 * k1 / dx;
                            else if (st == 2) rhs = dydx + dx * tb.d3 * dfdx + (tb.c31 * k1 + tb.c32 * k2) / dx;
                            else rhs = dydx + dx * tb.d4 * dfdx + (tb.c41 * k1 + tb.c42 * k2 + tb.c43 * k3) / dx;
                            if (w == 0) L.xv[lane] = pair_solve<N>(L, lane, ps, rd, plv, rhs);
                            __syncthreads();
                            const double kx = L.xv[lane];
                            // the next writer of xv is behind a barrier
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
        if (w == 0) {
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
        }
        if (!queue) break;
    }
}

// Device-level evaluations with this kernel's own physics and LU (one cell
// per workgroup; grid = nCells).
template <int N>
__global__ void __launch_bounds__(128) pair_eval_kernel(EvalArgs E) {
    __shared__ PairLDS<N> L;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int cell = blockIdx.x;
    if (cell >= E.nCells) return;
    const DeviceMech& m = E.m;
    const int nS = m.nS, n = nS + 1;
    const long ld = E.ldY;
    constexpr int P = PairLDS<N>::P;
    if (E.mode == 2) {
        const double* M = E.M + (long)cell * n * n;
        for (int e = threadIdx.x; e < N * P; e += 2 * kLanes) {
            const int i = e / P, k = e % P;
            L.J[e] = (i < n && k < n) ? M[i * n + k] : 0.0;
        }
        __syncthreads();
        int ps = 0, plv = 0;
        double rd = 1.0;
        pair_factorise<N>(L, n, lane, w, 0.0, false, ps, rd, plv);
        if (w == 0) {
            const double bl = (lane < n) ? E.b[lane * ld + cell] : 0.0;
            const double xl = pair_solve<N>(L, lane, ps, rd, plv, bl);
            if (lane < n) E.x[lane * ld + cell] = xl;
        }
        return;
    }
    const double y = (lane < n) ? E.Y[lane * ld + cell] : 0.0;
    const double p = E.Y[(long)(nS + 1) * ld + cell];
    const double T = readlane(y, nS);
    if (E.mode == 0) {
        const double d = pair_derivatives<N>(m, p, T, y, L, lane, w);
        if (w == 0 && lane < n) E.dYdt[lane * ld + cell] = d;
    } else {
        const double d = pair_jacobian<N>(m, p, T, y, L, lane, w);
        if (w == 0 && lane < n) E.dYdt[lane * ld + cell] = d;
        double* Jc = E.J + (long)cell * n * n;
        for (int e = threadIdx.x; e < n * n; e += 2 * kLanes) Jc[e] = L.J[(e / n) * P + e % n];
    }
}

template __global__ void pair_eval_kernel<54>(EvalArgs);
template __global__ void pair_eval_kernel<64>(EvalArgs);
template __global__ void pair_solve_kernel<54, true>(SolveArgs, int*);
template __global__ void pair_solve_kernel<64, false>(SolveArgs, int*);

} // namespace gfc
