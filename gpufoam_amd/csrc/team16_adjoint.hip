// Discrete adjoint with four cells per wavefront (16-lane teams, nSpecie+1
// <= 16; no counterpart in gpuFoam).
//
// The reverse sweep of adjoint.hip, with one 16-lane team per cell instead
// of one lane: each team takes cells from a queue and walks a cell's
// recorded accepted steps backwards, checkpoint segment by segment (a
// segment is re-integrated from its checkpoint first).  Per step it
// recomputes the Jacobian, LU and stages with the 16-lane solve's step code
// (team16_common.hpp; the states a re-integrated segment reaches agree with
// the recorded ones to the last bits, not bit for bit), then applies the
// reverse of the step as adjoint_math.hpp's step_reverse does:
//   * member i holds component i of every vector;
//   * transposed solves read the LU from LDS, rows in pivot order;
//   * the adjoint of the Jacobian (Jb) lives in LDS, row i updated by
//     member i, column sums taken by the column's member in row order;
//   * reaction sweeps run 16 reactions per pass, one per member, with the
//     one-lane form's per-reaction reverse (reaction_reverse_w) and LDS
//     atomics for the concentration adjoints;
//   * team sums are row_shr trees, so sums differ from the one-lane form's
//     sequential order in the last bits.
// All sweeps use the team solve's tables, reactions grouped by rate type so
// that the members of a pass mostly take the same branch; rate-parameter
// adjoints go to each reaction's reference index (rxnRef).
#include <hip/hip_runtime.h>

#include "adjoint_math.hpp"
#include "team16_common.hpp"

namespace gfc {

namespace adj {
// LDS accumulation shared by the members of a team (reaction_reverse_w)
struct LdsAcc {
    double* p;
};
__device__ __forceinline__ void acc_add(LdsAcc v, int i, double x) { atomicAdd(&v.p[i], x); }
} // namespace adj

namespace {

// A specie index read from the tables (checked in a bounds-check build).
__device__ __forceinline__ int spx(const DeviceMech& m, int i) {
    return bound(i, m.nS, m.bounds, kBoundSpecies);
}

using namespace t16;

template <int N>
struct T16AdjLDS {
    static constexpr int P = Team16LDS<N>::P;
    // The LU factors and Jbar live in the solve's dense third-body weights
    // (f.Wd) where they fit: the weights are used only inside t16_jacobian,
    // the factors and Jbar only after it, until the step's reverse ends.
    static constexpr bool kShare = 2 * N * P <= kSets16 * kTW;
    Team16LDS<N> f;     // the solve's arrays; f.J holds the step's Jacobian
    double raw[2 * kTW]; // per species row before the transform: sum_j
                         // (dN_i/dc_j) c_j, then dN_i/dT
    double own[kShare ? 1 : 2 * N * P];
    double cb[kTW];     // concentration adjoints
    double dNb[kTW];    // adjoints of dN_i/dt
    double vj[kTW];     // 1/rho_i of each specie
    double v[kTW];      // a vector being broadcast
    // LU factors, row j = pivot position j; then the adjoint of the Jacobian
    __device__ __forceinline__ double* LU() { return kShare ? f.Wd : own; }
    __device__ __forceinline__ const double* LU() const { return kShare ? f.Wd : own; }
    __device__ __forceinline__ double* Jb() { return LU() + N * P; }
};

// Rate-parameter adjoints of one cell while its steps are reversed: a
// contiguous per-cell block in the order the sweeps visit the reactions,
// acc[(pass*kParams + q)*16 + mem] for reaction pass*16 + mem, so a team's
// additions are coalesced; written to thBar in reference order when the
// cell ends.  Every access is an atomic at device scope, so the additions
// need not wait for the value they update (no-return adds at the L2) and
// the final read sees them.
struct T16Th {
    double* acc; // nullptr: no parameter adjoints
    __device__ __forceinline__ void add(int pass, int mem, const double* v) const {
        if (!acc) return;
#pragma unroll
        for (int q = 0; q < adj::kParams; ++q)
            if (v[q] != 0.0) atomicAdd(&acc[(pass * adj::kParams + q) * kTW + mem], v[q]);
    }
};

// Third-body terms of the reaction sweeps, by efficiency set: M of each set
// from third_body16, the reactions' M adjoints summed per set in f.gL (the
// specie Gibbs array, used only by the forward evaluations) and scattered
// over the species once.
static_assert(kSets16 <= kTW, "per-set sums in gL");
template <int N>
__device__ __forceinline__ void t16_sets_begin(const DeviceMech& mr, T16AdjLDS<N>& L, int mem) {
    third_body16<N>(mr, L.f, mem);
    if (mem < mr.nSets) L.f.gL[mem] = 0.0;
    wave_sync();
}
template <int N>
__device__ __forceinline__ double t16_sets_scatter(const DeviceMech& mr, T16AdjLDS<N>& L,
                                                   int mem) {
    double v = 0.0;
    if (mem < mr.nS)
        for (int s = 0; s < mr.nSets; ++s) v += mr.effSets[s * kLanes + mem] * L.f.gL[s];
    return v;
}
template <int N>
__device__ __forceinline__ void t16_reaction_reverse(const DeviceMech& mr, T16AdjLDS<N>& L,
                                                     int r, double T, bool clip,
                                                     const adj::RxnBarT<kTeamWidth>& in,
                                                     double& Tb, double* t6) {
    const int set = mr.effSet[r];
    double Mb = 0.0;
    adj::reaction_reverse_w<kTeamWidth, adj::LdsAcc, true>(
        mr, r, T, clip, lane::SV{L.f.c, 1}, in, adj::LdsAcc{L.cb}, Tb, t6,
        set >= 0 ? L.f.M[set] : 0.0, &Mb);
    if (set >= 0 && Mb != 0.0) atomicAdd(&L.f.gL[set], Mb);
}

// One step's recomputed quantities, this member's components.
struct T16Stages {
    double F, dydx0, k1, k2, k3, k4, f2, f3;
    int ps; // pivot position of this member's row
};

// The accepted attempt of a step of size dx from y, as the solve computes it
// (team16_kernel.hip): returns the new state.  With KEEP, also leaves the
// pre-transform rows in L.raw and the LU in L.LU().
template <int N, bool KEEP>
__device__ __forceinline__ double t16_step(const SolveArgs& A, T16AdjLDS<N>& L, double p,
                                           double y, double dx, int lane, int mem,
                                           T16Stages& S) {
    constexpr int P = T16AdjLDS<N>::P;
    const DeviceMech& m = A.m;
    const int nS = m.nS, n = nS + 1;
    const RosTableau& tb = A.tab;
    const double T = team_bcast(y, lane, nS);
    const double dfdx = t16_jacobian<N, KEEP>(m, p, T, y, L.f, lane, mem, L.raw);
    const double dydx0 = (T >= m.Tlow && T <= m.Thigh)
                             ? dfdx
                             : t16_derivatives<N>(m, p, T, y, L.f, lane, mem);
    double row[N];
    const double tdiag = 1.0 / (tb.gamma * dx);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double v = 0.0;
        if (k < n && mem < n) v = -L.f.J[mem * P + k];
        if (k == mem) v += (mem < n) ? tdiag : 1.0;
        row[k] = v;
    }
    int ps, plv;
    double rd;
    t16_lu<N>(row, n, lane, mem, ps, rd, plv);
    if constexpr (KEEP) {
        // (a NaN Jacobian can leave a row at a padding position >= n)
        if (mem < n && ps < n) {
#pragma unroll
            for (int k = 0; k < N; ++k)
                if (k < n) L.LU()[ps * P + k] = row[k];
        }
    }
    t16_pivot_order<N>(row, rd, lane, plv);
    double k1 = 0, k2 = 0, k3 = 0, k4 = 0, dydx = 0, f2 = 0, f3 = 0;
#pragma unroll 1
    for (int st = 0; st < tb.stages; ++st) {
        if (st == 1 || (st == 2 && tb.stages == 4)) {
            const double yt = (st == 1) ? y + tb.a21 * k1 : y + tb.a31 * k1 + tb.a32 * k2;
            dydx = t16_derivatives<N>(m, p, team_bcast(yt, lane, nS), yt, L.f, lane, mem);
            if (st == 1) f2 = dydx;
            else f3 = dydx;
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
    S.F = dfdx;
    S.dydx0 = dydx0;
    S.k1 = k1;
    S.k2 = k2;
    S.k3 = k3;
    S.k4 = k4;
    S.f2 = f2;
    S.f3 = f3;
    S.ps = ps;
    wave_sync();
    if (tb.stages == 2) return y + tb.b1 * k1 + tb.b2 * k2;
    if (tb.stages == 3) return y + tb.b1 * k1 + tb.b2 * k2 + tb.b3 * k3;
    return y + tb.b1 * k1 + tb.b2 * k2 + tb.b3 * k3 + tb.b4 * k4;
}

// Solves A^T x = r for the step's A = P^T L U (adjoint_math.hpp lu_solve_T):
// member i holds r_i and receives x_i.
template <int N>
__device__ __forceinline__ double t16_solveT(const T16AdjLDS<N>& L, int n, int ps, double r,
                                             int lane, int mem) {
    constexpr int P = T16AdjLDS<N>::P;
    // U^T w = r; L.f.gH holds the pivots' reciprocals (t16_pivot_reciprocals)
    double w = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (j < n) {
            const double wj = team_bcast_k(r, j) * L.f.gH[j];
            if (mem == j) w = wj;
            if (mem > j && mem < n) r -= L.LU()[j * P + mem] * wj;
        }
    }
    // L^T v = w (unit diagonal)
    double v = 0.0;
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
        if (k < n) {
            const double vk = team_bcast_k(w, k);
            if (mem == k) v = vk;
            if (mem < k) w -= L.LU()[k * P + mem] * vk;
        }
    }
    // x = P^T v: row i sits at position ps_i
    return team_bcast(v, lane, ps);
}

// 1 / U[j][j] of the step's factors into L.f.gH (the specie Gibbs array of
// the forward evaluations, unused until the next step's recompute).
template <int N>
__device__ __forceinline__ void t16_pivot_reciprocals(T16AdjLDS<N>& L, int n, int mem) {
    static_assert(N <= kTW, "a pivot per member");
    constexpr int P = T16AdjLDS<N>::P;
    if (mem < n) L.f.gH[mem] = 1.0 / L.LU()[mem * P + mem];
    wave_sync();
}

// Jb(i, j) += r_i k_j
template <int N>
__device__ __forceinline__ void t16_add_outer(T16AdjLDS<N>& L, int n, double r, double k,
                                              int mem) {
    constexpr int P = T16AdjLDS<N>::P;
    L.v[mem] = k;
    wave_sync();
    if (mem < n && r != 0.0)
        for (int j = 0; j < n; ++j) L.Jb()[mem * P + j] += r * L.v[j];
    wave_sync();
}

// Mixture quantities of the team's cell at y (adjoint_math.hpp mixture):
// team scalars, and this member's specie terms.
struct T16Mix {
    double S1, rho, CpM, dCpMdT, avRaw, av;
    double yc, vi, Cp, dCp, Ha, c;
};
__device__ __forceinline__ T16Mix t16_mixture(const DeviceMech& m, double p, double y, double T,
                                              int lane, int mem) {
    T16Mix x{};
    double S1p = 0, CpMp = 0, dCpp = 0, avp = 0, cu = 0;
    if (mem < m.nS) {
        x.yc = fmax(y, 0.0);
        const double rhoi = p / (m.R[mem] * T);
        x.vi = 1.0 / rhoi;
        const double* a = coeffs_of(m.high + 7 * mem, m.low + 7 * mem, m.Tcommon[mem], T);
        x.Cp = poly_Cp(a, T);
        x.dCp = poly_dCpdT(a, T);
        x.Ha = poly_Ha(a, T);
        S1p = x.yc / rhoi;
        CpMp = x.yc * x.Cp;
        dCpp = x.yc * x.dCp;
        avp = x.yc * x.vi * (1.0 / T);
        cu = x.yc / m.W[mem];
    }
    x.S1 = team_sum(S1p, lane);
    x.rho = 1.0 / x.S1;
    x.CpM = team_sum(CpMp, lane);
    x.dCpMdT = team_sum(dCpp, lane);
    x.avRaw = team_sum(avp, lane);
    x.av = x.avRaw * x.rho;
    x.c = cu * x.rho;
    return x;
}

// Reverse of t16_mixture (adjoint_math.hpp mixture_reverse): adds to this
// member's ybar; member nS receives the temperature part plus TbExtra.
__device__ __forceinline__ void t16_mixture_reverse(const DeviceMech& m, double p, double y,
                                                    double T, const T16Mix& x, double cb,
                                                    double Cpb, double Hab, double vb, bool full,
                                                    adj::MixBar mb, double TbExtra,
                                                    double& ybar, int lane, int mem) {
    const int nS = m.nS;
    const double rho = x.rho;
    const double arb = mb.alphav * rho;
    double rhoPart = 0.0;
    if (mem < nS) rhoPart = cb * (x.yc / m.W[mem]);
    const double rhob = mb.rho + mb.alphav * x.avRaw + team_sum(rhoPart, lane);
    const double S1b = -rhob * rho * rho;
    double TbPart = 0.0, ycb = 0.0;
    if (mem < nS) {
        ycb = cb * rho / m.W[mem] + S1b * x.vi + arb * x.vi * (1.0 / T) + mb.CpM * x.Cp;
        TbPart = Hab * x.Cp;
        if (full) {
            const double* a = coeffs_of(m.high + 7 * mem, m.low + 7 * mem, m.Tcommon[mem], T);
            ycb += mb.dCpMdT * x.dCp;
            TbPart += mb.dCpMdT * x.yc * adj::poly_d2CpdT2(a, T) + Cpb * x.dCp + vb * x.vi / T;
        }
    }
    const double Tb = S1b * x.S1 / T + mb.CpM * x.dCpMdT + team_sum(TbPart, lane);
    if (mem < nS && y >= 0.0) ybar += ycb; // right derivative of max(y, 0) at 0
    if (mem == nS) ybar = (ybar + TbExtra) + Tb;
}

// ybar += d<fbar, f(y)>/dy, th += d<fbar, f(y)>/dtheta (adjoint_math.hpp
// vjp_f); fy = f(y).
template <int N>
__device__ __forceinline__ void t16_vjp_f(const DeviceMech& mr, T16AdjLDS<N>& L, double p,
                                          double y, double fbar, double fy, double& ybar,
                                          const T16Th& th, int lane, int mem) {
    const int nS = mr.nS;
    const double T = team_bcast(y, lane, nS);
    const T16Mix x = t16_mixture(mr, p, y, T, lane, mem);
    const double g = team_bcast(fbar, lane, nS);
    const double fT = team_bcast(fy, lane, nS);
    adj::MixBar mb;
    mb.CpM = -g * fT / x.CpM;
    double Hab = 0.0, rhoPart = 0.0;
    if (mem < nS) {
        Hab = -g * fy / x.CpM;
        const double fbi = fbar - g * x.Ha / x.CpM;
        L.dNb[mem] = fbi * mr.W[mem] / x.rho;
        rhoPart = -fbi * fy / x.rho;
    }
    mb.rho = team_sum(rhoPart, lane);
    L.cb[mem] = 0.0;
    L.f.c[mem] = (mem < nS) ? x.c : 0.0;
    wave_sync();
    t16_sets_begin<N>(mr, L, mem);
    double Tb = 0.0;
    const int nPass = (mr.nR + kTW - 1) / kTW;
#pragma unroll 1
    for (int pass = 0; pass < nPass; ++pass) {
        const int r = pass * kTW + mem;
        if (r >= mr.nR) continue;
        const long o = (long)r * kWidth;
        adj::RxnBarT<kTeamWidth> in;
        const int nl = mr.nl[r], nr = mr.nr[r];
#pragma unroll
        for (int q = 0; q < kTeamWidth; ++q)
            if (q < nl) in.w -= mr.ls[o + q] * L.dNb[spx(mr, mr.li[o + q])];
#pragma unroll
        for (int q = 0; q < kTeamWidth; ++q)
            if (q < nr) in.w += mr.rs[o + q] * L.dNb[spx(mr, mr.ri[o + q])];
        if (in.w == 0.0) continue;
        double t6[adj::kParams] = {0, 0, 0, 0, 0, 0};
        t16_reaction_reverse<N>(mr, L, r, T, true, in, Tb, t6);
        th.add(pass, mem, t6);
    }
    Tb = team_sum(Tb, lane);
    wave_sync();
    const double cb = (mem < nS) ? L.cb[mem] + t16_sets_scatter<N>(mr, L, mem) : 0.0;
    t16_mixture_reverse(mr, p, y, T, x, cb, 0.0, Hab, 0.0, false, mb, Tb, ybar, lane, mem);
    wave_sync();
}

// ybar += d/dy and th += d/dtheta of <Jb, J(y)> + <Fb, F(y)> for the fast
// Jacobian J = L.f.J (pre-transform rows in L.raw) and F at y
// (adjoint_math.hpp jact).  Jb is consumed.
template <int N>
__device__ __forceinline__ void t16_jact(const DeviceMech& mr, T16AdjLDS<N>& L, double p,
                                         double y, double F, double Fb, double& ybar,
                                         const T16Th& th, int lane, int mem) {
    constexpr int P = T16AdjLDS<N>::P;
    const int nS = mr.nS, iT = nS;
    const double T = team_bcast(y, lane, nS);
    const T16Mix x = t16_mixture(mr, p, y, T, lane, mem);
    const double CpM = x.CpM, rho = x.rho, av = x.av;
    const double FT = team_bcast(F, lane, nS);
    const bool sp = mem < nS;
    L.f.c[mem] = sp ? x.c : 0.0;
    L.f.Ha[mem] = sp ? x.Ha : 0.0;
    L.vj[mem] = sp ? x.vi : 0.0;
    L.v[mem] = F;
    double Cpb = 0.0, Hab = 0.0, vb = 0.0, cbT = 0.0;
    adj::MixBar mb;
    // ddT_dtdT (gpuODESystem.H:129-146)
    {
        const double g = L.Jb()[iT * P + iT];
        if (sp) {
            Fb += -g * x.Cp / CpM;
            Cpb += -g * F / CpM;
            L.Jb()[mem * P + iT] += -g * x.Ha / CpM;
            Hab += -g * L.f.J[mem * P + iT] / CpM;
        }
        if (mem == iT) Fb += -g * x.dCpMdT / CpM;
        mb.dCpMdT += -g * FT / CpM;
        mb.CpM += -g * L.f.J[iT * P + iT] / CpM;
    }
    wave_sync();
    // ddT_dtdY (:108-126): member j walks the columns i of its row
    {
        if (sp) {
            for (int i = 0; i < nS; ++i) {
                const double g = L.Jb()[iT * P + i];
                if (g == 0.0) continue;
                L.Jb()[mem * P + i] += -g * x.Ha / CpM;
                Hab += -g * L.f.J[mem * P + i] / CpM;
            }
        }
        const double gi = sp ? L.Jb()[iT * P + mem] : 0.0;
        double fbT = 0.0, cpm = 0.0;
        if (sp && gi != 0.0) {
            Cpb += -gi * FT / CpM;
            fbT = -gi * x.Cp / CpM;
            cpm = -gi * L.f.J[iT * P + mem] / CpM;
        }
        fbT = team_sum(fbT, lane);
        mb.CpM += team_sum(cpm, lane);
        if (mem == iT) Fb += fbT;
    }
    wave_sync();
    // ddY_dcdY (:65-103).  Column sums first, from the entries before the
    // row transforms: vb_j (species columns) and the temperature column's
    // concentration adjoints.
    if (sp) {
        // sum_i g_i (W_i/rho) t_im with t the rows before the transform:
        // the transform made J_im = (W_i/rho) t_im (rho/W_m) + rho v_m F_i
        double gJ = 0.0, gF = 0.0;
        for (int i = 0; i < nS; ++i) {
            vb += L.Jb()[i * P + mem] * rho * L.v[i];
            const double g = L.Jb()[i * P + iT];
            gJ += g * L.f.J[i * P + mem];
            gF += g * L.v[i];
        }
        cbT = -(gJ - (rho * x.vi) * gF) / (rho / mr.W[mem]) * av;
    }
    wave_sync();
    double rhoPart = 0.0, avPart = 0.0;
    if (sp) {
        const double WiR = mr.W[mem] / rho;
        double* row = &L.Jb()[mem * P];
        for (int j = 0; j < nS; ++j) {
            const double g = row[j];
            rhoPart += g * L.vj[j] * F;
            Fb += g * rho * L.vj[j];
            row[j] = g * WiR * (rho / mr.W[j]);
        }
        // temperature column
        const double g = row[iT];
        const double sumc = L.raw[mem];
        const double X = L.raw[kTW + mem] - sumc * av;
        row[iT] = g * WiR;
        for (int j = 0; j < nS; ++j) row[j] += -g * WiR * L.f.c[j] * av;
        avPart = g * (F - WiR * sumc);
        rhoPart += -g * mr.W[mem] / (rho * rho) * X;
        Fb += g * av;
    }
    mb.rho += team_sum(rhoPart, lane);
    mb.alphav += team_sum(avPart, lane);
    // dfdx: F_T = -sum F_i Ha_i / CpM, F_i = dN_i W_i / rho
    {
        const double g = team_bcast(Fb, lane, iT);
        if (sp) {
            Fb += -g * x.Ha / CpM;
            Hab += -g * F / CpM;
        }
        mb.CpM += -g * FT / CpM;
        double part = 0.0;
        if (sp) {
            L.dNb[mem] = Fb * mr.W[mem] / rho;
            part = -Fb * F / rho;
        }
        mb.rho += team_sum(part, lane);
    }
    L.cb[mem] = sp ? cbT : 0.0;
    wave_sync();
    t16_sets_begin<N>(mr, L, mem);
    // reactions (gpuReaction.H:547-702): adjoints of omega, dOmega/dT,
    // kf dCf/dc_j, kr dCr/dc_j and D from the raw-entry adjoints
    double Tb = 0.0;
    const int nPass = (mr.nR + kTW - 1) / kTW;
#pragma unroll 1
    for (int pass = 0; pass < nPass; ++pass) {
        const int r = pass * kTW + mem;
        if (r >= mr.nR) continue;
        const long o = (long)r * kWidth;
        const int nl = mr.nl[r], nr = mr.nr[r];
        adj::RxnBarT<kTeamWidth> in;
        in.fast = true;
        auto u = [&](int col) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < kTeamWidth; ++q)
                if (q < nl) v -= mr.ls[o + q] * L.Jb()[spx(mr, mr.li[o + q]) * P + col];
#pragma unroll
            for (int q = 0; q < kTeamWidth; ++q)
                if (q < nr) v += mr.rs[o + q] * L.Jb()[spx(mr, mr.ri[o + q]) * P + col];
            return v;
        };
#pragma unroll
        for (int q = 0; q < kTeamWidth; ++q)
            if (q < nl) in.w -= mr.ls[o + q] * L.dNb[spx(mr, mr.li[o + q])];
#pragma unroll
        for (int q = 0; q < kTeamWidth; ++q)
            if (q < nr) in.w += mr.rs[o + q] * L.dNb[spx(mr, mr.ri[o + q])];
#pragma unroll
        for (int j = 0; j < kTeamWidth; ++j)
            if (j < nl) in.dCf[j] = u(spx(mr, mr.li[o + j]));
#pragma unroll
        for (int j = 0; j < kTeamWidth; ++j)
            if (j < nr) in.dCr[j] = -u(spx(mr, mr.ri[o + j]));
        in.dwdT = u(iT);
        if (mr.type[r] == GFC_RATE_THIRD_BODY) {
            const double* eff = mr.eff + (long)r * nS;
            double D = 0.0;
            for (int j = 0; j < nS; ++j)
                if (eff[j] != 0.0) D += eff[j] * u(j);
            in.D = D;
        }
        double t6[adj::kParams] = {0, 0, 0, 0, 0, 0};
        t16_reaction_reverse<N>(mr, L, r, T, false, in, Tb, t6);
        th.add(pass, mem, t6);
    }
    Tb = team_sum(Tb, lane);
    wave_sync();
    const double cb = sp ? L.cb[mem] + t16_sets_scatter<N>(mr, L, mem) : 0.0;
    t16_mixture_reverse(mr, p, y, T, x, cb, Cpb, Hab, vb, true, mb, Tb, ybar, lane, mem);
    wave_sync();
}

// Reverse of one accepted step of size h from y0 (adjoint_math.hpp
// step_reverse): takes this member's component of the adjoint of the step's
// result and returns that of y0.
template <int N>
__device__ __forceinline__ double t16_step_reverse(const SolveArgs& A, const DeviceMech& mr,
                                                   T16AdjLDS<N>& L, double p, double y0,
                                                   double h, double yb, const T16Th& th,
                                                   int lane, int mem) {
    const DeviceMech& m = A.m;
    const int nS = m.nS, n = nS + 1;
    const RosTableau& t = A.tab;
    const double T0 = team_bcast(y0, lane, nS);
    const bool inRange = T0 >= m.Tlow && T0 <= m.Thigh;
    T16Stages S;
    (void)t16_step<N, true>(A, L, p, y0, h, lane, mem, S);
    t16_pivot_reciprocals<N>(L, n, mem);
    double kb1 = t.b1 * yb, kb2 = t.b2 * yb, kb3 = t.b3 * yb, kb4 = t.b4 * yb;
    double Fb = 0.0, fb2 = 0.0, fb3 = 0.0;
    for (int e = mem; e < N * T16AdjLDS<N>::P; e += kTW) L.Jb()[e] = 0.0;
    wave_sync();
    if (t.stages == 4) {
        const double rb = t16_solveT<N>(L, n, S.ps, kb4, lane, mem);
        t16_add_outer<N>(L, n, rb, S.k4, mem);
        fb3 += rb;
        Fb += h * t.d4 * rb;
        kb1 += t.c41 / h * rb;
        kb2 += t.c42 / h * rb;
        kb3 += t.c43 / h * rb;
    }
    if (t.stages >= 3) {
        const double rb = t16_solveT<N>(L, n, S.ps, kb3, lane, mem);
        t16_add_outer<N>(L, n, rb, S.k3, mem);
        if (t.stages == 4) fb3 += rb;
        else fb2 += rb;
        Fb += h * t.d3 * rb;
        kb1 += t.c31 / h * rb;
        kb2 += t.c32 / h * rb;
    }
    if (t.stages == 4) {
        const double ys = y0 + t.a31 * S.k1 + t.a32 * S.k2;
        double ysb = 0.0;
        t16_vjp_f<N>(mr, L, p, ys, fb3, S.f3, ysb, th, lane, mem);
        yb += ysb;
        kb1 += t.a31 * ysb;
        kb2 += t.a32 * ysb;
    }
    {
        const double rb = t16_solveT<N>(L, n, S.ps, kb2, lane, mem);
        t16_add_outer<N>(L, n, rb, S.k2, mem);
        fb2 += rb;
        Fb += h * t.d2 * rb;
        kb1 += t.c21 / h * rb;
    }
    {
        const double ys = y0 + t.a21 * S.k1;
        double ysb = 0.0;
        t16_vjp_f<N>(mr, L, p, ys, fb2, S.f2, ysb, th, lane, mem);
        yb += ysb;
        kb1 += t.a21 * ysb;
    }
    const double rb = t16_solveT<N>(L, n, S.ps, kb1, lane, mem);
    t16_add_outer<N>(L, n, rb, S.k1, mem);
    Fb += h * t.d1 * rb;
    // within the reaction temperature limits f(y0) is the Jacobian call's F
    if (inRange) Fb += rb;
    else t16_vjp_f<N>(mr, L, p, y0, rb, S.dydx0, yb, th, lane, mem);
    t16_jact<N>(mr, L, p, y0, S.F, Fb, yb, th, lane, mem);
    return yb;
}

} // namespace

// Reverse sweep, one team per cell (adjoint.hip adjoint_reverse_kernel).
// A.m: the team solve's tables.  seg: the K-state segment buffers (segPer
// doubles per cell, wave-blocked).  Ybar/Y0bar: [n][ldb]; thBar:
// [6 nR][ldb], reactions in reference order.  pacc: the parameter
// accumulators, cell c at pacc + (c/64)*paccSlot + (c%64)*paccPer, each
// kParams*16*nPass doubles long.
template <int N>
__global__ void __launch_bounds__(64, 2)
team16_adjoint_reverse_kernel(SolveArgs A, T16Record R, double* seg, long segPer,
                              const double* Ybar, long ldb, double* Y0bar, double* thBar,
                              double* pacc, long paccSlot, long paccPer, int* queue) {
    __shared__ T16AdjLDS<N> Ls[kLanes / kTW];
    const int lane = threadIdx.x;
    const int team = lane >> 4, mem = lane & 15;
    T16AdjLDS<N>& L = Ls[team];
    const DeviceMech& mr = A.m;
    const int nS = A.m.nS, n = nS + 1;
    for (;;) {
        int c = 0;
        if (mem == 0) c = atomicAdd(queue, 1);
        int cell = team_bcast_i(c, lane, 0);
        if (cell >= A.nCells) break;
        cell = bound(cell, A.nCells, mr.bounds, kBoundCell);
        const double* y0c = rec_slot(R.y0c, cell, n + 1);
        const double p = y0c[(long)n * kRecStride];
        const double T0 = y0c[(long)nS * kRecStride];
        double yb = (mem < n) ? Ybar[mem * ldb + cell] : 0.0;
        const int nPass = (mr.nR + kTW - 1) / kTW;
        const T16Th th{thBar ? pacc + (cell / kLanes) * paccSlot + (cell % kLanes) * paccPer
                             : nullptr};
        if (th.acc)
            for (int q = mem; q < adj::kParams * kTW * nPass; q += kTW)
                __hip_atomic_store(&th.acc[q], 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (T0 > A.Treact && A.deltaT > kSmall) {
            const double* yE = rec_slot(R.yEnd, cell, n);
            if (mem < nS && yE[mem * kRecStride] < 0.0) yb = 0.0;
            const int nAcc = bound(R.nAcc[cell], R.maxSteps + 1, mr.bounds, kBoundStep);
            const double* ck = rec_slot(R.ck, cell, R.nck * n);
            const double* hs = rec_slot(R.hs, cell, R.maxSteps);
            double* sg = rec_slot(seg, cell, segPer);
            const int nseg = (nAcc + R.K - 1) / R.K;
            for (int g = nseg - 1; g >= 0; --g) {
                const int s0 = g * R.K, s1 = min(s0 + R.K, nAcc);
                double y = (mem < n) ? ck[((long)bound(g, R.nck, mr.bounds, kBoundSlot) * n + mem) *
                                             kRecStride]
                                        : 0.0;
                if (mem < n) sg[mem * kRecStride] = y;
                T16Stages S;
                for (int t = s0; t < s1 - 1; ++t) {
                    y = t16_step<N, false>(A, L, p, y, hs[(long)t * kRecStride], lane, mem, S);
                    if (mem < n) sg[((long)(t - s0 + 1) * n + mem) * kRecStride] = y;
                }
                for (int t = s1 - 1; t >= s0; --t) {
                    const double y0 = (mem < n) ? sg[((long)(t - s0) * n + mem) * kRecStride] : 0.0;
                    yb = t16_step_reverse<N>(A, mr, L, p, y0, hs[(long)t * kRecStride], yb, th,
                                             lane, mem);
                }
            }
        }
        if (mem < n) Y0bar[mem * ldb + cell] = yb;
        if (th.acc)
            for (int pass = 0; pass < nPass; ++pass) {
                const int r = pass * kTW + mem;
                if (r >= mr.nR) break;
                const long ref = mr.rxnRef[r];
#pragma unroll
                for (int q = 0; q < adj::kParams; ++q)
                    thBar[(adj::kParams * ref + q) * ldb + cell] =
                        __hip_atomic_load(&th.acc[(pass * adj::kParams + q) * kTW + mem],
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        wave_sync();
    }
}

template __global__ void team16_adjoint_reverse_kernel<11>(SolveArgs, T16Record, double*, long,
                                                           const double*, long, double*,
                                                           double*, double*, long, long, int*);
template __global__ void team16_adjoint_reverse_kernel<16>(SolveArgs, T16Record, double*, long,
                                                           const double*, long, double*,
                                                           double*, double*, long, long, int*);

} // namespace gfc
