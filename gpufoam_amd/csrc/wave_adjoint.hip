// Discrete adjoint with one wavefront per cell (nSpecie+1 <= 64; no
// counterpart in gpuFoam).
//
// The reverse sweep of adjoint.hip with the wavefront kernel's mapping: a
// wavefront takes cells from a queue and walks a cell's recorded accepted
// steps backwards, checkpoint segment by segment (a segment is
// re-integrated from its checkpoint first).  Per step it recomputes the
// Jacobian, LU and stages with the wavefront solve's step code
// (team_device.hpp), then applies the reverse of the step as
// adjoint_math.hpp's step_reverse does:
//   * lane i holds component i of every vector;
//   * the step's Jacobian (after the transform) lives in LDS; of the
//     species rows before the transform the Jacobian adjoint needs dN_i/dT,
//     each row's product with c (both kept when the Jacobian is assembled)
//     and one column combination, recovered from the transformed rows;
//   * the adjoint of the Jacobian (Jb) is kept as factors: it is the sum of
//     the stages' outer products r_s k_s^T, and the reverse of the
//     Jacobian's transforms adds three rank-one terms and scales rows and
//     columns, so Jb = sum_q a_q b_q^T with at most seven terms, held in the
//     region of the solve's dense third-body weights (dead once the
//     Jacobian is built); every Jb entry or row / column sum the reverse
//     needs is formed from the factors;
//   * the LU factors go through that region once, 16 rows in pivot order at
//     a time, so that lane k keeps column k of them in registers; the
//     transposed solves read their entries from registers;
//   * reaction sweeps run 64 reactions per pass, one per lane, with the
//     one-lane form's per-reaction reverse (reaction_reverse_w) and LDS
//     atomics for the concentration adjoints;
//   * wavefront sums are DPP trees, so sums differ from the one-lane form's
//     sequential order in the last bits.
// The sweeps use the wavefront solve's tables, reactions grouped by rate
// type; rate-parameter adjoints go to each reaction's reference index
// (rxnRef).  The LDS of a cell (the solve's arrays and five vectors, 40 KB
// for GRI-Mech's n = 54) holds four wavefronts per compute unit.
#include <hip/hip_runtime.h>

#include "adjoint_math.hpp"
#include "team16_common.hpp"
#include "team_device.hpp"

namespace gfc {

namespace adj {
// LDS accumulation shared by the lanes of a wavefront (reaction_reverse_w)
struct WaveAcc {
    double* p;
};
__device__ __forceinline__ void acc_add(WaveAcc v, int i, double x) { atomicAdd(&v.p[i], x); }
} // namespace adj

namespace {

// A specie index read from the tables (checked in a bounds-check build).
__device__ __forceinline__ int spx(const DeviceMech& m, int i) {
    return bound(i, m.nS, m.bounds, kBoundSpecies);
}

// Rate-parameter adjoints of one cell while its steps are reversed: a
// contiguous per-cell block in the order the sweeps visit the reactions,
// acc[(pass*kParams + q)*64 + lane] for reaction pass*64 + lane, so each
// sweep's additions are coalesced; written to thBar in reference order
// when the cell ends.  Every access is an atomic at device scope, so the
// additions need not wait for the value they update (no-return adds at
// the L2) and the final read sees them.
struct WTh {
    double* acc; // nullptr: no parameter adjoints
    __device__ __forceinline__ void add(int pass, int lane, const double* v) const {
        if (!acc) return;
#pragma unroll
        for (int q = 0; q < adj::kParams; ++q)
            if (v[q] != 0.0) atomicAdd(&acc[(pass * adj::kParams + q) * kLanes + lane], v[q]);
    }
};

using t16::kRecStride;
using t16::rec_slot;
using t16::T16Record;
using t16::wave_sync;

// Factors of the Jacobian's adjoint: Jb = sum_{q < rank} a_q b_q^T.
constexpr int kJbRank = 8;
static_assert(2 * kJbRank * kLanes <= kMaxSets * kLanes, "Jb factors in the weight region");

template <int N>
struct WaveAdjLDS {
    static constexpr int P = TeamLDS<N>::P;
    TeamLDS<N> f;       // the solve's arrays; f.J holds the step's Jacobian,
                        // f.Wd the LU's transpose, then the factors of Jb
    double raw[2 * kLanes]; // per species row before the transform: sum_j
                            // (dN_i/dc_j) c_j, then dN_i/dT
    double cb[kLanes];  // concentration adjoints
    double dNb[kLanes]; // adjoints of dN_i/dt
    double vj[kLanes];  // 1/rho_i of each specie
    double v[kLanes];   // a vector being broadcast
};

__device__ __forceinline__ double bcast_from(double v, int src) {
    const int lo = __shfl(__double2loint(v), src);
    const int hi = __shfl(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}

// One step's recomputed quantities, this lane's components.
template <int N>
struct WStages {
    double F, dydx0, k1, k2, k3, k4, f2, f3;
    int ps;       // pivot position of this lane's row
    double col[N]; // column `lane` of the LU factors, rows in pivot order
    double rdg;    // 1 / U[lane][lane] (1 on padding lanes)
};

// The accepted attempt of a step of size dx from y, as the solve computes it
// (team_kernel.hip): returns the new state.  With KEEP, also leaves the
// pre-transform row quantities in L.raw and this lane's column of the LU
// factors in S.col.
template <int N, bool KEEP>
__device__ __forceinline__ double w_step(const SolveArgs& A, WaveAdjLDS<N>& L, double p,
                                         double y, double dx, int lane, WStages<N>& S) {
    constexpr int P = WaveAdjLDS<N>::P;
    const DeviceMech& m = A.m;
    const int nS = m.nS, n = nS + 1;
    const RosTableau& tb = A.tab;
    const double T = readlane(y, nS);
    const double dfdx = team_jacobian<N, KEEP>(m, p, T, y, L.f, lane, L.raw);
    const double dydx0 = (T >= m.Tlow && T <= m.Thigh)
                             ? dfdx
                             : team_derivatives<N>(m, p, T, y, L.f, lane);
    // -J + I/(gamma dx) as the solve builds it
    double row[N];
    const double tdiag = 1.0 / (tb.gamma * dx);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double v = 0.0;
        if (k < n && lane < n) v = -L.f.J[lane * P + k];
        if (k == lane) v += (lane < n) ? tdiag : 1.0;
        row[k] = v;
    }
    int ps, plv;
    double rd;
    team_lu<N>(row, n, lane, ps, rd, plv);
    if constexpr (KEEP) {
        S.rdg = 1.0;
        // BR pivot positions at a time: their rows into the weight region,
        // then each lane takes its column (a NaN Jacobian can leave a row at
        // a padding position >= n)
        double* X = L.f.Wd;
        constexpr int BR = kMaxSets * kLanes / P;
#pragma unroll
        for (int b = 0; b < (N + BR - 1) / BR; ++b) {
            if (lane < n && ps < n && ps >= BR * b && ps < BR * b + BR) {
#pragma unroll
                for (int k = 0; k < N; ++k)
                    if (k < n) X[(ps - BR * b) * P + k] = row[k];
            }
            wave_sync();
#pragma unroll
            for (int t = 0; t < BR; ++t) {
                const int j = BR * b + t;
                if (j < N) {
                    S.col[j] = (j < n && lane < n) ? X[t * P + min(lane, N - 1)] : 0.0;
                    if (j == lane) S.rdg = S.col[j];
                }
            }
            wave_sync();
        }
        S.rdg = (lane < n) ? 1.0 / S.rdg : 1.0;
    }
    double k1 = 0, k2 = 0, k3 = 0, k4 = 0, dydx = 0, f2 = 0, f3 = 0;
#pragma unroll 1
    for (int st = 0; st < tb.stages; ++st) {
        if (st == 1 || (st == 2 && tb.stages == 4)) {
            const double yt = (st == 1) ? y + tb.a21 * k1 : y + tb.a31 * k1 + tb.a32 * k2;
            dydx = team_derivatives<N>(m, p, readlane(yt, nS), yt, L.f, lane);
            if (st == 1) f2 = dydx;
            else f3 = dydx;
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
// lane i holds r_i and receives x_i; lane k holds column k of the factors
// (col[j] = LU[j][k], rows in pivot order) and rdg = 1/U[k][k].
template <int N>
__device__ __forceinline__ double w_solveT(const double (&col)[N], double rdg, int n, int ps,
                                           double r, int lane) {
    // U^T w = r
    double w = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (j < n) {
            const double wj = readlane(r, j) * readlane(rdg, j);
            if (lane == j) w = wj;
            if (lane > j && lane < n) r -= col[j] * wj;
        }
    }
    // L^T v = w (unit diagonal)
    double v = 0.0;
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
        if (k >= n) continue;
        const double vk = readlane(w, k);
        if (lane == k) v = vk;
        if (lane < k) w -= col[k] * vk;
    }
    // x = P^T v: row i sits at position ps_i
    return bcast_from(v, ps);
}

template <int N>
__device__ __forceinline__ double* jb_a(WaveAdjLDS<N>& L, int q) {
    return L.f.Wd + q * kLanes;
}
template <int N>
__device__ __forceinline__ double* jb_b(WaveAdjLDS<N>& L, int q) {
    return L.f.Wd + (kJbRank + q) * kLanes;
}

// Jb += r k^T (one more factor pair; lane i holds r_i and k_i)
template <int N>
__device__ __forceinline__ void w_add_outer(WaveAdjLDS<N>& L, int n, double r, double k,
                                            int lane, int& rank) {
    jb_a(L, rank)[lane] = (lane < n) ? r : 0.0;
    jb_b(L, rank)[lane] = (lane < n) ? k : 0.0;
    ++rank;
    wave_sync();
}

// sum_q a_q[i] b_q[j] for this lane's row i = lane and a column j (uniform)
template <int N>
__device__ __forceinline__ double jb_entry(WaveAdjLDS<N>& L, int rank, int i, int j) {
    double v = 0.0;
    for (int q = 0; q < rank; ++q) v += jb_a(L, q)[i] * jb_b(L, q)[j];
    return v;
}

// Third-body terms of the reaction sweeps, by efficiency set: the sweeps
// take each set's M from L.f.M (third_body) and sum the reactions' M
// adjoints per set, then scatter them over the species once.  The per-set
// sums live in the specie Gibbs arrays, which only the forward evaluations
// use; the Jacobian adjoint's per-set products eff_s . b_q (sum_j eff_sj
// b_q[j], rank * nSets <= 128 of them) in the enthalpy arrays.
static_assert(kTeamMaxSets <= 2 * kLanes, "per-set sums in gL and gH");
template <int N>
__device__ __forceinline__ double* mb_set(WaveAdjLDS<N>& L, int s) {
    return s < kLanes ? &L.f.gL[s] : &L.f.gH[s - kLanes];
}
template <int N>
__device__ __forceinline__ double* eb_slot(WaveAdjLDS<N>& L, int t) {
    return t < kLanes ? &L.f.hL[t] : &L.f.hH[t - kLanes];
}

// M of every set at the c in L.f.c, and the per-set M adjoints cleared.
template <int N>
__device__ __forceinline__ void w_sets_begin(const DeviceMech& mr, WaveAdjLDS<N>& L, int lane) {
    third_body<N>(mr, L.f, lane);
    for (int s = lane; s < mr.nSets; s += kLanes) *mb_set(L, s) = 0.0;
    wave_sync();
}

// This specie's part of the per-set M adjoints: sum_s eff_s,i Mb_s.
template <int N>
__device__ __forceinline__ double w_sets_scatter(const DeviceMech& mr, WaveAdjLDS<N>& L,
                                                 int lane) {
    double v = 0.0;
    if (lane < mr.nS)
        for (int s = 0; s < mr.nSets; ++s) v += mr.effSets[s * kLanes + lane] * *mb_set(L, s);
    return v;
}

// One reaction's reverse with its set's M, its M adjoint summed per set.
template <int N>
__device__ __forceinline__ void w_reaction_reverse(const DeviceMech& mr, WaveAdjLDS<N>& L,
                                                   int r, double T, bool clip,
                                                   const adj::RxnBarT<kTeamWidth>& in,
                                                   double& Tb, double* t6) {
    const int set = mr.effSet[r];
    double Mb = 0.0;
    adj::reaction_reverse_w<kTeamWidth, adj::WaveAcc, true>(
        mr, r, T, clip, lane::SV{L.f.c, 1}, in, adj::WaveAcc{L.cb}, Tb, t6,
        set >= 0 ? L.f.M[set] : 0.0, &Mb);
    if (set >= 0 && Mb != 0.0) atomicAdd(mb_set(L, set), Mb);
}

// Mixture quantities of the cell at y (adjoint_math.hpp mixture): wavefront
// scalars, and this lane's specie terms.
struct WMix {
    double S1, rho, CpM, dCpMdT, avRaw, av;
    double yc, vi, Cp, dCp, Ha, c;
};
__device__ __forceinline__ WMix w_mixture(const DeviceMech& m, double p, double y, double T,
                                          int lane) {
    WMix x{};
    double S1p = 0, CpMp = 0, dCpp = 0, avp = 0, cu = 0;
    if (lane < m.nS) {
        x.yc = fmax(y, 0.0);
        const double rhoi = p / (m.R[lane] * T);
        x.vi = 1.0 / rhoi;
        const double* a = coeffs_of(m.high + 7 * lane, m.low + 7 * lane, m.Tcommon[lane], T);
        x.Cp = poly_Cp(a, T);
        x.dCp = poly_dCpdT(a, T);
        x.Ha = poly_Ha(a, T);
        S1p = x.yc / rhoi;
        CpMp = x.yc * x.Cp;
        dCpp = x.yc * x.dCp;
        avp = x.yc * x.vi * (1.0 / T);
        cu = x.yc / m.W[lane];
    }
    x.S1 = wave_sum(S1p);
    x.rho = 1.0 / x.S1;
    x.CpM = wave_sum(CpMp);
    x.dCpMdT = wave_sum(dCpp);
    x.avRaw = wave_sum(avp);
    x.av = x.avRaw * x.rho;
    x.c = cu * x.rho;
    return x;
}

// Reverse of w_mixture (adjoint_math.hpp mixture_reverse): adds to this
// lane's ybar; lane nS receives the temperature part plus TbExtra.
__device__ __forceinline__ void w_mixture_reverse(const DeviceMech& m, double p, double y,
                                                  double T, const WMix& x, double cb,
                                                  double Cpb, double Hab, double vb, bool full,
                                                  adj::MixBar mb, double TbExtra,
                                                  double& ybar, int lane) {
    const int nS = m.nS;
    const double rho = x.rho;
    const double arb = mb.alphav * rho;
    double rhoPart = 0.0;
    if (lane < nS) rhoPart = cb * (x.yc / m.W[lane]);
    const double rhob = mb.rho + mb.alphav * x.avRaw + wave_sum(rhoPart);
    const double S1b = -rhob * rho * rho;
    double TbPart = 0.0, ycb = 0.0;
    if (lane < nS) {
        ycb = cb * rho / m.W[lane] + S1b * x.vi + arb * x.vi * (1.0 / T) + mb.CpM * x.Cp;
        TbPart = Hab * x.Cp;
        if (full) {
            const double* a = coeffs_of(m.high + 7 * lane, m.low + 7 * lane, m.Tcommon[lane], T);
            ycb += mb.dCpMdT * x.dCp;
            TbPart += mb.dCpMdT * x.yc * adj::poly_d2CpdT2(a, T) + Cpb * x.dCp + vb * x.vi / T;
        }
    }
    const double Tb = S1b * x.S1 / T + mb.CpM * x.dCpMdT + wave_sum(TbPart);
    if (lane < nS && y >= 0.0) ybar += ycb; // right derivative of max(y, 0) at 0
    if (lane == nS) ybar = (ybar + TbExtra) + Tb;
}

// ybar += d<fbar, f(y)>/dy, th += d<fbar, f(y)>/dtheta (adjoint_math.hpp
// vjp_f); fy = f(y).
template <int N>
__device__ __forceinline__ void w_vjp_f(const DeviceMech& mr, WaveAdjLDS<N>& L, double p,
                                        double y, double fbar, double fy, double& ybar,
                                        const WTh& th, int lane) {
    const int nS = mr.nS;
    const double T = readlane(y, nS);
    const WMix x = w_mixture(mr, p, y, T, lane);
    const double g = readlane(fbar, nS);
    const double fT = readlane(fy, nS);
    adj::MixBar mb;
    mb.CpM = -g * fT / x.CpM;
    double Hab = 0.0, rhoPart = 0.0;
    if (lane < nS) {
        Hab = -g * fy / x.CpM;
        const double fbi = fbar - g * x.Ha / x.CpM;
        L.dNb[lane] = fbi * mr.W[lane] / x.rho;
        rhoPart = -fbi * fy / x.rho;
    }
    mb.rho = wave_sum(rhoPart);
    L.cb[lane] = 0.0;
    L.f.c[lane] = (lane < nS) ? x.c : 0.0;
    wave_sync();
    w_sets_begin<N>(mr, L, lane);
    double Tb = 0.0;
    const int nPass = (mr.nR + kLanes - 1) / kLanes;
#pragma unroll 1
    for (int pass = 0; pass < nPass; ++pass) {
        const int r = pass * kLanes + lane;
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
        w_reaction_reverse<N>(mr, L, r, T, true, in, Tb, t6);
        th.add(pass, lane, t6);
    }
    Tb = wave_sum(Tb);
    wave_sync();
    const double cb = (lane < nS) ? L.cb[lane] + w_sets_scatter<N>(mr, L, lane) : 0.0;
    w_mixture_reverse(mr, p, y, T, x, cb, 0.0, Hab, 0.0, false, mb, Tb, ybar, lane);
    wave_sync();
}

// ybar += d/dy and th += d/dtheta of <Jb, J(y)> + <Fb, F(y)> for the fast
// Jacobian J = L.f.J (pre-transform row quantities in L.raw) and F at y
// (adjoint_math.hpp jact).  Jb (its factors, `rank` pairs) is consumed: the
// reverse of each transform of gpuODESystem.H updates the factors (rows and
// columns scaled, rank-one terms added) instead of the entries.
template <int N>
__device__ __forceinline__ void w_jact(const DeviceMech& mr, WaveAdjLDS<N>& L, double p,
                                       double y, double F, double Fb, double& ybar,
                                       const WTh& th, int lane, int rank) {
    constexpr int P = WaveAdjLDS<N>::P;
    const int nS = mr.nS, iT = nS;
    const double T = readlane(y, nS);
    const WMix x = w_mixture(mr, p, y, T, lane);
    const double CpM = x.CpM, rho = x.rho, av = x.av;
    const double FT = readlane(F, nS);
    const bool sp = lane < nS;
    L.f.c[lane] = sp ? x.c : 0.0;
    L.f.Ha[lane] = sp ? x.Ha : 0.0;
    L.vj[lane] = sp ? x.vi : 0.0;
    L.v[lane] = F;
    double Cpb = 0.0, Hab = 0.0, vb = 0.0, cbT = 0.0;
    adj::MixBar mb;
    // ddT_dtdT (gpuODESystem.H:129-146): Jb(i, T) += -g Ha_i / CpM, i species
    {
        const double g = jb_entry(L, rank, iT, iT);
        if (sp) {
            Fb += -g * x.Cp / CpM;
            Cpb += -g * F / CpM;
            Hab += -g * L.f.J[lane * P + iT] / CpM;
        }
        if (lane == iT) Fb += -g * x.dCpMdT / CpM;
        mb.dCpMdT += -g * FT / CpM;
        mb.CpM += -g * L.f.J[iT * P + iT] / CpM;
        jb_a(L, rank)[lane] = sp ? -g * x.Ha / CpM : 0.0;
        jb_b(L, rank)[lane] = (lane == iT) ? 1.0 : 0.0;
        ++rank;
    }
    wave_sync();
    // ddT_dtdY (:108-126): Jb(i, j) += -Jb(T, j) Ha_i / CpM, i and j species
    {
        // this lane's column of the temperature row
        const double gi = sp ? jb_entry(L, rank, iT, lane) : 0.0;
        jb_a(L, rank)[lane] = sp ? -x.Ha / CpM : 0.0;
        jb_b(L, rank)[lane] = gi;
        wave_sync();
        if (sp) {
            const double* gT = jb_b(L, rank);
            double h = 0.0;
            for (int i = 0; i < nS; ++i) h += gT[i] * L.f.J[lane * P + i];
            Hab += -h / CpM;
        }
        ++rank;
        double fbT = 0.0, cpm = 0.0;
        if (sp && gi != 0.0) {
            Cpb += -gi * FT / CpM;
            fbT = -gi * x.Cp / CpM;
            cpm = -gi * L.f.J[iT * P + lane] / CpM;
        }
        fbT = wave_sum(fbT);
        mb.CpM += wave_sum(cpm);
        if (lane == iT) Fb += fbT;
    }
    wave_sync();
    // ddY_dcdY (:65-103).  Column sums first, from the entries before the
    // row transforms: vb_j (species columns) and the temperature column's
    // concentration adjoints.
    double aF[kJbRank];
#pragma unroll
    for (int q = 0; q < kJbRank; ++q)
        aF[q] = (q < rank) ? wave_sum(sp ? jb_a(L, q)[lane] * F : 0.0) : 0.0;
    // this lane's row entry in the temperature column
    const double gTi = sp ? jb_entry(L, rank, lane, iT) : 0.0;
    const double gF = wave_sum(sp ? gTi * F : 0.0);
    L.cb[lane] = gTi;
    wave_sync();
    if (sp) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < kJbRank; ++q)
            if (q < rank) s += jb_b(L, q)[lane] * aF[q];
        vb = s * rho;
        // sum_i g_i (W_i/rho) t_ij with t the rows before the transform:
        // the transform made J_ij = (W_i/rho) t_ij (rho/W_j) + rho v_j F_i
        double gJ = 0.0;
        for (int i = 0; i < nS; ++i) gJ += L.cb[i] * L.f.J[i * P + lane];
        cbT = -(gJ - (rho * x.vi) * gF) / (rho / mr.W[lane]) * av;
    }
    wave_sync();
    double rhoPart = 0.0, avPart = 0.0;
    {
        // row sums with v_j, then the row and column scalings and the
        // temperature column's rank-one term
        double bv[kJbRank];
#pragma unroll
        for (int q = 0; q < kJbRank; ++q)
            bv[q] = (q < rank) ? wave_sum(sp ? jb_b(L, q)[lane] * L.vj[lane] : 0.0) : 0.0;
        if (sp) {
            const double WiR = mr.W[lane] / rho;
            double rs = 0.0;
#pragma unroll
            for (int q = 0; q < kJbRank; ++q)
                if (q < rank) rs += jb_a(L, q)[lane] * bv[q];
            rhoPart += rs * F;
            Fb += rho * rs;
            const double g = gTi;
            const double sumc = L.raw[lane];
            const double X = L.raw[kLanes + lane] - sumc * av;
            avPart = g * (F - WiR * sumc);
            rhoPart += -g * mr.W[lane] / (rho * rho) * X;
            Fb += g * av;
            const double rW = rho / mr.W[lane];
            for (int q = 0; q < rank; ++q) {
                jb_a(L, q)[lane] *= WiR;
                jb_b(L, q)[lane] *= rW;
            }
            jb_a(L, rank)[lane] = -g * WiR * av;
            jb_b(L, rank)[lane] = L.f.c[lane];
        } else {
            jb_a(L, rank)[lane] = 0.0;
            jb_b(L, rank)[lane] = 0.0;
        }
        ++rank;
    }
    mb.rho += wave_sum(rhoPart);
    mb.alphav += wave_sum(avPart);
    // dfdx: F_T = -sum F_i Ha_i / CpM, F_i = dN_i W_i / rho
    {
        const double g = readlane(Fb, iT);
        if (sp) {
            Fb += -g * x.Ha / CpM;
            Hab += -g * F / CpM;
        }
        mb.CpM += -g * FT / CpM;
        double part = 0.0;
        if (sp) {
            L.dNb[lane] = Fb * mr.W[lane] / rho;
            part = -Fb * F / rho;
        }
        mb.rho += wave_sum(part);
    }
    wave_sync();
    L.cb[lane] = sp ? cbT : 0.0;
    wave_sync();
    w_sets_begin<N>(mr, L, lane);
    // eff_s . b_q for the third-body reactions' D = sum_j eff_j u(j)
    const bool setProducts = mr.nSets * kJbRank <= 2 * kLanes;
    if (setProducts) {
        for (int t = lane; t < mr.nSets * kJbRank; t += kLanes) {
            const int s = t / kJbRank, q = t % kJbRank;
            double v = 0.0;
            if (q < rank) {
                const double* e = mr.effSets + s * kLanes;
                const double* b = jb_b(L, q);
                for (int j = 0; j < nS; ++j) v += e[j] * b[j];
            }
            *eb_slot(L, t) = v;
        }
        wave_sync();
    }
    // reactions (gpuReaction.H:547-702): adjoints of omega, dOmega/dT,
    // kf dCf/dc_j, kr dCr/dc_j and D from the raw-entry adjoints; a
    // reaction's combination of Jb rows is sum_q (sum_i nu_i a_q[i]) b_q
    double Tb = 0.0;
    const int nPass = (mr.nR + kLanes - 1) / kLanes;
#pragma unroll 1
    for (int pass = 0; pass < nPass; ++pass) {
        const int r = pass * kLanes + lane;
        if (r >= mr.nR) continue;
        const long o = (long)r * kWidth;
        const int nl = mr.nl[r], nr = mr.nr[r];
        adj::RxnBarT<kTeamWidth> in;
        in.fast = true;
        double al[kJbRank];
#pragma unroll
        for (int q = 0; q < kJbRank; ++q) {
            double v = 0.0;
            if (q < rank) {
                const double* aq = jb_a(L, q);
                for (int t = 0; t < kTeamWidth && t < nl; ++t) v -= mr.ls[o + t] * aq[spx(mr, mr.li[o + t])];
                for (int t = 0; t < kTeamWidth && t < nr; ++t) v += mr.rs[o + t] * aq[spx(mr, mr.ri[o + t])];
            }
            al[q] = v;
        }
        auto u = [&](int col) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < kJbRank; ++q)
                if (q < rank) v += al[q] * jb_b(L, q)[col];
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
            double D = 0.0;
            if (setProducts) {
                const int set = mr.effSet[r];
#pragma unroll
                for (int q = 0; q < kJbRank; ++q)
                    if (q < rank) D += al[q] * *eb_slot(L, set * kJbRank + q);
            } else {
                const double* eff = mr.eff + (long)r * nS;
                for (int j = 0; j < nS; ++j)
                    if (eff[j] != 0.0) D += eff[j] * u(j);
            }
            in.D = D;
        }
        double t6[adj::kParams] = {0, 0, 0, 0, 0, 0};
        w_reaction_reverse<N>(mr, L, r, T, false, in, Tb, t6);
        th.add(pass, lane, t6);
    }
    Tb = wave_sum(Tb);
    wave_sync();
    const double cb = sp ? L.cb[lane] + w_sets_scatter<N>(mr, L, lane) : 0.0;
    w_mixture_reverse(mr, p, y, T, x, cb, Cpb, Hab, vb, true, mb, Tb, ybar, lane);
    wave_sync();
}

// Reverse of one accepted step of size h from y0 (adjoint_math.hpp
// step_reverse): takes this lane's component of the adjoint of the step's
// result and returns that of y0.
template <int N>
__device__ __forceinline__ double w_step_reverse(const SolveArgs& A, const DeviceMech& mr,
                                                 WaveAdjLDS<N>& L, double p, double y0,
                                                 double h, double yb, const WTh& th,
                                                 int lane) {
    const DeviceMech& m = A.m;
    const int nS = m.nS, n = nS + 1;
    const RosTableau& t = A.tab;
    const double T0 = readlane(y0, nS);
    const bool inRange = T0 >= m.Tlow && T0 <= m.Thigh;
    WStages<N> S;
    (void)w_step<N, true>(A, L, p, y0, h, lane, S);
    double kb1 = t.b1 * yb, kb2 = t.b2 * yb, kb3 = t.b3 * yb, kb4 = t.b4 * yb;
    double Fb = 0.0, fb2 = 0.0, fb3 = 0.0;
    int rank = 0; // factor pairs of Jb
    if (t.stages == 4) {
        const double rb = w_solveT<N>(S.col, S.rdg, n, S.ps, kb4, lane);
        w_add_outer<N>(L, n, rb, S.k4, lane, rank);
        fb3 += rb;
        Fb += h * t.d4 * rb;
        kb1 += t.c41 / h * rb;
        kb2 += t.c42 / h * rb;
        kb3 += t.c43 / h * rb;
    }
    if (t.stages >= 3) {
        const double rb = w_solveT<N>(S.col, S.rdg, n, S.ps, kb3, lane);
        w_add_outer<N>(L, n, rb, S.k3, lane, rank);
        if (t.stages == 4) fb3 += rb;
        else fb2 += rb;
        Fb += h * t.d3 * rb;
        kb1 += t.c31 / h * rb;
        kb2 += t.c32 / h * rb;
    }
    if (t.stages == 4) {
        const double ys = y0 + t.a31 * S.k1 + t.a32 * S.k2;
        double ysb = 0.0;
        w_vjp_f<N>(mr, L, p, ys, fb3, S.f3, ysb, th, lane);
        yb += ysb;
        kb1 += t.a31 * ysb;
        kb2 += t.a32 * ysb;
    }
    {
        const double rb = w_solveT<N>(S.col, S.rdg, n, S.ps, kb2, lane);
        w_add_outer<N>(L, n, rb, S.k2, lane, rank);
        fb2 += rb;
        Fb += h * t.d2 * rb;
        kb1 += t.c21 / h * rb;
    }
    {
        const double ys = y0 + t.a21 * S.k1;
        double ysb = 0.0;
        w_vjp_f<N>(mr, L, p, ys, fb2, S.f2, ysb, th, lane);
        yb += ysb;
        kb1 += t.a21 * ysb;
    }
    const double rb = w_solveT<N>(S.col, S.rdg, n, S.ps, kb1, lane);
    w_add_outer<N>(L, n, rb, S.k1, lane, rank);
    Fb += h * t.d1 * rb;
    // within the reaction temperature limits f(y0) is the Jacobian call's F
    if (inRange) Fb += rb;
    else w_vjp_f<N>(mr, L, p, y0, rb, S.dydx0, yb, th, lane);
    w_jact<N>(mr, L, p, y0, S.F, Fb, yb, th, lane, rank);
    return yb;
}

} // namespace

// Reverse sweep, one wavefront per cell (adjoint.hip adjoint_reverse_kernel).
// A.m: the wavefront solve's tables.  seg: the K-state segment buffers
// (segPer doubles per cell, wave-blocked).  Ybar/Y0bar: [n][ldb]; thBar:
// [6 nR][ldb], reactions in reference order.  pacc: the parameter
// accumulators, cell c at pacc + (c/64)*paccSlot + (c%64)*paccPer, each
// kParams*64*nPass doubles long.
template <int N>
__global__ void __launch_bounds__(64)
wave_adjoint_reverse_kernel(SolveArgs A, T16Record R, double* seg, long segPer,
                            const double* Ybar, long ldb, double* Y0bar, double* thBar,
                            double* pacc, long paccSlot, long paccPer, int* queue) {
    __shared__ WaveAdjLDS<N> L;
    const int lane = threadIdx.x;
    const DeviceMech& mr = A.m;
    const int nS = A.m.nS, n = nS + 1;
    for (;;) {
        int c = 0;
        if (lane == 0) c = atomicAdd(queue, 1);
        int cell = __builtin_amdgcn_readfirstlane(__shfl(c, 0));
        if (cell >= A.nCells) break;
        cell = bound(cell, A.nCells, mr.bounds, kBoundCell);
        const double* y0c = rec_slot(R.y0c, cell, n + 1);
        const double p = y0c[(long)n * kRecStride];
        const double T0 = y0c[(long)nS * kRecStride];
        double yb = (lane < n) ? Ybar[lane * ldb + cell] : 0.0;
        const int nAccP = adj::kParams * kLanes * mr.nPass;
        const WTh th{thBar ? pacc + (cell / kLanes) * paccSlot + (cell % kLanes) * paccPer
                           : nullptr};
        if (th.acc)
            for (int q = lane; q < nAccP; q += kLanes)
                __hip_atomic_store(&th.acc[q], 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (T0 > A.Treact && A.deltaT > kSmall) {
            const double* yE = rec_slot(R.yEnd, cell, n);
            if (lane < nS && yE[lane * kRecStride] < 0.0) yb = 0.0;
            const int nAcc = bound(R.nAcc[cell], R.maxSteps + 1, mr.bounds, kBoundStep);
            const double* ck = rec_slot(R.ck, cell, R.nck * n);
            const double* hs = rec_slot(R.hs, cell, R.maxSteps);
            double* sg = rec_slot(seg, cell, segPer);
            const int nseg = (nAcc + R.K - 1) / R.K;
            for (int g = nseg - 1; g >= 0; --g) {
                const int s0 = g * R.K, s1 = min(s0 + R.K, nAcc);
                double y = (lane < n) ? ck[((long)bound(g, R.nck, mr.bounds, kBoundSlot) * n + lane) *
                                             kRecStride]
                                        : 0.0;
                if (lane < n) sg[lane * kRecStride] = y;
                WStages<N> S;
                for (int t = s0; t < s1 - 1; ++t) {
                    y = w_step<N, false>(A, L, p, y, hs[(long)t * kRecStride], lane, S);
                    if (lane < n) sg[((long)(t - s0 + 1) * n + lane) * kRecStride] = y;
                }
                for (int t = s1 - 1; t >= s0; --t) {
                    const double y0 = (lane < n) ? sg[((long)(t - s0) * n + lane) * kRecStride] : 0.0;
                    yb = w_step_reverse<N>(A, mr, L, p, y0, hs[(long)t * kRecStride], yb, th,
                                           lane);
                }
            }
        }
        if (lane < n) Y0bar[lane * ldb + cell] = yb;
        if (th.acc)
            for (int pass = 0; pass < mr.nPass; ++pass) {
                const int r = pass * kLanes + lane;
                if (r >= mr.nR) break;
                const long ref = mr.rxnRef[r];
#pragma unroll
                for (int q = 0; q < adj::kParams; ++q)
                    thBar[(adj::kParams * ref + q) * ldb + cell] =
                        __hip_atomic_load(&th.acc[(pass * adj::kParams + q) * kLanes + lane],
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        wave_sync();
    }
}

template __global__ void wave_adjoint_reverse_kernel<32>(SolveArgs, T16Record, double*, long,
                                                         const double*, long, double*, double*,
                                                         double*, long, long, int*);
template __global__ void wave_adjoint_reverse_kernel<54>(SolveArgs, T16Record, double*, long,
                                                         const double*, long, double*, double*,
                                                         double*, long, long, int*);
template __global__ void wave_adjoint_reverse_kernel<64>(SolveArgs, T16Record, double*, long,
                                                         const double*, long, double*, double*,
                                                         double*, long, long, int*);

} // namespace gfc
