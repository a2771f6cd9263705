// Workgroup-per-cell chemistry kernel for large mechanisms (64 < n <= 208).
//
// One workgroup of 512 threads (8 wavefronts, two per SIMD) integrates one
// cell; a persistent grid takes cells from an atomic queue.  The n x n
// Jacobian (n = nSpecie+1; 201 for 200 species) does not fit in LDS, so it
// lives in registers, distributed 2D block-cyclically over a 16 x 32 grid of
// threads: thread (tr, tc) holds rows tr + 16a (a < NA) and columns
// tc + 32b (b < NC).  A wavefront holds four thread-columns, so every
// column (and the pivot search over it) is inside one wavefront.
//
// The steps of gpuFoam's per-cell solver (gpu_chemistry/src/gpuOde/*.H) map
// as follows:
//   * mixture and per-specie thermo: specie-parallel, block sums;
//   * reaction rates: reaction-parallel over the type-sorted records; the
//     net rates go to LDS and dN/dt is a specie-parallel gather over each
//     specie's reactions (no atomics, so the sums are deterministic);
//   * Jacobian (gpuODESystem.H:198-281 with gpuReaction.H:547-612): the
//     per-reaction concentration derivatives go to LDS; each thread gathers
//     its own entries from host-built lists (sparse part per entry, dense
//     third-body part per row), then applies the Y / T transforms; the
//     result is kept in global scratch and reloaded by every attempt;
//   * LU (ludecompose.H:9-92): right-looking elimination with the
//     reference's implicit row scaling and pivot rule (largest scaled value,
//     ties to the last row position, zero pivot -> DBL_EPSILON); rows are
//     not moved, a position map stands for the swaps.  One barrier per
//     column: the owner wavefront picks the pivot and writes the
//     multipliers to LDS, then every thread updates its entries;
//   * back-substitution (lusolve.H): blocked by 32 positions; a diagonal
//     block is solved by one wavefront from an LDS copy, the off-diagonal
//     products are per-thread partial sums reduced in a fixed order.
#include <hip/hip_runtime.h>

#include "team_common.hpp"

namespace gfc {

namespace {

using namespace wave;

constexpr int kBT = 512;          // threads per workgroup
constexpr int kBW = kBT / 64;     // wavefronts per workgroup
constexpr int kPR = 16;           // thread rows
constexpr int kPC = 32;           // thread columns

template <int NA, int NC>
struct BlockLDS {
    static constexpr int NV = kPR * NA; // vector capacity (>= n)
    static constexpr int RSZ = (NC * 32 * 33 > kPC * NV) ? NC * 32 * 33 : kPC * NV;
    // diagonal blocks of the factors [NC][32][33] / row partial sums [32][NV]
    double R[RSZ];
    double part[kPC * 32];  // solve partial sums [tc][position in block]
    double y[NV], dydx0[NV], dfdx[NV], dydx[NV], yt[NV], rhs[NV], xs[NV];
    double k1[NV], k2[NV], k3[NV], k4[NV];
    double c[NV], Ha[NV], Cp[NV], colA[NV], colB[NV], JT[NV], aux[NV], rowT[NV], vv[NV];
    double gL[NV], gH[NV], hL[NV], hH[NV];
    double w[kBlockMaxR];    // net rate of each reaction
    double dwdT[kBlockMaxR]; // its temperature derivative (Jacobian only)
    double G[kBlockMaxR];    // dense third-body weight k (Cf - Cr/Kc)
    double M[kBlockMaxSets];
    double colbuf[2][NV];
    double rdiag[NV];
    double rowbuf[kBW][4 * NC];
    double red[2][kBW * 4];
    unsigned long long vmax[NV];
    int perm[NV], pos[NV], poison[NV];
    int piv[2];
    int ppiv[2][4]; // pivot rows of the current / next 4-column panel
    int cell;
};

// Grid coordinates of this thread.
struct Coord {
    int tid, w, lane, g, tr, tc;
};

__device__ __forceinline__ Coord coord() {
    Coord q;
    q.tid = threadIdx.x;
    q.w = q.tid >> 6;
    q.lane = q.tid & 63;
    q.g = q.lane >> 4;
    q.tr = q.lane & 15;
    q.tc = q.w * 4 + q.g;
    return q;
}

// The thread coordinates behind an empty asm: values derived from them
// (per-entry addresses) are then recomputed where they are used instead of
// being hoisted out of the cell loop and kept live (spilled) throughout.
__device__ __forceinline__ Coord opaque(Coord q) {
    asm volatile("" : "+v"(q.tid), "+v"(q.w), "+v"(q.lane), "+v"(q.g), "+v"(q.tr), "+v"(q.tc));
    return q;
}

// Entry (aP, b) of the per-thread block for a uniform aP.  The row values
// are read unconditionally and the choice is a switch on the scalar index,
// so the register block stays in registers (a chain of conditional
// accesses would be merged into one access through a selected address,
// which forces the block into scratch memory).
template <int NA, int NC, int b>
__device__ __forceinline__ double row_pick(const double (&A)[NA][NC], int aP) {
    static_assert(NA <= 16, "row_pick handles up to 16 rows");
    double v[16];
#pragma unroll
    for (int a = 0; a < 16; ++a) v[a] = (a < NA) ? A[a < NA ? a : 0][b] : 0.0;
    double r;
    switch (aP) {
    case 0: r = v[0]; break;
    case 1: r = v[1]; break;
    case 2: r = v[2]; break;
    case 3: r = v[3]; break;
    case 4: r = v[4]; break;
    case 5: r = v[5]; break;
    case 6: r = v[6]; break;
    case 7: r = v[7]; break;
    case 8: r = v[8]; break;
    case 9: r = v[9]; break;
    case 10: r = v[10]; break;
    case 11: r = v[11]; break;
    case 12: r = v[12]; break;
    case 13: r = v[13]; break;
    case 14: r = v[14]; break;
    default: r = v[15]; break;
    }
    return r;
}

template <int NA, int NC, int b = 0>
__device__ __forceinline__ void row_out(const double (&A)[NA][NC], int aP, int bK, double* rb) {
    if constexpr (b < NC) {
        if (b >= bK) rb[b] = row_pick<NA, NC, b>(A, aP);
        row_out<NA, NC, b + 1>(A, aP, bK, rb);
    }
}

template <int NA, int NC, int b = 0>
__device__ __forceinline__ void row_vals(const double (&A)[NA][NC], int aP, double (&out)[NC]) {
    if constexpr (b < NC) {
        out[b] = row_pick<NA, NC, b>(A, aP);
        row_vals<NA, NC, b + 1>(A, aP, out);
    }
}

// Per-thread register images in global scratch, [wavefront][entry][lane]:
// the wavefront's base is uniform and an entry adds a constant, so the
// per-entry addresses need no registers of their own.
template <int NE>
__device__ __forceinline__ double* wave_slots(double* buf, const Coord& q) {
    return buf + (long)__builtin_amdgcn_readfirstlane(q.w) * NE * 64;
}

// Sums of K values over the workgroup (wavefront DPP sums, then the eight
// wavefront partials in a fixed order).  Alternating LDS buffers let two
// consecutive calls run without an extra barrier.
template <int NA, int NC, int K>
__device__ __forceinline__ void block_sum(double (&v)[K], BlockLDS<NA, NC>& L, int& tog,
                                          const Coord& q) {
    double* r = L.red[tog];
    tog ^= 1;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double s = wave_sum(v[k]);
        if (q.lane == 0) r[q.w * 4 + k] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBW; ++w) s += r[w * 4 + k];
        v[k] = s;
    }
}

template <int NA, int NC>
__device__ __forceinline__ double block_max(double v, BlockLDS<NA, NC>& L, int& tog,
                                            const Coord& q) {
    double* r = L.red[tog];
    tog ^= 1;
    const double s = wave_max(v);
    if (q.lane == 0) r[q.w * 4] = s;
    __syncthreads();
    double m = r[0];
#pragma unroll
    for (int w = 1; w < kBW; ++w) m = fmax(m, r[w * 4]);
    return m;
}

// Third-body concentrations of every efficiency set (specie order).
template <int NA, int NC>
__device__ __forceinline__ void third_body_sets(const DeviceMech& m, BlockLDS<NA, NC>& L,
                                                const Coord& q) {
    const int nS = m.nS, nSets = m.nSets;
    for (int s = q.tid; s < nSets; s += kBT) {
        const double* e = m.blk.effT + s;
        double acc = 0.0;
#pragma unroll 4
        for (int i = 0; i < nS; ++i) acc += e[(long)i * nSets] * L.c[i];
        L.M[s] = acc;
    }
}

template <int NA, int NC, bool ENTHALPY>
__device__ __forceinline__ void gibbs(const DeviceMech& m, BlockLDS<NA, NC>& L,
                                      const TFun& f, int i) {
    const double W = m.W[i];
    const double* lo = m.low + 7 * i;
    const double* hi = m.high + 7 * i;
    L.gL[i] = W * poly_Gstd(lo, f.T, f.logT);
    L.gH[i] = W * poly_Gstd(hi, f.T, f.logT);
    if (ENTHALPY) {
        L.hL[i] = W * poly_Ha(lo, f.T);
        L.hH[i] = W * poly_Ha(hi, f.T);
    }
}

// ---------------------------------------------------------- derivatives --
// gpuODESystem::derivatives (gpuODESystem.H:164-196) of the state yv (LDS)
// into out (LDS).  Ends with a barrier.
template <int NA, int NC>
__device__ __forceinline__ void blk_derivatives(const DeviceMech& m, double p, const double* yv,
                                double* out, BlockLDS<NA, NC>& L, int& tog, const Coord& q0) {
    const Coord q = opaque(q0);
    const int nS = m.nS, t = q.tid;
    const double T = yv[nS];
    double s[2] = {0.0, 0.0};
    double Ha = 0.0, cu = 0.0;
    if (t < nS) {
        const double yi = fmax(yv[t], 0.0);
        const double rhoi = p / (m.R[t] * T);
        const double* a = coeffs_of(m.high + 7 * t, m.low + 7 * t, m.Tcommon[t], T);
        s[0] = yi / rhoi;
        s[1] = yi * poly_Cp(a, T);
        Ha = poly_Ha(a, T);
        cu = yi / m.W[t];
    }
    block_sum(s, L, tog, q);
    const double rhoM = 1.0 / s[0];
    const double CpM = s[1];
    // omega() evaluates the rates at T clipped to [Tlow, Thigh]
    const TFun f = tfun(fmin(fmax(T, m.Tlow), m.Thigh));
    if (t < nS) {
        L.c[t] = cu * rhoM;
        gibbs<NA, NC, false>(m, L, f, t);
    }
    __syncthreads();
    third_body_sets(m, L, q);
    __syncthreads();
    for (int r = t; r < m.nR; r += kBT) {
        const Rxn x = load_rxn(m, r >> 6, r & 63);
        double kArr, ddT, pw[kTeamWidth];
        const double kf = forward_rate(x, f, L.M, kArr, ddT);
        const double kr = x.irr ? 0.0 : kf * exp(-ln_kc(x, f, L, nullptr));
        const double Cf = side_powers(x.li, x.lei, x.le, x.nl, L.c, pw);
        const double Cr = side_powers(x.ri, x.rei, x.re, x.nr, L.c, pw);
        L.w[r] = kf * Cf - kr * Cr;
    }
    __syncthreads();
    double d[1] = {0.0};
    double dY = 0.0;
    if (t < nS) {
        double dN = 0.0;
        const int e1 = m.blk.spPtr[t + 1];
#pragma unroll 4
        for (int e = m.blk.spPtr[t]; e < e1; ++e) dN += m.blk.spNu[e] * L.w[m.blk.spR[e]];
        dY = dN * (m.W[t] / rhoM);
        d[0] = dY * Ha;
    }
    block_sum(d, L, tog, q);
    if (t < nS) out[t] = dY;
    if (t == nS) out[nS] = -d[0] / CpM;
    __syncthreads();
}

// ------------------------------------------------------------- jacobian --
// gpuODESystem::jacobian (gpuODESystem.H:198-281): leaves J (n x n) in the
// per-thread layout of Jg and its derivative (the reference's dfdx) in
// L.dfdx.  The per-reaction derivatives D[q][r] (q < 6 specie positions)
// occupy L.R during the assembly.
template <int NA, int NC>
__device__ __forceinline__ void blk_jacobian(const DeviceMech& m, double p, BlockLDS<NA, NC>& L,
                             long nRp, double* Jg, int& tog, const Coord& q0) {
    const Coord q = opaque(q0);
    constexpr int NV = BlockLDS<NA, NC>::NV;
    const int nS = m.nS, t = q.tid;
    const double T = L.y[nS];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    double Ha = 0.0, Cp = 0.0, cu = 0.0;
    if (t < nS) {
        const double yi = fmax(L.y[t], 0.0);
        const double rhoi = p / (m.R[t] * T);
        const double vi = 1.0 / rhoi;
        const double* a = coeffs_of(m.high + 7 * t, m.low + 7 * t, m.Tcommon[t], T);
        Cp = poly_Cp(a, T);
        Ha = poly_Ha(a, T);
        s[0] = yi / rhoi;
        s[1] = yi * Cp;
        s[2] = yi * poly_dCpdT(a, T);
        s[3] = yi * vi * (1.0 / T);
        cu = yi / m.W[t];
    }
    block_sum(s, L, tog, q);
    const double rhoM = 1.0 / s[0];
    const double CpM = s[1];
    const double dCpMdT = s[2];
    const double alphavM = s[3] * rhoM;
    const TFun f = tfun(T);
    if (t < nS) {
        L.c[t] = cu * rhoM;
        L.Ha[t] = Ha;
        L.Cp[t] = Cp;
        L.colA[t] = rhoM / m.W[t];
        L.colB[t] = rhoM * (1.0 / (p / (m.R[t] * T)));
        gibbs<NA, NC, true>(m, L, f, t);
    }
    for (int i = t; i < NV; i += kBT) L.poison[i] = 0;
    __syncthreads();
    third_body_sets(m, L, q);
    __syncthreads();
    // reaction-parallel: rates, dN/dt terms and concentration derivatives
    double* D = L.R;
    for (int r = t; r < m.nR; r += kBT) {
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
        L.w[r] = w;
        L.dwdT[r] = dwdT;
        L.G[r] = (x.type == GFC_RATE_THIRD_BODY) ? kArr * temp : 0.0;
        // dw/dc of the specie at each position (left side kf dCf/dc, right
        // side -kr dCr/dc) in D[q][r] = L.R[q*nRp + r]
#pragma unroll
        for (int j = 0; j < kTeamWidth; ++j) {
            D[j * nRp + r] =
                (j < x.nl) ? side_dpow(x.li, x.lei, x.le, x.nl, j, L.c, pwl) * kf : 0.0;
            D[(kTeamWidth + j) * nRp + r] =
                (!x.irr && j < x.nr) ? -(side_dpow(x.ri, x.rei, x.re, x.nr, j, L.c, pwr) * kr)
                                     : 0.0;
        }
        // a fall-off reaction with a non-finite Cf - Cr/Kc makes the
        // reference's dense (zero) concentration derivative NaN for its rows
        if (x.type > GFC_RATE_THIRD_BODY && !isfinite(temp)) {
#pragma unroll
            for (int j = 0; j < kTeamWidth; ++j) {
                if (j < x.nl) L.poison[x.li[j]] = 1;
                if (j < x.nr) L.poison[x.ri[j]] = 1;
            }
        }
    }
    __syncthreads();
    // dN/dt and the temperature column, specie-parallel gathers
    double d = 0.0;
    double sT[1] = {0.0};
    if (t < nS) {
        double dN = 0.0, JT = 0.0;
        const int e1 = m.blk.spPtr[t + 1];
#pragma unroll 4
        for (int e = m.blk.spPtr[t]; e < e1; ++e) {
            const int r = m.blk.spR[e];
            const double nu = m.blk.spNu[e];
            dN += nu * L.w[r];
            JT += nu * L.dwdT[r];
        }
        d = dN * (m.W[t] / rhoM);
        L.dfdx[t] = d;
        L.JT[t] = JT;
        sT[0] = d * Ha;
    }
    block_sum(sT, L, tog, q);
    const double dfdxT = -sT[0] / CpM;
    if (t == nS) L.dfdx[nS] = dfdxT;
    // this thread's entries: the concentration derivatives gathered from
    // the per-reaction values in LDS (sparse part in reaction order, then
    // the dense third-body part), transformed (ddY_dcdY,
    // gpuODESystem.H:65-103) and stored; the row sums of the temperature
    // column and the column sums of the temperature row (ddT_dtdY, :108-126)
    // are accumulated on the way
    double* Jw = wave_slots<NA * NC>(Jg, q);
    double rp[NA], cs[NC];
#pragma unroll
    for (int b = 0; b < NC; ++b) cs[b] = 0.0;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int i = q.tr + kPR * a;
        rp[a] = 0.0;
        double v[NC];
#pragma unroll
        for (int b = 0; b < NC; ++b) v[b] = 0.0;
        if (i < nS) {
            int e0[NC], e1[NC];
#pragma unroll
            for (int b = 0; b < NC; ++b) {
                const int j = q.tc + kPC * b;
                const int* ep = m.blk.elPtr + (long)i * nS + (j < nS ? j : 0);
                e0[b] = ep[0];
                e1[b] = (j < nS) ? ep[1] : e0[b];
            }
#pragma unroll
            for (int b = 0; b < NC; ++b) {
#pragma unroll 2
                for (int e = e0[b]; e < e1[b]; ++e) {
                    const int rs = m.blk.elRS[e];
                    v[b] += m.blk.elNu[e] * D[(rs & 7) * nRp + (rs >> 3)];
                }
            }
            const int te = m.blk.tbPtr[i + 1];
            for (int e = m.blk.tbPtr[i]; e < te; ++e) {
                const double g = m.blk.tbNu[e] * L.G[m.blk.tbR[e]];
                const double* ef = m.blk.effB + (long)m.blk.tbSet[e] * nS;
#pragma unroll
                for (int b = 0; b < NC; ++b) {
                    const int j = q.tc + kPC * b;
                    if (j < nS) v[b] += g * ef[j];
                }
            }
            const bool bad = L.poison[i] != 0;
            const double WiByrhoM = m.W[i] / rhoM;
            const double di = L.dfdx[i];
            const double Hai = L.Ha[i];
#pragma unroll
            for (int b = 0; b < NC; ++b) {
                const int j = q.tc + kPC * b;
                if (j < nS) {
                    const double tj = bad ? __longlong_as_double(0x7ff8000000000000LL) : v[b];
                    rp[a] += tj * L.c[j];
                    v[b] = WiByrhoM * tj * L.colA[j] + L.colB[j] * di;
                    cs[b] += v[b] * Hai;
                }
            }
        }
#pragma unroll
        for (int b = 0; b < NC; ++b) Jw[(a * NC + b) * 64 + q.lane] = v[b];
    }
    // temperature row: column sums over the 16 thread-rows of the column
#pragma unroll
    for (int b = 0; b < NC; ++b) {
        const int j = q.tc + kPC * b;
        double v = cs[b];
        v += dpp<0x111>(0.0, v);
        v += dpp<0x112>(0.0, v);
        v += dpp<0x114>(0.0, v);
        v += dpp<0x118>(0.0, v);
        v = __shfl(v, (q.lane & ~15) | 15);
        if (j < nS && q.tr == (nS & 15)) {
#pragma unroll
            for (int a = 0; a < NA; ++a)
                if (q.tr + kPR * a == nS)
                    Jw[(a * NC + b) * 64 + q.lane] = (-v - L.Cp[j] * dfdxT) / CpM;
        }
    }
    __syncthreads();
    // ddNdT: the row sums reduced through LDS in thread-column order
#pragma unroll
    for (int a = 0; a < NA; ++a) L.R[q.tc * NV + q.tr + kPR * a] = rp[a];
    __syncthreads();
    if (t < nS) {
        double acc = 0.0;
#pragma unroll 8
        for (int c = 0; c < kPC; ++c) acc += L.R[c * NV + t];
        const double ddNdT = L.JT[t] - acc * alphavM;
        L.aux[t] = (m.W[t] / rhoM) * ddNdT + alphavM * L.dfdx[t];
    }
    __syncthreads();
    // temperature column: specie rows from aux, then ddT_dtdT (:129-146)
    const bool colT = q.tc == (nS & 31);
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int i = q.tr + kPR * a;
#pragma unroll
        for (int b = 0; b < NC; ++b)
            if (colT && b == (nS >> 5) && i < nS) Jw[(a * NC + b) * 64 + q.lane] = L.aux[i];
    }
    double sTT[1] = {0.0};
    if (t < nS) sTT[0] = L.dfdx[t] * L.Cp[t] + L.aux[t] * L.Ha[t];
    block_sum(sTT, L, tog, q);
    const double JTT = (-sTT[0] - dfdxT * dCpMdT) / CpM;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int i = q.tr + kPR * a;
#pragma unroll
        for (int b = 0; b < NC; ++b)
            if (colT && b == (nS >> 5) && i == nS) Jw[(a * NC + b) * 64 + q.lane] = JTT;
    }
}

// --------------------------------------------------------------- LU ----
// Columns 32 bK .. 32 bK + 31 of the factorisation (bK fixed at compile
// time so that every register entry is addressed statically).
template <int NA, int NC, int bK>
__device__ __forceinline__ void lu_cols(double (&A)[NA][NC], int n, BlockLDS<NA, NC>& L,
                                        const Coord& q0, unsigned& pivoted) {
    const Coord q = opaque(q0);
    for (int kk = 0; kk < kPC; ++kk) {
        const int k = kPC * bK + kk;
        if (k >= n) break;
        const int par = k & 1;
        if (q.tc == kk) {
            // pivot search over column k: (scaled value, position),
            // larger wins; rows already pivots and padding never win,
            // NaN keys lose to every number.
            double vvv[NA];
#pragma unroll
            for (int a = 0; a < NA; ++a) vvv[a] = L.vv[q.tr + kPR * a];
            // keys and their maximum over the 16 lanes (row_ror)
            double keyv[NA];
            double kmax = -2.0;
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int i = q.tr + kPR * a;
                const bool cand = i < n && !((pivoted >> a) & 1u);
                double key = vvv[a] * fabs(A[a][bK]);
                if (!(key >= 0.0)) key = -1.0;
                keyv[a] = cand ? key : -2.0;
                kmax = fmax(kmax, keyv[a]);
            }
            kmax = fmax(kmax, dpp<0x121>(kmax, kmax));
            kmax = fmax(kmax, dpp<0x122>(kmax, kmax));
            kmax = fmax(kmax, dpp<0x124>(kmax, kmax));
            kmax = fmax(kmax, dpp<0x128>(kmax, kmax));
            // (lane, row) pairs holding the maximum: the usual single one
            // is the pivot; several (a tie, or only NaN/zero keys) go to the
            // position rule below
            int nm = 0, aStar = 0;
            unsigned long long mStar = 0ull;
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const unsigned long long mk = __ballot(keyv[a] == kmax);
                nm += __builtin_popcountll(mk);
                if (mk) {
                    aStar = a;
                    mStar = mk;
                }
            }
            int Pw, p;
            double dj;
            if (nm == 1) {
                const int lStar = __builtin_ctzll(mStar);
                p = (lStar & 15) + kPR * aStar;
                dj = readlane(row_pick<NA, NC, bK>(A, aStar), lStar);
                Pw = -1; // read from the position map below
            } else {
                // largest (key, position); the row and value travel along
                double bkey = -2.0, bval = 0.0;
                int bpos = -1, brow = 0;
#pragma unroll
                for (int a = 0; a < NA; ++a) {
                    const int i = q.tr + kPR * a;
                    const int ps = (keyv[a] > -2.0) ? L.pos[i] : -1;
                    const bool better = keyv[a] > bkey || (keyv[a] == bkey && ps > bpos);
                    bkey = better ? keyv[a] : bkey;
                    bpos = better ? ps : bpos;
                    brow = better ? i : brow;
                    bval = better ? A[a][bK] : bval;
                }
#define GFC_ROR_STEP(CTRL)                                                   \
    {                                                                        \
        const double ok = dpp<CTRL>(bkey, bkey);                             \
        const int op = dppi<CTRL>(bpos, bpos);                               \
        const int orw = dppi<CTRL>(brow, brow);                              \
        const double ov = dpp<CTRL>(bval, bval);                             \
        const bool better = ok > bkey || (ok == bkey && op > bpos);          \
        bkey = better ? ok : bkey;                                           \
        bpos = better ? op : bpos;                                           \
        brow = better ? orw : brow;                                          \
        bval = better ? ov : bval;                                           \
    }
                GFC_ROR_STEP(0x121)
                GFC_ROR_STEP(0x122)
                GFC_ROR_STEP(0x124)
                GFC_ROR_STEP(0x128)
#undef GFC_ROR_STEP
                Pw = __builtin_amdgcn_readfirstlane(bpos);
                p = __builtin_amdgcn_readfirstlane(brow);
                dj = readlane(bval, __builtin_amdgcn_readfirstlane(q.lane));
            }
            const int aP = p >> 4;
            if (dj == 0.0) {
                dj = kSmall;
                const bool me = q.tr == (p & 15);
#pragma unroll
                for (int a = 0; a < NA; ++a) A[a][bK] = (me && a == aP) ? kSmall : A[a][bK];
            }
            const double rDiag = 1.0 / dj;
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int i = q.tr + kPR * a;
                double l = 0.0;
                if (i < n && i != p && !((pivoted >> a) & 1u)) {
                    l = A[a][bK] * rDiag;
                    A[a][bK] = l;
                }
                L.colbuf[par][q.tr * NA + a] = l;
            }
            if (q.tr == 0) {
                if (Pw < 0) Pw = L.pos[p];
                const int qrow = L.perm[k];
                L.perm[Pw] = qrow;
                L.pos[qrow] = Pw;
                L.perm[k] = p;
                L.pos[p] = k;
                L.piv[par] = p;
                L.rdiag[k] = rDiag;
            }
        }
        __syncthreads();
        const int p = __builtin_amdgcn_readfirstlane(L.piv[par]);
        const int aP = p >> 4;
        const bool own = q.tr == (p & 15);
        if (own) pivoted |= 1u << aP;
        // pivot row, broadcast inside each wavefront through LDS
        double* rb = L.rowbuf[q.w] + q.g * NC;
        if (own) row_out(A, aP, bK, rb);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double u[NC];
#pragma unroll
        for (int b = bK; b < NC; ++b) u[b] = rb[b];
        const double* lp = &L.colbuf[par][q.tr * NA];
        const bool colLive = q.tc > kk;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const double la = lp[a];
            if (colLive) A[a][bK] = fma(-la, u[bK], A[a][bK]);
#pragma unroll
            for (int b = bK + 1; b < NC; ++b) A[a][b] = fma(-la, u[b], A[a][b]);
        }
        // the row broadcast buffer is rewritten next step: keep the
        // wavefront's reads ahead of those writes
        __builtin_amdgcn_wave_barrier();
    }
}

// Memory ordering of LDS among the lanes of one wavefront.
__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

// A[i][j] -= sum_p L[i][p] U[p][j] for register columns b >= B0 (and
// b == bK only where liveK), one pivot after the other, so that every entry
// sees its four fmas in pivot order and only one pivot row's values are
// held at a time.
template <int NA, int NC, int bK, int B0>
__device__ __forceinline__ void rank4_valu(double (&A)[NA][NC], const double* Lb,
                                           const double* Ub, int np, bool liveK,
                                           const Coord& q) {
    constexpr int NV = kPR * NA;
    if constexpr (B0 < NC) {
        for (int p = 0; p < np; ++p) {
            double u[NC];
#pragma unroll
            for (int b = B0; b < NC; ++b) u[b] = Ub[(p * 4 + q.g) * NC + b];
            const double* lp = Lb + p * NV + q.tr;
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const double la = lp[kPR * a];
#pragma unroll
                for (int b = B0; b < NC; ++b) {
                    if (b == bK) A[a][b] = liveK ? fma(-la, u[b], A[a][b]) : A[a][b];
                    else A[a][b] = fma(-la, u[b], A[a][b]);
                }
            }
        }
    }
}

// Columns 32 bK .. 32 bK + 31 as eight 4-column panels.  Panel columns
// k0 .. k0+3 sit in register column bK of the four thread columns of one
// wavefront wP = (k0 % 32) / 4, so that wavefront factorises the panel
// alone, with wavefront-level exchanges; one workgroup barrier publishes
// the panel's pivots and multipliers, and every thread then computes the
// four pivot rows' trailing entries U (in its thread column, by the lane
// holding each pivot row) and applies the rank-4 update
// A[i][j] -= sum_p L[i][p] U[p][j].  The arithmetic of every entry is the
// per-column algorithm's (the same fma sequence in the same order), so with
// vector FMAs the factors are bit-identical to lu_cols'.  With MFMA the
// rank-4 update is one v_mfma_f64_16x16x4f64 per 16 x 16 tile: a
// wavefront's entries A[a][b0..b0+3] are the accumulator tile
// D[4r + g][tr] of the transposed update D -= U^T L^T (lane l holds
// D[4r + l/16][l%16] in register r), and each lane loads one U and one L
// value from LDS as the tile's operands.
// Scratch in L.R (free during the factorisation): panel multipliers
// Lb[par][p][row] and per-wavefront pivot-row values Ub[w][p][g][b].
template <int NA, int NC, int bK, bool MFMA>
__device__ __forceinline__ void lu_panels(double (&A)[NA][NC], int n, BlockLDS<NA, NC>& L,
                                          const Coord& q0, unsigned& pivoted, int& par) {
    const Coord q = opaque(q0);
    constexpr int NV = BlockLDS<NA, NC>::NV;
    static_assert(2 * 4 * NV + kBW * 16 * NC <= BlockLDS<NA, NC>::RSZ, "panel scratch");
    for (int wP = 0; wP < kBW; ++wP) {
        const int k0 = kPC * bK + 4 * wP;
        if (k0 >= n) break;
        const int np = min(4, n - k0);
        double* Lb = L.R + par * (4 * NV);
        double* Ub = L.R + 2 * 4 * NV + q.w * (16 * NC);
        // ---- the panel, by wavefront wP
        if (q.w == wP) {
            for (int p = 0; p < np; ++p) {
                const int k = k0 + p;
                if (q.g == p) {
                    double keyv[NA];
                    double kmax = -2.0;
#pragma unroll
                    for (int a = 0; a < NA; ++a) {
                        const int i = q.tr + kPR * a;
                        const bool cand = i < n && !((pivoted >> a) & 1u);
                        double key = L.vv[i] * fabs(A[a][bK]);
                        if (!(key >= 0.0)) key = -1.0;
                        keyv[a] = cand ? key : -2.0;
                        kmax = fmax(kmax, keyv[a]);
                    }
                    kmax = fmax(kmax, dpp<0x121>(kmax, kmax));
                    kmax = fmax(kmax, dpp<0x122>(kmax, kmax));
                    kmax = fmax(kmax, dpp<0x124>(kmax, kmax));
                    kmax = fmax(kmax, dpp<0x128>(kmax, kmax));
                    // the ballots cover the whole wavefront: keep this
                    // group's lanes (its DPP row) only
                    const unsigned long long rowMask = 0xffffull << (16 * p);
                    int nm = 0, aStar = 0;
                    unsigned long long mStar = 0ull;
#pragma unroll
                    for (int a = 0; a < NA; ++a) {
                        const unsigned long long mk = __ballot(keyv[a] == kmax) & rowMask;
                        nm += __builtin_popcountll(mk);
                        if (mk) {
                            aStar = a;
                            mStar = mk;
                        }
                    }
                    int Pw, prow;
                    double dj;
                    if (nm == 1) {
                        const int lStar = __builtin_ctzll(mStar);
                        prow = (lStar & 15) + kPR * aStar;
                        dj = readlane(row_pick<NA, NC, bK>(A, aStar), lStar);
                        Pw = -1;
                    } else {
                        double bkey = -2.0, bval = 0.0;
                        int bpos = -1, brow = 0;
#pragma unroll
                        for (int a = 0; a < NA; ++a) {
                            const int i = q.tr + kPR * a;
                            const int ps = (keyv[a] > -2.0) ? L.pos[i] : -1;
                            const bool better = keyv[a] > bkey || (keyv[a] == bkey && ps > bpos);
                            bkey = better ? keyv[a] : bkey;
                            bpos = better ? ps : bpos;
                            brow = better ? i : brow;
                            bval = better ? A[a][bK] : bval;
                        }
#define GFC_ROR_STEP(CTRL)                                                   \
    {                                                                        \
        const double ok = dpp<CTRL>(bkey, bkey);                             \
        const int op = dppi<CTRL>(bpos, bpos);                               \
        const int orw = dppi<CTRL>(brow, brow);                              \
        const double ov = dpp<CTRL>(bval, bval);                             \
        const bool better = ok > bkey || (ok == bkey && op > bpos);          \
        bkey = better ? ok : bkey;                                           \
        bpos = better ? op : bpos;                                           \
        brow = better ? orw : brow;                                          \
        bval = better ? ov : bval;                                           \
    }
                        GFC_ROR_STEP(0x121)
                        GFC_ROR_STEP(0x122)
                        GFC_ROR_STEP(0x124)
                        GFC_ROR_STEP(0x128)
#undef GFC_ROR_STEP
                        // every lane of the row holds the result
                        Pw = bpos;
                        prow = brow;
                        dj = bval;
                    }
                    const int aP = prow >> 4;
                    if (dj == 0.0) {
                        dj = kSmall;
                        const bool me = q.tr == (prow & 15);
#pragma unroll
                        for (int a = 0; a < NA; ++a)
                            A[a][bK] = (me && a == aP) ? kSmall : A[a][bK];
                    }
                    const double rDiag = 1.0 / dj;
#pragma unroll
                    for (int a = 0; a < NA; ++a) {
                        const int i = q.tr + kPR * a;
                        double l = 0.0;
                        if (i < n && i != prow && !((pivoted >> a) & 1u)) {
                            l = A[a][bK] * rDiag;
                            A[a][bK] = l;
                        }
                        Lb[p * NV + i] = l;
                    }
                    if (q.tr == 0) {
                        if (Pw < 0) Pw = L.pos[prow];
                        const int qrow = L.perm[k];
                        L.perm[Pw] = qrow;
                        L.pos[qrow] = Pw;
                        L.perm[k] = prow;
                        L.pos[prow] = k;
                        L.ppiv[par][p] = prow;
                        L.rdiag[k] = rDiag;
                    }
                }
                wave_fence();
                const int prow = __builtin_amdgcn_readfirstlane(L.ppiv[par][p]);
                const int aP = prow >> 4;
                const bool own = q.tr == (prow & 15);
                // the pivot row's entries of the later panel columns
                if (own && q.g > p) Ub[q.g] = row_pick<NA, NC, bK>(A, aP);
                wave_fence();
                if (q.g > p) {
                    const double u = Ub[q.g];
#pragma unroll
                    for (int a = 0; a < NA; ++a)
                        A[a][bK] = fma(-Lb[p * NV + q.tr + kPR * a], u, A[a][bK]);
                }
                if (own) pivoted |= 1u << aP;
                wave_fence();
            }
        }
        __syncthreads();
        int prw[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            prw[p] = p < np ? __builtin_amdgcn_readfirstlane(L.ppiv[par][p]) : -1;
            if (p < np && q.w != wP && q.tr == (prw[p] & 15)) pivoted |= 1u << (prw[p] >> 4);
        }
        // trailing columns of this thread: register columns after bK, and
        // bK itself in the thread columns after the panel's
        const bool liveK = q.tc > 4 * wP + 3;
        // U rows: U[p][j] = A[piv_p][j] - sum_{s<p} L[piv_p][s] U[s][j], by
        // the lane holding row piv_p, pivot after pivot
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            if (p < np) {
                const int prow = prw[p];
                if (q.tr == (prow & 15)) {
                    double ls[4];
#pragma unroll
                    for (int s = 0; s < 4; ++s) ls[s] = (s < p) ? Lb[s * NV + prow] : 0.0;
                    double ur[NC];
                    row_vals<NA, NC>(A, prow >> 4, ur);
#pragma unroll
                    for (int b = bK; b < NC; ++b) {
                        if (b == bK && !liveK) continue;
                        double u = ur[b];
#pragma unroll
                        for (int s = 0; s < 4; ++s)
                            if (s < p) u = fma(-ls[s], Ub[(s * 4 + q.g) * NC + b], u);
                        Ub[(p * 4 + q.g) * NC + b] = u;
                    }
                }
                wave_fence();
            }
        }
        // rank-4 update of every row
        if constexpr (MFMA) {
            // full 4-column tiles by MFMA: lane (g, tr) supplies
            // B[p = g][n = tr] = L[row][g] and A[m = tr][p = g] =
            // -U[g][column (tr / 4, tr % 4) of the tile]
            constexpr int kFull = NC / 4 * 4;
#pragma unroll
            for (int b0 = (bK / 4) * 4; b0 < kFull; b0 += 4) {
                const int bu = b0 + (q.tr >> 2);   // register column of this lane's U operand
                const int gu = q.tr & 3;           // its thread-column group
                const bool uLive = q.g < np && (bu > bK || (bu == bK && 4 * q.w + gu > 4 * wP + 3));
                const double aop = uLive ? -Ub[(q.g * 4 + gu) * NC + bu] : 0.0;
#pragma unroll
                for (int a = 0; a < NA; ++a) {
                    const double bop = (q.g < np) ? Lb[q.g * NV + q.tr + kPR * a] : 0.0;
                    dbl4 c;
#pragma unroll
                    for (int r = 0; r < 4; ++r) c[r] = A[a][b0 + r];
                    const dbl4 d = __builtin_amdgcn_mfma_f64_16x16x4f64(aop, bop, c, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int b = b0 + r;
                        if (b > bK) A[a][b] = d[r];
                        else if (b == bK) A[a][b] = liveK ? d[r] : A[a][b];
                    }
                }
            }
            // the columns beyond the last full tile with vector FMAs
            rank4_valu<NA, NC, bK, (bK > kFull ? bK : kFull)>(A, Lb, Ub, np, liveK, q);
        } else {
            rank4_valu<NA, NC, bK, bK>(A, Lb, Ub, np, liveK, q);
        }
        par ^= 1;
    }
}

template <int NA, int NC, int LUV, int B = 0>
__device__ __forceinline__ void lu_all(double (&A)[NA][NC], int n, BlockLDS<NA, NC>& L,
                                       const Coord& q, unsigned& pivoted, int& par) {
    if constexpr (B < NC) {
        if (kPC * B < n) {
            if constexpr (LUV == 0) lu_cols<NA, NC, B>(A, n, L, q, pivoted);
            else lu_panels<NA, NC, B, LUV == 2>(A, n, L, q, pivoted, par);
        }
        lu_all<NA, NC, LUV, B + 1>(A, n, L, q, pivoted, par);
    }
}

// Factorises the per-thread entries A in place and stores them to LUg.  On
// return L.pos / L.perm map rows to pivot positions, L.rdiag holds the
// reciprocal pivots and the diagonal 32 x 32 blocks of the factors
// (position order) are in L.R.
template <int NA, int NC, int LUV>
__device__ __forceinline__ void blk_lu(double (&A)[NA][NC], int n, BlockLDS<NA, NC>& L,
                                       double* LUg, const Coord& q0) {
    const Coord q = opaque(q0);
    constexpr int NV = BlockLDS<NA, NC>::NV;
    for (int i = q.tid; i < NV; i += kBT) {
        L.vmax[i] = 0ull;
        L.pos[i] = i;
        L.perm[i] = i;
    }
    __syncthreads();
    // implicit scaling: 1 / max_j |a_ij| of the unfactorised rows
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int i = q.tr + kPR * a;
        double mx = 0.0;
#pragma unroll
        for (int b = 0; b < NC; ++b) {
            const double v = fabs(A[a][b]);
            if (v > mx) mx = v;
        }
        if (i < n && mx > 0.0)
            atomicMax(&L.vmax[i], (unsigned long long)__double_as_longlong(mx));
    }
    __syncthreads();
    for (int i = q.tid; i < n; i += kBT) L.vv[i] = 1.0 / __longlong_as_double((long long)L.vmax[i]);
    __syncthreads();
    unsigned pivoted = 0u;
    int par = 0;
    lu_all<NA, NC, LUV>(A, n, L, q, pivoted, par);
    __syncthreads();
    // diagonal blocks of the factors in position order: R[(b*32 + r)*33 + c]
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int i = q.tr + kPR * a;
        if (i < n) {
            const int P = L.pos[i];
#pragma unroll
            for (int b = 0; b < NC; ++b)
                if ((P >> 5) == b) L.R[(b * 32 + (P & 31)) * 33 + q.tc] = A[a][b];
        }
    }
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NC; ++b) wave_slots<NA * NC>(LUg, q)[(a * NC + b) * 64 + q.lane] = A[a][b];
    __syncthreads();
}

// One 32-position block of the forward substitution L y = P b (y in L.xs,
// position order): the block's diagonal part by wavefront 0, then every
// thread's off-diagonal products for the rows below, published as the next
// block's partial sums.
template <int NA, int NC, int B>
__device__ __forceinline__ void fwd_block(const double (&A)[NA][NC], int n,
                                          BlockLDS<NA, NC>& L, const double* bv,
                                          const int (&P)[NA], double (&acc)[NA],
                                          const Coord& q) {
    if (q.w == 0 && q.lane < 32) {
        const int r = q.lane;
        const int Pr = kPC * B + r;
        double s = 0.0;
        if (Pr < n) {
            s = bv[L.perm[Pr]];
            if (B > 0) {
                double ps = 0.0;
#pragma unroll 8
                for (int c = 0; c < kPC; ++c) ps += L.part[c * 32 + r];
                s -= ps;
            }
        }
        const double* D = L.R + (B * 32 + r) * 33;
        const int cn = min(32, n - kPC * B);
#pragma unroll
        for (int c0 = 0; c0 < 32; c0 += 8) {
            if (c0 < cn) {
                double d8[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) d8[u] = D[c0 + u];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int c = c0 + u;
                    if (c < cn) {
                        const double yc = readlane(s, c);
                        if (r > c) s = fma(-d8[u], yc, s);
                    }
                }
            }
        }
        if (Pr < n) L.xs[Pr] = s;
    }
    __syncthreads();
    const int j = q.tc + kPC * B;
    const double yj = (j < n) ? L.xs[j] : 0.0;
#pragma unroll
    for (int a = 0; a < NA; ++a)
        if (P[a] >= kPC * (B + 1)) acc[a] = fma(A[a][B], yj, acc[a]);
    if (B + 1 < NC) {
#pragma unroll
        for (int a = 0; a < NA; ++a)
            if ((P[a] >> 5) == B + 1) L.part[q.tc * 32 + (P[a] & 31)] = acc[a];
    }
    __syncthreads();
}

// One block of the backward substitution U x = y (x by column).
template <int NA, int NC, int B>
__device__ __forceinline__ void bwd_block(const double (&A)[NA][NC], int n, int nBk,
                                          BlockLDS<NA, NC>& L, double* x,
                                          const int (&P)[NA], double (&acc)[NA],
                                          const Coord& q) {
    if (q.w == 0 && q.lane < 32) {
        const int r = q.lane;
        const int Pr = kPC * B + r;
        double s = 0.0;
        if (Pr < n) {
            s = L.xs[Pr];
            if (B < nBk - 1) {
                double ps = 0.0;
#pragma unroll 8
                for (int c = 0; c < kPC; ++c) ps += L.part[c * 32 + r];
                s -= ps;
            }
        }
        const double* D = L.R + (B * 32 + r) * 33;
        const int cn = min(32, n - kPC * B);
        const double rd = (Pr < n) ? L.rdiag[Pr] : 0.0;
        double xr = 0.0;
#pragma unroll
        for (int c0 = 24; c0 >= 0; c0 -= 8) {
            if (c0 < cn) {
                double d8[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) d8[u] = D[c0 + u];
#pragma unroll
                for (int u = 7; u >= 0; --u) {
                    const int c = c0 + u;
                    if (c < cn) {
                        const double xc = readlane(s * rd, c);
                        if (r < c) s = fma(-d8[u], xc, s);
                        if (r == c) xr = xc;
                    }
                }
            }
        }
        if (Pr < n) x[Pr] = xr;
    }
    __syncthreads();
    const int j = q.tc + kPC * B;
    const double xj = (j < n) ? x[j] : 0.0;
#pragma unroll
    for (int a = 0; a < NA; ++a)
        if (P[a] >= 0 && P[a] < kPC * B) acc[a] = fma(A[a][B], xj, acc[a]);
    if (B > 0) {
#pragma unroll
        for (int a = 0; a < NA; ++a)
            if (P[a] >= 0 && (P[a] >> 5) == B - 1) L.part[q.tc * 32 + (P[a] & 31)] = acc[a];
    }
    __syncthreads();
}

template <int NA, int NC, int B = 0>
__device__ __forceinline__ void fwd_all(const double (&A)[NA][NC], int n, BlockLDS<NA, NC>& L,
                                        const double* bv, const int (&P)[NA],
                                        double (&acc)[NA], const Coord& q) {
    if constexpr (B < NC) {
        if (kPC * B < n) fwd_block<NA, NC, B>(A, n, L, bv, P, acc, q);
        fwd_all<NA, NC, B + 1>(A, n, L, bv, P, acc, q);
    }
}

template <int NA, int NC, int B = NC - 1>
__device__ __forceinline__ void bwd_all(const double (&A)[NA][NC], int n, int nBk,
                                        BlockLDS<NA, NC>& L, double* x, const int (&P)[NA],
                                        double (&acc)[NA], const Coord& q) {
    if constexpr (B >= 0) {
        if (B < nBk) bwd_block<NA, NC, B>(A, n, nBk, L, x, P, acc, q);
        bwd_all<NA, NC, B - 1>(A, n, nBk, L, x, P, acc, q);
    }
}

// Solves (LU) x = bv (bv indexed by row, x by column; both LDS) with the
// factors of blk_lu, read back from LUg (the derivative evaluations between
// the stages need the registers), blocked by 32 positions.  Ends with a
// barrier.
template <int NA, int NC>
__device__ __forceinline__ void blk_solve(const double* LUg, int n, BlockLDS<NA, NC>& L,
                                          const double* bv, double* x, const Coord& q0) {
    const Coord q = opaque(q0);
    double A[NA][NC];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NC; ++b) A[a][b] = wave_slots<NA * NC>(const_cast<double*>(LUg), q)[(a * NC + b) * 64 + q.lane];
    int P[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int i = q.tr + kPR * a;
        P[a] = (i < n) ? L.pos[i] : -1;
    }
    double acc[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[a] = 0.0;
    fwd_all(A, n, L, bv, P, acc, q);
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[a] = 0.0;
    bwd_all(A, n, (n + kPC - 1) / kPC, L, x, P, acc, q);
}

} // namespace

// ---------------------------------------------------------------- kernel --

template <int NA, int NC, int LUV>
__global__ void __launch_bounds__(kBT)
block_solve_kernel(SolveArgs A, int* queue, double* scratch, long perWG) {
    __shared__ BlockLDS<NA, NC> L;
    const Coord q = coord();
    const int t = q.tid;
    const DeviceMech& m = A.m;
    const int nS = m.nS, n = nS + 1;
    const long ld = A.ldY;
    const RosTableau& tb = A.tab;
    const long nRp = ((long)m.nR + 63) / 64 * 64;
    double* Jg = scratch + (long)blockIdx.x * perWG;
    double* LUg = Jg + (long)NA * NC * kBT;
    int tog = 0;
    // without a queue (the drop-in launch) the cells are dealt round-robin:
    // block b integrates cells b, b + gridDim.x, ...
    int next = blockIdx.x;

    for (;;) {
        int cell;
        if (queue) {
            if (t == 0) L.cell = atomicAdd(queue, 1);
            __syncthreads();
            cell = __builtin_amdgcn_readfirstlane(L.cell);
        } else {
            cell = next;
            next += gridDim.x;
        }
        if (cell >= A.nCells) break;
        cell = bound(cell, A.nCells, m.bounds, kBoundCell);
        if (t < n) L.y[t] = A.Y0[t * ld + cell];
        const double p = A.Y0[(long)(nS + 1) * ld + cell];
        double dtChem = A.dtChem[cell];
        int status = 0, attempts = 0, accepted = 0;
        __syncthreads();
        if (L.y[nS] > A.Treact) {
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
                    const double T = L.y[nS];
                    blk_jacobian(m, p, L, nRp, Jg, tog, q);
                    __syncthreads();
                    // derivative at y0: the Jacobian's own unless T is outside
                    // the reaction temperature limits (omega() clips it there)
                    if (T >= m.Tlow && T <= m.Thigh) {
                        if (t < n) L.dydx0[t] = L.dfdx[t];
                        __syncthreads();
                    } else {
                        blk_derivatives(m, p, L.y, L.dydx0, L, tog, q);
                    }
                    double err;
                    for (;;) {
                        {
                            // -J + I/(gamma dx), factorised in registers
                            double Am[NA][NC];
                            const double tdiag = 1.0 / (tb.gamma * dx);
#pragma unroll
                            for (int a = 0; a < NA; ++a) {
                                const int i = q.tr + kPR * a;
#pragma unroll
                                for (int b = 0; b < NC; ++b) {
                                    const int j = q.tc + kPC * b;
                                    double v = -wave_slots<NA * NC>(Jg, q)[(a * NC + b) * 64 + q.lane];
                                    if (i == j && i < n) v += tdiag;
                                    Am[a][b] = v;
                                }
                            }
                            blk_lu<NA, NC, LUV>(Am, n, L, LUg, q);
                        }
                        // Rosenbrock stages (gpuRosenbrock{12,23,34}.H::solve)
#pragma unroll 1
                        for (int st = 0; st < tb.stages; ++st) {
                            if (st == 1 || (st == 2 && tb.stages == 4)) {
                                if (t < n)
                                    L.yt[t] = (st == 1) ? L.y[t] + tb.a21 * L.k1[t]
                                                        : L.y[t] + tb.a31 * L.k1[t] + tb.a32 * L.k2[t];
                                __syncthreads();
                                blk_derivatives(m, p, L.yt, L.dydx, L, tog, q);
                            }
                            if (t < n) {
                                const double df = L.dfdx[t];
                                double rhs;
                                if (st == 0) rhs = L.dydx0[t] + dx * tb.d1 * df;
                                else if (st == 1) rhs = L.dydx[t] + dx * tb.d2 * df + tb.c21 * L.k1[t] / dx;
                                else if (st == 2) rhs = L.dydx[t] + dx * tb.d3 * df + (tb.c31 * L.k1[t] + tb.c32 * L.k2[t]) / dx;
                                else rhs = L.dydx[t] + dx * tb.d4 * df + (tb.c41 * L.k1[t] + tb.c42 * L.k2[t] + tb.c43 * L.k3[t]) / dx;
                                L.rhs[t] = rhs;
                            }
                            __syncthreads();
                            double* kx = (st == 0) ? L.k1 : (st == 1) ? L.k2 : (st == 2) ? L.k3 : L.k4;
                            blk_solve(LUg, n, L, L.rhs, kx, q);
                        }
                        double ne = 0.0;
                        if (t < n) {
                            const double yv = L.y[t];
                            double yNew, e;
                            if (tb.stages == 2) {
                                yNew = yv + tb.b1 * L.k1[t] + tb.b2 * L.k2[t];
                                e = tb.e1 * L.k1[t] + tb.e2 * L.k2[t];
                            } else if (tb.stages == 3) {
                                yNew = yv + tb.b1 * L.k1[t] + tb.b2 * L.k2[t] + tb.b3 * L.k3[t];
                                e = tb.e1 * L.k1[t] + tb.e2 * L.k2[t] + tb.e3 * L.k3[t];
                            } else {
                                yNew = yv + tb.b1 * L.k1[t] + tb.b2 * L.k2[t] + tb.b3 * L.k3[t] + tb.b4 * L.k4[t];
                                e = tb.e1 * L.k1[t] + tb.e2 * L.k2[t] + tb.e4 * L.k4[t];
                            }
                            L.yt[t] = yNew;
                            const double tol = A.absTol + A.relTol * fmax(fabs(yv), fabs(yNew));
                            ne = fabs(e) / tol;
                        }
                        err = block_max(ne, L, tog, q);
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
                    if (t < n) L.y[t] = L.yt[t];
                    __syncthreads();
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
                if (t < nS) L.y[t] = fmax(0.0, L.y[t]);
                __syncthreads();
                timeLeft -= xEnd;
                if (status & GFC_CELL_STEP_UNDERFLOW) break;
            }
        } else {
            dtChem = A.deltaT;
            status = GFC_CELL_SKIPPED;
        }
        if (A.RR && t < nS)
            A.RR[t * ld + cell] = A.rho[cell] * (L.y[t] - A.Y0[t * ld + cell]) / A.deltaT;
        if (A.Ynew) {
            if (t < n) A.Ynew[t * ld + cell] = L.y[t];
            if (t == n) A.Ynew[(long)n * ld + cell] = p;
        }
        if (t == 0) {
            if (A.dtMinBits) atomic_min_positive(A.dtMinBits, dtChem);
            A.dtChem[cell] = fmin(dtChem, A.deltaTChemMax);
            if (A.stats) {
                A.stats[cell] = status;
                A.stats[ld + cell] = attempts;
                A.stats[2 * ld + cell] = accepted;
            }
        }
        __syncthreads();
    }
}

// Device-level evaluations with the workgroup kernel's own physics and LU;
// block b takes cells b, b + gridDim.x, ... and its scratch slot as in the
// solve.
template <int NA, int NC>
__global__ void __launch_bounds__(kBT)
block_eval_kernel(EvalArgs E, double* scratch, long perWG) {
    __shared__ BlockLDS<NA, NC> L;
    const Coord q = coord();
    const int t = q.tid;
    const DeviceMech& m = E.m;
    const int nS = m.nS, n = nS + 1;
    const long ld = E.ldY;
    const long nRp = ((long)m.nR + 63) / 64 * 64;
    double* Jg = scratch + (long)blockIdx.x * perWG;
    double* LUg = Jg + (long)NA * NC * kBT;
    int tog = 0;
    for (int cell = blockIdx.x; cell < E.nCells; cell += gridDim.x) {
        if (E.mode == 2) {
            const double* M = E.M + (long)cell * n * n;
            double Am[NA][NC];
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int i = q.tr + kPR * a;
#pragma unroll
                for (int b = 0; b < NC; ++b) {
                    const int j = q.tc + kPC * b;
                    Am[a][b] = (i < n && j < n) ? M[(long)i * n + j] : 0.0;
                }
            }
            if (E.lu == GFC_LU_PANEL) blk_lu<NA, NC, 1>(Am, n, L, LUg, q);
            else if (E.lu == GFC_LU_PANEL_MFMA) blk_lu<NA, NC, 2>(Am, n, L, LUg, q);
            else blk_lu<NA, NC, 0>(Am, n, L, LUg, q);
            if (t < n) L.rhs[t] = E.b[t * ld + cell];
            __syncthreads();
            blk_solve(LUg, n, L, L.rhs, L.k1, q);
            if (t < n) E.x[t * ld + cell] = L.k1[t];
            __syncthreads();
            continue;
        }
        if (t < n) L.y[t] = E.Y[t * ld + cell];
        const double p = E.Y[(long)(nS + 1) * ld + cell];
        __syncthreads();
        if (E.mode == 0) {
            blk_derivatives(m, p, L.y, L.dydx, L, tog, q);
            if (t < n) E.dYdt[t * ld + cell] = L.dydx[t];
        } else {
            blk_jacobian(m, p, L, nRp, Jg, tog, q);
            __syncthreads();
            if (t < n) E.dYdt[t * ld + cell] = L.dfdx[t];
            double* Jc = E.J + (long)cell * n * n;
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int i = q.tr + kPR * a;
#pragma unroll
                for (int b = 0; b < NC; ++b) {
                    const int j = q.tc + kPC * b;
                    if (i < n && j < n)
                        Jc[(long)i * n + j] = wave_slots<NA * NC>(Jg, q)[(a * NC + b) * 64 + q.lane];
                }
            }
        }
        __syncthreads();
    }
}

template __global__ void block_eval_kernel<8, 4>(EvalArgs, double*, long);
template __global__ void block_eval_kernel<13, 7>(EvalArgs, double*, long);

static_assert(BlockLDS<8, 4>::RSZ == block_lds_doubles(128), "LDS region of n <= 128");
static_assert(BlockLDS<13, 7>::RSZ == block_lds_doubles(208), "LDS region of n <= 208");
// factorisation variants: 0 column by column, 1 four-column panels with
// vector FMAs, 2 four-column panels with MFMA trailing updates
template __global__ void block_solve_kernel<8, 4, 0>(SolveArgs, int*, double*, long);
template __global__ void block_solve_kernel<8, 4, 1>(SolveArgs, int*, double*, long);
template __global__ void block_solve_kernel<8, 4, 2>(SolveArgs, int*, double*, long);
template __global__ void block_solve_kernel<13, 7, 0>(SolveArgs, int*, double*, long);
template __global__ void block_solve_kernel<13, 7, 1>(SolveArgs, int*, double*, long);
template __global__ void block_solve_kernel<13, 7, 2>(SolveArgs, int*, double*, long);

} // namespace gfc
