// Per-cell chemistry helpers shared by the wavefront-per-cell and
// workgroup-per-cell kernels: wavefront reductions (DPP), the per-cell
// temperature functions, the field-major reaction record and the rate /
// equilibrium-constant evaluation in log form.
#pragma once

#include <hip/hip_runtime.h>

#include "kernel_common.hpp"

namespace gfc {
namespace wave {

constexpr int kLanes = 64;
constexpr int kMaxSets = 16;                     // efficiency vectors per LDS chunk
constexpr double kLnRootVGreat = 354.25825000766; // ln(sqrt(DBL_MAX/10))

__device__ __forceinline__ double readlane(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ double dpp(double old, double v) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v),
                                               CTRL, ROW_MASK, BANK_MASK, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v),
                                               CTRL, ROW_MASK, BANK_MASK, false);
    return __hiloint2double(hi, lo);
}

template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ int dppi(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW_MASK, BANK_MASK, false);
}

// Sum over the 64 lanes in a fixed tree (row_shr 1,2,4,8 inside rows of 16,
// then row_bcast 15 and 31); the result is uniform.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp<0x111>(0.0, v);
    v += dpp<0x112>(0.0, v);
    v += dpp<0x114>(0.0, v);
    v += dpp<0x118>(0.0, v);
    v += dpp<0x142, 0xa>(0.0, v);
    v += dpp<0x143, 0xc>(0.0, v);
    return readlane(v, 63);
}

__device__ __forceinline__ double wave_max(double v) {
    const double lo = -HUGE_VAL;
    v = fmax(v, dpp<0x111>(lo, v));
    v = fmax(v, dpp<0x112>(lo, v));
    v = fmax(v, dpp<0x114>(lo, v));
    v = fmax(v, dpp<0x118>(lo, v));
    v = fmax(v, dpp<0x142, 0xa>(lo, v));
    v = fmax(v, dpp<0x143, 0xc>(lo, v));
    return readlane(v, 63);
}

__device__ __forceinline__ int wave_max_i(int v) {
    const int lo = -2147483647 - 1;
    v = max(v, dppi<0x111>(lo, v));
    v = max(v, dppi<0x112>(lo, v));
    v = max(v, dppi<0x114>(lo, v));
    v = max(v, dppi<0x118>(lo, v));
    v = max(v, dppi<0x142, 0xa>(lo, v));
    v = max(v, dppi<0x143, 0xc>(lo, v));
    return __builtin_amdgcn_readlane(v, 63);
}

// Per-cell temperature functions shared by every reaction.
struct TFun {
    double T, logT, invT, invRRT, L1; // L1 = ln(pStd/(RR T))
};

__device__ __forceinline__ TFun tfun(double T) {
    TFun f;
    f.T = T;
    f.logT = log(T);
    f.invT = 1.0 / T;
    f.invRRT = 1.0 / (kRR * T);
    f.L1 = log(kPstd * f.invRRT);
    return f;
}

// One reaction of a pass, loaded from the field-major record.
struct Rxn {
    int type, irr, set, nl, nr;
    int li[kTeamWidth], ri[kTeamWidth], lei[kTeamWidth], rei[kTeamWidth];
    double lnA, beta, Ta, sgnA, lnAi, betai, Tai, sgnAi;
    double troe[4];
    double dn, Tcommon;
    double ls[kTeamWidth], rs[kTeamWidth], le[kTeamWidth], re[kTeamWidth];
};

__device__ __forceinline__ Rxn load_rxn(const DeviceMech& m, int pass, int lane) {
    const double* D = m.recD + (long)pass * kRecD * kLanes + lane;
    const int* I = m.recI + (long)pass * kRecI * kLanes + lane;
    Rxn x;
    x.type = I[riType * kLanes];
    x.irr = I[riIrr * kLanes];
    x.set = I[riSet * kLanes];
    x.nl = I[riNl * kLanes];
    x.nr = I[riNr * kLanes];
#pragma unroll
    for (int q = 0; q < kTeamWidth; ++q) {
        x.li[q] = I[(riLi + q) * kLanes];
        x.ri[q] = I[(riRi + q) * kLanes];
        x.lei[q] = I[(riLei + q) * kLanes];
        x.rei[q] = I[(riRei + q) * kLanes];
        x.ls[q] = D[(rdLs + q) * kLanes];
        x.rs[q] = D[(rdRs + q) * kLanes];
        x.le[q] = D[(rdLe + q) * kLanes];
        x.re[q] = D[(rdRe + q) * kLanes];
    }
    x.lnA = D[rdLnA * kLanes];
    x.beta = D[rdBeta * kLanes];
    x.Ta = D[rdTa * kLanes];
    x.sgnA = D[rdSgnA * kLanes];
    // the fall-off fields only for the fall-off rates (a pass of Arrhenius
    // or third-body reactions does not load them)
    x.lnAi = x.betai = x.Tai = 0.0;
    x.sgnAi = 1.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) x.troe[q] = 0.0;
    if (x.type > GFC_RATE_THIRD_BODY) {
        x.lnAi = D[rdLnAi * kLanes];
        x.betai = D[rdBetai * kLanes];
        x.Tai = D[rdTai * kLanes];
        x.sgnAi = D[rdSgnAi * kLanes];
#pragma unroll
        for (int q = 0; q < 4; ++q) x.troe[q] = D[(rdAlpha + q) * kLanes];
    }
    x.dn = D[rdDn * kLanes];
    x.Tcommon = D[rdTcommon * kLanes];
    if (kBoundsChecked) {
#pragma unroll
        for (int q = 0; q < kTeamWidth; ++q) {
            if (q < x.nl) x.li[q] = bound(x.li[q], m.nS, m.bounds, kBoundSpecies);
            if (q < x.nr) x.ri[q] = bound(x.ri[q], m.nS, m.bounds, kBoundSpecies);
        }
        if (x.set >= 0) x.set = bound(x.set, m.nSets, m.bounds, kBoundSet);
    }
    return x;
}

// x^e for the integer exponents of the mechanism, with the multiplication
// order of integerPow (gpuSpecieExponent.H:21-36); real exponents use pow.
__device__ __forceinline__ double ipow(double x, int e, double er) {
    if (e == 1) return x;
    if (e == 2) return x * x;
    if (e == 3) return x * (x * x);
    if (e == 0) return 1.0;
    return specie_pow(x, e, er);
}
__device__ __forceinline__ double ipow_m1(double x, int e, double er) {
    if (e == 1) return 1.0;
    if (e == 2) return x;
    if (e == 3) return x * x;
    return specie_pow_m1(x, e, er);
}

// Concentration powers of one side (lhsPowers, gpuReaction.H:348-372) and
// their product in written order (calcCf, :374-386).
__device__ __forceinline__ double side_powers(const int* idx, const int* ei,
                                              const double* er, int n,
                                              const double* c, double* pw) {
    double C = 1.0;
#pragma unroll
    for (int q = 0; q < kTeamWidth; ++q) {
        pw[q] = (q < n) ? ipow(c[idx[q]], ei[q], er[q]) : 1.0;
        if (q < n) C *= pw[q];
    }
    return C;
}

// d(product)/dc at position j (calcdCfdcj, :388-429)
__device__ __forceinline__ double side_dpow(const int* idx, const int* ei,
                                            const double* er, int n, int j,
                                            const double* c, const double* pw) {
    double v = 1;
#pragma unroll
    for (int i = 0; i < kTeamWidth; ++i) {
        if (i >= n) break;
        if (i == j) v *= er[i] * ipow_m1(c[idx[i]], ei[i], er[i]);
        else v *= pw[i];
    }
    return v;
}

// ln Kc (gpuThermo.H:80-113 in log form) and optionally dKcdTbyKc
// (gpuThermo.H:139-169, with S + Gstd/T = Ha/T).
// G holds the per-specie W Gstd / W Ha arrays gL, gH, hL, hH.
template <class G>
__device__ __forceinline__ double ln_kc(const Rxn& x, const TFun& f, const G& L,
                                        double* dKc) {
    const bool low = f.T < x.Tcommon;
    const double* g = low ? L.gL : L.gH;
    double sG = 0.0;
#pragma unroll
    for (int q = 0; q < kTeamWidth; ++q) {
        if (q < x.nr) sG += x.rs[q] * g[x.ri[q]];
        if (q < x.nl) sG -= x.ls[q] * g[x.li[q]];
    }
    const double arg = -sG * f.invRRT;
    const double lnK = ((arg < 600) ? arg : kLnRootVGreat) + x.dn * f.L1;
    if (dKc) {
        const double* h = low ? L.hL : L.hH;
        double sH = 0.0;
#pragma unroll
        for (int q = 0; q < kTeamWidth; ++q) {
            if (q < x.nr) sH += x.rs[q] * h[x.ri[q]];
            if (q < x.nl) sH -= x.ls[q] * h[x.li[q]];
        }
        *dKc = sH * f.invRRT * f.invT - x.dn * f.invT;
    }
    return lnK;
}

// Forward rate constant (gpuReactionRate.H:164-173; everything(),
// :237-245) in log form.  kArr: the factor of the concentration derivative
// (third body: the Arrhenius part; fall-off: none); ddT: dkf/dT.
__device__ __forceinline__ double forward_rate(const Rxn& x, const TFun& f,
                                               const double* Mset, double& kArr,
                                               double& ddT) {
    const double Ak = x.sgnA * exp(x.lnA + x.beta * f.logT - x.Ta * f.invT);
    if (x.type == GFC_RATE_ARRHENIUS) {
        kArr = Ak;
        ddT = Ak * (x.beta + x.Ta * f.invT) * f.invT;
        return Ak;
    }
    const double M = Mset[x.set];
    if (x.type == GFC_RATE_THIRD_BODY) {
        kArr = Ak;
        ddT = M * (Ak * (x.beta + x.Ta * f.invT) * f.invT);
        return M * Ak;
    }
    // fall-off (gpuFallOffReactionRate.H:39-158): the Jacobian keeps dkInf/dT only
    const double kInf = x.sgnAi * exp(x.lnAi + x.betai * f.logT - x.Tai * f.invT);
    const double Pr = Ak / kInf * M;
    const double F = (x.type == GFC_RATE_TROE) ? troe_F(x.troe, f.T, Pr) : 1.0;
    const double temp = (Pr / (1.0 + Pr)) * F;
    kArr = 0.0;
    ddT = (kInf * (x.betai + x.Tai * f.invT) * f.invT) * temp;
    return kInf * temp;
}

// One column's efficiencies of the first kSetsT sets (a uniform address,
// read with vector loads so that the loads of the next column stay in
// flight across the LDS traffic of this one).
__device__ __forceinline__ void dense_row_load(const double* e, double (&out)[kSetsT]) {
    int z = 0;  // a per-lane (zero) offset keeps these global vector loads
    asm volatile("" : "+v"(z));
    const double2* e2 = reinterpret_cast<const double2*>(e) + z;
#pragma unroll
    for (int q = 0; q < kSetsT / 2; ++q) {
        const double2 v = e2[q];
        out[2 * q] = v.x;
        out[2 * q + 1] = v.y;
    }
}

// sum_q w[q] e[q] as two interleaved partial sums (even and odd q)
__device__ __forceinline__ double dense_dot(const double (&w)[kSetsT],
                                            const double (&e)[kSetsT]) {
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int q = 0; q < kSetsT; q += 2) {
        a0 += w[q] * e[q];
        a1 += w[q + 1] * e[q + 1];
    }
    return a0 + a1;
}

} // namespace wave
} // namespace gfc
