// Device-resident mechanism tables and the per-specie / per-reaction
// physics shared by every kernel.  All arithmetic is FP64 and follows the
// operation order of gpuFoam's device code (paths below are relative to the
// gpuFoam repository root).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gpufoam_amd.h"
#include "bounds.hpp"
#include "constants.hpp"

namespace gfc {

constexpr int kWidth = 10; // gpuReaction.H:12 MAX_REACTION_WIDTH
constexpr int kNoIntExp = 214748364;

// Gather lists of the workgroup-per-cell kernel (reaction indices in the
// stored order).  Empty unless that kernel can run the mechanism.
struct BlockTables {
    // reactions of each specie with its net stoichiometric coefficient:
    // dN_i/dt = sum_e spNu[e] w[spR[e]], e in [spPtr[i], spPtr[i+1])
    const int* spPtr;
    const int* spR;
    const double* spNu;
    // contributions to the concentration Jacobian entry (i, j), row-major
    // [nS*nS+1]: J(i,j) = sum_e elNu[e] D[elRS[e] & 7][elRS[e] >> 3], where
    // D[q][r] is dw_r/dc of the specie at position q of reaction r
    // (0-2 left side, 3-5 right side)
    const int* elPtr;
    const int* elRS;
    const double* elNu;
    // third-body reactions of each specie row: J(i,j) += tbNu g_r effB[set][j]
    const int* tbPtr;
    const int* tbR;
    const int* tbSet;
    const double* tbNu;
    const double* effB; // [nSets*nS] efficiency sets, set-major
    const double* effT; // [nS*nSets] specie-major
};

// Flat device copy of gfc_mech_view plus host-precomputed constants.  Every
// array is in one hipMalloc'd block owned by the evaluator.
struct DeviceMech {
    int nS;
    int nR;
    double Tlow, Thigh; // reaction temperature clip (gpuReaction.H:472-488)
    // species [nS]
    const double* W;
    const double* R;       // RR / W
    const double* Tcommon;
    const double* high;    // [nS*7] R-scaled NASA-7
    const double* low;     // [nS*7]
    // reactions [nR]
    const int* type;
    const int* irreversible;
    const double* k;    // [nR*3]
    const double* kInf; // [nR*3]
    const double* troe; // [nR*4]
    const double* eff;  // [nR*nS]
    const int* nl;
    const int* nr;
    const int* li;  // [nR*10]
    const double* ls;
    const double* le;
    const int* lei;
    const int* ri;
    const double* rs;
    const double* re;
    const int* rei;
    const double* rY;
    const double* rW;
    const double* rTcommon;
    const double* rhigh; // [nR*7]
    const double* rlow;  // [nR*7]
    // wavefront kernel only: reactions are grouped in passes of 64;
    // effP[(pass*nS + i)*64 + lane] is the efficiency of specie i for the
    // reaction evaluated by `lane` in `pass`
    const double* effP;
    int nPass;
    // natural-log form of the rate constants: lnA[2r] = ln|A| of k,
    // lnA[2r+1] = ln|A| of kInf, sgnA likewise (+1/-1)
    const double* lnA;
    const double* sgnA;
    // distinct third-body efficiency vectors: effSets[s*64 + i], and the
    // set of each reaction (-1 for plain Arrhenius)
    const double* effSets;
    // the first kSetsT sets transposed: effSetsT[i*kSetsT + s] (specie-major,
    // so one specie's efficiencies of all sets are one scalar load)
    const double* effSetsT;
    const int* effSet;
    int nSets;
    // reference index of the reaction stored at r (the identity for the
    // reference-order tables)
    const int* rxnRef;
    // wavefront kernel: one record per reaction, field-major within a pass
    // so that the 64 lanes of a pass load each field in one coalesced access:
    // recD[(pass*kRecD + f)*64 + lane], recI[(pass*kRecI + f)*64 + lane]
    const double* recD;
    const int* recI;
    int teamWidth; // widest reaction side (the record holds kTeamWidth)
    BlockTables blk;
    // parameters of the rate types gpuFoam rejects, [nR*GFC_EXTRA_PARAMS],
    // and whether the mechanism has any (only the one-lane and query
    // kernels evaluate them)
    const double* extra;
    int extended;
    // the bounds-check word of a library built with GFC_BOUNDS_CHECK (null
    // otherwise)
    unsigned* bounds;
};

// Rates that depend on a third-body concentration M.
GFC_HD __forceinline__ bool uses_third_body(int type) {
    return type >= GFC_RATE_THIRD_BODY && type <= GFC_RATE_SRI_CA;
}

constexpr int kSetsT = 16;

// Record layout of the wavefront kernel (at most kTeamWidth species per side).
constexpr int kTeamWidth = 3;
// Third-body concentrations the wavefront kernel holds in LDS (one per
// distinct efficiency set); mechanisms with more sets take another kernel.
constexpr int kTeamMaxSets = 128;
enum RecD {
    rdLnA, rdBeta, rdTa, rdSgnA,       // k (Arrhenius / third body) or k0
    rdLnAi, rdBetai, rdTai, rdSgnAi,   // kInf of fall-off rates
    rdAlpha, rdTsss, rdTs, rdTss,      // Troe
    rdDn,                              // moles of products - reactants
    rdTcommon,                         // Tcommon of the composite thermo
    rdLs, rdRs = rdLs + kTeamWidth,    // stoichiometric coefficients
    rdLe = rdRs + kTeamWidth,          // real exponents
    rdRe = rdLe + kTeamWidth,
    kRecD = rdRe + kTeamWidth
};
enum RecI {
    riType, riIrr, riSet, riNl, riNr,
    riLi, riRi = riLi + kTeamWidth,    // specie indices
    riLei = riRi + kTeamWidth,         // integer exponents
    riRei = riLei + kTeamWidth,
    kRecI = riRei + kTeamWidth
};

// ---------------------------------------------------------------- thermo --
// gpuThermo.H:29-32
GFC_HD __forceinline__ const double* coeffs_of(const double* high,
                                                   const double* low,
                                                   double Tcommon, double T) {
    return (T < Tcommon) ? low : high;
}
// gpuThermo.H:177-184
GFC_HD __forceinline__ double poly_Cp(const double* a, double T) {
    return ((((a[4] * T + a[3]) * T + a[2]) * T + a[1]) * T + a[0]);
}
// gpuThermo.H:210-221
GFC_HD __forceinline__ double poly_Ha(const double* a, double T) {
    return (((((a[4] / 5.0 * T + a[3] / 4.0) * T + a[2] / 3.0) * T +
              a[1] / 2.0) * T + a[0]) * T + a[5]);
}
// gpuThermo.H:333-338
GFC_HD __forceinline__ double poly_dCpdT(const double* a, double T) {
    return (((4.0 * a[4] * T + 3.0 * a[3]) * T + 2.0 * a[2]) * T + a[1]);
}
// gpuThermo.H:261-271 (the perfect-gas entropy term at pStd is zero)
GFC_HD __forceinline__ double poly_S(const double* a, double T,
                                         double logT) {
    return ((((a[4] / 4.0 * T + a[3] / 3.0) * T + a[2] / 2.0) * T + a[1]) *
                T + a[0] * logT + a[6]);
}
// gpuThermo.H:287-299
GFC_HD __forceinline__ double poly_Gstd(const double* a, double T,
                                            double logT) {
    return ((a[0] * (1.0 - logT) -
             (((a[4] / 20.0 * T + a[3] / 12.0) * T + a[2] / 6.0) * T +
              a[1] / 2.0) * T - a[6]) * T + a[5]);
}

// ----------------------------------------------------------------- rates --
// gpuArrheniusReactionRate.H:16-26
GFC_HD __forceinline__ double arrhenius(const double* k, double T) {
    double ak = k[0];
    if (fabs(k[1]) > kVSmall) ak *= pow(T, k[1]);
    if (fabs(k[2]) > kVSmall) ak *= exp(-k[2] / T);
    return ak;
}

// gpuTroeFallOffFunction.H:27-47
GFC_HD __forceinline__ double troe_F(const double* tr, double T,
                                         double Pr) {
    const double alpha = tr[0], Tsss = tr[1], Ts = tr[2], Tss = tr[3];
    const double logPr = log10(fmax(Pr, kSmall));
    const double Fcent = (1.0 - alpha) * exp(-T / Tsss) +
                         alpha * exp(-T / Ts) + exp(-Tss / T);
    const double logFcent = log10(fmax(Fcent, kSmall));
    const double c = -0.4 - 0.67 * logFcent;
    const double d = 0.14;
    const double n = 0.75 - 1.27 * logFcent;
    const double x1 = n - d * (logPr + c);
    const double x2 = (logPr + c) / x1;
    const double x3 = 1.0 + x2 * x2;
    const double x4 = logFcent / x3;
    return pow(10.0, x4);
}

// ------------------------------------- rate types gpuFoam rejects --
// Published forms as OpenFOAM's rate classes take them
// (makeMoreGpuChemistryModels.C:59-95 registers them on the CPU side only).

// SRI fall-off function (Stewart, Larson & Golden 1989), x = a b c d e:
// F = d [a exp(-b/T) + exp(-T/c)]^X T^e, X = 1/(1 + log10(max(Pr,small))^2)
GFC_HD __forceinline__ double sri_F(const double* x, double T, double Pr) {
    const double lp = log10(fmax(Pr, kSmall));
    const double X = 1.0 / (1.0 + lp * lp);
    return x[3] * pow(x[0] * exp(-x[1] / T) + exp(-T / x[2]), X) * pow(T, x[4]);
}

// Landau-Teller, Janev and power-series rates and dk/dT (ddT may be null):
//   LT     k = A T^b exp(-Ta/T + B T^(-1/3) + C T^(-2/3))
//   Janev  k = A T^b exp(-Ta/T) exp(sum_{n=0..8} b_n (ln T)^n)
//   power  k = A T^b exp(-Ta/T + sum_{n=1..4} c_n T^(-n))
GFC_HD __forceinline__ double tdep_rate(int type, const double* k, const double* x,
                                        double T, double* ddT) {
    double v = k[0];
    if (fabs(k[1]) > kVSmall) v *= pow(T, k[1]);
    double arg = 0.0, darg = 0.0;
    if (fabs(k[2]) > kVSmall) {
        arg -= k[2] / T;
        darg += k[2] / (T * T);
    }
    if (type == GFC_RATE_LANDAU_TELLER) {
        const double t13 = pow(T, 1.0 / 3.0);
        if (fabs(x[0]) > kVSmall) {
            arg += x[0] / t13;
            darg -= x[0] / (3.0 * T * t13);
        }
        if (fabs(x[1]) > kVSmall) {
            arg += x[1] / (t13 * t13);
            darg -= 2.0 * x[1] / (3.0 * T * t13 * t13);
        }
    } else if (type == GFC_RATE_JANEV) {
        const double lnT = log(T);
        double Tn = 1.0, dsum = 0.0;
        for (int n = 0; n < 9; ++n) {
            arg += x[n] * Tn;
            if (n + 1 < 9) dsum += (n + 1) * x[n + 1] * Tn;
            Tn *= lnT;
        }
        darg += dsum / T;
    } else {
        double Tn = 1.0;
        for (int n = 1; n <= 4; ++n) {
            Tn *= T;
            arg += x[n - 1] / Tn;
            darg -= n * x[n - 1] / (Tn * T);
        }
    }
    if (fabs(arg) > kVSmall) v *= exp(arg);
    if (ddT) *ddT = v * (k[1] / T + darg);
    return v;
}

// Blending function F(T, Pr) of a fall-off or chemically activated rate.
GFC_HD __forceinline__ double blend_F(const DeviceMech& m, int r, double T, double Pr) {
    const int t = m.type[r];
    if (t == GFC_RATE_TROE || t == GFC_RATE_TROE_CA) return troe_F(m.troe + 4 * r, T, Pr);
    if (t == GFC_RATE_SRI || t == GFC_RATE_SRI_CA)
        return sri_F(m.extra + (long)GFC_EXTRA_PARAMS * r, T, Pr);
    return 1.0;
}

// gpuSpecieExponent.H:21-36
GFC_HD __forceinline__ double integer_pow(double x, int e) {
    if (e == 0) return 1;
    double xx = e > 0 ? x : 1 / x;
    double y = 1;
    for (int i = e; i != 0; i /= 2) {
        if (i % 2) y *= xx;
        xx *= xx;
    }
    return y;
}
// gpuSpecieExponent.H:103-110
GFC_HD __forceinline__ double specie_pow(double x, int ei, double er) {
    if (ei != kNoIntExp) return integer_pow(x, ei);
    return pow(x, er);
}
GFC_HD __forceinline__ double specie_pow_m1(double x, int ei, double er) {
    if (ei != kNoIntExp) return integer_pow(x, ei - 1);
    return pow(x, er + (-1.0));
}

// Kc of reaction r at T (gpuThermo.H:80-113), as omega() evaluates it.
GFC_HD __forceinline__ double reaction_Kc(const DeviceMech& m, int r,
                                              double T) {
    const double Y = m.rY[r], W = m.rW[r];
    const double* a = coeffs_of(m.rhigh + 7 * r, m.rlow + 7 * r,
                                m.rTcommon[r], T);
    const double nm = Y / W;
    const double arg = -Y * poly_Gstd(a, T, log(T)) / (kRR * T);
    const double K = (arg < 600) ? exp(arg) : sqrt(kVGreat);
    if (fabs(nm - kSmall) < kVSmall) return K;
    return K * pow(kPstd / (kRR * T), nm);
}

// Intermediate terms of reaction_Kc_dKc, for callers that also need the
// temperature derivatives of Kc.
struct KcTerms {
    const double* a; // the reaction's NASA-7 coefficients at T
    double Y, nm, RRT, Gstd, S, arg;
    bool nmSmall;
};

// KcAnddKcTbyKc (gpuThermo.H:139-169)
GFC_HD __forceinline__ void reaction_Kc_dKc(const DeviceMech& m, int r,
                                                double T, double& Kc,
                                                double& dKcdTbyKc,
                                                KcTerms* terms = nullptr) {
    const double Y = m.rY[r], W = m.rW[r];
    const double* a = coeffs_of(m.rhigh + 7 * r, m.rlow + 7 * r,
                                m.rTcommon[r], T);
    const double nm = Y / W;
    const double logT = log(T);
    const double Gstd = poly_Gstd(a, T, logT);
    const bool nm_small = fabs(nm - kSmall) < kVSmall;
    const double RRT = kRR * T;
    const double arg = -Y * Gstd / RRT;
    const double K = (arg < 600) ? exp(arg) : sqrt(kVGreat);
    Kc = nm_small ? K : K * pow(kPstd / RRT, nm);
    const double S = poly_S(a, T, logT);
    double d = (S + Gstd / T) * Y / RRT;
    if (!nm_small) d = d - nm / T;
    dKcdTbyKc = d;
    if (terms) *terms = KcTerms{a, Y, nm, RRT, Gstd, S, arg, nm_small};
}

} // namespace gfc
