// Per-lane chemistry physics: one lane integrates one cell, with its
// vectors and matrices in a strided (structure-of-arrays) buffer.  Used by
// the plain per-cell kernel and by the discrete adjoint; the functions are
// host-and-device so the same code can be exercised on the CPU.
//
// The arithmetic follows gpuFoam's device code operation by operation
// (paths relative to the gpuFoam repository root); translation units that
// include this header disable fused multiply-add contraction so that the
// only differences from the reference built for the host are the last-ulp
// differences of exp/log/pow.
#pragma once

#include "kernel_common.hpp"

#pragma clang fp contract(off)

namespace gfc {
namespace lane {

// A per-cell array in SoA scratch.
struct SV {
    double* p;
    long s;
    GFC_HD __forceinline__ double& operator[](int e) const {
        return p[e * s];
    }
};
struct SI {
    int* p;
    long s;
    GFC_HD __forceinline__ int& operator[](int e) const { return p[e * s]; }
};

// Per-cell scratch of one lane: `base` points at the lane's first element,
// consecutive elements are `s` doubles apart; the arrays follow each other
// at offsets that are multiples of n.
struct LaneBuf {
    double* base;
    long s;
    int n;
    // Optional copies kept for the discrete adjoint (null otherwise), laid
    // out like J: element (i, j) at [(i*n + j)*s].  `raw` receives the
    // species rows of the reaction-assembled Jacobian before the Y/T
    // transforms, `jt` the transformed Jacobian before the LU overwrites it.
    double* raw = nullptr;
    double* jt = nullptr;
    // optional copy of the first stage's right-hand side f(y2), which the
    // step overwrites with f(y3)
    double* f2 = nullptr;
    // take the step's f(y0) from the Jacobian call's F instead of the caller
    // (the same function of y0 within the reaction temperature limits)
    bool dydx0FromF = false;
    GFC_HD __forceinline__ SV at(int k) const { return SV{base + (long)k * n * s, s}; }
    GFC_HD __forceinline__ SV y() const { return at(0); }
    GFC_HD __forceinline__ SV dydx0() const { return at(1); }
    GFC_HD __forceinline__ SV yTemp() const { return at(2); }
    GFC_HD __forceinline__ SV dydx() const { return at(3); }
    GFC_HD __forceinline__ SV dfdx() const { return at(4); }
    GFC_HD __forceinline__ SV k1() const { return at(5); }
    GFC_HD __forceinline__ SV k2() const { return at(6); }
    GFC_HD __forceinline__ SV k3() const { return at(7); }
    GFC_HD __forceinline__ SV k4() const { return at(8); }
    GFC_HD __forceinline__ SV c() const { return at(9); }
    GFC_HD __forceinline__ SV Ha() const { return at(10); }
    GFC_HD __forceinline__ SV Cp() const { return at(11); }
    GFC_HD __forceinline__ SV ddc() const { return at(12); }
    GFC_HD __forceinline__ SV vv() const { return at(13); }
    GFC_HD __forceinline__ SV J() const { return at(14); }
    GFC_HD __forceinline__ SI piv() const {
        return SI{reinterpret_cast<int*>(base + (long)(14 + n) * n * s), 2 * s};
    }
    GFC_HD __forceinline__ double& Jat(int i, int j) const {
        return base[((long)(14 + i) * n + j) * s];
    }
};

// Per-cell scratch is laid out wave-blocked: the 64 cells of a wavefront
// share one contiguous block of per*64 doubles, element e of cell c at
// [e*64 + c%64] of block c/64, so every access of a wavefront is one
// coalesced 512-byte line and a wavefront's whole scratch is one compact
// region (few pages, good L2 reuse).
constexpr long kSlotStride = 64;
GFC_HD __forceinline__ double* wave_slot(double* base, long cell, long per) {
    return base + (cell / kSlotStride) * (per * kSlotStride) + (cell % kSlotStride);
}

// Number of doubles of scratch per slot for nS species.
GFC_HD inline long lane_scratch_doubles(int nS) {
    const long n = nS + 1;
    return 14 * n + n * n + n;
}

// ------------------------------------------------------------- physics --

// k(p,T,c) dispatch (gpuReactionRate.H:164-173)
GFC_HD inline double rate_k(const DeviceMech& m, int r, double T, SV c) {
    const double* k = m.k + 3 * r;
    const int t = m.type[r];
    if (t == GFC_RATE_ARRHENIUS) return arrhenius(k, T);
    if (t > GFC_RATE_SRI_CA)
        return tdep_rate(t, k, m.extra + (long)GFC_EXTRA_PARAMS * r, T, nullptr);
    const double* eff = m.eff + (long)r * m.nS;
    double M = 0;
    for (int i = 0; i < m.nS; ++i) M += eff[i] * c[i];
    if (t == GFC_RATE_THIRD_BODY) return M * arrhenius(k, T);
    const double k0 = arrhenius(k, T);
    const double kInf = arrhenius(m.kInf + 3 * r, T);
    const double Pr = k0 / kInf * M;
    const double F = blend_F(m, r, T, Pr);
    if (t >= GFC_RATE_LINDEMANN_CA) return k0 * (1.0 / (1.0 + Pr)) * F; // chemically activated
    return kInf * (Pr / (1.0 + Pr)) * F;
}

GFC_HD __forceinline__ double side_product(const int* idx, const int* ei,
                                               const double* er, int n, SV c,
                                               double* pw) {
    for (int q = 0; q < n; ++q) pw[q] = specie_pow(c[idx[q]], ei[q], er[q]);
    double C = 1.0;
    for (int q = 0; q < n; ++q) C *= pw[q];
    return C;
}

// gpuReaction::omega, gpuReaction.H:472-488
GFC_HD inline double omega(const DeviceMech& m, int r, double T, SV c) {
    const double Tc = fmin(fmax(T, m.Tlow), m.Thigh);
    const double kf = rate_k(m, r, Tc, c);
    const double kr = m.irreversible[r] ? 0.0 : kf / reaction_Kc(m, r, Tc);
    double pw[kWidth];
    const long b = (long)r * kWidth;
    const double Cf = side_product(m.li + b, m.lei + b, m.le + b, m.nl[r], c, pw);
    const double Cr = side_product(m.ri + b, m.rei + b, m.re + b, m.nr[r], c, pw);
    return kf * Cf - kr * Cr;
}

// gpuODESystem::derivatives (gpuODESystem.H:164-196) with
// gpuMixture::computeDerivativeRequired (gpuMixture.H:88-121).
GFC_HD __forceinline__ void derivatives(const DeviceMech& m, double p, SV y, SV dydx,
                            LaneBuf b) {
    const int nS = m.nS;
    for (int i = 0; i <= nS; ++i) dydx[i] = 0;
    const double T = y[nS];
    double rhoM = 0.0, CpM = 0.0;
    for (int i = 0; i < nS; ++i) {
        const double yi = fmax(y[i], 0.0);
        const double rhoi = p / (m.R[i] * T);
        const double* a = coeffs_of(m.high + 7 * i, m.low + 7 * i, m.Tcommon[i], T);
        rhoM += yi / rhoi;
        CpM += yi * poly_Cp(a, T);
        b.Ha()[i] = poly_Ha(a, T);
        b.c()[i] = yi / m.W[i];
    }
    rhoM = 1.0 / rhoM;
    for (int i = 0; i < nS; ++i) b.c()[i] *= rhoM;
    for (int r = 0; r < m.nR; ++r) {
        const double w = omega(m, r, T, b.c());
        const long o = (long)r * kWidth;
        for (int q = 0; q < m.nl[r]; ++q) dydx[m.li[o + q]] -= m.ls[o + q] * w;
        for (int q = 0; q < m.nr[r]; ++q) dydx[m.ri[o + q]] += m.rs[o + q] * w;
    }
    for (int i = 0; i < nS; ++i) dydx[i] *= m.W[i] / rhoM;
    double ret = 0.0;
    for (int i = 0; i < nS; ++i) ret -= dydx[i] * b.Ha()[i];
    dydx[nS] = ret / CpM;
}

// gpuODESystem::jacobian (gpuODESystem.H:198-281) with
// computeReactionParameters / ddNdtByVdcTp (gpuReaction.H:217-233,547-702)
// and the Y/T transforms (gpuODESystem.H:65-146).
GFC_HD __forceinline__ void jacobian(const DeviceMech& m, double p, SV y, LaneBuf b) {
    const int nS = m.nS, n = nS + 1;
    SV dfdx = b.dfdx();
    for (int i = 0; i < n; ++i) dfdx[i] = 0.0;
    for (int i = 0; i < n * n; ++i) b.J()[i] = 0.0;
    const double T = y[nS];
    double rhoM = 0.0, CpM = 0.0, dCpMdT = 0.0, alphav = 0.0;
    for (int i = 0; i < nS; ++i) {
        const double yi = fmax(y[i], 0.0);
        const double rhoi = p / (m.R[i] * T);
        const double vi = 1.0 / rhoi;
        const double* a = coeffs_of(m.high + 7 * i, m.low + 7 * i, m.Tcommon[i], T);
        const double Cpi = poly_Cp(a, T);
        rhoM += yi / rhoi;
        CpM += yi * Cpi;
        dCpMdT += yi * poly_dCpdT(a, T);
        alphav += yi * vi * (1.0 / T);
        b.Cp()[i] = Cpi;
        b.Ha()[i] = poly_Ha(a, T);
        b.c()[i] = yi / m.W[i];
    }
    rhoM = 1.0 / rhoM;
    const double alphavM = alphav * rhoM;
    for (int i = 0; i < nS; ++i) b.c()[i] *= rhoM;
    SV c = b.c();

    for (int r = 0; r < m.nR; ++r) {
        const double* k = m.k + 3 * r;
        const int t = m.type[r];
        const long o = (long)r * kWidth;
        const int nl = m.nl[r], nr = m.nr[r];
        // rate and its derivatives ("everything", gpuReactionRate.H:237-245)
        double kf, ddT;
        double M = 0.0, ddcScale = 0.0; // ddc[j] = eff[j]*ddcScale*(Cf-Cr/Kc)
        if (t == GFC_RATE_ARRHENIUS) {
            const double Ak = arrhenius(k, T);
            kf = Ak;
            ddT = Ak * (k[1] + k[2] / T) / T;
        } else if (t == GFC_RATE_THIRD_BODY) {
            const double Ak = arrhenius(k, T);
            const double dT = Ak * (k[1] + k[2] / T) / T;
            const double* eff = m.eff + (long)r * nS;
            for (int i = 0; i < nS; ++i) M += eff[i] * c[i];
            kf = M * Ak;
            ddT = M * dT;
            ddcScale = Ak;
        } else if (t > GFC_RATE_SRI_CA) {
            // Landau-Teller, Janev, power series: k(T) only
            kf = tdep_rate(t, k, m.extra + (long)GFC_EXTRA_PARAMS * r, T, &ddT);
        } else if (t >= GFC_RATE_LINDEMANN_CA) {
            // chemically activated: the fall-off convention of gpuFoam
            // applied to its leading factor k0 (no dPr/dT, dF/dT, dM/dc)
            const double k0 = arrhenius(k, T);
            const double kInf = arrhenius(m.kInf + 3 * r, T);
            const double* eff = m.eff + (long)r * nS;
            for (int i = 0; i < nS; ++i) M += eff[i] * c[i];
            const double Pr = k0 / kInf * M;
            const double temp = (1.0 / (1.0 + Pr)) * blend_F(m, r, T, Pr);
            kf = k0 * temp;
            ddT = (arrhenius(k, T) * (k[1] + k[2] / T) / T) * temp;
        } else {
            // gpuFallOffReactionRate.H:303-335 (kInf derivative only)
            const double k0 = arrhenius(k, T);
            const double* ki = m.kInf + 3 * r;
            const double kInf = arrhenius(ki, T);
            const double* eff = m.eff + (long)r * nS;
            for (int i = 0; i < nS; ++i) M += eff[i] * c[i];
            const double Pr = k0 / kInf * M;
            const double F = blend_F(m, r, T, Pr);
            const double temp = (Pr / (1.0 + Pr)) * F;
            kf = kInf * temp;
            ddT = (arrhenius(ki, T) * (ki[1] + ki[2] / T) / T) * temp;
        }
        double pw[kWidth], dCf[kWidth], dCr[kWidth];
        const double Cf = side_product(m.li + o, m.lei + o, m.le + o, nl, c, pw);
        // calcdCfdcj (gpuReaction.H:388-429)
        for (int j = 0; j < nl; ++j) {
            double v = 1;
            for (int i = 0; i < nl; ++i) {
                if (i == j) v *= m.le[o + i] * specie_pow_m1(c[m.li[o + i]], m.lei[o + i], m.le[o + i]);
                else v *= pw[i];
            }
            dCf[j] = v * kf;
        }
        double dwdT = Cf * ddT;
        double w = Cf * kf;
        double Cr = 0.0, Kc = 1.0;
        const bool rev = !m.irreversible[r];
        if (rev) {
            double dKcdTbyV;
            reaction_Kc_dKc(m, r, T, Kc, dKcdTbyV);
            const double kr = kf / Kc;
            Cr = side_product(m.ri + o, m.rei + o, m.re + o, nr, c, pw);
            const double dkrdT = ddT / Kc - kr * dKcdTbyV;
            dwdT -= Cr * dkrdT;
            w -= Cr * kr;
            for (int j = 0; j < nr; ++j) {
                double v = 1;
                for (int i = 0; i < nr; ++i) {
                    if (i == j) v *= m.re[o + i] * specie_pow_m1(c[m.ri[o + i]], m.rei[o + i], m.re[o + i]);
                    else v *= pw[i];
                }
                dCr[j] = v * kr;
            }
        }
        // dNdtByV(params, dfdx)
        for (int q = 0; q < nl; ++q) dfdx[m.li[o + q]] -= m.ls[o + q] * w;
        for (int q = 0; q < nr; ++q) dfdx[m.ri[o + q]] += m.rs[o + q] * w;
        // jac_dCfdcj_contribution
        for (int j = 0; j < nl; ++j) {
            const int sj = m.li[o + j];
            for (int i = 0; i < nl; ++i) b.Jat(m.li[o + i], sj) -= m.ls[o + i] * dCf[j];
            for (int i = 0; i < nr; ++i) b.Jat(m.ri[o + i], sj) += m.rs[o + i] * dCf[j];
        }
        // jac_dCrdcj_contribution (zero for irreversible reactions)
        if (rev) {
            for (int j = 0; j < nr; ++j) {
                const int sj = m.ri[o + j];
                for (int i = 0; i < nl; ++i) b.Jat(m.li[o + i], sj) += m.ls[o + i] * dCr[j];
                for (int i = 0; i < nr; ++i) b.Jat(m.ri[o + i], sj) -= m.rs[o + i] * dCr[j];
            }
        }
        // jac_dCdT_contribution
        for (int i = 0; i < nl; ++i) b.Jat(m.li[o + i], nS) -= m.ls[o + i] * dwdT;
        for (int i = 0; i < nr; ++i) b.Jat(m.ri[o + i], nS) += m.rs[o + i] * dwdT;
        // jac_dCdC_contribution (rates of a third-body concentration only)
        if (uses_third_body(t)) {
            const double temp = (Cf - Cr / Kc);
            if (t == GFC_RATE_THIRD_BODY) {
                const double* eff = m.eff + (long)r * nS;
                for (int i = 0; i < nl; ++i) {
                    const int si = m.li[o + i];
                    const double sl = m.ls[o + i];
                    for (int j = 0; j < nS; ++j) b.Jat(si, j) -= sl * ((eff[j] * ddcScale) * temp);
                }
                for (int i = 0; i < nr; ++i) {
                    const int si = m.ri[o + i];
                    const double sr = m.rs[o + i];
                    for (int j = 0; j < nS; ++j) b.Jat(si, j) += sr * ((eff[j] * ddcScale) * temp);
                }
            } else if (!(fabs(temp) <= DBL_MAX)) {  // not finite
                // fall-off rates have ddc = 0; 0*inf poisons the rows as in
                // the reference
                const double nanv = 0.0 * temp;
                for (int i = 0; i < nl; ++i)
                    for (int j = 0; j < nS; ++j) b.Jat(m.li[o + i], j) -= m.ls[o + i] * nanv;
                for (int i = 0; i < nr; ++i)
                    for (int j = 0; j < nS; ++j) b.Jat(m.ri[o + i], j) += m.rs[o + i] * nanv;
            }
        }
    }
    // dYdT and dTdt of the derivative computed alongside
    for (int i = 0; i < nS; ++i) dfdx[i] *= m.W[i] / rhoM;
    {
        double ret = 0.0;
        for (int i = 0; i < nS; ++i) ret -= dfdx[i] * b.Ha()[i];
        dfdx[nS] = ret / CpM;
    }
    // ddY_dcdY (gpuODESystem.H:65-103); row i copied to ddc as temp
    SV temp = b.ddc();
    for (int i = 0; i < nS; ++i) {
        const double WiByrhoM = m.W[i] / rhoM;
        for (int j = 0; j <= nS; ++j) temp[j] = b.Jat(i, j);
        if (b.raw)
            for (int j = 0; j <= nS; ++j) b.raw[((long)i * n + j) * b.s] = temp[j];
        for (int j = 0; j < nS; ++j) {
            const double rhoMByWi = rhoM / m.W[j];
            const double vj = 1.0 / (p / (m.R[j] * T));
            b.Jat(i, j) = WiByrhoM * temp[j] * rhoMByWi + rhoM * vj * dfdx[i];
        }
        double ddNidtByVdT = temp[nS];
        for (int j = 0; j < nS; ++j) ddNidtByVdT -= temp[j] * c[j] * alphavM;
        b.Jat(i, nS) = WiByrhoM * ddNidtByVdT + alphavM * dfdx[i];
    }
    // ddT_dtdY (:108-126)
    for (int i = 0; i < nS; ++i) {
        double v = 0.0;
        for (int j = 0; j < nS; ++j) v -= b.Jat(j, i) * b.Ha()[j];
        v -= b.Cp()[i] * dfdx[nS];
        b.Jat(nS, i) = v / CpM;
    }
    // ddT_dtdT (:129-146)
    {
        double v = 0.0;
        for (int i = 0; i < nS; ++i) v -= dfdx[i] * b.Cp()[i] + b.Jat(i, nS) * b.Ha()[i];
        v -= dfdx[nS] * dCpMdT;
        b.Jat(nS, nS) = v / CpM;
    }
}

// ---------------------------------------------------------------- LU --

// ludecompose.H:9-92 (Crout, implicit scaling, last maximal row wins)
GFC_HD __forceinline__ void lu_decompose(LaneBuf b) {
    const int m = b.n;
    for (int i = 0; i < m; i++) {
        double largest = 0.0;
        for (int j = 0; j < m; j++) {
            const double t = fabs(b.Jat(i, j));
            if (t > largest) largest = t;
        }
        b.vv()[i] = 1.0 / largest;
    }
    for (int j = 0; j < m; j++) {
        for (int i = 0; i < j; i++) {
            double sum = b.Jat(i, j);
            for (int k = 0; k < i; k++) sum -= b.Jat(i, k) * b.Jat(k, j);
            b.Jat(i, j) = sum;
        }
        int iMax = 0;
        double largest = 0.0;
        for (int i = j; i < m; i++) {
            double sum = b.Jat(i, j);
            for (int k = 0; k < j; k++) sum -= b.Jat(i, k) * b.Jat(k, j);
            b.Jat(i, j) = sum;
            const double t = b.vv()[i] * fabs(sum);
            if (t >= largest) {
                largest = t;
                iMax = i;
            }
        }
        b.piv()[j] = iMax;
        if (j != iMax) {
            for (int k = 0; k < m; k++) {
                const double s = b.Jat(j, k);
                b.Jat(j, k) = b.Jat(iMax, k);
                b.Jat(iMax, k) = s;
            }
            b.vv()[iMax] = b.vv()[j];
        }
        if (b.Jat(j, j) == 0.0) b.Jat(j, j) = kSmall;
        if (j != m - 1) {
            const double rDiag = 1.0 / b.Jat(j, j);
            for (int i = j + 1; i < m; i++) b.Jat(i, j) *= rDiag;
        }
    }
}

// ludecompose.H:94-130
GFC_HD __forceinline__ void lu_backsub(LaneBuf b, SV x) {
    const int m = b.n;
    int ii = 0;
    for (int i = 0; i < m; i++) {
        const int ip = b.piv()[i];
        double sum = x[ip];
        x[ip] = x[i];
        if (ii != 0) {
            for (int j = ii - 1; j < i; j++) sum -= b.Jat(i, j) * x[j];
        } else if (sum != 0.0) {
            ii = i + 1;
        }
        x[i] = sum;
    }
    for (int i = m - 1; i >= 0; i--) {
        double sum = x[i];
        for (int j = i + 1; j < m; j++) sum -= b.Jat(i, j) * x[j];
        x[i] = sum / b.Jat(i, i);
    }
}

// ------------------------------------------------------------ solvers --

// One Rosenbrock step (gpuRosenbrock{12,23,34}.H::solve); returns the
// normalised error (normaliseError.H:6-29).
GFC_HD inline double ros_step(const SolveArgs& A, double p, SV y0, double dx,
                           SV y, LaneBuf b) {
    const int n = b.n;
    const RosTableau& t = A.tab;
    jacobian(A.m, p, y0, b);
    if (b.jt)
        for (int i = 0; i < n * n; ++i) b.jt[(long)i * b.s] = b.J()[i];
    if (b.dydx0FromF)
        for (int i = 0; i < n; ++i) b.dydx0()[i] = b.dfdx()[i];
    // computeDecomposedJacobian (gpuAdaptiveSolver.H:51-73)
    const double temp = 1.0 / (t.gamma * dx);
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) b.Jat(i, j) = -b.Jat(i, j);
        b.Jat(i, i) += temp;
    }
    lu_decompose(b);
    SV k1 = b.k1(), k2 = b.k2(), k3 = b.k3(), k4 = b.k4(), dfdx = b.dfdx(), dydx = b.dydx();
    for (int i = 0; i < n; ++i) k1[i] = b.dydx0()[i] + dx * t.d1 * dfdx[i];
    lu_backsub(b, k1);
    for (int i = 0; i < n; ++i) y[i] = y0[i] + t.a21 * k1[i];
    derivatives(A.m, p, y, dydx, b);
    if (b.f2)
        for (int i = 0; i < n; ++i) b.f2[(long)i * b.s] = dydx[i];
    for (int i = 0; i < n; ++i) k2[i] = dydx[i] + dx * t.d2 * dfdx[i] + t.c21 * k1[i] / dx;
    lu_backsub(b, k2);
    double maxErr = 0.0;
    if (t.stages == 2) {
        for (int i = 0; i < n; ++i) {
            y[i] = y0[i] + t.b1 * k1[i] + t.b2 * k2[i];
            const double err = t.e1 * k1[i] + t.e2 * k2[i];
            const double tol = A.absTol + A.relTol * fmax(fabs(y0[i]), fabs(y[i]));
            maxErr = fmax(maxErr, fabs(err) / tol);
        }
        return maxErr;
    }
    if (t.stages == 4) {
        for (int i = 0; i < n; ++i) y[i] = y0[i] + t.a31 * k1[i] + t.a32 * k2[i];
        derivatives(A.m, p, y, dydx, b);
    }
    for (int i = 0; i < n; ++i)
        k3[i] = dydx[i] + dx * t.d3 * dfdx[i] + (t.c31 * k1[i] + t.c32 * k2[i]) / dx;
    lu_backsub(b, k3);
    if (t.stages == 3) {
        for (int i = 0; i < n; ++i) {
            y[i] = y0[i] + t.b1 * k1[i] + t.b2 * k2[i] + t.b3 * k3[i];
            const double err = t.e1 * k1[i] + t.e2 * k2[i] + t.e3 * k3[i];
            const double tol = A.absTol + A.relTol * fmax(fabs(y0[i]), fabs(y[i]));
            maxErr = fmax(maxErr, fabs(err) / tol);
        }
        return maxErr;
    }
    for (int i = 0; i < n; ++i)
        k4[i] = dydx[i] + dx * t.d4 * dfdx[i] +
                (t.c41 * k1[i] + t.c42 * k2[i] + t.c43 * k3[i]) / dx;
    lu_backsub(b, k4);
    for (int i = 0; i < n; ++i) {
        y[i] = y0[i] + t.b1 * k1[i] + t.b2 * k2[i] + t.b3 * k3[i] + t.b4 * k4[i];
        const double err = t.e1 * k1[i] + t.e2 * k2[i] + t.e4 * k4[i];
        const double tol = A.absTol + A.relTol * fmax(fabs(y0[i]), fabs(y[i]));
        maxErr = fmax(maxErr, fabs(err) / tol);
    }
    return maxErr;
}

// Accepted-step observer of the step controller (no-op for the solve):
// accepted(y0, dx) is called with the state at the start of every accepted
// step and the step size it was taken with.
struct NoRecord {
    GFC_HD __forceinline__ void accepted(SV, double) const {}
};

// adaptiveSolve (gpuAdaptiveSolver.H:83-134)
template <class Rec>
GFC_HD inline int adaptive(const SolveArgs& A, double p, double& x, double& dxTry,
                    LaneBuf b, int& attempts, int& accepted, Rec& rec) {
    const int n = b.n;
    double dx = dxTry, err = 0.0;
    derivatives(A.m, p, b.y(), b.dydx0(), b);
    do {
        err = ros_step(A, p, b.y(), dx, b.yTemp(), b);
        ++attempts;
        if (err > 1) {
            const double scale = fmax(A.safeScale * pow(err, -A.alphaDec), A.minScale);
            dx *= scale;
            if (dx < kVSmall) return GFC_CELL_STEP_UNDERFLOW;
        }
    } while (err > 1.0);
    ++accepted;
    rec.accepted(b.y(), dx);
    x += dx;
    for (int i = 0; i < n; ++i) b.y()[i] = b.yTemp()[i];
    if (err > A.incThreshold) {
        dxTry = fmin(fmax(A.safeScale * pow(err, -A.alphaInc), A.minScale), A.maxScale) * dx;
    } else {
        dxTry = A.safeScale * A.maxScale * dx;
    }
    return 0;
}

// odeSolve (gpuODESolverBase.H:67-111)
template <class Rec>
GFC_HD inline int ode_solve(const SolveArgs& A, double p, double xEnd, double& dxTry,
                     LaneBuf b, int& attempts, int& accepted, Rec& rec) {
    const double xStart = 0.0;
    double stepDx = dxTry;
    bool last = false;
    double x = xStart;
    for (int nStep = 0; nStep < A.maxSteps; nStep++) {
        const double dxTry0 = stepDx;
        if ((x + stepDx - xEnd) * (x + stepDx - xStart) > 0) {
            last = true;
            stepDx = xEnd - x;
        }
        const int st = adaptive(A, p, x, stepDx, b, attempts, accepted, rec);
        if (st) return st;
        if ((x - xEnd) * (xEnd - xStart) >= 0) {
            if (nStep > 0 && last) stepDx = dxTry0;
            dxTry = stepDx;
            return 0;
        }
    }
    return GFC_CELL_MAX_STEPS;
}

} // namespace lane
} // namespace gfc
