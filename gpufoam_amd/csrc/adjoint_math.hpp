// Discrete adjoint of one accepted Rosenbrock step for one cell (one lane).
//
// The forward map of an accepted step of size h from y0 is (n = nSpecie+1,
// A = I/(gamma h) - J(y0), F = the dfdx vector of the Jacobian call):
//   k1 = A^-1 (f(y0) + h d1 F)
//   k2 = A^-1 (f(y2) + h d2 F + c21 k1/h),           y2 = y0 + a21 k1
//   k3 = A^-1 (f(y3) + h d3 F + (c31 k1 + c32 k2)/h), y3 = y0 + a31 k1 + a32 k2
//   k4 = A^-1 (f(y3) + h d4 F + (c41 k1 + c42 k2 + c43 k3)/h)
//   y1 = y0 + sum_s b_s k_s
// (gpuRosenbrock34.H:67-139; Rosenbrock23 uses f(y2) for k3 and stops
// there, Rosenbrock12 stops after k2).  With the step sizes of the forward
// solve held fixed, the reverse of this map needs
//   * transposed solves with A (the LU of the forward step),
//   * vector-Jacobian products of the exact right-hand side f
//     (gpuODESystem.H:164-196), and
//   * the gradient of <Jbar, J(y0)> + <Fbar, F(y0)> for the reference's
//     "fast" analytic Jacobian (gpuODESystem.H:198-281 with
//     gpuReaction.H:547-702), i.e. the derivative of the Jacobian exactly as
//     the reference assembles it, including its simplified fall-off terms.
// Every derivative below is the derivative of the arithmetic the forward
// code performs (the functions of lane_physics.hpp), so the gradients are
// those of the discrete map the solver computed.
//
// Rate parameters: six per reaction, (ln|A|, beta, Ta) of the rate
// expression (k0 for fall-off rates) and of kInf (fall-off rates only),
// where k = A T^beta exp(-Ta/T) (gpuArrheniusReactionRate.H:16-26; the
// activation temperature Ta = Ea/R, makeGpuReactionRate.H:9-14).
#pragma once

#include "lane_physics.hpp"

namespace gfc {
namespace adj {

using lane::LaneBuf;
using lane::SV;

constexpr int kParams = 6; // per reaction

// d2Cp/dT2 of a NASA-7 polynomial
GFC_HD __forceinline__ double poly_d2CpdT2(const double* a, double T) {
    return (12.0 * a[4] * T + 6.0 * a[3]) * T + 2.0 * a[2];
}

// Adjoint work arrays of one lane.  Vectors have n elements, matrices n*n
// (element (i, j) at [(i*n + j)*s]); `b` is the forward scratch with its
// raw/jt copies pointing at the raw and jt matrices below.
struct AdjBuf {
    LaneBuf b;
    double* base;
    long s;
    int n;
    GFC_HD __forceinline__ SV vec(int k) const { return SV{base + (long)k * n * s, s}; }
    GFC_HD __forceinline__ SV ybar() const { return vec(0); }
    GFC_HD __forceinline__ SV kb(int q) const { return vec(1 + q); } // q = 0..3
    GFC_HD __forceinline__ SV rb() const { return vec(5); }
    GFC_HD __forceinline__ SV Fb() const { return vec(6); }
    GFC_HD __forceinline__ SV fb2() const { return vec(7); }
    GFC_HD __forceinline__ SV fb3() const { return vec(8); }
    GFC_HD __forceinline__ SV yst() const { return vec(9); }
    GFC_HD __forceinline__ SV ysb() const { return vec(10); }
    GFC_HD __forceinline__ SV fval() const { return vec(11); }
    GFC_HD __forceinline__ SV cb() const { return vec(12); }
    GFC_HD __forceinline__ SV Hab() const { return vec(13); }
    GFC_HD __forceinline__ SV Cpb() const { return vec(14); }
    GFC_HD __forceinline__ SV vb() const { return vec(15); }
    GFC_HD __forceinline__ SV dNb() const { return vec(16); }
    GFC_HD __forceinline__ SV dCp() const { return vec(17); }
    GFC_HD __forceinline__ SV f2() const { return vec(18); }
    static constexpr int kVectors = 19;
    GFC_HD __forceinline__ double* mat(int k) const {
        return base + ((long)kVectors * n + (long)k * n * n) * s;
    }
    GFC_HD __forceinline__ double& Jb(int i, int j) const { return mat(0)[((long)i * n + j) * s]; }
    GFC_HD __forceinline__ double& raw(int i, int j) const { return mat(1)[((long)i * n + j) * s]; }
    GFC_HD __forceinline__ double& jt(int i, int j) const { return mat(2)[((long)i * n + j) * s]; }
    static constexpr int kMatrices = 3;
};

GFC_HD inline long adj_scratch_doubles(int nS) {
    const long n = nS + 1;
    return (long)AdjBuf::kVectors * n + (long)AdjBuf::kMatrices * n * n;
}

// Per-cell parameter adjoints: th[(kParams*r + q)*s].
struct ThView {
    double* p;
    long s;
    GFC_HD __forceinline__ void add(int r, const double* v) const {
        if (!p) return;
        for (int q = 0; q < kParams; ++q)
            if (v[q] != 0.0) p[((long)kParams * r + q) * s] += v[q];
    }
};

// ------------------------------------------------------------- mixture --

struct Mix {
    double rho, S1, CpM, dCpMdT, alphavRaw, alphav;
};

// gpuMixture::computeJacobianRequired (gpuMixture.H:33-77); the derivative
// path (gpuMixture.H:88-121) computes the same rho, CpM, Ha and c.
GFC_HD inline Mix mixture(const DeviceMech& m, double p, SV y, SV c, SV Ha,
                          SV Cp, SV dCp) {
    const int nS = m.nS;
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
        Cp[i] = Cpi;
        dCp[i] = poly_dCpdT(a, T);
        Ha[i] = poly_Ha(a, T);
        c[i] = yi / m.W[i];
    }
    Mix x;
    x.S1 = rhoM;
    x.rho = 1.0 / rhoM;
    x.CpM = CpM;
    x.dCpMdT = dCpMdT;
    x.alphavRaw = alphav;
    x.alphav = alphav * x.rho;
    for (int i = 0; i < nS; ++i) c[i] *= x.rho;
    return x;
}

// Reverse of mixture(): adds to ybar the contributions of the adjoints of
// c, rho, Cp_i, Ha_i, CpM, dCpMdT, alphav and v_i = 1/rho_i.
struct MixBar {
    double rho = 0, CpM = 0, dCpMdT = 0, alphav = 0;
};
GFC_HD inline void mixture_reverse(const DeviceMech& m, double p, SV y, const Mix& x,
                                   SV Cp, SV dCp, SV cb, SV Cpb, SV Hab, SV vb,
                                   bool full, MixBar mb, SV ybar) {
    const int nS = m.nS;
    const double T = y[nS];
    const double rho = x.rho;
    // alphav = alphavRaw * rho; alphavRaw = sum yi vi / T does not depend on T
    double rhob = mb.rho + mb.alphav * x.alphavRaw;
    const double arb = mb.alphav * rho;
    for (int i = 0; i < nS; ++i) rhob += cb[i] * (fmax(y[i], 0.0) / m.W[i]);
    const double S1b = -rhob * rho * rho;
    double Tb = S1b * x.S1 / T + mb.CpM * x.dCpMdT;
    for (int i = 0; i < nS; ++i) {
        const double yc = fmax(y[i], 0.0);
        const double vi = 1.0 / (p / (m.R[i] * T));
        double ycb = cb[i] * rho / m.W[i] + S1b * vi + arb * vi * (1.0 / T) + mb.CpM * Cp[i];
        Tb += Hab[i] * Cp[i];
        if (full) {
            const double* a = coeffs_of(m.high + 7 * i, m.low + 7 * i, m.Tcommon[i], T);
            ycb += mb.dCpMdT * dCp[i];
            Tb += mb.dCpMdT * yc * poly_d2CpdT2(a, T) + Cpb[i] * dCp[i] + vb[i] * vi / T;
        }
        if (y[i] >= 0.0) ybar[i] += ycb; // right derivative of max(y, 0) at 0
    }
    ybar[nS] += Tb;
}

// ------------------------------------------------------------ reactions --

// Adjoint of an Arrhenius expression k = A T^b exp(-Ta/T) with respect to
// T and (ln|A|, b, Ta), given the adjoints of its value k and of its
// temperature derivative k q, q = (b + Ta/T)/T.  Returned by value (not
// added through a pointer), so that a caller's per-reaction accumulators
// stay in registers.
struct ArrBar {
    double T, lnA, beta, Ta;
};
GFC_HD __forceinline__ ArrBar arr_adjoint(const double* kk, double iT, double lnT, double k,
                                          double kb, double kqb) {
    // iT = 1/T: the adjoint's divisions by T are products with it
    const double be = kk[1], Ta = kk[2];
    const double q = (be + Ta * iT) * iT;
    const double dq = -(be + 2.0 * Ta * iT) * (iT * iT);
    ArrBar a;
    a.T = kb * k * q + kqb * (k * q * q + k * dq);
    a.lnA = kb * k + kqb * k * q;
    a.beta = kb * k * lnT + kqb * (k * lnT * q + k * iT);
    a.Ta = -kb * k * iT + kqb * (k * (iT * iT) - k * q * iT);
    return a;
}

// Troe blending function F(T, Pr) (gpuTroeFallOffFunction.H:27-47) and its
// partial derivatives.
struct TroeVal {
    double F, dT, dPr;
};
GFC_HD inline TroeVal troe_eval(const double* tr, double T, double Pr) {
    const double alpha = tr[0], Tsss = tr[1], Ts = tr[2], Tss = tr[3];
    const double logPr = log10(fmax(Pr, kSmall));
    const double dlogPr = Pr > kSmall ? 1.0 / (Pr * kLogTen) : 0.0;
    const double e1 = exp(-T / Tsss), e2 = exp(-T / Ts), e3 = exp(-Tss / T);
    const double Fcent = (1.0 - alpha) * e1 + alpha * e2 + e3;
    const double dFcent = -(1.0 - alpha) * e1 / Tsss - alpha * e2 / Ts + e3 * Tss / (T * T);
    const double logFcent = log10(fmax(Fcent, kSmall));
    const double dlogFc = Fcent > kSmall ? dFcent / (Fcent * kLogTen) : 0.0;
    const double c = -0.4 - 0.67 * logFcent;
    const double d = 0.14;
    const double nn = 0.75 - 1.27 * logFcent;
    const double u = logPr + c;
    const double x1 = nn - d * u;
    const double x2 = u / x1;
    const double x3 = 1.0 + x2 * x2;
    const double x4 = logFcent / x3;
    TroeVal r;
    r.F = pow(10.0, x4);
    const double dx2_dlogPr = (x1 + d * u) / (x1 * x1);
    const double du = -0.67, dx1 = -1.27 - d * du;
    const double dx2_dlogFc = (du * x1 - u * dx1) / (x1 * x1);
    const double dx4_dlogPr = -logFcent / (x3 * x3) * 2.0 * x2 * dx2_dlogPr;
    const double dx4_dlogFc = 1.0 / x3 - logFcent / (x3 * x3) * 2.0 * x2 * dx2_dlogFc;
    r.dPr = r.F * kLogTen * dx4_dlogPr * dlogPr;
    r.dT = r.F * kLogTen * dx4_dlogFc * dlogFc;
    return r;
}

// Kc, dKcdTbyKc (gpuThermo.H:80-169) and their temperature derivatives.
struct KcVal {
    double Kc, dK, dlnKc, ddK;
};
GFC_HD inline KcVal kc_eval(const DeviceMech& m, int r, double T) {
    KcVal v;
    KcTerms t;
    reaction_Kc_dKc(m, r, T, v.Kc, v.dK, &t);
    const double iT = 1.0 / T;
    const double HT = t.S + t.Gstd * iT; // H/T
    const double YR = t.Y / t.RRT;
    v.dlnKc = (t.arg < 600 ? HT * YR : 0.0) - (t.nmSmall ? 0.0 : t.nm * iT);
    v.ddK = YR * (poly_Cp(t.a, T) * iT - 2.0 * HT * iT) +
            (t.nmSmall ? 0.0 : t.nm * (iT * iT));
    return v;
}

GFC_HD __forceinline__ double specie_pow_m2(double x, int ei, double er) {
    if (ei != kNoIntExp) return integer_pow(x, ei - 2);
    return pow(x, er - 2.0);
}

// Accumulation into an adjoint vector (one lane per cell: a plain add; a
// team of lanes sharing one cell supplies an atomic overload).
GFC_HD __forceinline__ void acc_add(SV v, int i, double x) { v[i] += x; }

// One side of a reaction: C = prod c^e, dC[j] = dC/dc_j (product without j).
// W bounds the number of species on a side.  The loops run to W with the
// entries beyond n masked, so that they unroll and the arrays stay in
// registers.
template <int W>
struct SideT {
    int n;
    const int* idx;
    const int* ei;
    const double* er;
    double pw[W], d1[W], dC[W];
    double C;
    GFC_HD void eval(SV c) {
#pragma unroll
        for (int q = 0; q < W; ++q) {
            pw[q] = 1.0;
            d1[q] = 0.0;
            if (q < n) {
                pw[q] = specie_pow(c[idx[q]], ei[q], er[q]);
                d1[q] = er[q] * specie_pow_m1(c[idx[q]], ei[q], er[q]);
            }
        }
        C = 1.0;
#pragma unroll
        for (int q = 0; q < W; ++q)
            if (q < n) C *= pw[q];
#pragma unroll
        for (int j = 0; j < W; ++j) {
            double v = 1;
#pragma unroll
            for (int i = 0; i < W; ++i)
                if (i < n) v *= (i == j) ? d1[i] : pw[i];
            dC[j] = v;
        }
    }
    // cbar[idx[l]] += Cb dC/dc_l + sum_j dCb[j] d2C/dc_j dc_l
    template <class CB>
    GFC_HD void reverse(SV c, double Cb, const double* dCb, CB cbar) const {
#pragma unroll
        for (int l = 0; l < W; ++l) {
            if (l >= n) continue;
            double acc = Cb * dC[l];
#pragma unroll
            for (int j = 0; j < W; ++j) {
                if (j >= n || dCb[j] == 0.0) continue;
                double h;
                if (j == l) {
                    const double coef = er[j] * (er[j] - 1.0);
                    if (coef == 0.0) continue;
                    h = coef * specie_pow_m2(c[idx[j]], ei[j], er[j]);
#pragma unroll
                    for (int q = 0; q < W; ++q)
                        if (q < n && q != j) h *= pw[q];
                } else {
                    h = d1[j] * d1[l];
#pragma unroll
                    for (int q = 0; q < W; ++q)
                        if (q < n && q != j && q != l) h *= pw[q];
                }
                acc += dCb[j] * h;
            }
            acc_add(cbar, idx[l], acc);
        }
    }
};

// Adjoints of the quantities one reaction contributes: omega, and for the
// Jacobian its fast dOmega/dT, kf*dCf/dc_j, kr*dCr/dc_j and the third-body
// scalar D (ddc_j = eff_j*D).
template <int W>
struct RxnBarT {
    double w = 0, dwdT = 0, D = 0;
    double dCf[W], dCr[W];
    bool fast = false;
};
using RxnBar = RxnBarT<kWidth>;

// Reverse of reaction r at temperature T (clipped to [Tlow, Thigh] when
// `clip`, as gpuReaction::omega does; the Jacobian uses T unclipped).
// Adds to cbar, Tb and th[0..5].  W bounds the species per reaction side.
// With MSET the caller supplies the third-body concentration (Mset) and
// receives its adjoint in *MbOut instead of the scatter eff_i * Mb into
// cbar (a caller that sums the reactions of one efficiency set first).
template <int W, class CB, bool MSET = false>
GFC_HD __forceinline__ void reaction_reverse_w(const DeviceMech& m, int r, double T, bool clip,
                                            SV c, const RxnBarT<W>& in, CB cbar, double& Tb,
                                            double* th, double Mset = 0.0,
                                            double* MbOut = nullptr) {
    const double Te = clip ? fmin(fmax(T, m.Tlow), m.Thigh) : T;
    const double lnT = log(Te);
    const int t = m.type[r];
    const double* kk = m.k + 3 * r;
    const long o = (long)r * kWidth;
    const bool rev = !m.irreversible[r];
    const double* eff = m.eff + (long)r * m.nS;
    double M = Mset;
    if (!MSET && t != GFC_RATE_ARRHENIUS)
        for (int i = 0; i < m.nS; ++i) M += eff[i] * c[i];
    const double kA = arrhenius(kk, Te);
    const double iTe = 1.0 / Te;
    const double qA = (kk[1] + kk[2] * iTe) * iTe;
    double kf, dkfdT, kI = 0.0, qI = 0.0, temp = 0.0, g = 0.0, Pr = 0.0;
    TroeVal F{1.0, 0.0, 0.0};
    const double* ki = m.kInf + 3 * r;
    if (t == GFC_RATE_ARRHENIUS) {
        kf = kA;
        dkfdT = kA * qA;
    } else if (t == GFC_RATE_THIRD_BODY) {
        kf = M * kA;
        dkfdT = M * (kA * qA);
    } else {
        kI = arrhenius(ki, Te);
        qI = (ki[1] + ki[2] * iTe) * iTe;
        Pr = kA / kI * M;
        if (t == GFC_RATE_TROE) F = troe_eval(m.troe + 4 * r, Te, Pr);
        g = Pr / (1.0 + Pr);
        temp = g * F.F;
        kf = kI * temp;
        dkfdT = (kI * qI) * temp;
    }
    SideT<W> L{m.nl[r], m.li + o, m.lei + o, m.le + o};
    L.eval(c);
    SideT<W> R{m.nr[r], m.ri + o, m.rei + o, m.re + o};
    double Cr = 0.0;
    KcVal K{1.0, 0.0, 0.0, 0.0};
    if (rev) {
        R.eval(c);
        Cr = R.C;
        K = kc_eval(m, r, Te);
    }
    const double Cf = L.C;
    const double iKc = 1.0 / K.Kc;
    const double kr = rev ? kf * iKc : 0.0;
    const double dkrdT = rev ? dkfdT * iKc - kr * K.dK : 0.0;

    double Cfb = in.w * kf, Crb = -in.w * kr, kfb = in.w * Cf, krb = -in.w * Cr;
    double dkfdTb = 0.0, dkrdTb = 0.0, Kcb = 0.0, kAb = 0.0;
    double dCLb[W], dCRb[W];
    for (int j = 0; j < W; ++j) dCLb[j] = dCRb[j] = 0.0;
    if (in.fast) {
        Cfb += in.dwdT * dkfdT;
        Crb += -in.dwdT * dkrdT;
        dkfdTb = in.dwdT * Cf;
        dkrdTb = -in.dwdT * Cr;
#pragma unroll
        for (int j = 0; j < W; ++j)
            if (j < L.n) {
                kfb += in.dCf[j] * L.dC[j];
                dCLb[j] = in.dCf[j] * kf;
            }
        if (rev)
#pragma unroll
            for (int j = 0; j < W; ++j)
                if (j < R.n) {
                    krb += in.dCr[j] * R.dC[j];
                    dCRb[j] = in.dCr[j] * kr;
                }
        if (t == GFC_RATE_THIRD_BODY) {
            Cfb += in.D * kA;
            Crb += -in.D * kA * iKc;
            Kcb += in.D * kA * Cr * (iKc * iKc);
            kAb += in.D * (Cf - Cr * iKc);
        }
    }
    double TbE = 0.0;
    if (rev) {
        dkfdTb += dkrdTb * iKc;
        Kcb += -dkrdTb * dkfdT * (iKc * iKc);
        krb += -dkrdTb * K.dK;
        const double dKb = -dkrdTb * kr;
        kfb += krb * iKc;
        Kcb += -krb * kr * iKc;
        TbE += Kcb * K.Kc * K.dlnKc + dKb * K.ddK;
    }
    L.reverse(c, Cfb, dCLb, cbar);
    if (rev) R.reverse(c, Crb, dCRb, cbar);
    double Mb = 0.0;
    ArrBar a0{0.0, 0.0, 0.0, 0.0}, a1{0.0, 0.0, 0.0, 0.0};
    if (t == GFC_RATE_ARRHENIUS) {
        a0 = arr_adjoint(kk, iTe, lnT, kA, kfb, dkfdTb);
    } else if (t == GFC_RATE_THIRD_BODY) {
        Mb = kfb * kA + dkfdTb * (kA * qA);
        a0 = arr_adjoint(kk, iTe, lnT, kA, kfb * M + kAb, dkfdTb * M);
    } else {
        double kIb = kfb * temp;
        const double tempb = kfb * kI + dkfdTb * (kI * qI);
        const double kqIb = dkfdTb * temp;
        const double gp = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
        const double Prb = tempb * (gp * F.F + g * F.dPr);
        TbE += tempb * g * F.dT;
        const double ikI = 1.0 / kI;
        const double k0b = Prb * M * ikI;
        Mb = Prb * kA * ikI;
        kIb += -Prb * Pr * ikI;
        a0 = arr_adjoint(kk, iTe, lnT, kA, k0b, 0.0);
        a1 = arr_adjoint(ki, iTe, lnT, kI, kIb, kqIb);
    }
    TbE += a0.T;
    TbE += a1.T;
    th[0] += a0.lnA;
    th[1] += a0.beta;
    th[2] += a0.Ta;
    th[3] += a1.lnA;
    th[4] += a1.beta;
    th[5] += a1.Ta;
    if (MSET) *MbOut = Mb;
    else if (Mb != 0.0)
        for (int i = 0; i < m.nS; ++i) acc_add(cbar, i, eff[i] * Mb);
    if (!clip || (T >= m.Tlow && T <= m.Thigh)) Tb += TbE;
}

GFC_HD __forceinline__ void reaction_reverse(const DeviceMech& m, int r, double T, bool clip,
                                          SV c, const RxnBar& in, SV cbar, double& Tb,
                                          double* th) {
    reaction_reverse_w<kWidth>(m, r, T, clip, c, in, cbar, Tb, th);
}

// ---------------------------------------------------- right-hand side --

// ybar += d<fbar, f(y)>/dy, th += d<fbar, f(y)>/dtheta.  fy: f(y) when the
// caller has it (fy.p non-null), else it is evaluated.
GFC_HD inline __noinline__ void vjp_f(const DeviceMech& m, double p, SV y, SV fbar, SV ybar,
                               ThView th, const AdjBuf& a, SV fy) {
    const int nS = m.nS;
    const LaneBuf& b = a.b;
    SV f = fy.p ? fy : a.fval();
    SV c = b.c(), Ha = b.Ha(), Cp = b.Cp(), dCp = a.dCp();
    if (!fy.p) lane::derivatives(m, p, y, f, b);
    const Mix x = mixture(m, p, y, c, Ha, Cp, dCp);
    SV cb = a.cb(), Hab = a.Hab(), dNb = a.dNb();
    MixBar mb;
    const double g = fbar[nS];
    mb.CpM = -g * f[nS] / x.CpM;
    for (int i = 0; i < nS; ++i) {
        Hab[i] = -g * f[i] / x.CpM;
        const double fbi = fbar[i] - g * Ha[i] / x.CpM;
        dNb[i] = fbi * m.W[i] / x.rho;
        mb.rho += -fbi * f[i] / x.rho;
        cb[i] = 0.0;
    }
    const double T = y[nS];
    double Tb = 0.0;
    for (int r = 0; r < m.nR; ++r) {
        const long o = (long)r * kWidth;
        RxnBar in;
        for (int q = 0; q < m.nl[r]; ++q) in.w -= m.ls[o + q] * dNb[m.li[o + q]];
        for (int q = 0; q < m.nr[r]; ++q) in.w += m.rs[o + q] * dNb[m.ri[o + q]];
        if (in.w == 0.0) continue;
        double t6[kParams] = {0, 0, 0, 0, 0, 0};
        reaction_reverse(m, r, T, true, c, in, cb, Tb, t6);
        th.add(r, t6);
    }
    ybar[nS] += Tb;
    mixture_reverse(m, p, y, x, Cp, dCp, cb, a.Cpb(), Hab, a.vb(), false, mb, ybar);
}

// ------------------------------------------------------------- Jacobian --

// ybar += d/dy and th += d/dtheta of <Jb, J(y)> + <Fb, F(y)>, where J and F
// are the fast Jacobian and the dfdx vector of lane::jacobian at y; a.raw
// and a.jt hold that call's pre-transform species rows and its result, and
// b.dfdx() its F.  Jb is consumed.
GFC_HD inline __noinline__ void jact(const DeviceMech& m, double p, SV y, SV ybar, ThView th,
                              const AdjBuf& a) {
    const int nS = m.nS, iT = nS;
    const LaneBuf& b = a.b;
    SV c = b.c(), Ha = b.Ha(), Cp = b.Cp(), dCp = a.dCp(), F = b.dfdx();
    const Mix x = mixture(m, p, y, c, Ha, Cp, dCp);
    const double T = y[nS];
    SV cb = a.cb(), Hab = a.Hab(), Cpb = a.Cpb(), vb = a.vb(), Fb = a.Fb(), dNb = a.dNb();
    for (int i = 0; i < nS; ++i) cb[i] = Hab[i] = Cpb[i] = vb[i] = 0.0;
    MixBar mb;
    const double CpM = x.CpM, rho = x.rho, av = x.alphav;
    // ddT_dtdT (gpuODESystem.H:129-146)
    {
        const double g = a.Jb(iT, iT);
        for (int i = 0; i < nS; ++i) {
            Fb[i] += -g * Cp[i] / CpM;
            Cpb[i] += -g * F[i] / CpM;
            a.Jb(i, iT) += -g * Ha[i] / CpM;
            Hab[i] += -g * a.jt(i, iT) / CpM;
        }
        Fb[iT] += -g * x.dCpMdT / CpM;
        mb.dCpMdT += -g * F[iT] / CpM;
        mb.CpM += -g * a.jt(iT, iT) / CpM;
    }
    // ddT_dtdY (:108-126)
    for (int i = 0; i < nS; ++i) {
        const double g = a.Jb(iT, i);
        if (g == 0.0) continue;
        for (int j = 0; j < nS; ++j) {
            a.Jb(j, i) += -g * Ha[j] / CpM;
            Hab[j] += -g * a.jt(j, i) / CpM;
        }
        Cpb[i] += -g * F[iT] / CpM;
        Fb[iT] += -g * Cp[i] / CpM;
        mb.CpM += -g * a.jt(iT, i) / CpM;
    }
    // ddY_dcdY (:65-103), species columns, in place: Jb(i, j) becomes the
    // adjoint of the raw (reaction-assembled) entry
    for (int i = 0; i < nS; ++i) {
        const double WiR = m.W[i] / rho;
        for (int j = 0; j < nS; ++j) {
            const double g = a.Jb(i, j);
            const double vj = 1.0 / (p / (m.R[j] * T));
            mb.rho += g * vj * F[i];
            vb[j] += g * rho * F[i];
            Fb[i] += g * rho * vj;
            a.Jb(i, j) = g * WiR * (rho / m.W[j]);
        }
    }
    // ddY_dcdY, temperature column
    for (int i = 0; i < nS; ++i) {
        const double g = a.Jb(i, iT);
        const double WiR = m.W[i] / rho;
        double sumc = 0.0;
        for (int j = 0; j < nS; ++j) sumc += a.raw(i, j) * c[j];
        const double X = a.raw(i, iT) - sumc * av;
        a.Jb(i, iT) = g * WiR;
        for (int j = 0; j < nS; ++j) {
            a.Jb(i, j) += -g * WiR * c[j] * av;
            cb[j] += -g * WiR * a.raw(i, j) * av;
        }
        mb.alphav += g * (F[i] - WiR * sumc);
        mb.rho += -g * m.W[i] / (rho * rho) * X;
        Fb[i] += g * av;
    }
    // dfdx: F_T = -sum F_i Ha_i / CpM, F_i = dN_i W_i / rho
    {
        const double g = Fb[iT];
        for (int i = 0; i < nS; ++i) {
            Fb[i] += -g * Ha[i] / CpM;
            Hab[i] += -g * F[i] / CpM;
        }
        mb.CpM += -g * F[iT] / CpM;
        for (int i = 0; i < nS; ++i) {
            dNb[i] = Fb[i] * m.W[i] / rho;
            mb.rho += -Fb[i] * F[i] / rho;
        }
    }
    // reactions (gpuReaction.H:547-702): adjoints of omega, dOmega/dT,
    // kf dCf/dc_j, kr dCr/dc_j and D from the raw-entry adjoints
    double Tb = 0.0;
    for (int r = 0; r < m.nR; ++r) {
        const long o = (long)r * kWidth;
        const int nl = m.nl[r], nr = m.nr[r];
        RxnBar in;
        in.fast = true;
        // u(col) = sum over the reaction's entries of nu * Jb(species, col)
        auto u = [&](int col) {
            double v = 0.0;
            for (int q = 0; q < nl; ++q) v -= m.ls[o + q] * a.Jb(m.li[o + q], col);
            for (int q = 0; q < nr; ++q) v += m.rs[o + q] * a.Jb(m.ri[o + q], col);
            return v;
        };
        for (int q = 0; q < nl; ++q) in.w -= m.ls[o + q] * dNb[m.li[o + q]];
        for (int q = 0; q < nr; ++q) in.w += m.rs[o + q] * dNb[m.ri[o + q]];
        for (int j = 0; j < nl; ++j) in.dCf[j] = u(m.li[o + j]);
        for (int j = 0; j < nr; ++j) in.dCr[j] = -u(m.ri[o + j]);
        in.dwdT = u(iT);
        if (m.type[r] == GFC_RATE_THIRD_BODY) {
            const double* eff = m.eff + (long)r * nS;
            double D = 0.0;
            for (int j = 0; j < nS; ++j)
                if (eff[j] != 0.0) D += eff[j] * u(j);
            in.D = D;
        }
        double t6[kParams] = {0, 0, 0, 0, 0, 0};
        reaction_reverse(m, r, T, false, c, in, cb, Tb, t6);
        th.add(r, t6);
    }
    ybar[nS] += Tb;
    mixture_reverse(m, p, y, x, Cp, dCp, cb, Cpb, Hab, vb, true, mb, ybar);
}

// -------------------------------------------------------------- linear --

// Solves A^T x = rhs in place, with A's LU and pivots in b (ludecompose.H:
// P A = L U, L unit lower, the row swaps in piv).
GFC_HD inline __noinline__ void lu_solve_T(const LaneBuf& b, SV x) {
    const int m = b.n;
    for (int i = 0; i < m; ++i) {
        double sum = x[i];
        for (int k = 0; k < i; ++k) sum -= b.Jat(k, i) * x[k];
        x[i] = sum / b.Jat(i, i);
    }
    for (int i = m - 1; i >= 0; --i) {
        double sum = x[i];
        for (int k = i + 1; k < m; ++k) sum -= b.Jat(k, i) * x[k];
        x[i] = sum;
    }
    for (int j = m - 1; j >= 0; --j) {
        const int ip = b.piv()[j];
        if (ip != j) {
            const double t = x[j];
            x[j] = x[ip];
            x[ip] = t;
        }
    }
}

// ----------------------------------------------------------------- step --

// Reverse of one accepted step of size h from y0: on entry a.ybar() holds
// the adjoint of the step's result, on exit that of y0; th accumulates the
// parameter adjoints.  Recomputes the step (one Jacobian, LU and two right-
// hand sides) to obtain its stages.
GFC_HD inline __noinline__ void step_reverse(const SolveArgs& A, double p, SV y0, double h,
                                      ThView th, const AdjBuf& a) {
    const DeviceMech& m = A.m;
    const int n = a.n;
    const LaneBuf& b = a.b;
    const RosTableau& t = A.tab;
    // recompute the step; within the reaction temperature limits f(y0) is
    // the Jacobian call's F, so the separate right-hand side is skipped
    const double T0 = y0[m.nS];
    const bool inRange = T0 >= m.Tlow && T0 <= m.Thigh;
    LaneBuf bs = b;
    bs.dydx0FromF = inRange;
    if (!inRange) lane::derivatives(m, p, y0, b.dydx0(), b);
    (void)lane::ros_step(A, p, y0, h, b.yTemp(), bs);
    SV yb = a.ybar(), rb = a.rb(), Fb = a.Fb(), fb2 = a.fb2(), fb3 = a.fb3();
    SV k1 = b.k1(), k2 = b.k2(), k3 = b.k3(), k4 = b.k4();
    SV kb1 = a.kb(0), kb2 = a.kb(1), kb3 = a.kb(2), kb4 = a.kb(3);
    for (int i = 0; i < n; ++i) {
        kb1[i] = t.b1 * yb[i];
        kb2[i] = t.b2 * yb[i];
        kb3[i] = t.b3 * yb[i];
        kb4[i] = t.b4 * yb[i];
        Fb[i] = fb2[i] = fb3[i] = 0.0;
    }
    for (int i = 0; i < n * n; ++i) a.mat(0)[(long)i * a.s] = 0.0;
    auto add_outer = [&](SV r, SV k) {
        for (int i = 0; i < n; ++i) {
            const double ri = r[i];
            if (ri == 0.0) continue;
            for (int j = 0; j < n; ++j) a.Jb(i, j) += ri * k[j];
        }
    };
    if (t.stages == 4) {
        for (int i = 0; i < n; ++i) rb[i] = kb4[i];
        lu_solve_T(b, rb);
        add_outer(rb, k4);
        for (int i = 0; i < n; ++i) {
            fb3[i] += rb[i];
            Fb[i] += h * t.d4 * rb[i];
            kb1[i] += t.c41 / h * rb[i];
            kb2[i] += t.c42 / h * rb[i];
            kb3[i] += t.c43 / h * rb[i];
        }
    }
    if (t.stages >= 3) {
        for (int i = 0; i < n; ++i) rb[i] = kb3[i];
        lu_solve_T(b, rb);
        add_outer(rb, k3);
        SV fb = t.stages == 4 ? fb3 : fb2;
        for (int i = 0; i < n; ++i) {
            fb[i] += rb[i];
            Fb[i] += h * t.d3 * rb[i];
            kb1[i] += t.c31 / h * rb[i];
            kb2[i] += t.c32 / h * rb[i];
        }
    }
    SV ys = a.yst(), ysb = a.ysb();
    if (t.stages == 4) {
        for (int i = 0; i < n; ++i) {
            ys[i] = y0[i] + t.a31 * k1[i] + t.a32 * k2[i];
            ysb[i] = 0.0;
        }
        vjp_f(m, p, ys, fb3, ysb, th, a, b.dydx()); // the step left f(y3) there
        for (int i = 0; i < n; ++i) {
            yb[i] += ysb[i];
            kb1[i] += t.a31 * ysb[i];
            kb2[i] += t.a32 * ysb[i];
        }
    }
    for (int i = 0; i < n; ++i) rb[i] = kb2[i];
    lu_solve_T(b, rb);
    add_outer(rb, k2);
    for (int i = 0; i < n; ++i) {
        fb2[i] += rb[i];
        Fb[i] += h * t.d2 * rb[i];
        kb1[i] += t.c21 / h * rb[i];
    }
    for (int i = 0; i < n; ++i) {
        ys[i] = y0[i] + t.a21 * k1[i];
        ysb[i] = 0.0;
    }
    vjp_f(m, p, ys, fb2, ysb, th, a, t.stages == 4 ? a.f2() : b.dydx());
    for (int i = 0; i < n; ++i) {
        yb[i] += ysb[i];
        kb1[i] += t.a21 * ysb[i];
    }
    for (int i = 0; i < n; ++i) rb[i] = kb1[i];
    lu_solve_T(b, rb);
    add_outer(rb, k1);
    for (int i = 0; i < n; ++i) Fb[i] += h * t.d1 * rb[i];
    // rb is now the adjoint of f(y0) (the dydx0 of the step).  Within the
    // reaction temperature limits f(y0) is the same function as the
    // Jacobian call's F (omega clips T, the Jacobian does not), so its
    // adjoint joins Fbar and one reverse sweep serves both.
    if (inRange) {
        for (int i = 0; i < n; ++i) Fb[i] += rb[i];
    } else {
        vjp_f(m, p, y0, rb, yb, th, a, b.dydx0());
    }
    jact(m, p, y0, yb, th, a);
}

} // namespace adj
} // namespace gfc

namespace gfc {
namespace adj {

// --------------------------------------------------------------- one cell --

// Records every accepted step's size (hs[t*ld]) and every K-th step's
// initial state (ck[(slot*n + e)*ld]); pointers are offset to the cell.
struct Recorder {
    double* ck;
    double* hs;
    long ld;
    int n;
    int K;
    int count;
    GFC_HD __forceinline__ void accepted(SV y0, double dx) {
        hs[(long)count * ld] = dx;
        if (count % K == 0) {
            const long slot = count / K;
            for (int e = 0; e < n; ++e) ck[(slot * n + e) * ld] = y0[e];
        }
        ++count;
    }
};

// singleCellSolver (singleCellSolver.H:27-51) with recording: b.y() holds
// the initial (Y, T) on entry and the final state on exit; yEnd receives
// the state before the final non-negativity clamp.  The reference's time
// loop runs once (dt = timeLeft), so one odeSolve call is made.
GFC_HD inline int cell_forward(const SolveArgs& A, double p, const LaneBuf& b, double& dtChem,
                               Recorder& rec, SV yEnd, int& attempts, int& accepted) {
    const int nS = A.m.nS, n = nS + 1;
    if (!(b.y()[nS] > A.Treact)) {
        dtChem = A.deltaT;
        return GFC_CELL_SKIPPED;
    }
    int status = 0;
    if (A.deltaT > kSmall) {
        status = lane::ode_solve(A, p, A.deltaT, dtChem, b, attempts, accepted, rec);
        for (int k = 0; k < n; ++k) yEnd[k] = b.y()[k];
        for (int i = 0; i < nS; i++) b.y()[i] = fmax(0.0, b.y()[i]);
    }
    return status;
}

// Reverse of cell_forward: a.ybar() holds dL/d(final state) on entry and
// dL/d(initial state) on exit; th accumulates dL/dparams.  `seg` is a
// K-state buffer; ck/hs as recorded, N accepted steps, T0 the initial
// temperature.
GFC_HD inline void cell_reverse(const SolveArgs& A, double p, double T0, const AdjBuf& a,
                                double* seg, const double* ck, const double* hs, long ld,
                                int K, int N, SV yEnd, ThView th) {
    const int nS = A.m.nS, n = nS + 1;
    if (!(T0 > A.Treact) || !(A.deltaT > kSmall)) return;
    SV yb = a.ybar();
    for (int i = 0; i < nS; ++i)
        if (yEnd[i] < 0.0) yb[i] = 0.0;
    auto segv = [&](int t) { return SV{seg + (long)t * n * a.s, a.s}; };
    const int nseg = (N + K - 1) / K;
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * K, s1 = (s0 + K < N) ? s0 + K : N;
        SV y = segv(0);
        for (int e = 0; e < n; ++e) y[e] = ck[((long)sg * n + e) * ld];
        for (int t = s0; t < s1 - 1; ++t) {
            SV yt = segv(t - s0);
            lane::derivatives(A.m, p, yt, a.b.dydx0(), a.b);
            (void)lane::ros_step(A, p, yt, hs[(long)t * ld], segv(t - s0 + 1), a.b);
        }
        for (int t = s1 - 1; t >= s0; --t) step_reverse(A, p, segv(t - s0), hs[(long)t * ld], th, a);
    }
}

} // namespace adj
} // namespace gfc
