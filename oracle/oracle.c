/*
 * CPU oracle: plain-C restatement of gpuFoam's per-cell chemistry path.
 * TEST INFRASTRUCTURE ONLY (checker, CPU baseline); see oracle.h.
 *
 * Build with a compiler that does not contract a*b+c into fused
 * multiply-adds (gcc -ffp-contract=off), so that every rounding step is the
 * one the reference performs on the host.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* gpu_utils/common/gpu_constants.H:12-82 */
#define RR_ (6.02214e+26 * 1.38065e-23)
#define PSTD_ 1e5
#define TSTD_ 298.15
#define SMALL_ DBL_EPSILON
#define VSMALL_ DBL_MIN
#define VGREAT_ (DBL_MAX / 10.0)
#define LOGTEN_ 2.302585092994046
#define NOINT_ GFC_NO_INT_EXPONENT
#define WMAX_ GFC_MAX_REACTION_WIDTH

/* ------------------------------------------------------------ thermo --- */

/* gpuThermo::coeffs, gpuThermo.H:29-32 */
static inline const double* coeffs_of(const double* high, const double* low,
                                      double Tcommon, double T) {
    return (T < Tcommon) ? low : high;
}

/* gpuThermo::Cp, gpuThermo.H:177-184 (perfect-gas part is 0) */
static inline double poly_Cp(const double* a, double T) {
    return ((((a[4] * T + a[3]) * T + a[2]) * T + a[1]) * T + a[0]);
}

/* gpuThermo::Ha, gpuThermo.H:210-221 */
static inline double poly_Ha(const double* a, double T) {
    return (((((a[4] / 5.0 * T + a[3] / 4.0) * T + a[2] / 3.0) * T +
              a[1] / 2.0) * T + a[0]) * T + a[5]);
}

/* gpuThermo::dCpdT, gpuThermo.H:333-338 */
static inline double poly_dCpdT(const double* a, double T) {
    return (((4.0 * a[4] * T + 3.0 * a[3]) * T + 2.0 * a[2]) * T + a[1]);
}

/* gpuThermo::S without the perfect-gas term, gpuThermo.H:261-271 */
static inline double poly_S(const double* a, double T, double logT) {
    return ((((a[4] / 4.0 * T + a[3] / 3.0) * T + a[2] / 2.0) * T + a[1]) *
                T + a[0] * logT + a[6]);
}

/* gpuThermo::Gstd, gpuThermo.H:287-299 */
static inline double poly_Gstd(const double* a, double T, double logT) {
    return ((a[0] * (1.0 - logT) -
             (((a[4] / 20.0 * T + a[3] / 12.0) * T + a[2] / 6.0) * T +
              a[1] / 2.0) * T - a[6]) * T + a[5]);
}

/* gpuThermo::Hf, gpuThermo.H:243-254 (low coefficients at Tstd) */
static inline double poly_Hf(const double* a) {
    const double T = TSTD_;
    return (((((a[4] / 5.0 * T + a[3] / 4.0) * T + a[2] / 3.0) * T +
              a[1] / 2.0) * T + a[0]) * T + a[5]);
}

void orc_species_thermo(const gfc_mech_view* m, int i, double p, double T,
                        double out[6]) {
    const double* a = coeffs_of(m->high + 7 * i, m->low + 7 * i,
                                m->sTcommon[i], T);
    const double R = RR_ / m->W[i];
    const double logT = log(T);
    out[0] = poly_Cp(a, T);
    out[1] = poly_Ha(a, T);
    out[2] = poly_Hf(m->low + 7 * i);
    /* gpuPerfectGas::Sp, gpuPerfectGas.H:210-214 */
    out[3] = poly_S(a, T, logT) + (-R * log(p / PSTD_));
    out[4] = poly_Gstd(a, T, logT);
    out[5] = poly_dCpdT(a, T);
}

/* gpuThermo::Kc (gpuThermo.H:80-113), the form omega() uses */
static double reaction_Kc_only(const gfc_mech_view* m, int r, double T) {
    const double Y = m->rY[r], W = m->rW[r];
    const double* a = coeffs_of(m->rhigh + 7 * r, m->rlow + 7 * r,
                                m->rTcommon[r], T);
    const double nm = Y / W;
    const double arg = -Y * poly_Gstd(a, T, log(T)) / (RR_ * T);
    const double K = (arg < 600) ? exp(arg) : sqrt(VGREAT_);
    if (fabs(nm - SMALL_) < VSMALL_) return K;
    return K * pow(PSTD_ / (RR_ * T), nm);
}

/* gpuThermo::KcAnddKcTbyKc, gpuThermo.H:139-169 */
void orc_reaction_Kc(const gfc_mech_view* m, int r, double p, double T,
                     double out[2]) {
    (void)p;
    const double Y = m->rY[r], W = m->rW[r];
    const double* a = coeffs_of(m->rhigh + 7 * r, m->rlow + 7 * r,
                                m->rTcommon[r], T);
    const double nm = Y / W;
    const double logT = log(T);
    const double Gstd = poly_Gstd(a, T, logT);
    const int nm_small = fabs(nm - SMALL_) < VSMALL_;
    const double RRT = RR_ * T;
    const double arg = -Y * Gstd / RRT;
    const double K = (arg < 600) ? exp(arg) : sqrt(VGREAT_);
    const double Kc = nm_small ? K : K * pow(PSTD_ / RRT, nm);
    /* S(pStd, T) carries -R*log(pStd/pStd) = -0, which adds nothing */
    const double S = poly_S(a, T, logT);
    double d = (S + Gstd / T) * Y / RRT;
    if (!nm_small) d = d - nm / T;
    out[0] = Kc;
    out[1] = d;
}

/* ------------------------------------------------------------ rates --- */

/* gpuArrheniusReactionRate::computeAk, gpuArrheniusReactionRate.H:16-26 */
static inline double arrhenius(const double* k, double T) {
    double ak = k[0];
    if (fabs(k[1]) > VSMALL_) ak *= pow(T, k[1]);
    if (fabs(k[2]) > VSMALL_) ak *= exp(-k[2] / T);
    return ak;
}

/* gpuArrheniusReactionRate::ddT, :58-64 */
static inline double arrhenius_ddT(const double* k, double T) {
    return arrhenius(k, T) * (k[1] + k[2] / T) / T;
}

/* gpuThirdBodyEfficiencies::M, gpuThirdBodyEfficiencies.H:21-27 */
static inline double third_body_M(const gfc_mech_view* m, int r,
                                  const double* c) {
    const double* eff = m->eff + (long)r * m->n_species;
    double M = 0;
    for (int i = 0; i < m->n_species; ++i) M += eff[i] * c[i];
    return M;
}

/* gpuTroeFallOffFunction::operator(), gpuTroeFallOffFunction.H:27-47 */
static double troe_F(const double* tr, double T, double Pr) {
    const double alpha = tr[0], Tsss = tr[1], Ts = tr[2], Tss = tr[3];
    const double logPr = log10(fmax(Pr, SMALL_));
    const double Fcent = (1.0 - alpha) * exp(-T / Tsss) +
                         alpha * exp(-T / Ts) + exp(-Tss / T);
    const double logFcent = log10(fmax(Fcent, SMALL_));
    const double c = -0.4 - 0.67 * logFcent;
    const double d = 0.14;
    const double n = 0.75 - 1.27 * logFcent;
    const double x1 = n - d * (logPr + c);
    const double x2 = (logPr + c) / x1;
    const double x3 = 1.0 + x2 * x2;
    const double x4 = logFcent / x3;
    return pow(10.0, x4);
}

/* rates that depend on a third-body concentration M (these carry the
 * reference's dense concentration-derivative row, zero for fall-off) */
static inline int uses_third_body(int type) {
    return type >= GFC_RATE_THIRD_BODY && type <= GFC_RATE_SRI_CA;
}

/* ---- rate types gpuFoam rejects (makeMoreGpuChemistryModels.C:59-95);
 * restated from their published forms, as OpenFOAM's rate classes take
 * them (no gpuFoam counterpart, so no reference vectors: pinned by closed
 * forms and finite differences, tests/test_extended_rates.py). */

/* SRI fall-off function (Stewart, Larson & Golden 1989):
 * F = d [a exp(-b/T) + exp(-T/c)]^X T^e, X = 1/(1 + log10(max(Pr,small))^2) */
static double sri_F(const double* x, double T, double Pr) {
    const double lp = log10(fmax(Pr, SMALL_));
    const double X = 1.0 / (1.0 + lp * lp);
    return x[3] * pow(x[0] * exp(-x[1] / T) + exp(-T / x[2]), X) * pow(T, x[4]);
}

/* fall-off / chemically-activated blending function F(T, Pr) */
static double blend_F(const gfc_mech_view* m, int r, double T, double Pr) {
    const int t = m->type[r];
    if (t == GFC_RATE_TROE || t == GFC_RATE_TROE_CA) return troe_F(m->troe + 4 * r, T, Pr);
    if (t == GFC_RATE_SRI || t == GFC_RATE_SRI_CA)
        return sri_F(m->extra + (long)GFC_EXTRA_PARAMS * r, T, Pr);
    return 1.0;
}

/* Landau-Teller, Janev and power-series rates and their dk/dT:
 *   LT     k = A T^b exp(-Ta/T + B T^(-1/3) + C T^(-2/3))
 *   Janev  k = A T^b exp(-Ta/T) exp(sum_{n=0..8} b_n (ln T)^n)
 *   power  k = A T^b exp(-Ta/T + sum_{n=1..4} c_n T^(-n))
 * each factor skipped when its parameter is below vSmall in magnitude, as
 * Arrhenius does. */
static double tdep_rate(const gfc_mech_view* m, int r, double T, double* ddT) {
    const double* k = m->k + 3 * r;
    const double* x = m->extra + (long)GFC_EXTRA_PARAMS * r;
    double v = k[0];
    if (fabs(k[1]) > VSMALL_) v *= pow(T, k[1]);
    double arg = 0.0, darg = 0.0; /* exponent and its T derivative */
    if (fabs(k[2]) > VSMALL_) {
        arg -= k[2] / T;
        darg += k[2] / (T * T);
    }
    switch (m->type[r]) {
    case GFC_RATE_LANDAU_TELLER: {
        const double t13 = pow(T, 1.0 / 3.0);
        if (fabs(x[0]) > VSMALL_) {
            arg += x[0] / t13;
            darg -= x[0] / (3.0 * T * t13);
        }
        if (fabs(x[1]) > VSMALL_) {
            arg += x[1] / (t13 * t13);
            darg -= 2.0 * x[1] / (3.0 * T * t13 * t13);
        }
        break;
    }
    case GFC_RATE_JANEV: {
        const double lnT = log(T);
        double Tn = 1.0, dsum = 0.0;
        for (int n = 0; n < 9; ++n) {
            arg += x[n] * Tn;
            if (n + 1 < 9) dsum += (n + 1) * x[n + 1] * Tn;
            Tn *= lnT;
        }
        darg += dsum / T;
        break;
    }
    default: { /* power series */
        double Tn = 1.0;
        for (int n = 1; n <= 4; ++n) {
            Tn *= T;
            arg += x[n - 1] / Tn;
            darg -= n * x[n - 1] / (Tn * T);
        }
        break;
    }
    }
    if (fabs(arg) > VSMALL_) v *= exp(arg);
    if (ddT) *ddT = v * (k[1] / T + darg);
    return v;
}

/* gpuReactionRate::operator() dispatch (gpuReactionRate.H:135-173) */
static double rate_k(const gfc_mech_view* m, int r, double T, const double* c) {
    const double* k = m->k + 3 * r;
    switch (m->type[r]) {
    case GFC_RATE_LANDAU_TELLER:
    case GFC_RATE_JANEV:
    case GFC_RATE_POWER_SERIES:
        return tdep_rate(m, r, T, NULL);
    case GFC_RATE_LINDEMANN_CA:
    case GFC_RATE_TROE_CA:
    case GFC_RATE_SRI_CA: {
        /* chemically activated: k = k0 / (1 + Pr) F */
        const double k0 = arrhenius(k, T);
        const double kInf = arrhenius(m->kInf + 3 * r, T);
        const double Pr = k0 / kInf * third_body_M(m, r, c);
        return k0 * (1.0 / (1.0 + Pr)) * blend_F(m, r, T, Pr);
    }
    case GFC_RATE_ARRHENIUS: return arrhenius(k, T);
    case GFC_RATE_THIRD_BODY:
        /* gpuThirdBodyArrheniusReactionRate.H:39-44 */
        return third_body_M(m, r, c) * arrhenius(k, T);
    default: {
        /* gpuFallOffReactionRate.H:216-227 */
        const double k0 = arrhenius(k, T);
        const double kInf = arrhenius(m->kInf + 3 * r, T);
        const double M = third_body_M(m, r, c);
        const double Pr = k0 / kInf * M;
        const double F = blend_F(m, r, T, Pr);
        return kInf * (Pr / (1.0 + Pr)) * F;
    }
    }
}

/* gpuReactionRate::everything (gpuReactionRate.H:237-245): {k, dk/dT};
 * ddc is written for the concentration-dependent rates. */
static void rate_everything(const gfc_mech_view* m, int r, double T,
                            const double* c, double out[2], double* ddc) {
    const double* k = m->k + 3 * r;
    const int nS = m->n_species;
    switch (m->type[r]) {
    case GFC_RATE_ARRHENIUS: {
        const double Ak = arrhenius(k, T);
        out[0] = Ak;
        out[1] = Ak * (k[1] + k[2] / T) / T;
        return;
    }
    case GFC_RATE_THIRD_BODY: {
        /* gpuThirdBodyArrheniusReactionRate.H:105-121 */
        const double Ak = arrhenius(k, T);
        const double dT = Ak * (k[1] + k[2] / T) / T;
        const double* eff = m->eff + (long)r * nS;
        for (int i = 0; i < nS; ++i) ddc[i] = eff[i] * Ak;
        const double M = third_body_M(m, r, c);
        out[0] = M * Ak;
        out[1] = M * dT;
        return;
    }
    case GFC_RATE_LANDAU_TELLER:
    case GFC_RATE_JANEV:
    case GFC_RATE_POWER_SERIES: {
        double d;
        out[0] = tdep_rate(m, r, T, &d);
        out[1] = d;
        return;
    }
    case GFC_RATE_LINDEMANN_CA:
    case GFC_RATE_TROE_CA:
    case GFC_RATE_SRI_CA: {
        /* gpuFoam's fall-off convention carried over: only the leading
         * Arrhenius factor (k0 here) is differentiated in T, and there is
         * no concentration derivative */
        const double k0 = arrhenius(k, T);
        const double kInf = arrhenius(m->kInf + 3 * r, T);
        const double Pr = k0 / kInf * third_body_M(m, r, c);
        const double temp = (1.0 / (1.0 + Pr)) * blend_F(m, r, T, Pr);
        out[0] = k0 * temp;
        out[1] = arrhenius_ddT(k, T) * temp;
        for (int i = 0; i < nS; ++i) ddc[i] = 0.0;
        return;
    }
    default: {
        /* gpuFallOffReactionRate.H:303-335: the Jacobian keeps only the
         * kInf temperature derivative and no concentration derivative. */
        const double k0 = arrhenius(k, T);
        const double kInf = arrhenius(m->kInf + 3 * r, T);
        const double M = third_body_M(m, r, c);
        const double Pr = k0 / kInf * M;
        const double F = blend_F(m, r, T, Pr);
        const double temp = (Pr / (1.0 + Pr)) * F;
        const double kk = kInf * temp;
        const double dkInfdT = arrhenius_ddT(m->kInf + 3 * r, T);
        out[0] = kk;
        out[1] = dkInfdT * temp;
        for (int i = 0; i < nS; ++i) ddc[i] = 0.0;
        return;
    }
    }
}

/* gpuSpecieExponent.H:21-36 integerPow */
static inline double integer_pow(double x, int e) {
    if (e == 0) return 1;
    double xx = e > 0 ? x : 1 / x;
    double y = 1;
    for (int i = e; i != 0; i /= 2) {
        if (i % 2) y *= xx;
        xx *= xx;
    }
    return y;
}

/* speciePow, gpuSpecieExponent.H:103-110 */
static inline double specie_pow(double x, int eint, double ereal) {
    if (eint != NOINT_) return integer_pow(x, eint);
    return pow(x, ereal);
}

/* speciePow(x, e - 1): integer exponents stay integer
 * (gpuSpecieExponent.H:78-99) */
static inline double specie_pow_minus_one(double x, int eint, double ereal) {
    if (eint != NOINT_) return integer_pow(x, eint - 1);
    return pow(x, ereal + (-1.0));
}

static inline int rcoef(const gfc_mech_view* m, int r, int side, int q,
                        int* idx, double* st, int* ei, double* er) {
    const long o = (long)r * WMAX_ + q;
    if (side == 0) {
        *idx = m->lhs_index[o]; *st = m->lhs_stoich[o];
        *ei = m->lhs_exp_int[o]; *er = m->lhs_exp[o];
        return m->n_lhs[r];
    }
    *idx = m->rhs_index[o]; *st = m->rhs_stoich[o];
    *ei = m->rhs_exp_int[o]; *er = m->rhs_exp[o];
    return m->n_rhs[r];
}

/* lhsPowers/rhsPowers + calcCf/calcCr (gpuReaction.H:348-386) */
static double side_powers(const gfc_mech_view* m, int r, int side,
                          const double* c, double* pw) {
    const int n = side == 0 ? m->n_lhs[r] : m->n_rhs[r];
    for (int q = 0; q < n; ++q) {
        int idx, ei; double st, er;
        rcoef(m, r, side, q, &idx, &st, &ei, &er);
        pw[q] = specie_pow(c[idx], ei, er);
    }
    double C = 1.0;
    for (int q = 0; q < n; ++q) C *= pw[q];
    return C;
}

/* calcdCfdcj / calcdCrdcj: product without j (gpuReaction.H:388-470) */
static void side_derivs(const gfc_mech_view* m, int r, int side,
                        const double* c, const double* pw, double* d) {
    const int n = side == 0 ? m->n_lhs[r] : m->n_rhs[r];
    for (int j = 0; j < n; ++j) {
        double v = 1;
        for (int i = 0; i < n; ++i) {
            int idx, ei; double st, er;
            rcoef(m, r, side, i, &idx, &st, &ei, &er);
            if (i == j) v *= er * specie_pow_minus_one(c[idx], ei, er);
            else v *= pw[i];
        }
        d[j] = v;
    }
}

/* gpuReaction::omega, gpuReaction.H:472-488 */
void orc_reaction_rates(const gfc_mech_view* m, int r, double p, double T,
                        const double* c, double out[3]) {
    (void)p;
    const double Tc = fmin(fmax(T, m->Tlow), m->Thigh);
    const double kf = rate_k(m, r, Tc, c);
    const double kr = m->irreversible[r] ? 0.0 : kf / reaction_Kc_only(m, r, Tc);
    double pw[WMAX_];
    const double Cf = side_powers(m, r, 0, c, pw);
    const double Cr = side_powers(m, r, 1, c, pw);
    out[0] = kf;
    out[1] = kr;
    out[2] = kf * Cf - kr * Cr;
}

/* computeReactionParameters, gpuReaction.H:650-702 (unclipped T) */
void orc_reaction_parameters(const gfc_mech_view* m, int r, double p,
                             double T, const double* c, double out[2],
                             double* dCf, double* dCr, double* ddc) {
    const int nS = m->n_species;
    double ev[2];
    rate_everything(m, r, T, c, ev, ddc);
    const double kf = ev[0], ddT = ev[1];
    double pw[WMAX_];
    const double Cf = side_powers(m, r, 0, c, pw);
    side_derivs(m, r, 0, c, pw, dCf);
    double dwdT = Cf * ddT;
    double omega = Cf * kf;
    for (int j = 0; j < m->n_lhs[r]; ++j) dCf[j] *= kf;
    for (int j = m->n_lhs[r]; j < WMAX_; ++j) dCf[j] = 0.0;
    double Cr = 0.0, Kc = 1.0;
    for (int j = 0; j < WMAX_; ++j) dCr[j] = 0.0;
    if (!m->irreversible[r]) {
        double kk[2];
        orc_reaction_Kc(m, r, p, T, kk);
        Kc = kk[0];
        const double dKcdTbyV = kk[1];
        const double kr = kf / Kc;
        Cr = side_powers(m, r, 1, c, pw);
        const double dkrdT = ddT / Kc - kr * dKcdTbyV;
        dwdT -= Cr * dkrdT;
        omega -= Cr * kr;
        side_derivs(m, r, 1, c, pw, dCr);
        for (int j = 0; j < m->n_rhs[r]; ++j) dCr[j] *= kr;
    }
    if (uses_third_body(m->type[r])) {
        const double temp = (Cf - Cr / Kc);
        for (int j = 0; j < nS; ++j) ddc[j] *= temp;
    }
    out[0] = omega;
    out[1] = dwdT;
}

/* ------------------------------------------------------- ODE system --- */

long orc_workspace_doubles(int nS) {
    const long nE = nS + 2;
    /* c, Ha, Cp, ddc/temp, 13 vectors, J, vv, pivots (as doubles) */
    return 4 * nE + 13 * nE + nE * nE + nE + nE + 64;
}

/* gpuMixture::computeDerivativeRequired, gpuMixture.H:88-121 */
static void mixture_derivative(const gfc_mech_view* m, const double* y,
                               double* Ha, double* c, double* rhoM_,
                               double* CpM_) {
    const int nS = m->n_species;
    const double T = y[nS], p = y[nS + 1];
    double rhoM = 0.0, CpM = 0.0;
    for (int i = 0; i < nS; ++i) {
        const double yi = fmax(y[i], 0.0);
        const double rhoi = p / ((RR_ / m->W[i]) * T);
        const double* a = coeffs_of(m->high + 7 * i, m->low + 7 * i,
                                    m->sTcommon[i], T);
        const double Cpi = poly_Cp(a, T);
        const double Hai = poly_Ha(a, T);
        rhoM += yi / rhoi;
        CpM += yi * Cpi;
        Ha[i] = Hai;
        c[i] = yi / m->W[i];
    }
    rhoM = 1.0 / rhoM;
    for (int i = 0; i < nS; ++i) c[i] *= rhoM;
    *rhoM_ = rhoM;
    *CpM_ = CpM;
}

/* gpuMixture::computeJacobianRequired, gpuMixture.H:33-77 */
static void mixture_jacobian(const gfc_mech_view* m, const double* y,
                             double* Cp, double* Ha, double* c, double out[4]) {
    const int nS = m->n_species;
    const double T = y[nS], p = y[nS + 1];
    double rhoM = 0.0, CpM = 0.0, dCpMdT = 0.0, alphav = 0.0;
    for (int i = 0; i < nS; ++i) {
        const double yi = fmax(y[i], 0.0);
        const double rhoi = p / ((RR_ / m->W[i]) * T);
        const double vi = 1.0 / rhoi;
        const double* a = coeffs_of(m->high + 7 * i, m->low + 7 * i,
                                    m->sTcommon[i], T);
        const double Cpi = poly_Cp(a, T);
        const double Hai = poly_Ha(a, T);
        const double dCpMdTi = poly_dCpdT(a, T);
        rhoM += yi / rhoi;
        CpM += yi * Cpi;
        dCpMdT += yi * dCpMdTi;
        alphav += yi * vi * (1.0 / T);
        Cp[i] = Cpi;
        Ha[i] = Hai;
        c[i] = yi / m->W[i];
    }
    rhoM = 1.0 / rhoM;
    alphav = alphav * rhoM;
    for (int i = 0; i < nS; ++i) c[i] *= rhoM;
    out[0] = rhoM;
    out[1] = CpM;
    out[2] = dCpMdT;
    out[3] = alphav;
}

/* gpuReaction::dNdtByV(omega, .), gpuReaction.H:490-508 */
static void add_dNdtByV(const gfc_mech_view* m, int r, double omega,
                        double* dNdtByV) {
    for (int q = 0; q < m->n_lhs[r]; ++q) {
        const long o = (long)r * WMAX_ + q;
        dNdtByV[m->lhs_index[o]] -= m->lhs_stoich[o] * omega;
    }
    for (int q = 0; q < m->n_rhs[r]; ++q) {
        const long o = (long)r * WMAX_ + q;
        dNdtByV[m->rhs_index[o]] += m->rhs_stoich[o] * omega;
    }
}

/* gpuODESystem::derivatives, gpuODESystem.H:164-196 */
void orc_derivatives(const gfc_mech_view* m, const double* y, double* dydx,
                     double* work) {
    const int nS = m->n_species, nE = nS + 2;
    double* Ha = work;
    double* c = work + nE;
    for (int i = 0; i < nE; ++i) dydx[i] = 0;
    const double T = y[nS], p = y[nS + 1];
    double rhoM, CpM;
    mixture_derivative(m, y, Ha, c, &rhoM, &CpM);
    for (int r = 0; r < m->n_reactions; ++r) {
        double o[3];
        orc_reaction_rates(m, r, p, T, c, o);
        add_dNdtByV(m, r, o[2], dydx);
    }
    /* dYdT, :40-48 */
    for (int i = 0; i < nS; ++i) dydx[i] *= m->W[i] / rhoM;
    /* dTdt, :53-63 */
    double ret = 0.0;
    for (int i = 0; i < nS; ++i) ret -= dydx[i] * Ha[i];
    dydx[nS] = ret / CpM;
    dydx[nS + 1] = 0.0;
}

/* gpuODESystem::jacobian, gpuODESystem.H:198-281 with ddY_dcdY (:65-103),
 * ddT_dtdY (:108-126) and ddT_dtdT (:129-146) */
void orc_jacobian(const gfc_mech_view* m, const double* y, double* dfdx,
                  double* J, double* work) {
    const int nS = m->n_species, nE = nS + 2;
    double* Cp = work;
    double* Ha = work + nE;
    double* c = work + 2 * nE;
    double* ddc = work + 3 * nE; /* also "temp" of ddY_dcdY */
#define J_(i, j) J[(long)(i) * nE + (j)]
    for (int i = 0; i < nE; ++i) dfdx[i] = 0.0;
    for (int i = 0; i < nE * nE; ++i) J[i] = 0.0;
    const double T = y[nS], p = y[nS + 1];
    double mix[4];
    mixture_jacobian(m, y, Cp, Ha, c, mix);
    const double rhoM = mix[0], CpM = mix[1], dCpMdT = mix[2],
                 alphavM = mix[3];
    for (int r = 0; r < m->n_reactions; ++r) {
        double o[2], dCf[WMAX_], dCr[WMAX_];
        orc_reaction_parameters(m, r, p, T, c, o, dCf, dCr, ddc);
        add_dNdtByV(m, r, o[0], dfdx);
        const int nl = m->n_lhs[r], nr = m->n_rhs[r];
        const long b = (long)r * WMAX_;
        /* jac_dCfdcj_contribution, gpuReaction.H:547-569 */
        for (int j = 0; j < nl; ++j) {
            const int sj = m->lhs_index[b + j];
            for (int i = 0; i < nl; ++i)
                J_(m->lhs_index[b + i], sj) -= m->lhs_stoich[b + i] * dCf[j];
            for (int i = 0; i < nr; ++i)
                J_(m->rhs_index[b + i], sj) += m->rhs_stoich[b + i] * dCf[j];
        }
        /* jac_dCrdcj_contribution, :571-593 */
        for (int j = 0; j < nr; ++j) {
            const int sj = m->rhs_index[b + j];
            for (int i = 0; i < nl; ++i)
                J_(m->lhs_index[b + i], sj) += m->lhs_stoich[b + i] * dCr[j];
            for (int i = 0; i < nr; ++i)
                J_(m->rhs_index[b + i], sj) -= m->rhs_stoich[b + i] * dCr[j];
        }
        /* jac_dCdT_contribution, :595-612 */
        for (int i = 0; i < nl; ++i)
            J_(m->lhs_index[b + i], nS) -= m->lhs_stoich[b + i] * o[1];
        for (int i = 0; i < nr; ++i)
            J_(m->rhs_index[b + i], nS) += m->rhs_stoich[b + i] * o[1];
        /* jac_dCdC_contribution, :614-648 */
        if (uses_third_body(m->type[r])) {
            for (int i = 0; i < nl; ++i) {
                double* row = &J_(m->lhs_index[b + i], 0);
                const double sl = m->lhs_stoich[b + i];
                for (int j = 0; j < nS; ++j) row[j] -= sl * ddc[j];
            }
            for (int i = 0; i < nr; ++i) {
                double* row = &J_(m->rhs_index[b + i], 0);
                const double sr = m->rhs_stoich[b + i];
                for (int j = 0; j < nS; ++j) row[j] += sr * ddc[j];
            }
        }
    }
    for (int i = 0; i < nS; ++i) dfdx[i] *= m->W[i] / rhoM;
    {
        double ret = 0.0;
        for (int i = 0; i < nS; ++i) ret -= dfdx[i] * Ha[i];
        dfdx[nS] = ret / CpM;
    }
    dfdx[nS + 1] = 0.0;

    /* ddY_dcdY */
    double* temp = ddc;
    for (int i = 0; i < nS; ++i) {
        const double WiByrhoM = m->W[i] / rhoM;
        for (int j = 0; j < nS + 2; ++j) temp[j] = J_(i, j);
        for (int j = 0; j < nS; ++j) {
            const double rhoMByWi = rhoM / m->W[j];
            const double vj = 1.0 / (p / ((RR_ / m->W[j]) * T));
            J_(i, j) = WiByrhoM * temp[j] * rhoMByWi + rhoM * vj * dfdx[i];
        }
        double ddNidtByVdT = temp[nS];
        for (int j = 0; j < nS; ++j) ddNidtByVdT -= temp[j] * c[j] * alphavM;
        J_(i, nS) = WiByrhoM * ddNidtByVdT + alphavM * dfdx[i];
        J_(i, nS + 1) = 0.0;
    }
    /* ddT_dtdY */
    for (int i = 0; i < nS; ++i) {
        J_(nS, i) = 0.0;
        for (int j = 0; j < nS; ++j) J_(nS, i) -= J_(j, i) * Ha[j];
        J_(nS, i) -= Cp[i] * dfdx[nS];
        J_(nS, i) /= CpM;
    }
    /* ddT_dtdT */
    J_(nS, nS) = 0.0;
    for (int i = 0; i < nS; ++i)
        J_(nS, nS) -= dfdx[i] * Cp[i] + J_(i, nS) * Ha[i];
    J_(nS, nS) -= dfdx[nS] * dCpMdT;
    J_(nS, nS) /= CpM;
#undef J_
}

/* --------------------------------------------------------------- LU --- */

/* Crout LU with implicit row scaling, ludecompose.H:9-92 */
void orc_lu_decompose(int m, double* a, int* piv, double* vv) {
#define A_(i, j) a[(long)(i) * m + (j)]
    for (int i = 0; i < m; i++) {
        double largest = 0.0, t;
        for (int j = 0; j < m; j++)
            if ((t = fabs(A_(i, j))) > largest) largest = t;
        vv[i] = 1.0 / largest;
    }
    for (int j = 0; j < m; j++) {
        for (int i = 0; i < j; i++) {
            double sum = A_(i, j);
            for (int k = 0; k < i; k++) sum -= A_(i, k) * A_(k, j);
            A_(i, j) = sum;
        }
        int iMax = 0;
        double largest = 0.0;
        for (int i = j; i < m; i++) {
            double sum = A_(i, j);
            for (int k = 0; k < j; k++) sum -= A_(i, k) * A_(k, j);
            A_(i, j) = sum;
            double t;
            if ((t = vv[i] * fabs(sum)) >= largest) {
                largest = t;
                iMax = i;
            }
        }
        piv[j] = iMax;
        if (j != iMax) {
            for (int k = 0; k < m; k++) {
                const double s = A_(j, k);
                A_(j, k) = A_(iMax, k);
                A_(iMax, k) = s;
            }
            vv[iMax] = vv[j];
        }
        if (A_(j, j) == 0.0) A_(j, j) = SMALL_;
        if (j != m - 1) {
            const double rDiag = 1.0 / A_(j, j);
            for (int i = j + 1; i < m; i++) A_(i, j) *= rDiag;
        }
    }
#undef A_
}

/* ludecompose.H:94-130 */
void orc_lu_backsubstitute(int m, const double* a, const int* piv, double* b) {
    int ii = 0;
    for (int i = 0; i < m; i++) {
        const int ip = piv[i];
        double sum = b[ip];
        b[ip] = b[i];
        const double* ai = a + (long)i * m;
        if (ii != 0) {
            for (int j = ii - 1; j < i; j++) sum -= ai[j] * b[j];
        } else if (sum != 0.0) {
            ii = i + 1;
        }
        b[i] = sum;
    }
    for (int i = m - 1; i >= 0; i--) {
        double sum = b[i];
        const double* ai = a + (long)i * m;
        for (int j = i + 1; j < m; j++) sum -= ai[j] * b[j];
        b[i] = sum / ai[i];
    }
}

/* ---------------------------------------------------------- solvers --- */

/* normaliseError, normaliseError.H:6-29 */
static double normalise_error(int n, const double* y0, const double* y,
                              const double* err, double absTol,
                              double relTol) {
    double maxErr = 0.0;
    for (int i = 0; i < n; ++i) {
        const double tol = absTol + relTol * fmax(fabs(y0[i]), fabs(y[i]));
        maxErr = fmax(maxErr, fabs(err[i]) / tol);
    }
    return maxErr;
}

/* Work layout: [0,4nE) system scratch | vectors | J | vv | pivots */
typedef struct {
    double *sys, *dydx0, *yTemp, *dydx, *dfdx, *k1, *k2, *k3, *k4, *err,
        *J, *vv;
    int* piv;
} orc_buf;

static orc_buf split_work(int nS, double* w) {
    const long nE = nS + 2;
    orc_buf b;
    b.sys = w;
    w += 4 * nE;
    b.dydx0 = w; w += nE;
    b.yTemp = w; w += nE;
    b.dydx = w; w += nE;
    b.dfdx = w; w += nE;
    b.k1 = w; w += nE;
    b.k2 = w; w += nE;
    b.k3 = w; w += nE;
    b.k4 = w; w += nE;
    b.err = w; w += nE;
    b.J = w; w += nE * nE;
    b.vv = w; w += nE;
    b.piv = (int*)w;
    return b;
}

/* gpuAdaptiveSolver::computeDecomposedJacobian, gpuAdaptiveSolver.H:51-73 */
static void decomposed_jacobian(const gfc_mech_view* m, double dx,
                                double gamma, const double* y0, orc_buf* b) {
    const int nE = m->n_species + 2;
    orc_jacobian(m, y0, b->dfdx, b->J, b->sys);
    const double temp = 1.0 / (gamma * dx);
    for (int i = 0; i < nE; i++) {
        for (int j = 0; j < nE; j++) b->J[(long)i * nE + j] = -b->J[(long)i * nE + j];
        b->J[(long)i * nE + i] += temp;
    }
    orc_lu_decompose(nE, b->J, b->piv, b->vv);
}

/* gpuRosenbrock34.H:13-46 (Shampine) */
static const double R34_a21 = 2.0, R34_a31 = 48.0 / 25.0, R34_a32 = 6.0 / 25.0;
static const double R34_c21 = -8.0, R34_c31 = 372.0 / 25.0, R34_c32 = 12.0 / 5.0;
static const double R34_c41 = -112.0 / 125.0, R34_c42 = -54.0 / 125.0,
                    R34_c43 = -2.0 / 5.0;
static const double R34_b1 = 19.0 / 9.0, R34_b2 = 1.0 / 2.0,
                    R34_b3 = 25.0 / 108.0, R34_b4 = 125.0 / 108.0;
static const double R34_e1 = 34.0 / 108.0, R34_e2 = 7.0 / 36.0,
                    R34_e4 = 125.0 / 108.0;
static const double R34_gamma = 1.0 / 2.0;
static const double R34_d1 = 1.0 / 2.0, R34_d2 = -3.0 / 2.0,
                    R34_d3 = 605.0 / 250.0, R34_d4 = 29.0 / 250.0;

/* gpuRosenbrock23.H:13-37 */
static const double R23_a21 = 1, R23_a31 = 1, R23_a32 = 0;
static const double R23_c21 = -1.0156171083877702091975600115545;
static const double R23_c31 = 4.0759956452537699824805835358067;
static const double R23_c32 = 9.2076794298330791242156818474003;
static const double R23_b1 = 1, R23_b2 = 6.1697947043828245592553615689730,
                    R23_b3 = -0.4277225654321857332623837380651;
static const double R23_e1 = 0.5, R23_e2 = -2.9079558716805469821718236208017,
                    R23_e3 = 0.2235406989781156962736090927619;
static const double R23_gamma = 0.43586652150845899941601945119356;
static const double R23_d1 = 0.43586652150845899941601945119356,
                    R23_d2 = 0.24291996454816804366592249683314,
                    R23_d3 = 2.1851380027664058511513169485832;

/* gpuRosenbrock12.H:14-31 */
#define R12_gamma (1.0 + 1.0 / 1.414213562373095)
#define R12_a21 (1.0 / R12_gamma)
#define R12_c21 (-2.0 / R12_gamma)
#define R12_b1 ((3.0 / 2.0) / R12_gamma)
#define R12_b2 ((1.0 / 2.0) / R12_gamma)
#define R12_e1 (R12_b1 - 1.0 / R12_gamma)
#define R12_e2 R12_b2
#define R12_d1 R12_gamma
#define R12_d2 (-R12_gamma)

static double step_impl(const gfc_mech_view* m, const gfc_ode_inputs* ode,
                        const double* y0, const double* dydx0, double dx,
                        double* y, orc_buf* b) {
    const int nE = m->n_species + 2;
    double *k1 = b->k1, *k2 = b->k2, *k3 = b->k3, *k4 = b->k4,
           *dydx = b->dydx, *dfdx = b->dfdx, *err = b->err;
    if (ode->solver == GFC_ROSENBROCK34) {
        /* gpuRosenbrock34.H:67-139 */
        decomposed_jacobian(m, dx, R34_gamma, y0, b);
        for (int i = 0; i < nE; ++i) k1[i] = dydx0[i] + dx * R34_d1 * dfdx[i];
        orc_lu_backsubstitute(nE, b->J, b->piv, k1);
        for (int i = 0; i < nE; ++i) y[i] = y0[i] + R34_a21 * k1[i];
        orc_derivatives(m, y, dydx, b->sys);
        for (int i = 0; i < nE; ++i)
            k2[i] = dydx[i] + dx * R34_d2 * dfdx[i] + R34_c21 * k1[i] / dx;
        orc_lu_backsubstitute(nE, b->J, b->piv, k2);
        for (int i = 0; i < nE; ++i)
            y[i] = y0[i] + R34_a31 * k1[i] + R34_a32 * k2[i];
        orc_derivatives(m, y, dydx, b->sys);
        for (int i = 0; i < nE; ++i)
            k3[i] = dydx[i] + dx * R34_d3 * dfdx[i] +
                    (R34_c31 * k1[i] + R34_c32 * k2[i]) / dx;
        orc_lu_backsubstitute(nE, b->J, b->piv, k3);
        for (int i = 0; i < nE; ++i)
            k4[i] = dydx[i] + dx * R34_d4 * dfdx[i] +
                    (R34_c41 * k1[i] + R34_c42 * k2[i] + R34_c43 * k3[i]) / dx;
        orc_lu_backsubstitute(nE, b->J, b->piv, k4);
        for (int i = 0; i < nE; ++i) {
            y[i] = y0[i] + R34_b1 * k1[i] + R34_b2 * k2[i] + R34_b3 * k3[i] +
                   R34_b4 * k4[i];
            err[i] = R34_e1 * k1[i] + R34_e2 * k2[i] + R34_e4 * k4[i];
        }
    } else if (ode->solver == GFC_ROSENBROCK23) {
        /* gpuRosenbrock23.H:58-116 */
        decomposed_jacobian(m, dx, R23_gamma, y0, b);
        for (int i = 0; i < nE; ++i) k1[i] = dydx0[i] + dx * R23_d1 * dfdx[i];
        orc_lu_backsubstitute(nE, b->J, b->piv, k1);
        for (int i = 0; i < nE; ++i) y[i] = y0[i] + R23_a21 * k1[i];
        orc_derivatives(m, y, dydx, b->sys);
        for (int i = 0; i < nE; ++i)
            k2[i] = dydx[i] + dx * R23_d2 * dfdx[i] + R23_c21 * k1[i] / dx;
        orc_lu_backsubstitute(nE, b->J, b->piv, k2);
        for (int i = 0; i < nE; ++i)
            k3[i] = dydx[i] + dx * R23_d3 * dfdx[i] +
                    (R23_c31 * k1[i] + R23_c32 * k2[i]) / dx;
        orc_lu_backsubstitute(nE, b->J, b->piv, k3);
        for (int i = 0; i < nE; ++i) {
            y[i] = y0[i] + R23_b1 * k1[i] + R23_b2 * k2[i] + R23_b3 * k3[i];
            err[i] = R23_e1 * k1[i] + R23_e2 * k2[i] + R23_e3 * k3[i];
        }
        (void)R23_a31; (void)R23_a32;
    } else {
        /* gpuRosenbrock12.H:52-103 */
        decomposed_jacobian(m, dx, R12_gamma, y0, b);
        for (int i = 0; i < nE; ++i) k1[i] = dydx0[i] + dx * R12_d1 * dfdx[i];
        orc_lu_backsubstitute(nE, b->J, b->piv, k1);
        for (int i = 0; i < nE; ++i) y[i] = y0[i] + R12_a21 * k1[i];
        orc_derivatives(m, y, dydx, b->sys);
        for (int i = 0; i < nE; ++i)
            k2[i] = dydx[i] + dx * R12_d2 * dfdx[i] + R12_c21 * k1[i] / dx;
        orc_lu_backsubstitute(nE, b->J, b->piv, k2);
        for (int i = 0; i < nE; ++i) {
            y[i] = y0[i] + R12_b1 * k1[i] + R12_b2 * k2[i];
            err[i] = R12_e1 * k1[i] + R12_e2 * k2[i];
        }
    }
    return normalise_error(nE, y0, y, err, ode->absTol, ode->relTol);
}

double orc_rosenbrock_step(const gfc_mech_view* m, const gfc_ode_inputs* ode,
                           const double* y0, const double* dydx0, double dx,
                           double* y, double* work) {
    orc_buf b = split_work(m->n_species, work);
    return step_impl(m, ode, y0, dydx0, dx, y, &b);
}

/* adaptiveSolve, gpuAdaptiveSolver.H:83-134; returns status bits */
static int adaptive_solve(const gfc_mech_view* m, const gfc_ode_inputs* ode,
                          double* x, double* y, double* dxTry, orc_buf* b,
                          int* stats) {
    const int nE = m->n_species + 2;
    double dx = *dxTry, err = 0.0;
    int status = 0;
    orc_derivatives(m, y, b->dydx0, b->sys);
    do {
        err = step_impl(m, ode, y, b->dydx0, dx, b->yTemp, b);
        if (stats) stats[0]++;
        if (err > 1) {
            const double scale =
                fmax(ode->safeScale * pow(err, -ode->alphaDec), ode->minScale);
            dx *= scale;
            if (dx < VSMALL_) {
                /* the reference prints and asserts; report and stop */
                status |= GFC_CELL_STEP_UNDERFLOW;
                return status;
            }
        }
    } while (err > 1.0);
    if (stats) stats[1]++;
    *x += dx;
    for (int i = 0; i < nE; ++i) y[i] = b->yTemp[i];
    if (err > pow(ode->maxScale / ode->safeScale, -1.0 / ode->alphaInc)) {
        *dxTry = fmin(fmax(ode->safeScale * pow(err, -ode->alphaInc),
                           ode->minScale),
                      ode->maxScale) * dx;
    } else {
        *dxTry = ode->safeScale * ode->maxScale * dx;
    }
    return status;
}

/* odeSolve, gpuODESolverBase.H:67-111 with stepState (stepState.H:8-26) */
int orc_ode_solve(const gfc_mech_view* m, const gfc_ode_inputs* ode,
                  double xStart, double xEnd, double* y, double* dxTry,
                  double* work, int* stats) {
    orc_buf b = split_work(m->n_species, work);
    double stepDxTry = *dxTry;
    int last = 0;
    double x = xStart;
    for (int nStep = 0; nStep < ode->maxSteps; nStep++) {
        const double dxTry0 = stepDxTry;
        if ((x + stepDxTry - xEnd) * (x + stepDxTry - xStart) > 0) {
            last = 1;
            stepDxTry = xEnd - x;
        }
        const int st = adaptive_solve(m, ode, &x, y, &stepDxTry, &b, stats);
        if (st) return st;
        if ((x - xEnd) * (xEnd - xStart) >= 0) {
            if (nStep > 0 && last) stepDxTry = dxTry0;
            *dxTry = stepDxTry;
            return 0;
        }
    }
    return GFC_CELL_MAX_STEPS;
}

/* singleCellSolver (singleCellSolver.H:27-51) and the computeRR post
 * processing (gpuKernelEvaluator.cu:65-97) */
int orc_solve_cells(const gfc_mech_view* m, const gfc_ode_inputs* ode,
                    double Treact, double deltaT, double deltaTChemMax,
                    int nCells, int ldY, const double* rho, const double* Y0,
                    double* Ynew, double* RR, double* deltaTChem,
                    double* deltaTMin, int* cellStats, int nThreads) {
    const int nS = m->n_species, nE = nS + 2;
    const long wsz = orc_workspace_doubles(nS);
    double minDt = DBL_MAX;
#ifdef _OPENMP
    if (nThreads <= 0) nThreads = omp_get_max_threads();
#pragma omp parallel num_threads(nThreads) reduction(min : minDt)
#endif
    {
        double* work = (double*)malloc(sizeof(double) * (size_t)(wsz + 2 * nE));
        double* y = work + wsz;
        double* y0 = y + nE;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int cell = 0; cell < nCells; ++cell) {
            for (int k = 0; k < nE; ++k) {
                y0[k] = Y0[(long)k * ldY + cell];
                y[k] = y0[k];
            }
            double dtChem = deltaTChem[cell];
            int stats[2] = {0, 0};
            int status = 0;
            if (y[nS] > Treact) {
                double timeLeft = deltaT;
                while (timeLeft > SMALL_) {
                    const double dt = timeLeft;
                    status |= orc_ode_solve(m, ode, 0, dt, y, &dtChem, work,
                                            stats);
                    for (int i = 0; i < nS; i++) y[i] = fmax(0.0, y[i]);
                    timeLeft -= dt;
                    if (status & GFC_CELL_STEP_UNDERFLOW) break;
                }
            } else {
                dtChem = deltaT;
                status |= GFC_CELL_SKIPPED;
            }
            if (Ynew)
                for (int k = 0; k < nE; ++k) Ynew[(long)k * ldY + cell] = y[k];
            if (RR)
                for (int i = 0; i < nS; ++i)
                    RR[(long)i * ldY + cell] =
                        rho[cell] * (y[i] - y0[i]) / deltaT;
            if (dtChem < minDt) minDt = dtChem;
            deltaTChem[cell] = fmin(dtChem, deltaTChemMax);
            if (cellStats) {
                cellStats[cell] = status;
                cellStats[(long)ldY + cell] = stats[0];
                cellStats[2L * ldY + cell] = stats[1];
            }
        }
        free(work);
    }
    if (deltaTMin) *deltaTMin = minDt;
    return 0;
}

/* ---------------------------------------------------------- queries --- */

/* OpenFOAM's Reaction::omega(p, T, c, li, omegaf, omegar), the function
 * gpuChemistryModel's CPU queries call; its net value is gpuReaction::omega
 * (gpuReaction.H:472-488), which the reference's tests hold equal to
 * OpenFOAM's (Test-gpuChemistry.C:384-388). */
void orc_reaction_omega_fr(const gfc_mech_view* m, int r, double p, double T,
                           const double* c, double out[2]) {
    (void)p;
    const double Tc = fmin(fmax(T, m->Tlow), m->Thigh);
    const double kf = rate_k(m, r, Tc, c);
    const double kr = m->irreversible[r] ? 0.0 : kf / reaction_Kc_only(m, r, Tc);
    double pw[WMAX_];
    const double Cf = side_powers(m, r, 0, c, pw);
    const double Cr = side_powers(m, r, 1, c, pw);
    out[0] = kf * Cf;
    out[1] = kr * Cr;
}

/* the concentrations the queries use: c_i = rho Y_i / W_i
 * (gpuChemistryModel.C:168-171) */
static void query_c(const gfc_mech_view* m, int cell, int ldY, const double* rho,
                    const double* Y, double* c) {
    for (int i = 0; i < m->n_species; ++i)
        c[i] = rho[cell] * Y[(long)i * ldY + cell] / m->W[i];
}

/* gpuChemistryModel::calculate, gpuChemistryModel.C:146-183 */
void orc_calculate(const gfc_mech_view* m, int nCells, int ldY,
                   const double* rho, const double* Y, double* RR) {
    const int nS = m->n_species;
    double* c = (double*)malloc(sizeof(double) * 2 * (size_t)(nS + 2));
    double* dN = c + nS + 2;
    for (int cell = 0; cell < nCells; ++cell) {
        const double T = Y[(long)nS * ldY + cell];
        const double p = Y[(long)(nS + 1) * ldY + cell];
        query_c(m, cell, ldY, rho, Y, c);
        for (int i = 0; i < nS + 2; ++i) dN[i] = 0.0;
        for (int r = 0; r < m->n_reactions; ++r) {
            double out[3];
            orc_reaction_rates(m, r, p, T, c, out);
            add_dNdtByV(m, r, out[2], dN);
        }
        for (int i = 0; i < nS; ++i) RR[(long)i * ldY + cell] = dN[i] * m->W[i];
    }
    free(c);
}

/* gpuChemistryModel::reactionRR, gpuChemistryModel.C:186-240 */
void orc_reaction_rr(const gfc_mech_view* m, int r, int nCells, int ldY,
                     const double* rho, const double* Y, double* out) {
    const int nS = m->n_species;
    double* c = (double*)malloc(sizeof(double) * (size_t)(nS + 2));
    for (int cell = 0; cell < nCells; ++cell) {
        query_c(m, cell, ldY, rho, Y, c);
        double o[3];
        orc_reaction_rates(m, r, Y[(long)(nS + 1) * ldY + cell],
                           Y[(long)nS * ldY + cell], c, o);
        out[cell] = o[2];
    }
    free(c);
}

/* gpuChemistryModel::specieReactionRR, gpuChemistryModel.C:242-317 */
void orc_specie_reaction_rr(const gfc_mech_view* m, int r, int nCells,
                            int ldY, const double* rho, const double* Y,
                            double* RR) {
    const int nS = m->n_species;
    double* c = (double*)malloc(sizeof(double) * 2 * (size_t)(nS + 2));
    double* dN = c + nS + 2;
    for (int cell = 0; cell < nCells; ++cell) {
        query_c(m, cell, ldY, rho, Y, c);
        for (int i = 0; i < nS + 2; ++i) dN[i] = 0.0;
        double o[3];
        orc_reaction_rates(m, r, Y[(long)(nS + 1) * ldY + cell],
                           Y[(long)nS * ldY + cell], c, o);
        add_dNdtByV(m, r, o[2], dN);
        for (int i = 0; i < nS; ++i) RR[(long)i * ldY + cell] = dN[i] * m->W[i];
    }
    free(c);
}

/* gpuChemistryModel::tc, gpuChemistryModel.C:366-443: the reciprocal of the
 * reactions' molar production rates per total mole, weighted by those
 * rates, forward and reverse parts taken separately */
void orc_tc(const gfc_mech_view* m, int nCells, int ldY, const double* rho,
            const double* Y, double* tc) {
    const int nS = m->n_species;
    double* c = (double*)malloc(sizeof(double) * (size_t)(nS + 2));
    for (int cell = 0; cell < nCells; ++cell) {
        const double T = Y[(long)nS * ldY + cell];
        const double p = Y[(long)(nS + 1) * ldY + cell];
        query_c(m, cell, ldY, rho, Y, c);
        double sumW = 0, sumWRateByCTot = 0;
        for (int r = 0; r < m->n_reactions; ++r) {
            double o[2];
            orc_reaction_omega_fr(m, r, p, T, c, o);
            double wf = 0;
            for (int s = 0; s < m->n_rhs[r]; ++s)
                wf += m->rhs_stoich[(long)r * WMAX_ + s] * o[0];
            sumW += wf;
            sumWRateByCTot += wf * wf;
            double wr = 0;
            for (int s = 0; s < m->n_lhs[r]; ++s)
                wr += m->lhs_stoich[(long)r * WMAX_ + s] * o[1];
            sumW += wr;
            sumWRateByCTot += wr * wr;
        }
        double ctot = 0;
        for (int i = 0; i < nS; ++i) ctot += c[i];
        tc[cell] = sumWRateByCTot == 0 ? VGREAT_ : sumW / sumWRateByCTot * ctot;
    }
    free(c);
}

/* gpuChemistryModel::Qdot, gpuChemistryModel.C:445-466, with
 * hf = gpuThermo::Hf (the reference's tests hold it equal to OpenFOAM's
 * hf(), Test-gpuChemistry.C:173-180) */
void orc_qdot(const gfc_mech_view* m, int nCells, int ldY, const double* RR,
              double* Qdot) {
    for (int cell = 0; cell < nCells; ++cell) Qdot[cell] = 0.0;
    for (int i = 0; i < m->n_species; ++i) {
        const double hi = poly_Hf(m->low + 7 * i);
        for (int cell = 0; cell < nCells; ++cell)
            Qdot[cell] -= hi * RR[(long)i * ldY + cell];
    }
}
