// Host driver around gpuFoam's own header-only chemistry code.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file against the
// reference headers where they lie (gpu_chemistry/src/{gpuOde,gpuReaction,
// gpuThermo,gpuEquationOfState,memoryResource,gpuKernelEvaluator} and
// gpu_utils/common) into oracle/_ref/libgfref.so.  This is the
// configuration the reference's own [CPU] tests run (CpuTestKernels,
// gpu_chemistry/unittest/cpuTestKernels/cpu_test_kernels.C), so the library
// is the reference path itself, built for the host.
//
// The reference builds its gpuThermo/gpuReaction objects from OpenFOAM
// dictionaries (makeGpuThermo.H, makeGpuReactions.H), which need OpenFOAM.
// This driver builds them from the parsed raw mechanism data instead, with
// the reference's own constructors and its host thermo operators (+=, *, ==)
// for the composite reaction thermo, restating the few lines of converter
// arithmetic (makeGpuThermo.H:41-44, makeGpuReactions.H:18-59,61-88).
#include <cstring>
#include <stdexcept>
#include <vector>

#include "../include/gpufoam_amd.h"

#include "cpuMemoryResource.H"
#include "gpuODESystem.H"
#include "makeGpuOdeSolver.H"
#include "singleCellSolver.H"

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace FoamGpu;

struct gfr_mech {
    int nS = 0;
    int nR = 0;
    std::vector<gpuThermo> thermos;
    std::vector<gpuReaction> reactions;
};

namespace {

gpuSpecieExponent makeExponent(double e) {
    if (isInteger(e)) return gpuSpecieExponent(gLabel(e));
    return gpuSpecieExponent(gScalar(e));
}

gpuReaction::specieCoeffArray makeSide(const int* idx, const double* st,
                                       const double* ex, int n) {
    gpuReaction::specieCoeffArray ret{};
    for (int i = 0; i < n; ++i) {
        ret[i].index = idx[i];
        ret[i].stoichCoeff = st[i];
        ret[i].exponent = makeExponent(ex[i]);
    }
    return ret;
}

gpuThirdBodyEfficiencies makeEff(const double* eff, int nS) {
    gpuThirdBodyEfficiencies::effArray a{};
    for (int i = 0; i < nS; ++i) a[i] = eff[i];
    return gpuThirdBodyEfficiencies(nS, a);
}

gpuODESolverInputs makeInputs(const gfc_ode_inputs* o) {
    gpuODESolverInputs i;
    i.name = o->solver == GFC_ROSENBROCK12   ? "Rosenbrock12"
             : o->solver == GFC_ROSENBROCK23 ? "Rosenbrock23"
                                             : "Rosenbrock34";
    i.absTol = o->absTol;
    i.relTol = o->relTol;
    i.maxSteps = o->maxSteps;
    i.safeScale = o->safeScale;
    i.alphaInc = o->alphaInc;
    i.alphaDec = o->alphaDec;
    i.minScale = o->minScale;
    i.maxScale = o->maxScale;
    return i;
}

} // namespace

extern "C" {

gfr_mech* gfr_create(const gfc_mech_view* v, const double* rawHigh,
                     const double* rawLow) {
    auto* m = new gfr_mech;
    m->nS = v->n_species;
    m->nR = v->n_reactions;
    for (int i = 0; i < m->nS; ++i) {
        // makeGpuThermo.H:41-44: coefficients times the specie R = RR/W
        const double R = gpuPerfectGas(1.0, v->W[i]).R();
        gpuThermo::coeffArray hi{}, lo{};
        for (int c = 0; c < 7; ++c) {
            hi[c] = R * rawHigh[7 * i + c];
            lo[c] = R * rawLow[7 * i + c];
        }
        m->thermos.push_back(gpuThermo(1.0, v->W[i], v->sTlow[i],
                                       v->sThigh[i], v->sTcommon[i], hi, lo));
    }
    const int W = GFC_MAX_REACTION_WIDTH;
    for (int r = 0; r < m->nR; ++r) {
        const double* k = v->k + 3 * r;
        gpu::Arrhenius arr(k[0], k[1], k[2]);
        gpuReactionRate rate;
        const bool irr = v->irreversible[r] != 0;
        switch (v->type[r]) {
        case GFC_RATE_ARRHENIUS: rate = gpuReactionRate(arr, irr); break;
        case GFC_RATE_THIRD_BODY:
            rate = gpuReactionRate(
                gpu::ThirdBodyArrhenius(arr, makeEff(v->eff + (long)r * m->nS,
                                                     m->nS)),
                irr);
            break;
        case GFC_RATE_LINDEMANN: {
            const double* ki = v->kInf + 3 * r;
            rate = gpuReactionRate(
                gpu::ArrheniusLindemannFallOff(
                    arr, gpu::Arrhenius(ki[0], ki[1], ki[2]),
                    gpuLindemannFallOffFunction(),
                    makeEff(v->eff + (long)r * m->nS, m->nS)),
                irr);
            break;
        }
        default: {
            const double* ki = v->kInf + 3 * r;
            const double* t = v->troe + 4 * r;
            rate = gpuReactionRate(
                gpu::ArrheniusTroeFallOff(
                    arr, gpu::Arrhenius(ki[0], ki[1], ki[2]),
                    gpuTroeFallOffFunction(t[0], t[1], t[2], t[3]),
                    makeEff(v->eff + (long)r * m->nS, m->nS)),
                irr);
            break;
        }
        }
        const int nl = v->n_lhs[r], nr = v->n_rhs[r];
        auto lhs = makeSide(v->lhs_index + r * W, v->lhs_stoich + r * W,
                            v->lhs_exp + r * W, nl);
        auto rhs = makeSide(v->rhs_index + r * W, v->rhs_stoich + r * W,
                            v->rhs_exp + r * W, nr);
        gpuReaction reaction(m->thermos[0], lhs, rhs, nl, nr, rate, v->Tlow,
                             v->Thigh);
        // makeGpuReactions.H:18-59 with the reference's thermo operators
        const auto& th = m->thermos;
        gpuThermo rhsThermo(rhs[0].stoichCoeff * th[rhs[0].index].W() *
                            th[rhs[0].index]);
        for (int i = 1; i < nr; ++i)
            rhsThermo += rhs[i].stoichCoeff * th[rhs[i].index].W() *
                         th[rhs[i].index];
        gpuThermo lhsThermo(lhs[0].stoichCoeff * th[lhs[0].index].W() *
                            th[lhs[0].index]);
        for (int i = 1; i < nl; ++i)
            lhsThermo += lhs[i].stoichCoeff * th[lhs[i].index].W() *
                         th[lhs[i].index];
        reaction.get_thermo() = (lhsThermo == rhsThermo);
        m->reactions.push_back(reaction);
    }
    return m;
}

void gfr_destroy(gfr_mech* m) { delete m; }

// Y, W, Tcommon, Tlow, Thigh, high[7], low[7]
void gfr_composite_thermo(const gfr_mech* m, int r, double* out) {
    const gpuReaction& R = m->reactions[r];
    const gpuThermo& t = R;
    out[0] = t.Y();
    out[1] = t.W();
    out[2] = t.Tcommon_;
    out[3] = t.Tlow_;
    out[4] = t.Thigh_;
    for (int c = 0; c < 7; ++c) {
        out[5 + c] = t.highCpCoeffs()[c];
        out[12 + c] = t.lowCpCoeffs()[c];
    }
}

void gfr_species_thermo(const gfr_mech* m, int i, double p, double T,
                        double out[6]) {
    const gpuThermo& t = m->thermos[i];
    out[0] = t.Cp(p, T);
    out[1] = t.Ha(p, T);
    out[2] = t.Hf();
    out[3] = t.S(p, T);
    out[4] = t.Gstd(T);
    out[5] = t.dCpdT(p, T);
}

void gfr_reaction_Kc(const gfr_mech* m, int r, double p, double T,
                     double out[2]) {
    const auto a = m->reactions[r].KcAnddKcTbyKc(p, T);
    out[0] = a[0];
    out[1] = a[1];
}

void gfr_reaction_rates(const gfr_mech* m, int r, double p, double T,
                        const double* c, double out[3]) {
    const gpuReaction& R = m->reactions[r];
    mdspan<const gScalar, 1> cs(c, extents<1>{size_t(m->nS)});
    const double Tc = fmin(fmax(T, R.Tlow()), R.Thigh());
    const double kf = R.kf(p, Tc, cs);
    out[0] = kf;
    out[1] = R.kr(kf, p, Tc, R.Kc(p, Tc), cs);
    out[2] = R.omega(p, T, cs);
}

void gfr_reaction_parameters(const gfr_mech* m, int r, double p, double T,
                             const double* c, double out[2], double* dCf,
                             double* dCr, double* ddc) {
    mdspan<const gScalar, 1> cs(c, extents<1>{size_t(m->nS)});
    mdspan<gScalar, 1> dd(ddc, extents<1>{size_t(m->nS)});
    const auto rp = computeReactionParameters(m->reactions[r], cs, p, T, dd);
    out[0] = rp.omega;
    out[1] = rp.dwdT;
    for (int j = 0; j < GFC_MAX_REACTION_WIDTH; ++j) {
        dCf[j] = rp.dCfdjs[j];
        dCr[j] = rp.dCrdjs[j];
    }
}

void gfr_derivatives(const gfr_mech* m, const double* y, double* dydx) {
    const int nE = m->nS + 2;
    gpuODESystem sys(nE, m->nR, m->thermos.data(), m->reactions.data());
    cpuMemoryResource mr(1, m->nS);
    auto bufs = splitToBuffers(mr);
    sys.derivatives(mdspan<const gScalar, 1>(y, extents<1>{size_t(nE)}),
                    mdspan<gScalar, 1>(dydx, extents<1>{size_t(nE)}), bufs[0]);
}

void gfr_jacobian(const gfr_mech* m, const double* y, double* dfdx,
                  double* J) {
    const int nE = m->nS + 2;
    gpuODESystem sys(nE, m->nR, m->thermos.data(), m->reactions.data());
    cpuMemoryResource mr(1, m->nS);
    auto bufs = splitToBuffers(mr);
    sys.jacobian(mdspan<const gScalar, 1>(y, extents<1>{size_t(nE)}),
                 mdspan<gScalar, 1>(dfdx, extents<1>{size_t(nE)}),
                 mdspan<gScalar, 2>(J, extents<2>{size_t(nE), size_t(nE)}),
                 bufs[0]);
}

void gfr_lu_decompose(int n, double* a, int* piv) {
    std::vector<double> vv(n);
    LUDecompose(mdspan<gScalar, 2>(a, extents<2>{size_t(n), size_t(n)}),
                mdspan<gLabel, 1>(piv, extents<1>{size_t(n)}),
                mdspan<gScalar, 1>(vv.data(), extents<1>{size_t(n)}));
}

void gfr_lu_backsubstitute(int n, const double* a, const int* piv,
                           double* b) {
    LUBacksubstitute(
        mdspan<const gScalar, 2>(a, extents<2>{size_t(n), size_t(n)}),
        mdspan<const gLabel, 1>(piv, extents<1>{size_t(n)}),
        mdspan<gScalar, 1>(b, extents<1>{size_t(n)}));
}

int gfr_ode_solve(const gfr_mech* m, const gfc_ode_inputs* o, double xStart,
                  double xEnd, double* y, double* dxTry) {
    const int nE = m->nS + 2;
    gpuODESystem sys(nE, m->nR, m->thermos.data(), m->reactions.data());
    try {
        auto solver = make_gpuODESolver(sys, makeInputs(o));
        cpuMemoryResource mr(1, m->nS);
        auto bufs = splitToBuffers(mr);
        solver.solve(xStart, xEnd, mdspan<gScalar, 1>(y, extents<1>{size_t(nE)}),
                     *dxTry, bufs[0]);
    } catch (const std::exception&) {
        return 1;
    }
    return 0;
}

// singleCellSolver over SoA cells plus computeRR's post-processing
// (gpuKernelEvaluator.cu:65-97).  One singleCellSolver per thread, each on
// its own one-cell AoS row and scratch buffer.
int gfr_solve_cells(const gfr_mech* m, const gfc_ode_inputs* o, double Treact,
                    double deltaT, double deltaTChemMax, int nCells, int ldY,
                    const double* rho, const double* Y0, double* Ynew,
                    double* RR, double* deltaTChem, double* deltaTMin,
                    int nThreads) {
    const int nS = m->nS, nE = nS + 2;
    gpuODESystem sys(nE, m->nR, m->thermos.data(), m->reactions.data());
    gpuODESolver solver = make_gpuODESolver(sys, makeInputs(o));
    double minDt = 1e300;
#ifdef _OPENMP
    if (nThreads <= 0) nThreads = omp_get_max_threads();
#pragma omp parallel num_threads(nThreads) reduction(min : minDt)
#endif
    {
        cpuMemoryResource mr(1, nS);
        auto bufs = splitToBuffers(mr);
        std::vector<double> row(nE), y0(nE);
        double dtChem = 0;
        singleCellSolver op(
            deltaT, nS, mdspan<gScalar, 1>(&dtChem, extents<1>{1}),
            mdspan<gScalar, 2>(row.data(), extents<2>{1, size_t(nE)}),
            mdspan<gpuBuffer, 1>(bufs.data(), extents<1>{1}), solver, Treact);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int cell = 0; cell < nCells; ++cell) {
            for (int k = 0; k < nE; ++k) {
                y0[k] = Y0[(long)k * ldY + cell];
                row[k] = y0[k];
            }
            dtChem = deltaTChem[cell];
            op(0);
            if (Ynew)
                for (int k = 0; k < nE; ++k) Ynew[(long)k * ldY + cell] = row[k];
            if (RR)
                for (int i = 0; i < nS; ++i)
                    RR[(long)i * ldY + cell] =
                        rho[cell] * (row[i] - y0[i]) / deltaT;
            if (dtChem < minDt) minDt = dtChem;
            deltaTChem[cell] = std::min(dtChem, deltaTChemMax);
        }
    }
    if (deltaTMin) *deltaTMin = minDt;
    return 0;
}

} // extern "C"
