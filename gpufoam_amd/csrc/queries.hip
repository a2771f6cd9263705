// Chemistry queries at the current state of every cell: the specie
// reaction rates, one reaction's rate and its specie rates, the chemical
// time scale and the heat release rate.  gpuFoam computes these on the CPU
// with OpenFOAM's Reaction objects, cell by cell
// (gpu_chemistry/src/gpuChemistryModels/gpuChemistryModel/gpuChemistryModel.C:146-466);
// here they run on the device over the same structure-of-arrays fields as
// gfc_solve.
//
// Concentrations are c_i = rho Y_i / W_i with the caller's density (the
// thermo's rho, :168-171), not clipped at zero.  Rates follow
// gpuReaction::omega (gpuReaction.H:472-488): the rate constants at T
// clipped to the reaction limits, kr = kf / Kc, and the concentration
// products with the reference's integer/real exponents.
//
// Mappings:
//   * calculate and tc need every reaction of a cell: one wavefront per
//     cell, concentrations in LDS, reactions spread over the lanes,
//     specie sums through LDS atomics (calculate) or wavefront sums (tc);
//   * reactionRR, specieReactionRR and Qdot touch one reaction or a
//     weighted sum over species: one lane per cell, every field access a
//     coalesced row of the SoA layout.
#include <hip/hip_runtime.h>

#include <string>

#include "evaluator.hpp"
#include "team_common.hpp"

namespace gfc {

namespace {

constexpr int kQMaxSpecies = 320; // >= GFC_MAX_SPECIES
constexpr int kQLanes = 64;

struct QueryArgs {
    DeviceMech m; // reference-order tables
    int nCells;
    long ld;
    int reaction;
    const double* rho;
    const double* Y;
    const double* RRin;
    double* out;
};

// Concentrations held in LDS (one cell per wavefront).
struct LdsConc {
    const double* c;
    __device__ __forceinline__ double operator()(int i) const { return c[i]; }
};

// Concentrations read from the SoA fields (one cell per lane).
struct FieldConc {
    const double* Y;
    const double* W;
    long ld;
    long cell;
    double rho;
    __device__ __forceinline__ double operator()(int i) const {
        return rho * Y[i * ld + cell] / W[i];
    }
};

// Forward and reverse parts of the rate of reaction r,
// wf = kf prod c^e (left side), wr = kr prod c^e (right side): the split of
// OpenFOAM's Reaction::omega(p, T, c, li, omegaf, omegar) whose difference
// is gpuReaction::omega.
template <class Conc>
__device__ __forceinline__ void omega_fr(const DeviceMech& m, int r, double T,
                                         const Conc& c, double& wf, double& wr) {
    const double Tc = fmin(fmax(T, m.Tlow), m.Thigh);
    const int type = m.type[r];
    double kf = arrhenius(m.k + 3 * r, Tc);
    if (type > GFC_RATE_SRI_CA) {
        // Landau-Teller, Janev, power series
        kf = tdep_rate(type, m.k + 3 * r, m.extra + (long)GFC_EXTRA_PARAMS * r, Tc, nullptr);
    } else if (type != GFC_RATE_ARRHENIUS) {
        // gpuThirdBodyEfficiencies.H:21-27, summed in specie order
        const double* eff = m.eff + (long)r * m.nS;
        double M = 0;
        for (int i = 0; i < m.nS; ++i) M += eff[i] * c(i);
        if (type == GFC_RATE_THIRD_BODY) {
            kf = M * kf; // gpuThirdBodyArrheniusReactionRate.H:39-44
        } else {
            // gpuFallOffReactionRate.H:216-227; chemically activated
            // rates blend from k0 instead: k0 / (1 + Pr) F
            const double kInf = arrhenius(m.kInf + 3 * r, Tc);
            const double Pr = kf / kInf * M;
            const double F = blend_F(m, r, Tc, Pr);
            kf = (type >= GFC_RATE_LINDEMANN_CA) ? kf * (1.0 / (1.0 + Pr)) * F
                                                 : kInf * (Pr / (1.0 + Pr)) * F;
        }
    }
    const double kr = m.irreversible[r] ? 0.0 : kf / reaction_Kc(m, r, Tc);
    const long o = (long)r * kWidth;
    double Cf = 1.0, Cr = 1.0;
    for (int q = 0; q < m.nl[r]; ++q) Cf *= specie_pow(c(m.li[o + q]), m.lei[o + q], m.le[o + q]);
    for (int q = 0; q < m.nr[r]; ++q) Cr *= specie_pow(c(m.ri[o + q]), m.rei[o + q], m.re[o + q]);
    wf = kf * Cf;
    wr = kr * Cr;
}

// calculate (MODE 0, gpuChemistryModel.C:146-183) and tc (MODE 1, :366-443).
template <int MODE>
__global__ void __launch_bounds__(kQLanes) mechanism_query_kernel(QueryArgs Q) {
    __shared__ double c[kQMaxSpecies];
    __shared__ double dN[kQMaxSpecies];
    const int lane = threadIdx.x;
    const long cell = blockIdx.x;
    const DeviceMech& m = Q.m;
    const int nS = m.nS;
    const long ld = Q.ld;
    const double rho = Q.rho[cell];
    const double T = Q.Y[nS * ld + cell];
    double csum = 0.0;
    for (int i = lane; i < nS; i += kQLanes) {
        const double ci = rho * Q.Y[i * ld + cell] / m.W[i];
        c[i] = ci;
        dN[i] = 0.0;
        csum += ci;
    }
    __syncthreads();
    double sumW = 0.0, sumW2 = 0.0;
    for (int r = lane; r < m.nR; r += kQLanes) {
        double wf, wr;
        omega_fr(m, r, T, LdsConc{c}, wf, wr);
        const long o = (long)r * kWidth;
        if (MODE == 0) {
            // gpuReaction::dNdtByV (gpuReaction.H:490-508)
            const double w = wf - wr;
            for (int q = 0; q < m.nl[r]; ++q) atomicAdd(&dN[m.li[o + q]], -(m.ls[o + q] * w));
            for (int q = 0; q < m.nr[r]; ++q) atomicAdd(&dN[m.ri[o + q]], m.rs[o + q] * w);
        } else {
            // forward molar production counted on the right side, reverse
            // on the left (gpuChemistryModel.C:426-434)
            double a = 0.0, b = 0.0;
            for (int q = 0; q < m.nr[r]; ++q) a += m.rs[o + q] * wf;
            for (int q = 0; q < m.nl[r]; ++q) b += m.ls[o + q] * wr;
            sumW += a;
            sumW += b;
            sumW2 += a * a;
            sumW2 += b * b;
        }
    }
    if (MODE == 0) {
        __syncthreads();
        for (int i = lane; i < nS; i += kQLanes) Q.out[i * ld + cell] = dN[i] * m.W[i];
    } else {
        const double sW = wave::wave_sum(sumW);
        const double sW2 = wave::wave_sum(sumW2);
        const double ctot = wave::wave_sum(csum);
        if (lane == 0) Q.out[cell] = (sW2 == 0.0) ? kVGreat : sW / sW2 * ctot;
    }
}

// reactionRR (MODE 0, gpuChemistryModel.C:186-240) and specieReactionRR
// (MODE 1, :242-317) of reaction Q.reaction.
template <int MODE>
__global__ void __launch_bounds__(256) reaction_query_kernel(QueryArgs Q) {
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= Q.nCells) return;
    const DeviceMech& m = Q.m;
    const int nS = m.nS, r = Q.reaction;
    const long ld = Q.ld;
    const FieldConc c{Q.Y, m.W, ld, cell, Q.rho[cell]};
    double wf, wr;
    omega_fr(m, r, Q.Y[nS * ld + cell], c, wf, wr);
    const double w = wf - wr;
    if (MODE == 0) {
        Q.out[cell] = w;
        return;
    }
    for (int i = 0; i < nS; ++i) Q.out[i * ld + cell] = 0.0;
    // dNdtByV of one reaction, left side then right side; a specie on both
    // sides accumulates in that order
    const long o = (long)r * kWidth;
    const int nl = m.nl[r], nr = m.nr[r];
    for (int s = 0; s < nl + nr; ++s) {
        const int sp = s < nl ? m.li[o + s] : m.ri[o + s - nl];
        bool first = true; // write each specie once, at its first position
        for (int q = 0; q < s; ++q)
            first = first && sp != (q < nl ? m.li[o + q] : m.ri[o + q - nl]);
        if (!first) continue;
        double v = 0.0;
        for (int q = 0; q < nl; ++q)
            if (m.li[o + q] == sp) v -= m.ls[o + q] * w;
        for (int q = 0; q < nr; ++q)
            if (m.ri[o + q] == sp) v += m.rs[o + q] * w;
        Q.out[sp * ld + cell] = v * m.W[sp];
    }
}

// Qdot (gpuChemistryModel.C:445-466): -sum_i hf_i RR_i in specie order,
// hf_i = gpuThermo::Hf (gpuThermo.H:243-254, low coefficients at Tstd).
__global__ void __launch_bounds__(256) qdot_kernel(QueryArgs Q) {
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= Q.nCells) return;
    const DeviceMech& m = Q.m;
    double q = 0.0;
    for (int i = 0; i < m.nS; ++i) {
        const double* a = m.low + 7 * i;
        const double T = kTstd;
        const double hf = (((((a[4] / 5.0 * T + a[3] / 4.0) * T + a[2] / 3.0) * T +
                             a[1] / 2.0) * T + a[0]) * T + a[5]);
        q -= hf * Q.RRin[i * Q.ld + cell];
    }
    Q.out[cell] = q;
}

bool hipFail(hipError_t e, const char* what) {
    if (e == hipSuccess) return false;
    setLastError(std::string(what) + ": " + hipGetErrorString(e));
    return true;
}

enum class Query { Calculate, Tc, ReactionRR, SpecieReactionRR, Qdot };

// Validates, stages host buffers (GFC_HOST_POINTERS) through the
// evaluator's device buffers, launches and copies results back.
int run_query(gfc_evaluator* ev, Query q, int reaction, int nCells, int ldY,
              const double* rho, const double* Y, const double* RRin, double* out,
              void* stream, unsigned flags, const char* name) {
    const bool needState = q != Query::Qdot;
    if (!ev || nCells < 0 || ldY < nCells ||
        (nCells > 0 && (!out || (needState && (!rho || !Y)) || (!needState && !RRin)))) {
        setLastError(std::string(name) + ": bad argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    if ((q == Query::ReactionRR || q == Query::SpecieReactionRR) &&
        (reaction < 0 || reaction >= ev->nR)) {
        setLastError(std::string(name) + ": reaction index out of range");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    if (ev->nS > kQMaxSpecies) {
        setLastError(std::string(name) + ": too many species");
        return GFC_ERR_TOO_MANY_SPECIES;
    }
    if (nCells > ev->maxCells) {
        setLastError(std::string(name) + ": nCells exceeds the evaluator's maxCells");
        return GFC_ERR_CAPACITY;
    }
    if (nCells == 0) return GFC_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool host = flags & GFC_HOST_POINTERS;
    const int nS = ev->nS, nE = nS + 2;
    const bool rowsOut = q == Query::Calculate || q == Query::SpecieReactionRR;
    QueryArgs Q{};
    Q.m = ev->dm;
    Q.nCells = nCells;
    Q.reaction = reaction;
    if (host) {
        const long ld = nCells;
        if (needState) {
            for (int k = 0; k < nE; ++k)
                if (hipFail(hipMemcpyAsync(ev->sY0 + k * ld, Y + (long)k * ldY,
                                           sizeof(double) * nCells, hipMemcpyHostToDevice, s),
                            "hipMemcpyAsync"))
                    return GFC_ERR_HIP;
            if (hipFail(hipMemcpyAsync(ev->sRho, rho, sizeof(double) * nCells,
                                       hipMemcpyHostToDevice, s), "hipMemcpyAsync"))
                return GFC_ERR_HIP;
        } else {
            for (int k = 0; k < nS; ++k)
                if (hipFail(hipMemcpyAsync(ev->sRR + k * ld, RRin + (long)k * ldY,
                                           sizeof(double) * nCells, hipMemcpyHostToDevice, s),
                            "hipMemcpyAsync"))
                    return GFC_ERR_HIP;
        }
        Q.ld = ld;
        Q.Y = ev->sY0;
        Q.rho = ev->sRho;
        Q.RRin = ev->sRR;
        Q.out = rowsOut ? ev->sYnew : ev->sDt;
    } else {
        Q.ld = ldY;
        Q.Y = Y;
        Q.rho = rho;
        Q.RRin = RRin;
        Q.out = out;
    }
    const int lanesGrid = (nCells + 255) / 256;
    switch (q) {
    case Query::Calculate:
        hipLaunchKernelGGL(mechanism_query_kernel<0>, dim3(nCells), dim3(kQLanes), 0, s, Q);
        break;
    case Query::Tc:
        hipLaunchKernelGGL(mechanism_query_kernel<1>, dim3(nCells), dim3(kQLanes), 0, s, Q);
        break;
    case Query::ReactionRR:
        hipLaunchKernelGGL(reaction_query_kernel<0>, dim3(lanesGrid), dim3(256), 0, s, Q);
        break;
    case Query::SpecieReactionRR:
        hipLaunchKernelGGL(reaction_query_kernel<1>, dim3(lanesGrid), dim3(256), 0, s, Q);
        break;
    case Query::Qdot:
        hipLaunchKernelGGL(qdot_kernel, dim3(lanesGrid), dim3(256), 0, s, Q);
        break;
    }
    if (hipFail(hipGetLastError(), name)) return GFC_ERR_HIP;
    if (host) {
        const long ld = nCells;
        if (rowsOut) {
            for (int k = 0; k < nS; ++k)
                if (hipFail(hipMemcpyAsync(out + (long)k * ldY, ev->sYnew + k * ld,
                                           sizeof(double) * nCells, hipMemcpyDeviceToHost, s),
                            "hipMemcpyAsync"))
                    return GFC_ERR_HIP;
        } else if (hipFail(hipMemcpyAsync(out, ev->sDt, sizeof(double) * nCells,
                                          hipMemcpyDeviceToHost, s), "hipMemcpyAsync")) {
            return GFC_ERR_HIP;
        }
        if (hipFail(hipStreamSynchronize(s), "hipStreamSynchronize")) return GFC_ERR_HIP;
    }
    return GFC_OK;
}

} // namespace
} // namespace gfc

using gfc::Query;

extern "C" {

int gfc_calculate(gfc_evaluator* ev, int nCells, int ldY, const double* rho,
                  const double* Y, double* RR, void* stream, unsigned flags) {
    return gfc::run_query(ev, Query::Calculate, 0, nCells, ldY, rho, Y, nullptr, RR, stream,
                          flags, "gfc_calculate");
}

int gfc_tc(gfc_evaluator* ev, int nCells, int ldY, const double* rho, const double* Y,
           double* tc, void* stream, unsigned flags) {
    return gfc::run_query(ev, Query::Tc, 0, nCells, ldY, rho, Y, nullptr, tc, stream, flags,
                          "gfc_tc");
}

int gfc_reaction_rr(gfc_evaluator* ev, int reaction, int nCells, int ldY, const double* rho,
                    const double* Y, double* rr, void* stream, unsigned flags) {
    return gfc::run_query(ev, Query::ReactionRR, reaction, nCells, ldY, rho, Y, nullptr, rr,
                          stream, flags, "gfc_reaction_rr");
}

int gfc_specie_reaction_rr(gfc_evaluator* ev, int reaction, int nCells, int ldY,
                           const double* rho, const double* Y, double* RR, void* stream,
                           unsigned flags) {
    return gfc::run_query(ev, Query::SpecieReactionRR, reaction, nCells, ldY, rho, Y, nullptr,
                          RR, stream, flags, "gfc_specie_reaction_rr");
}

int gfc_qdot(gfc_evaluator* ev, int nCells, int ldY, const double* RR, double* Qdot,
             void* stream, unsigned flags) {
    return gfc::run_query(ev, Query::Qdot, 0, nCells, ldY, nullptr, nullptr, RR, Qdot, stream,
                          flags, "gfc_qdot");
}

} // extern "C"
