/* C++ face of libgpufoam_amd with the interface of gpuFoam's host evaluator
 * (gpu_chemistry/src/gpuKernelEvaluator/gpuKernelEvaluator/gpuKernelEvaluator.H:18-85),
 * so that an OpenFOAM chemistry model written against FoamGpu::GpuKernelEvaluator
 * can switch by changing the type it holds.  Header-only, over the C ABI in
 * gpufoam_amd.h.
 *
 * Differences from the reference, all deliberate:
 *   - the mechanism is compiled from the OpenFOAM thermo and reactions
 *     dictionaries (Mechanism), in place of makeGpuThermos/makeGpuReactions
 *     over OpenFOAM's in-memory objects (makeGpuThermo.H:16-66,
 *     makeGpuReactions.H:90-142), which need OpenFOAM to compile;
 *   - errors throw gpufoam_amd::Error (status code + message) instead of
 *     exit(1) / device assert (error_handling.H), and a cell that runs out of
 *     maxSteps is reported (cellStats) instead of silently accepted.
 */
#ifndef GPUFOAM_AMD_HPP
#define GPUFOAM_AMD_HPP

#include <algorithm>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "gpufoam_amd.h"

namespace gpufoam_amd {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(int status) {
    if (status != GFC_OK) {
        const char* msg = gfc_last_error();
        throw Error(status, msg ? msg : "gpufoam_amd error");
    }
}

// A compiled mechanism (makeGpuThermos + makeGpuReactions).
class Mechanism {
public:
    Mechanism(const std::string& thermoDict, const std::string& reactionsDict) {
        check(gfc_mechanism_from_foam(thermoDict.c_str(), reactionsDict.c_str(), &m_));
    }
    Mechanism(const Mechanism&) = delete;
    Mechanism& operator=(const Mechanism&) = delete;
    ~Mechanism() { gfc_mechanism_destroy(m_); }
    const gfc_mechanism* get() const { return m_; }
    gfc_mech_view view() const {
        gfc_mech_view v;
        check(gfc_mechanism_view(m_, &v));
        return v;
    }
    int nSpecie() const { return view().n_species; }
    int nReaction() const { return view().n_reactions; }

private:
    gfc_mechanism* m_ = nullptr;
};

// read_gpuODESolverInputs (readGpuOdeInputs.H:8-26) over a
// chemistryProperties dictionary text; Treact as gpuChemistryModel reads it.
inline std::pair<gfc_ode_inputs, double> readOdeInputs(const std::string& chemistryProperties) {
    gfc_ode_inputs in;
    double Treact = 0.0;
    check(gfc_ode_inputs_from_foam(chemistryProperties.c_str(), &in, &Treact));
    return {in, Treact};
}

// FoamGpu::GpuKernelEvaluator: Y is cell-major [nCells][nEqns] (Y..., T, p),
// as the reference's computeYNew/computeRR take it.
class GpuKernelEvaluator {
public:
    GpuKernelEvaluator(int nCells, int nEqns, int nSpecie, const Mechanism& mech,
                       gfc_ode_inputs odeInputs, double Treact = 0.0)
        : nEqns_(nEqns), nSpecie_(nSpecie), mech_(&mech), inputs_(odeInputs), Treact_(Treact) {
        if (nEqns != nSpecie + 2 || nSpecie != mech.nSpecie())
            throw Error(GFC_ERR_INVALID_ARGUMENT, "GpuKernelEvaluator: nEqns must be nSpecie+2 "
                                                  "and nSpecie the mechanism's");
        reserve(std::max(nCells, 1));
    }
    GpuKernelEvaluator(const GpuKernelEvaluator&) = delete;
    GpuKernelEvaluator& operator=(const GpuKernelEvaluator&) = delete;
    ~GpuKernelEvaluator() { gfc_evaluator_destroy(ev_); }

    // gpuKernelEvaluator.cu:35-63: new solution vectors and (unclamped)
    // chemistry time steps.
    std::pair<std::vector<double>, std::vector<double>>
    computeYNew(double deltaT, double deltaTChemMax, const std::vector<double>& deltaTChem,
                const std::vector<double>& Y) {
        (void)deltaTChemMax;
        const int nCells = static_cast<int>(deltaTChem.size());
        check_size(Y, nCells);
        std::vector<double> soa = to_soa(Y, nCells), out(soa.size()), dt = deltaTChem;
        reserve(nCells);
        check(gfc_solve(ev_, deltaT, 1e300, nCells, nCells, nullptr, soa.data(), out.data(),
                        nullptr, dt.data(), nullptr, nullptr, nullptr, GFC_HOST_POINTERS));
        return {to_aos(out, nCells), dt};
    }

    // gpuKernelEvaluator.cu:65-97: RR [nCells][nSpecie] = rho (Ynew - Y)/deltaT,
    // chemistry time steps clamped to deltaTChemMax, and their minimum
    // before clamping.
    std::tuple<std::vector<double>, std::vector<double>, double>
    computeRR(double deltaT, double deltaTChemMax, const std::vector<double>& rho,
              const std::vector<double>& deltaTChem, const std::vector<double>& Y) {
        const int nCells = static_cast<int>(rho.size());
        check_size(Y, nCells);
        if (deltaTChem.size() != rho.size())
            throw Error(GFC_ERR_INVALID_ARGUMENT, "computeRR: rho and deltaTChem sizes differ");
        std::vector<double> soa = to_soa(Y, nCells), rr((size_t)nSpecie_ * nCells),
                            dt = deltaTChem;
        double dtMin = 0.0;
        reserve(nCells);
        check(gfc_solve(ev_, deltaT, deltaTChemMax, nCells, nCells, rho.data(), soa.data(),
                        nullptr, rr.data(), dt.data(), &dtMin, nullptr, nullptr,
                        GFC_HOST_POINTERS));
        std::vector<double> RR((size_t)nCells * nSpecie_);
        for (int j = 0; j < nCells; ++j)
            for (int i = 0; i < nSpecie_; ++i)
                RR[(size_t)j * nSpecie_ + i] = rr[(size_t)i * nCells + j];
        return std::make_tuple(std::move(RR), std::move(dt), dtMin);
    }

    gfc_evaluator* handle() { return ev_; }

private:
    // the reference resizes its per-cell memory on every call
    // (gpuKernelEvaluator.cu:44); here the evaluator is re-created only when
    // the cell count grows
    void reserve(int nCells) {
        if (ev_ && nCells <= cap_) return;
        gfc_evaluator_destroy(ev_);
        ev_ = nullptr;
        check(gfc_evaluator_create(mech_->get(), &inputs_, Treact_, nCells, &ev_));
        cap_ = nCells;
    }
    void check_size(const std::vector<double>& Y, int nCells) const {
        if (Y.size() != (size_t)nCells * nEqns_)
            throw Error(GFC_ERR_INVALID_ARGUMENT, "Y must hold nCells*nEqns values");
    }
    std::vector<double> to_soa(const std::vector<double>& Y, int nCells) const {
        std::vector<double> s(Y.size());
        for (int j = 0; j < nCells; ++j)
            for (int k = 0; k < nEqns_; ++k) s[(size_t)k * nCells + j] = Y[(size_t)j * nEqns_ + k];
        return s;
    }
    std::vector<double> to_aos(const std::vector<double>& s, int nCells) const {
        std::vector<double> Y(s.size());
        for (int j = 0; j < nCells; ++j)
            for (int k = 0; k < nEqns_; ++k) Y[(size_t)j * nEqns_ + k] = s[(size_t)k * nCells + j];
        return Y;
    }

    int nEqns_, nSpecie_;
    const Mechanism* mech_;
    gfc_ode_inputs inputs_;
    double Treact_;
    gfc_evaluator* ev_ = nullptr;
    int cap_ = 0;
};

} // namespace gpufoam_amd

#endif /* GPUFOAM_AMD_HPP */
