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

// Thrown where the reference throws std::logic_error: the LTS solve
// (gpuChemistryModel.C:328-333).  Carries the library's status code.
struct LogicError : std::logic_error {
    int status;
    LogicError(int s, const std::string& what) : std::logic_error(what), status(s) {}
};

// The surface of gpuFoam's chemistry-model plugin
// (gpuChemistryModel.H:26-185, gpuChemistryModel.C) without OpenFOAM: the
// model owns RR (per specie, [nSpecie][nCells]) and deltaTChem
// ([nCells]), reads its switches from chemistryProperties, and takes the
// cell fields as structure-of-arrays host arrays, one row per field
// (Y_0..Y_nS-1, T, p), which is how an OpenFOAM adapter holds them (one
// volScalarField per specie).  An adapter forwards its virtual overrides
// here: solve(deltaT) with the old-time fields, calculate/tc/reactionRR/
// specieReactionRR with the current ones.
class ChemistryModel {
public:
    ChemistryModel(const Mechanism& mech, const std::string& chemistryProperties, int nCells)
        : mech_(&mech), nCells_(nCells) {
        if (nCells < 0)
            throw Error(GFC_ERR_INVALID_ARGUMENT, "ChemistryModel: negative cell count");
        const gfc_mech_view v = mech.view();
        nSpecie_ = v.n_species;
        nReaction_ = v.n_reactions;
        int on = 0;
        double dt0 = 0.0;
        check(gfc_chemistry_switches_from_foam(chemistryProperties.c_str(), &on, &dt0,
                                               &deltaTChemMax_));
        chemistry_ = on != 0;
        check(gfc_ode_inputs_from_foam(chemistryProperties.c_str(), &inputs_, &Treact_));
        check(gfc_evaluator_create(mech.get(), &inputs_, Treact_, std::max(nCells, 1), &ev_));
        RR_.assign((size_t)nSpecie_ * nCells, 0.0);
        deltaTChem_.assign(nCells, dt0);
    }
    ChemistryModel(const ChemistryModel&) = delete;
    ChemistryModel& operator=(const ChemistryModel&) = delete;
    ~ChemistryModel() { gfc_evaluator_destroy(ev_); }

    int nSpecie() const { return nSpecie_; }
    int nReaction() const { return nReaction_; }
    int nCells() const { return nCells_; }
    int nEqns() const { return nSpecie_ + 2; }
    bool chemistry() const { return chemistry_; }
    double Treact() const { return Treact_; }
    double deltaTChemMax() const { return deltaTChemMax_; }
    // the reaction string as written in the dictionary
    std::string reactionName(int r) const {
        const char* s = gfc_mechanism_reaction_equation(mech_->get(), r);
        if (!s) throw Error(GFC_ERR_INVALID_ARGUMENT, "reactionName: index out of range");
        return s;
    }
    // RR[i*nCells + cell] [kg/m^3/s] of the last solve or calculate
    const std::vector<double>& RR() const { return RR_; }
    const std::vector<double>& deltaTChem() const { return deltaTChem_; }

    // solve(deltaT) (gpuChemistryModel.C:320-326): one chemistry step of
    // every cell from the old-time fields (rho0 [nCells], Y0 SoA
    // [(nSpecie+2)*nCells]); stores RR and the clamped deltaTChem and
    // returns min(deltaTMin, 2 deltaT), deltaTMin taken before clamping
    // (computeRRAndChemDeltaT, :336-364).  With chemistry off nothing is
    // computed and the step limit is min(great, 2 deltaT).
    double solve(double deltaT, const double* rho0, const double* Y0) {
        double dtMin = kGreat;
        if (chemistry_ && nCells_ > 0)
            check(gfc_solve(ev_, deltaT, deltaTChemMax_, nCells_, nCells_, rho0, Y0, nullptr,
                            RR_.data(), deltaTChem_.data(), &dtMin, nullptr, nullptr,
                            GFC_HOST_POINTERS | GFC_BATCHED));
        return std::min(dtMin, 2 * deltaT);
    }
    // solve(scalarField) (gpuChemistryModel.C:328-333): local time stepping
    // is not supported, as in the reference.
    double solve(const std::vector<double>& deltaT, const double* rho0, const double* Y0) {
        (void)deltaT;
        (void)rho0;
        (void)Y0;
        throw LogicError(GFC_ERR_LTS, "LTS time step not yet supported");
    }

    // calculate() (:146-183): RR from the current state.
    void calculate(const double* rho, const double* Y) {
        if (!chemistry_ || nCells_ == 0) return;
        check(gfc_calculate(ev_, nCells_, nCells_, rho, Y, RR_.data(), nullptr,
                            GFC_HOST_POINTERS));
    }
    // tc() (:366-443) [s]; `small` in every cell with chemistry off.
    std::vector<double> tc(const double* rho, const double* Y) const {
        std::vector<double> t(nCells_, kSmall);
        if (chemistry_ && nCells_ > 0)
            check(gfc_tc(ev_, nCells_, nCells_, rho, Y, t.data(), nullptr, GFC_HOST_POINTERS));
        return t;
    }
    // Qdot() (:445-466) [W/m^3] from the stored RR.
    std::vector<double> Qdot() const {
        std::vector<double> q(nCells_, 0.0);
        if (chemistry_ && nCells_ > 0)
            check(gfc_qdot(ev_, nCells_, nCells_, RR_.data(), q.data(), nullptr,
                           GFC_HOST_POINTERS));
        return q;
    }
    // reactionRR(r) (:186-240) [kmol/m^3/s].
    std::vector<double> reactionRR(int r, const double* rho, const double* Y) const {
        std::vector<double> out(nCells_, 0.0);
        if (chemistry_ && nCells_ > 0)
            check(gfc_reaction_rr(ev_, r, nCells_, nCells_, rho, Y, out.data(), nullptr,
                                  GFC_HOST_POINTERS));
        return out;
    }
    // specieReactionRR(r) (:242-317), [nSpecie][nCells] [kg/m^3/s].
    std::vector<double> specieReactionRR(int r, const double* rho, const double* Y) const {
        std::vector<double> out((size_t)nSpecie_ * nCells_, 0.0);
        if (chemistry_ && nCells_ > 0)
            check(gfc_specie_reaction_rr(ev_, r, nCells_, nCells_, rho, Y, out.data(), nullptr,
                                         GFC_HOST_POINTERS));
        return out;
    }

    gfc_evaluator* handle() { return ev_; }

    static constexpr double kGreat = 1.0 / 2.220446049250313e-16; // 1/DBL_EPSILON
    static constexpr double kSmall = 2.220446049250313e-16;       // DBL_EPSILON

private:
    const Mechanism* mech_;
    int nCells_ = 0, nSpecie_ = 0, nReaction_ = 0;
    bool chemistry_ = true;
    double Treact_ = 0.0, deltaTChemMax_ = 0.0;
    gfc_ode_inputs inputs_{};
    gfc_evaluator* ev_ = nullptr;
    std::vector<double> RR_, deltaTChem_;
};

} // namespace gpufoam_amd

#endif /* GPUFOAM_AMD_HPP */
