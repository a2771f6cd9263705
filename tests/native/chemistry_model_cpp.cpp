// Drives gpufoam_amd::ChemistryModel (include/gpufoam_amd.hpp) through the
// calls an OpenFOAM adapter forwards from gpuChemistryModel's virtual
// overrides (gpuChemistryModel.C:146-466).  Test-only.
//
//   chemistry_model_cpp <thermo> <reactions> <chemistryProperties> <in.bin> <out.bin>
//
// in.bin:  int32 nCells, int32 nSpecie, int32 reaction, f64 deltaT,
//          f64 Y[(nSpecie+2)*nCells] (SoA rows Y_i, T, p), f64 rho[nCells]
// out.bin: f64 solveReturn, f64 RR[nSpecie*nCells], f64 deltaTChem[nCells],
//          f64 Qdot[nCells] (of the solve's RR),
//          f64 RRcalc[nSpecie*nCells] (calculate at Y), f64 tc[nCells],
//          f64 reactionRR[nCells], f64 specieReactionRR[nSpecie*nCells]
// stdout:  "lts: status 12" when solve(scalarField) throws as the reference.
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

#include "gpufoam_amd.hpp"

static std::string slurp(const char* path) {
    std::ifstream f(path);
    std::stringstream s;
    s << f.rdbuf();
    return s.str();
}

template <class T>
static void put(std::ofstream& out, const std::vector<T>& v) {
    out.write(reinterpret_cast<const char*>(v.data()), sizeof(T) * v.size());
}

int main(int argc, char** argv) {
    if (argc != 6) {
        std::cerr << "usage: chemistry_model_cpp thermo reactions chemistryProperties in out\n";
        return 2;
    }
    try {
        gpufoam_amd::Mechanism mech(slurp(argv[1]), slurp(argv[2]));
        std::ifstream in(argv[4], std::ios::binary);
        int nCells = 0, nSpecie = 0, reaction = 0;
        double deltaT = 0;
        in.read(reinterpret_cast<char*>(&nCells), 4);
        in.read(reinterpret_cast<char*>(&nSpecie), 4);
        in.read(reinterpret_cast<char*>(&reaction), 4);
        in.read(reinterpret_cast<char*>(&deltaT), 8);
        std::vector<double> Y((size_t)(nSpecie + 2) * nCells), rho(nCells);
        in.read(reinterpret_cast<char*>(Y.data()), 8 * Y.size());
        in.read(reinterpret_cast<char*>(rho.data()), 8 * rho.size());

        gpufoam_amd::ChemistryModel model(mech, slurp(argv[3]), nCells);
        if (model.nSpecie() != nSpecie) {
            std::cerr << "specie count mismatch\n";
            return 1;
        }
        std::ofstream out(argv[5], std::ios::binary);
        const double ret = model.solve(deltaT, rho.data(), Y.data());
        out.write(reinterpret_cast<const char*>(&ret), 8);
        put(out, model.RR());
        put(out, model.deltaTChem());
        put(out, model.Qdot());
        model.calculate(rho.data(), Y.data());
        put(out, model.RR());
        put(out, model.tc(rho.data(), Y.data()));
        put(out, model.reactionRR(reaction, rho.data(), Y.data()));
        put(out, model.specieReactionRR(reaction, rho.data(), Y.data()));
        std::cout << "reaction " << reaction << ": " << model.reactionName(reaction) << "\n";
        try {
            std::vector<double> dts(nCells, deltaT);
            model.solve(dts, rho.data(), Y.data());
            std::cerr << "LTS solve did not throw\n";
            return 1;
        } catch (const std::logic_error& e) {
            const auto* le = dynamic_cast<const gpufoam_amd::LogicError*>(&e);
            std::cout << "lts: status " << (le ? le->status : -1) << " (" << e.what() << ")\n";
        }
    } catch (const gpufoam_amd::Error& e) {
        std::cerr << "gpufoam_amd error " << e.status << ": " << e.what() << "\n";
        return 1;
    }
    return 0;
}
