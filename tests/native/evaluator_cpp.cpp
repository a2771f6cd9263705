// Drives gpufoam_amd::GpuKernelEvaluator (include/gpufoam_amd.hpp) the way
// gpuChemistryModel drives FoamGpu::GpuKernelEvaluator
// (gpuChemistryModel.C:320-364): cell-major Y in, RR / deltaTChem /
// deltaTMin out.  Test-only.
//
//   evaluator_cpp <thermo> <reactions> <chemistryProperties> <in.bin> <out.bin>
//
// in.bin:  int32 nCells, int32 nEqns, f64 deltaT, f64 deltaTChemMax,
//          f64 Y[nCells*nEqns], f64 rho[nCells], f64 deltaTChem[nCells]
// out.bin: f64 RR[nCells*nSpecie], f64 deltaTChemRR[nCells], f64 deltaTMin,
//          f64 YNew[nCells*nEqns], f64 deltaTChemYNew[nCells]
#include <cstdio>
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

int main(int argc, char** argv) {
    if (argc != 6) {
        std::cerr << "usage: evaluator_cpp thermo reactions chemistryProperties in out\n";
        return 2;
    }
    try {
        gpufoam_amd::Mechanism mech(slurp(argv[1]), slurp(argv[2]));
        const auto [ode, Treact] = gpufoam_amd::readOdeInputs(slurp(argv[3]));
        std::ifstream in(argv[4], std::ios::binary);
        int nCells = 0, nEqns = 0;
        double deltaT = 0, deltaTChemMax = 0;
        in.read(reinterpret_cast<char*>(&nCells), 4);
        in.read(reinterpret_cast<char*>(&nEqns), 4);
        in.read(reinterpret_cast<char*>(&deltaT), 8);
        in.read(reinterpret_cast<char*>(&deltaTChemMax), 8);
        std::vector<double> Y((size_t)nCells * nEqns), rho(nCells), dt(nCells);
        in.read(reinterpret_cast<char*>(Y.data()), 8 * Y.size());
        in.read(reinterpret_cast<char*>(rho.data()), 8 * rho.size());
        in.read(reinterpret_cast<char*>(dt.data()), 8 * dt.size());
        const int nSpecie = nEqns - 2;
        gpufoam_amd::GpuKernelEvaluator ev(nCells, nEqns, nSpecie, mech, ode, Treact);
        auto [RR, dtRR, dtMin] = ev.computeRR(deltaT, deltaTChemMax, rho, dt, Y);
        auto [YNew, dtY] = ev.computeYNew(deltaT, deltaTChemMax, dt, Y);
        std::ofstream out(argv[5], std::ios::binary);
        out.write(reinterpret_cast<const char*>(RR.data()), 8 * RR.size());
        out.write(reinterpret_cast<const char*>(dtRR.data()), 8 * dtRR.size());
        out.write(reinterpret_cast<const char*>(&dtMin), 8);
        out.write(reinterpret_cast<const char*>(YNew.data()), 8 * YNew.size());
        out.write(reinterpret_cast<const char*>(dtY.data()), 8 * dtY.size());
        // error behaviour: a wrong size is reported, not an exit
        try {
            std::vector<double> bad(3);
            ev.computeYNew(deltaT, deltaTChemMax, dt, bad);
            std::cerr << "size error not raised\n";
            return 1;
        } catch (const gpufoam_amd::Error& e) {
            std::cout << "size error: status " << e.status << "\n";
        }
    } catch (const gpufoam_amd::Error& e) {
        std::cerr << "gpufoam_amd error " << e.status << ": " << e.what() << "\n";
        return 1;
    }
    return 0;
}
