// Host mechanism compiler: OpenFOAM thermo + reactions dictionaries ->
// structure-of-arrays tables (gfc_mech_view).
#pragma once

#include <string>
#include <vector>

#include "../../include/gpufoam_amd.h"

struct gfc_mechanism {
    std::vector<std::string> species;
    std::vector<std::string> reactionNames;
    std::vector<std::string> equations;
    int nS = 0;
    int nR = 0;
    double Tlow = 0;
    double Thigh = 0;
    // species
    std::vector<double> W, sTlow, sThigh, sTcommon, high, low, rawHigh, rawLow;
    // reactions
    std::vector<int> type, irreversible, nLhs, nRhs, lhsIndex, rhsIndex,
        lhsExpInt, rhsExpInt;
    std::vector<double> k, kInf, troe, eff, lhsStoich, lhsExp, rhsStoich,
        rhsExp;
    std::vector<double> rY, rW, rTcommon, rhigh, rlow;
    std::vector<double> extra; // [nR*GFC_EXTRA_PARAMS]

    gfc_mech_view view() const;
};

namespace gfc {

// Thrown by the compiler; carries a gfc_status code.
struct MechError {
    int status;
    std::string message;
};

gfc_mechanism* compileMechanism(const std::string& thermoText,
                                const std::string& reactionsText);

void setLastError(const std::string& msg);

} // namespace gfc
