// Physical and numerical constants shared by host and device code.
// The values are OpenFOAM's SI set as gpuFoam uses them
// (gpu_utils/common/gpu_constants.H:12-82): results depend on them bit for
// bit, so they are kept identical.
#pragma once

#include <cfloat>

#if defined(__HIPCC__)
#define GFC_HD __host__ __device__
#else
#define GFC_HD
#endif

namespace gfc {

constexpr double kNNA = 6.02214e+26;  // Avogadro number per kmol
constexpr double kBoltzmann = 1.38065e-23;
constexpr double kRR = kNNA * kBoltzmann; // universal gas constant [J/kmol/K]
constexpr double kPstd = 1e5;
constexpr double kTstd = 298.15;
constexpr double kSmall = DBL_EPSILON;
constexpr double kVSmall = DBL_MIN;
constexpr double kVGreat = DBL_MAX / 10.0;
constexpr double kGreat = 1.0 / kSmall;
constexpr double kLogTen = 2.302585092994046;
constexpr int kLabelMax = 2147483647 / 10;

} // namespace gfc
