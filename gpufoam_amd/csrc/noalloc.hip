// gfc_solve_noalloc: the persistent batched solve for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving
// loops).  No counterpart in the reference.
//
// The persistent wavefront kernels need two words of scratch besides the
// caller's inputs and outputs: the work-queue counter the wavefronts take
// cells from, and the bit pattern of the running minimum deltaTChem.
// gfc_solve takes them from the evaluator's own buffers; here they live in
// `scratch`, a file-scope device array of kScratchBytes that every GPU the
// library runs on holds until the process ends.  Its device address is
// resolved once per device with hipGetSymbolAddress when an evaluator is
// created on that device (host code never passes the array's name to a
// kernel: on the host the name is only a stand-in for the symbol).
//
// Every call on one GPU shares that GPU's array and leaves its last
// contents behind; each call clears what must start at zero (the queue
// counter, the minimum) with small kernels on its own stream before the
// solve.  Calls that overlap on one GPU race for the array.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "evaluator.hpp"

namespace gfc {

namespace {

constexpr size_t kScratchBytes = 256;
constexpr int kMaxDevices = 64;

// [0, 4): work-queue counter (int); [8, 16): minimum deltaTChem bits
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex gMutex;
void* gScratch[kMaxDevices] = {}; // device address of `scratch` per device

__global__ void static_init_min_kernel(unsigned long long* bits) {
    *bits = static_cast<unsigned long long>(__double_as_longlong(HUGE_VAL));
}

__global__ void static_finish_min_kernel(const unsigned long long* bits, double* out) {
    *out = __longlong_as_double(static_cast<long long>(*bits));
}

} // namespace

int noalloc_init_device(int device) {
    if (device < 0 || device >= kMaxDevices) {
        setLastError("gfc_evaluator_create: device index beyond the caller-free "
                     "entry point's table");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(gMutex);
    if (gScratch[device]) return GFC_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        setLastError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return GFC_ERR_HIP;
    }
    gScratch[device] = ptr;
    return GFC_OK;
}

} // namespace gfc

extern "C" int gfc_solve_noalloc(gfc_evaluator* ev, double deltaT, double deltaTChemMax,
                                 int nCells, int ldY, const double* rho, const double* Y0,
                                 double* Ynew, double* RR, double* deltaTChem,
                                 double* deltaTMin, int* cellStats, void* stream) {
    using namespace gfc;
    if (!ev || nCells < 0 || ldY < nCells || !(deltaT > 0) ||
        (nCells > 0 && (!Y0 || !deltaTChem || (RR && !rho)))) {
        setLastError("gfc_solve_noalloc: bad argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    if (nCells > ev->maxCells) {
        setLastError("gfc_solve_noalloc: nCells exceeds the evaluator's maxCells");
        return GFC_ERR_CAPACITY;
    }
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(gMutex);
        if (ev->device >= 0 && ev->device < kMaxDevices) base = gScratch[ev->device];
    }
    if (!base) {
        setLastError("gfc_solve_noalloc: the device array was not resolved on this "
                     "evaluator's device");
        return GFC_ERR_HIP;
    }
    int* queue = static_cast<int*>(base);
    auto* minBits = reinterpret_cast<unsigned long long*>(static_cast<char*>(base) + 8);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (nCells == 0) {
        if (deltaTMin) {
            hipLaunchKernelGGL(static_init_min_kernel, dim3(1), dim3(1), 0, s, minBits);
            hipLaunchKernelGGL(static_finish_min_kernel, dim3(1), dim3(1), 0, s, minBits,
                               deltaTMin);
        }
        return GFC_OK;
    }
    SolveArgs A = solve_args(ev, deltaT, deltaTChemMax, nCells);
    A.ldY = ldY;
    A.rho = rho;
    A.Y0 = Y0;
    A.Ynew = Ynew;
    A.RR = RR;
    A.dtChem = deltaTChem;
    A.stats = cellStats;
    A.dtMinBits = deltaTMin ? minBits : nullptr;
    if (deltaTMin) hipLaunchKernelGGL(static_init_min_kernel, dim3(1), dim3(1), 0, s, minBits);
    const int st = batched_launch_queue(ev, A, s, queue);
    if (st != GFC_OK) return st;
    if (deltaTMin)
        hipLaunchKernelGGL(static_finish_min_kernel, dim3(1), dim3(1), 0, s, minBits,
                           deltaTMin);
    if (hipGetLastError() != hipSuccess) {
        setLastError("gfc_solve_noalloc: kernel launch failed");
        return GFC_ERR_HIP;
    }
    return GFC_OK;
}
