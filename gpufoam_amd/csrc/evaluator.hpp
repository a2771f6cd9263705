// Internal definition of gfc_evaluator (the C ABI keeps it opaque).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gpufoam_amd.h"
#include "kernel_common.hpp"
#include "mechanism.hpp"

struct gfc_evaluator {
    int device = 0;
    int nS = 0;
    int nR = 0;
    int maxCells = 0;
    gfc_ode_inputs ode{};
    double Treact = 0;
    gfc::DeviceMech dm{};
    void* mechBlock = nullptr;
    gfc::DeviceMech dmTeam{}; // reactions grouped by rate type
    void* mechBlockTeam = nullptr;
    // plain kernel: per-cell SoA scratch
    double* laneScratch = nullptr;
    long laneStride = 0;  // cells of scratch (maxCells rounded to 64)
    long lanePer = 0;     // scratch doubles per cell (wave-blocked layout)
    size_t laneScratchBytes = 0;
    unsigned long long* minBits = nullptr;
    unsigned* boundsWord = nullptr; // bounds-check build: failed-check bits
    // staging for GFC_HOST_POINTERS
    double* sY0 = nullptr;
    double* sYnew = nullptr;
    double* sRR = nullptr;
    double* sRho = nullptr;
    double* sDt = nullptr;
    double* sDtMin = nullptr;
    int* sStats = nullptr;
    // persistent batched form
    int cellsPerWave = 0;  // 0: chosen automatically
    int wavesPerBlock = 0;
    int chunk = 0;
    int factorization = GFC_LU_COLUMN; // workgroup kernel's LU variant (set by batched_init)
    void* batched = nullptr; // owned state of the batched launcher
    void* adjoint = nullptr; // owned state of the discrete adjoint
    gfc_mechanism mech;      // host copy of the mechanism (rate parameters can be updated)
};

namespace gfc {
SolveArgs solve_args(const gfc_evaluator* ev, double deltaT, double deltaTChemMax,
                     int nCells);
int evaluator_upload(gfc_evaluator* ev);
// Bounds-check build: waits for the stream and returns GFC_ERR_BOUNDS (and
// clears the word) when a kernel's check failed; GFC_OK otherwise and in the
// default build, without waiting.
int bounds_report(gfc_evaluator* ev, hipStream_t s);
bool block_eligible(const gfc_mech_view& v);
void adjoint_free(gfc_evaluator* ev);
int batched_init(gfc_evaluator* ev);
void batched_free(gfc_evaluator* ev);
int batched_launch(gfc_evaluator* ev, const SolveArgs& A, hipStream_t s,
                   bool graph);
// The drop-in (non-persistent) launch of the same kernels: a wavefront per
// cell over a grid of nCells, four cells per wavefront, or the workgroup
// kernel dealing cells round-robin; lane = true: one lane per cell.
int plain_launch(gfc_evaluator* ev, const SolveArgs& A, hipStream_t s, bool lane);
// The persistent wavefront kernels with a caller-given work-queue counter
// (no graph); GFC_ERR_CAPACITY for mechanisms that need the workgroup or
// one-lane kernels.
int batched_launch_queue(gfc_evaluator* ev, const SolveArgs& A, hipStream_t s, int* queue);
// Resolves the caller-free entry point's device array on the current
// device (once per device).
int noalloc_init_device(int device);
// Whether a launch of nCells cells runs four cells per wavefront (the
// evaluator's shape, or the automatic choice); N: the kernels' size.
bool team16_selected(const gfc_evaluator* ev, int nCells, int& N);
// Row capacity N of the wavefront-per-cell kernel that can run this
// evaluator's mechanism (0: none; nSpecie+1 <= 64, at most 3 species per
// reaction side and 128 efficiency sets, gpuFoam's rate types only).
int wave_kernel_rows(const gfc_evaluator* ev);
// The work-queue chunking of the persistent wavefront kernel for a grid of
// `waves` wavefronts (16-cell tickets, the last cells one by one).
void queue_chunks_for(const gfc_evaluator* ev, SolveArgs& A, int waves);
} // namespace gfc
