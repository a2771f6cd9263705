// Batched forms of the solve.
//
//  * Persistent work-stealing launch: a grid sized to the number of
//    wavefronts the device can hold resident; every wavefront takes the
//    next cell from one atomic counter until the cells run out, so cells
//    with hundreds of sub-steps do not hold up a statically assigned tile.
//  * hipGraph: with GFC_GRAPH the queue reset and the persistent kernel
//    are recorded once (on a private stream) and the executable graph is
//    replayed on the caller's stream on later calls with identical
//    arguments; changed arguments update the graph in place.  Called inside
//    the caller's own stream capture, the launches join that graph instead.
//  * Launch shape (autotuned by gpufoam_amd/autotune.py): cells per
//    wavefront 1 (the wavefront kernel) or 64 (one lane per cell), and the
//    number of wavefronts kept resident per compute unit.
#include <hip/hip_runtime.h>

#include <cstring>

#include "evaluator.hpp"
#include "mechanism.hpp"

namespace gfc {

template <int N>
__global__ void team_solve_kernel(SolveArgs A, int* queue);
__global__ void lane_solve_kernel(SolveArgs A, double* scratch, long per);

namespace {

__global__ void reset_queue_kernel(int* q) { *q = 0; }

struct BatchedState {
    int* queue = nullptr;
    int numCUs = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t captureStream = nullptr;  // private stream the graph is recorded on
    SolveArgs lastArgs{};
    int lastGrid = 0;
    int lastN = 0;
    bool valid = false;
};

using KernelFn = void (*)(SolveArgs, int*);

// Row capacity of the wavefront kernel for n = nSpecie+1: exact sizes for
// the bundled mechanisms, padded capacities otherwise.
KernelFn team_kernel_for(int n, int& N) {
    if (n == 11) { N = 11; return team_solve_kernel<11>; }
    if (n == 54) { N = 54; return team_solve_kernel<54>; }
    if (n <= 32) { N = 32; return team_solve_kernel<32>; }
    if (n <= 64) { N = 64; return team_solve_kernel<64>; }
    N = 0;
    return nullptr;
}

bool hipFail(hipError_t e, const char* what) {
    if (e == hipSuccess) return false;
    setLastError(std::string(what) + ": " + hipGetErrorString(e));
    return true;
}

} // namespace

int batched_init(gfc_evaluator* ev) {
    auto* b = new BatchedState();
    ev->batched = b;
    if (hipFail(hipMalloc(&b->queue, 256), "hipMalloc(queue)")) return GFC_ERR_HIP;
    hipDeviceProp_t prop;
    if (hipFail(hipGetDeviceProperties(&prop, ev->device), "hipGetDeviceProperties"))
        return GFC_ERR_HIP;
    b->numCUs = prop.multiProcessorCount;
    // Graphs are recorded on a private stream and launched on the caller's:
    // the caller's stream may be the null stream, which cannot be captured.
    if (hipFail(hipStreamCreateWithFlags(&b->captureStream, hipStreamNonBlocking),
                "hipStreamCreateWithFlags"))
        return GFC_ERR_HIP;
    return GFC_OK;
}

void batched_free(gfc_evaluator* ev) {
    auto* b = static_cast<BatchedState*>(ev->batched);
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->queue) (void)hipFree(b->queue);
    if (b->captureStream) (void)hipStreamDestroy(b->captureStream);
    delete b;
    ev->batched = nullptr;
}

// Wavefronts per CU the wavefront kernel can keep resident.
static int resident_blocks(KernelFn fn) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void*)fn, 64, 0) !=
            hipSuccess ||
        blocks <= 0)
        blocks = 1;
    return blocks;
}

static int enqueue(gfc_evaluator* ev, BatchedState* b, const SolveArgs& A,
                   hipStream_t s, KernelFn fn, int grid) {
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, s, b->queue);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(64), 0, s, A, b->queue);
    if (hipFail(hipGetLastError(), "persistent launch")) return GFC_ERR_HIP;
    return GFC_OK;
}

int batched_launch(gfc_evaluator* ev, const SolveArgs& A, hipStream_t s,
                   bool graph) {
    auto* b = static_cast<BatchedState*>(ev->batched);
    const int n = A.m.nS + 1;
    if (ev->cellsPerWave == 64) {
        // one lane per cell
        const int blocks = (A.nCells + 63) / 64;
        hipLaunchKernelGGL(lane_solve_kernel, dim3(blocks), dim3(64), 0, s, A,
                           ev->laneScratch, ev->lanePer);
        return hipFail(hipGetLastError(), "lane launch") ? GFC_ERR_HIP : GFC_OK;
    }
    int N = 0;
    KernelFn fn = team_kernel_for(n, N);
    if (!fn || ev->dmTeam.teamWidth > kTeamWidth) {
        // mechanism outside the wavefront kernel's capacity: one lane per cell
        const int blocks = (A.nCells + 63) / 64;
        hipLaunchKernelGGL(lane_solve_kernel, dim3(blocks), dim3(64), 0, s, A,
                           ev->laneScratch, ev->lanePer);
        return hipFail(hipGetLastError(), "lane launch") ? GFC_ERR_HIP : GFC_OK;
    }
    SolveArgs At = A;
    At.m = ev->dmTeam;
    int perCU = resident_blocks(fn);
    if (ev->wavesPerBlock > 0 && ev->wavesPerBlock < perCU) perCU = ev->wavesPerBlock;
    int grid = b->numCUs * perCU;
    if (grid > A.nCells) grid = A.nCells > 0 ? A.nCells : 1;
    if (graph) {
        // inside a caller's own capture the launches simply join that graph
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (s && hipStreamIsCapturing(s, &cs) == hipSuccess &&
            cs != hipStreamCaptureStatusNone)
            graph = false;
    }
    if (!graph) return enqueue(ev, b, At, s, fn, grid);

    // graph path: reuse the executable graph when nothing changed
    const bool same = b->valid && b->lastGrid == grid && b->lastN == N &&
                      std::memcmp(&b->lastArgs, &At, sizeof(SolveArgs)) == 0;
    if (!same) {
        hipGraph_t g = nullptr;
        hipStream_t cs = b->captureStream;
        if (hipFail(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal),
                    "hipStreamBeginCapture"))
            return GFC_ERR_HIP;
        int st = enqueue(ev, b, At, cs, fn, grid);
        hipError_t e = hipStreamEndCapture(cs, &g);
        if (st != GFC_OK) return st;
        if (hipFail(e, "hipStreamEndCapture")) return GFC_ERR_HIP;
        bool updated = false;
        if (b->exec) {
            hipGraphExecUpdateResult res;
            hipGraphNode_t errNode = nullptr;
            if (hipGraphExecUpdate(b->exec, g, &errNode, &res) == hipSuccess &&
                res == hipGraphExecUpdateSuccess)
                updated = true;
        }
        if (!updated) {
            if (b->exec) (void)hipGraphExecDestroy(b->exec);
            b->exec = nullptr;
            if (hipFail(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0),
                        "hipGraphInstantiate")) {
                (void)hipGraphDestroy(g);
                return GFC_ERR_HIP;
            }
        }
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->graph = g;
        b->lastArgs = At;
        b->lastGrid = grid;
        b->lastN = N;
        b->valid = true;
    }
    if (hipFail(hipGraphLaunch(b->exec, s), "hipGraphLaunch")) return GFC_ERR_HIP;
    return GFC_OK;
}

} // namespace gfc

extern "C" {
int gfc_evaluator_set_batched_shape(gfc_evaluator* ev, int cellsPerWave,
                                    int wavesPerBlock, int chunk) {
    if (!ev || (cellsPerWave != 0 && cellsPerWave != 1 && cellsPerWave != 64) ||
        wavesPerBlock < 0 || chunk < 0 || chunk > 1) {
        gfc::setLastError("gfc_evaluator_set_batched_shape: unsupported shape "
                          "(cellsPerWave must be 0, 1 or 64, chunk 0 or 1)");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    ev->cellsPerWave = cellsPerWave;
    ev->wavesPerBlock = wavesPerBlock;
    ev->chunk = chunk;
    return GFC_OK;
}
int gfc_evaluator_get_batched_shape(const gfc_evaluator* ev, int* cellsPerWave,
                                    int* wavesPerBlock, int* chunk) {
    if (!ev) return GFC_ERR_INVALID_ARGUMENT;
    if (cellsPerWave) *cellsPerWave = ev->cellsPerWave;
    if (wavesPerBlock) *wavesPerBlock = ev->wavesPerBlock;
    if (chunk) *chunk = ev->chunk;
    return GFC_OK;
}
}
