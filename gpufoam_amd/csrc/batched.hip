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
//    wavefront 1 (the wavefront kernel), 4 (16-lane teams), 64 (one lane
//    per cell) or code 2 (two wavefronts per cell), and the number of
//    workgroups kept resident per compute unit.
#include <hip/hip_runtime.h>

#include <cstring>

#include "evaluator.hpp"
#include "mechanism.hpp"

namespace gfc {

template <int N>
__global__ void team_solve_kernel(SolveArgs A, int* queue);
template <int N>
__global__ void team16_solve_kernel(SolveArgs A, int* queue);
template <int N, bool EXACT>
__global__ void pair_solve_kernel(SolveArgs A, int* queue);
template <int N>
__global__ void pair_eval_kernel(EvalArgs E);
template <int NA, int NC, int LUV>
__global__ void block_solve_kernel(SolveArgs A, int* queue, double* scratch, long perWG);
__global__ void lane_solve_kernel(SolveArgs A, double* scratch, long per);
template <int N>
__global__ void team_eval_kernel(EvalArgs E);
template <int N>
__global__ void team16_eval_kernel(EvalArgs E);
template <int NA, int NC>
__global__ void block_eval_kernel(EvalArgs E, double* scratch, long perWG);

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
    // workgroup-per-cell kernel (mechanisms with 64 < n <= kBlockMaxN)
    int blkNA = 0;             // rows per thread (0: not available)
    int blkGrid = 0;           // resident workgroups
    long blkPerWG = 0;         // scratch doubles per workgroup
    double* blkScratch = nullptr;
};

using KernelFn = void (*)(SolveArgs, int*);
using BlockFn = void (*)(SolveArgs, int*, double*, long);

constexpr int kBlockThreads = 512;
constexpr int kLanesPerWave = 64;
// Cells per work-queue ticket of the wavefront kernel: 16 doubles are one
// 128-byte line of every state row, so a chunk's lines are fetched into one
// XCD's L2 instead of into each L2 whose wavefronts take a neighbouring cell.
constexpr int kQueueChunk = kQueueChunkMax;

// Chunked tickets for all but the last kQueueChunk cells per wavefront,
// which go out one by one so that the launch ends evenly.
void set_queue_chunks(const gfc_evaluator* ev, SolveArgs& A, int waves) {
    const int G = ev->chunk == 1 ? 1 : kQueueChunk;
    const long bulk = (long)A.nCells - (long)G * waves;
    A.qChunk = G;
    A.qBulk = (G > 1 && bulk > 0) ? (int)(bulk / G * G) : 0;
}

BlockFn block_kernel_for(int NA, int lu) {
    static const BlockFn k8[3] = {block_solve_kernel<8, 4, 0>, block_solve_kernel<8, 4, 1>,
                                  block_solve_kernel<8, 4, 2>};
    static const BlockFn k13[3] = {block_solve_kernel<13, 7, 0>, block_solve_kernel<13, 7, 1>,
                                   block_solve_kernel<13, 7, 2>};
    if (lu < 0 || lu > 2) return nullptr;
    if (NA == 8) return k8[lu];
    if (NA == 13) return k13[lu];
    return nullptr;
}

// One launch of the batched form: the wavefront kernel or the workgroup
// kernel, with its persistent grid.
struct Plan {
    KernelFn team = nullptr;
    BlockFn blk = nullptr;
    int grid = 0;
    int threads = 64; // workgroup size of `team` (128: two wavefronts per cell)
    int key = 0;      // identifies the kernel for graph reuse
};

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

// Four cells per wavefront (16-lane teams): n <= 16, at most kSetsT
// efficiency sets, at most kTeamWidth species per reaction side.
KernelFn team16_kernel_for(const DeviceMech& m, int& N) {
    const int n = m.nS + 1;
    N = 0;
    if (m.nSets > kSetsT || m.teamWidth > kTeamWidth) return nullptr;
    if (n == 11) { N = 11; return team16_solve_kernel<11>; }
    if (n <= 16) { N = 16; return team16_solve_kernel<16>; }
    return nullptr;
}

// Two wavefronts per cell: 32 < n <= 64, at most kSetsT efficiency sets (the
// dense third-body part is fused into the row transform), at most
// kTeamWidth species per reaction side.
KernelFn pair_kernel_for(const DeviceMech& m, int& N) {
    const int n = m.nS + 1;
    N = 0;
    if (m.nSets > kSetsT || m.teamWidth > kTeamWidth || m.extended) return nullptr;
    if (n == 54) { N = 54; return pair_solve_kernel<54, true>; }
    if (n > 32 && n <= 64) { N = 64; return pair_solve_kernel<64, false>; }
    return nullptr;
}

bool hipFail(hipError_t e, const char* what) {
    if (e == hipSuccess) return false;
    setLastError(std::string(what) + ": " + hipGetErrorString(e));
    return true;
}

} // namespace

void queue_chunks_for(const gfc_evaluator* ev, SolveArgs& A, int waves) {
    set_queue_chunks(ev, A, waves);
}

int batched_init(gfc_evaluator* ev) {
    auto* b = new BatchedState();
    ev->batched = b;
    if (hipFail(hipMalloc(&b->queue, 256), "hipMalloc(queue)")) return GFC_ERR_HIP;
    hipDeviceProp_t prop;
    if (hipFail(hipGetDeviceProperties(&prop, ev->device), "hipGetDeviceProperties"))
        return GFC_ERR_HIP;
    b->numCUs = prop.multiProcessorCount;
    // workgroup-per-cell kernel: scratch for its resident grid
    const int n = ev->nS + 1;
    if (block_eligible(ev->mech.view()) && ev->dmTeam.nSets <= kBlockMaxSets) {
        const int NA = n <= 128 ? 8 : 13;
        const int NC = NA == 8 ? 4 : 7;
        // resident workgroups: the smallest over the factorisation variants
        int occ = 1 << 30;
        for (int lu = 0; lu < 3; ++lu) {
            int o = 0;
            if (hipFail(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                            &o, (const void*)block_kernel_for(NA, lu), kBlockThreads, 0),
                        "hipOccupancyMaxActiveBlocksPerMultiprocessor"))
                return GFC_ERR_HIP;
            occ = o < occ ? o : occ;
        }
        if (occ > 0) {
            b->blkNA = NA;
            // measured default: the MFMA rank-4 update on the 8 x 4 register
            // block (4 % faster on 100 species), the column scheme on the
            // 13 x 7 one (the MFMA tuples spill there)
            ev->factorization = NA == 8 ? GFC_LU_PANEL_MFMA : GFC_LU_COLUMN;
            b->blkGrid = b->numCUs * occ;
            b->blkPerWG = 2L * NA * NC * kBlockThreads;
            if (hipFail(hipMalloc(&b->blkScratch, sizeof(double) * b->blkPerWG * b->blkGrid),
                        "hipMalloc(workgroup scratch)"))
                return GFC_ERR_HIP;
        }
    }
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
    if (b->blkScratch) (void)hipFree(b->blkScratch);
    if (b->captureStream) (void)hipStreamDestroy(b->captureStream);
    delete b;
    ev->batched = nullptr;
}

// Workgroups of `threads` per CU the kernel can keep resident.
static int resident_blocks(KernelFn fn, int threads = 64) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void*)fn, threads, 0) !=
            hipSuccess ||
        blocks <= 0)
        blocks = 1;
    return blocks;
}

static int enqueue(BatchedState* b, const SolveArgs& A, hipStream_t s, const Plan& pl) {
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, s, b->queue);
    if (pl.blk)
        hipLaunchKernelGGL(pl.blk, dim3(pl.grid), dim3(kBlockThreads), 0, s, A, b->queue,
                           b->blkScratch, b->blkPerWG);
    else
        hipLaunchKernelGGL(pl.team, dim3(pl.grid), dim3(pl.threads), 0, s, A, b->queue);
    if (hipFail(hipGetLastError(), "persistent launch")) return GFC_ERR_HIP;
    return GFC_OK;
}

// Batch size from which the default shape puts four cells in a wavefront.
constexpr int kTeam16MinCells = 16384;

// Four cells per wavefront when asked for, and by default for small
// mechanisms once the batch fills the chip four times over: on MI355X the
// batched solve takes 0.39x the time of the one-cell form at 65536 H2 cells,
// 0.41x to 0.90x at 16384 and 1.13x at 4096, where 1024 wavefronts leave
// most SIMDs idle.
int wave_kernel_rows(const gfc_evaluator* ev) {
    int N = 0;
    const DeviceMech& m = ev->dmTeam;
    if (m.extended || m.teamWidth > kTeamWidth || m.nSets > kTeamMaxSets ||
        !team_kernel_for(m.nS + 1, N))
        return 0;
    return N;
}

bool team16_selected(const gfc_evaluator* ev, int nCells, int& N) {
    const bool want = ev->cellsPerWave == 4 ||
                      (ev->cellsPerWave == 0 && nCells >= kTeam16MinCells);
    return want && team16_kernel_for(ev->dmTeam, N) != nullptr;
}

// The kernel and persistent grid of a launch of nCells cells: four cells
// per wavefront, a wavefront per cell, a workgroup per cell, or (pl.team
// and pl.blk both null) one lane per cell.
static void choose_plan(gfc_evaluator* ev, int nCells, Plan& pl, int& cellsPerBlock) {
    auto* b = static_cast<BatchedState*>(ev->batched);
    const int n = ev->nS + 1;
    pl = Plan{};
    cellsPerBlock = 1;
    // one lane per cell when asked for, and for mechanisms with rate types
    // beyond gpuFoam's four (the wavefront and workgroup kernels evaluate
    // those four from their packed records)
    if (ev->cellsPerWave == 64 || ev->dmTeam.extended) return;
    int N16 = 0, N = 0, N2 = 0;
    KernelFn t16 = team16_selected(ev, nCells, N16) ? team16_kernel_for(ev->dmTeam, N16)
                                                    : nullptr;
    KernelFn team = team_kernel_for(n, N);
    KernelFn pair = ev->cellsPerWave == 2 ? pair_kernel_for(ev->dmTeam, N2) : nullptr;
    if (pair) {
        int perCU = resident_blocks(pair, 2 * kLanesPerWave);
        if (ev->wavesPerBlock > 0 && ev->wavesPerBlock < perCU) perCU = ev->wavesPerBlock;
        pl.team = pair;
        pl.threads = 2 * kLanesPerWave;
        pl.grid = b->numCUs * perCU;
        pl.key = 3000 + N2;
    } else if (t16) {
        int perCU = resident_blocks(t16);
        if (ev->wavesPerBlock > 0 && ev->wavesPerBlock < perCU) perCU = ev->wavesPerBlock;
        pl.team = t16;
        pl.grid = b->numCUs * perCU;
        pl.key = 2000 + N16;
        cellsPerBlock = kLanesPerWave / 16;
    } else if (team && ev->dmTeam.teamWidth <= kTeamWidth &&
               ev->dmTeam.nSets <= kTeamMaxSets) {
        int perCU = resident_blocks(team);
        if (ev->wavesPerBlock > 0 && ev->wavesPerBlock < perCU) perCU = ev->wavesPerBlock;
        pl.team = team;
        pl.grid = b->numCUs * perCU;
        pl.key = N;
    } else if (b->blkNA > 0) {
        pl.blk = block_kernel_for(b->blkNA, ev->factorization);
        pl.grid = b->blkGrid;
        pl.key = 1000 + 10 * b->blkNA + ev->factorization;
    }
}

static int lane_launch(gfc_evaluator* ev, const SolveArgs& A, hipStream_t s) {
    const int blocks = (A.nCells + 63) / 64;
    hipLaunchKernelGGL(lane_solve_kernel, dim3(blocks), dim3(64), 0, s, A,
                       ev->laneScratch, ev->lanePer);
    return hipFail(hipGetLastError(), "lane launch") ? GFC_ERR_HIP : GFC_OK;
}

int plain_launch(gfc_evaluator* ev, const SolveArgs& A, hipStream_t s, bool lane) {
    Plan pl;
    int perBlock = 1;
    if (!lane) choose_plan(ev, A.nCells, pl, perBlock);
    if (lane || (!pl.team && !pl.blk)) return lane_launch(ev, A, s);
    auto* b = static_cast<BatchedState*>(ev->batched);
    SolveArgs At = A;
    At.m = ev->dmTeam;
    if (pl.blk) {
        // one workgroup per cell: each resident workgroup owns a scratch
        // slot, so the grid stays at the resident count and deals the cells
        // round-robin
        const int grid = pl.grid < A.nCells ? pl.grid : A.nCells;
        hipLaunchKernelGGL(pl.blk, dim3(grid), dim3(kBlockThreads), 0, s, At, nullptr,
                           b->blkScratch, b->blkPerWG);
    } else {
        // one wavefront per cell (or per four cells): the grid covers the
        // cells and the hardware dispatcher fills the CUs
        const int grid = (A.nCells + perBlock - 1) / perBlock;
        hipLaunchKernelGGL(pl.team, dim3(grid), dim3(pl.threads), 0, s, At, nullptr);
    }
    return hipFail(hipGetLastError(), "drop-in launch") ? GFC_ERR_HIP : GFC_OK;
}

int batched_launch_queue(gfc_evaluator* ev, const SolveArgs& A, hipStream_t s, int* queue) {
    Plan pl;
    int perBlock = 1;
    choose_plan(ev, A.nCells, pl, perBlock);
    if (!pl.team) {
        setLastError("the persistent form of this mechanism needs the workgroup-per-cell or "
                     "one-lane-per-cell kernel, whose per-cell scratch the caller-free "
                     "entry point does not hold (wavefront kernels only: nSpecie+1 <= 64)");
        return GFC_ERR_CAPACITY;
    }
    SolveArgs At = A;
    At.m = ev->dmTeam;
    const int blocksNeeded = (A.nCells + perBlock - 1) / perBlock;
    int grid = pl.grid < blocksNeeded ? pl.grid : blocksNeeded;
    if (grid < 1) grid = 1;
    if (perBlock == 1) set_queue_chunks(ev, At, grid);
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, s, queue);
    hipLaunchKernelGGL(pl.team, dim3(grid), dim3(pl.threads), 0, s, At, queue);
    return hipFail(hipGetLastError(), "persistent launch") ? GFC_ERR_HIP : GFC_OK;
}

int batched_launch(gfc_evaluator* ev, const SolveArgs& A, hipStream_t s,
                   bool graph) {
    auto* b = static_cast<BatchedState*>(ev->batched);
    Plan pl;
    int perBlock = 1;
    choose_plan(ev, A.nCells, pl, perBlock);
    if (!pl.team && !pl.blk) return lane_launch(ev, A, s);
    const bool t16 = perBlock > 1;
    SolveArgs At = A;
    At.m = ev->dmTeam;
    if (t16) {
        const int waves = (A.nCells + perBlock - 1) / perBlock;
        if (pl.grid > waves) pl.grid = waves > 0 ? waves : 1;
    } else if (pl.grid > A.nCells) {
        pl.grid = A.nCells > 0 ? A.nCells : 1;
    }
    const int grid = pl.grid;
    if (pl.team && !t16) set_queue_chunks(ev, At, grid);
    if (graph) {
        // inside a caller's own capture the launches simply join that graph
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (s && hipStreamIsCapturing(s, &cs) == hipSuccess &&
            cs != hipStreamCaptureStatusNone)
            graph = false;
    }
    if (!graph) return enqueue(b, At, s, pl);

    // graph path: reuse the executable graph when nothing changed
    const bool same = b->valid && b->lastGrid == grid && b->lastN == pl.key &&
                      std::memcmp(&b->lastArgs, &At, sizeof(SolveArgs)) == 0;
    if (!same) {
        hipGraph_t g = nullptr;
        hipStream_t cs = b->captureStream;
        if (hipFail(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal),
                    "hipStreamBeginCapture"))
            return GFC_ERR_HIP;
        int st = enqueue(b, At, cs, pl);
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
        b->lastN = pl.key;
        b->valid = true;
    }
    if (hipFail(hipGraphLaunch(b->exec, s), "hipGraphLaunch")) return GFC_ERR_HIP;
    return GFC_OK;
}

// Device-level evaluation (derivative, Jacobian, LU solve) with the
// functions of one kernel family: shape 0 = the family the batched solve
// uses for this evaluator, 1 = wavefront per cell, 4 = 16-lane teams,
// 512 = workgroup per cell.
int device_eval(gfc_evaluator* ev, int shape, EvalArgs E, hipStream_t s) {
    auto* b = static_cast<BatchedState*>(ev->batched);
    const int n = ev->nS + 1;
    if (ev->dmTeam.extended) {
        setLastError("device evaluation: the batched kernels evaluate gpuFoam's four rate "
                     "types only (this mechanism runs one lane per cell)");
        return GFC_ERR_UNSUPPORTED;
    }
    if (shape == 0) {
        Plan pl;
        int perBlock = 1;
        choose_plan(ev, ev->maxCells, pl, perBlock);
        shape = pl.blk ? 512 : perBlock > 1 ? 4 : pl.threads == 128 ? 2 : pl.team ? 1 : 64;
    }
    E.m = ev->dmTeam;
    if (E.nCells == 0) return GFC_OK;
    if (shape == 1) {
        int N = 0;
        if (!team_kernel_for(n, N) || ev->dmTeam.teamWidth > kTeamWidth ||
            ev->dmTeam.nSets > kTeamMaxSets) {
            setLastError("device evaluation: no wavefront kernel for this mechanism "
                         "(nSpecie+1 <= 64, at most 3 species per reaction side, at most "
                         "128 distinct third-body efficiency sets)");
            return GFC_ERR_UNSUPPORTED;
        }
        void (*k)(EvalArgs) = N == 11 ? team_eval_kernel<11> : N == 32 ? team_eval_kernel<32>
                            : N == 54 ? team_eval_kernel<54> : team_eval_kernel<64>;
        hipLaunchKernelGGL(k, dim3(E.nCells), dim3(64), 0, s, E);
    } else if (shape == 2) {
        int N = 0;
        if (!pair_kernel_for(ev->dmTeam, N)) {
            setLastError("device evaluation: no two-wavefront kernel for this mechanism "
                         "(32 < nSpecie+1 <= 64, at most 16 efficiency sets, 3 species per "
                         "side)");
            return GFC_ERR_UNSUPPORTED;
        }
        void (*k)(EvalArgs) = N == 54 ? pair_eval_kernel<54> : pair_eval_kernel<64>;
        hipLaunchKernelGGL(k, dim3(E.nCells), dim3(2 * kLanesPerWave), 0, s, E);
    } else if (shape == 4) {
        int N = 0;
        if (!team16_kernel_for(ev->dmTeam, N)) {
            setLastError("device evaluation: no 16-lane kernel for this mechanism "
                         "(nSpecie+1 <= 16, at most 16 efficiency sets, 3 species per side)");
            return GFC_ERR_UNSUPPORTED;
        }
        void (*k)(EvalArgs) = N == 11 ? team16_eval_kernel<11> : team16_eval_kernel<16>;
        hipLaunchKernelGGL(k, dim3((E.nCells + 3) / 4), dim3(64), 0, s, E);
    } else if (shape == 512) {
        if (b->blkNA == 0) {
            setLastError("device evaluation: no workgroup kernel for this mechanism "
                         "(64 < nSpecie+1 <= 208)");
            return GFC_ERR_UNSUPPORTED;
        }
        void (*k)(EvalArgs, double*, long) =
            b->blkNA == 8 ? block_eval_kernel<8, 4> : block_eval_kernel<13, 7>;
        const int grid = b->blkGrid < E.nCells ? b->blkGrid : E.nCells;
        hipLaunchKernelGGL(k, dim3(grid), dim3(kBlockThreads), 0, s, E, b->blkScratch,
                           b->blkPerWG);
    } else {
        setLastError("device evaluation: shape must be 0, 1, 2, 4 or 512");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    return hipFail(hipGetLastError(), "device evaluation launch") ? GFC_ERR_HIP : GFC_OK;
}

} // namespace gfc

extern "C" {
int gfc_eval_derivatives(gfc_evaluator* ev, int shape, int nCells, int ldY, const double* Y,
                         double* dYdt, void* stream) {
    if (!ev || nCells < 0 || ldY < nCells || (nCells > 0 && (!Y || !dYdt))) {
        gfc::setLastError("gfc_eval_derivatives: bad argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    gfc::EvalArgs E{};
    E.mode = 0;
    E.nCells = nCells;
    E.ldY = ldY;
    E.Y = Y;
    E.dYdt = dYdt;
    return gfc::device_eval(ev, shape, E, static_cast<hipStream_t>(stream));
}
int gfc_eval_jacobian(gfc_evaluator* ev, int shape, int nCells, int ldY, const double* Y,
                      double* dYdt, double* J, void* stream) {
    if (!ev || nCells < 0 || ldY < nCells || (nCells > 0 && (!Y || !dYdt || !J))) {
        gfc::setLastError("gfc_eval_jacobian: bad argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    gfc::EvalArgs E{};
    E.mode = 1;
    E.nCells = nCells;
    E.ldY = ldY;
    E.Y = Y;
    E.dYdt = dYdt;
    E.J = J;
    return gfc::device_eval(ev, shape, E, static_cast<hipStream_t>(stream));
}
int gfc_eval_lu_solve(gfc_evaluator* ev, int shape, int nCells, int ldB, const double* M,
                      const double* b, double* x, void* stream) {
    if (!ev || nCells < 0 || ldB < nCells || (nCells > 0 && (!M || !b || !x))) {
        gfc::setLastError("gfc_eval_lu_solve: bad argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    gfc::EvalArgs E{};
    E.mode = 2;
    E.nCells = nCells;
    E.ldY = ldB;
    E.M = M;
    E.b = b;
    E.x = x;
    E.lu = ev->factorization;
    return gfc::device_eval(ev, shape, E, static_cast<hipStream_t>(stream));
}

int gfc_evaluator_set_batched_shape(gfc_evaluator* ev, int cellsPerWave,
                                    int wavesPerBlock, int chunk) {
    if (!ev || (cellsPerWave != 0 && cellsPerWave != 1 && cellsPerWave != 2 &&
                cellsPerWave != 4 && cellsPerWave != 64) ||
        wavesPerBlock < 0 || chunk < 0 || chunk > 1) {
        gfc::setLastError("gfc_evaluator_set_batched_shape: unsupported shape "
                          "(cellsPerWave must be 0, 1, 2 (two wavefronts per cell), 4 or "
                          "64, chunk 0 (automatic) or 1 (one cell per work-queue ticket))");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    if (cellsPerWave == 2) {
        int N = 0;
        if (!gfc::pair_kernel_for(ev->dmTeam, N)) {
            gfc::setLastError("gfc_evaluator_set_batched_shape: two wavefronts per cell "
                              "needs 32 < nSpecie+1 <= 64, at most 16 third-body "
                              "efficiency sets and at most 3 species per reaction side");
            return GFC_ERR_INVALID_ARGUMENT;
        }
    }
    if (cellsPerWave == 4) {
        int N = 0;
        if (!gfc::team16_kernel_for(ev->dmTeam, N) || ev->dmTeam.extended) {
            gfc::setLastError("gfc_evaluator_set_batched_shape: 4 cells per wavefront "
                              "needs nSpecie+1 <= 16, at most 16 third-body efficiency "
                              "sets and at most 3 species per reaction side");
            return GFC_ERR_INVALID_ARGUMENT;
        }
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
int gfc_evaluator_set_factorization(gfc_evaluator* ev, int variant) {
    if (!ev || variant < GFC_LU_COLUMN || variant > GFC_LU_PANEL_MFMA) {
        gfc::setLastError("gfc_evaluator_set_factorization: variant must be GFC_LU_COLUMN, "
                          "GFC_LU_PANEL or GFC_LU_PANEL_MFMA");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    ev->factorization = variant;
    return GFC_OK;
}
int gfc_evaluator_get_factorization(const gfc_evaluator* ev, int* variant) {
    if (!ev || !variant) return GFC_ERR_INVALID_ARGUMENT;
    *variant = ev->factorization;
    return GFC_OK;
}
}
