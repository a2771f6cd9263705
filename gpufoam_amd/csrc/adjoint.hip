// Discrete adjoint of the per-cell chemistry solve (no counterpart in
// gpuFoam).
//
//  1. adjoint_forward_kernel: the plain per-cell solve (one lane per cell,
//     singleCellSolver.H:27-51) that also records, for every accepted step,
//     its step size, and every K-th accepted step its initial state
//     (checkpoints), plus the final state before the non-negativity clamp.
//  2. adjoint_reverse_kernel: walks the accepted steps backwards.  For each
//     segment of K steps it re-integrates the segment from its checkpoint
//     with the recorded step sizes, then applies the reverse of each step
//     (adjoint_math.hpp), which recomputes the step's Jacobian, LU and
//     stages.  Gradients are those of the discrete map with the accepted
//     step sizes held fixed.
//  3. adjoint_replay_kernel: re-integrates the recorded step sequence from
//     a (perturbed) initial state with the current rate parameters: the
//     map the adjoint differentiates, for finite-difference checks.
//
// Memory: step sizes maxSteps x cells, checkpoints ceil(maxSteps/K) x n x
// cells, per-cell scratch (forward + adjoint arrays + a K-state segment
// buffer).  K is chosen automatically as the smallest interval whose whole
// workspace fits in 80 % of the free device memory, or set by the caller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "adjoint_math.hpp"
#include "evaluator.hpp"
#include "mechanism.hpp"
#include "team16_common.hpp"

namespace gfc {

// Four cells per wavefront (team16_kernel.hip, team16_adjoint.hip): used
// when the batched solve would use that shape (nSpecie+1 <= 16).
template <int N>
__global__ void team16_record_kernel(SolveArgs A, int* queue, t16::T16Record R);
template <int N>
__global__ void team16_adjoint_reverse_kernel(SolveArgs A, t16::T16Record R, double* seg,
                                              long segPer, const double* Ybar, long ldb,
                                              double* Y0bar, double* thBar, double* pacc,
                                              long paccSlot, long paccPer, int* queue);
// One wavefront per cell (team_kernel.hip, wave_adjoint.hip): used for
// 16 < nSpecie+1 <= 64 unless one lane per cell is asked for.
template <int N>
__global__ void team_record_kernel(SolveArgs A, int* queue, t16::T16Record R);
template <int N>
__global__ void wave_adjoint_reverse_kernel(SolveArgs A, t16::T16Record R, double* seg,
                                            long segPer, const double* Ybar, long ldb,
                                            double* Y0bar, double* thBar, double* pacc,
                                            long paccSlot, long paccPer, int* queue);

namespace {

__global__ void adj_reset_queue_kernel(int* q) { *q = 0; }

bool hipFail(hipError_t e, const char* what) {
    if (e == hipSuccess) return false;
    setLastError(std::string(what) + ": " + hipGetErrorString(e));
    return true;
}

struct AdjState {
    long cap = 0;   // cells of workspace (multiple of 64)
    long maxSteps = 0;
    int K = 0;      // checkpoint interval in use
    int Kreq = 0;   // requested interval (0: automatic)
    long laneD = 0, adjD = 0;
    double* scratch = nullptr;
    double* ck = nullptr;
    double* hs = nullptr;
    double* y0 = nullptr;
    double* yEnd = nullptr;
    int* nAcc = nullptr;
    int* queue = nullptr; // work queue of the four-cells-per-wavefront kernels
    int nCells = 0;
    double deltaT = 0;
    bool valid = false;
    int teamN = 0;        // > 0: recorded by the four-cells-per-wavefront kernel
    int waveN = 0;        // > 0: recorded by the wavefront-per-cell kernel
    int numCUs = 0;
    void release() {
        void* ptrs[] = {scratch, ck, hs, y0, yEnd, nAcc, queue};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        scratch = ck = hs = y0 = yEnd = nullptr;
        nAcc = queue = nullptr;
        cap = 0;
        valid = false;
    }
};

AdjState* state(gfc_evaluator* ev) {
    if (!ev->adjoint) ev->adjoint = new AdjState();
    return static_cast<AdjState*>(ev->adjoint);
}

} // namespace

// Sizes of the per-cell adjoint workspace.  Every per-cell array is
// wave-blocked (lane::wave_slot): the 64 cells of a wavefront share one
// contiguous block, element e of a cell at [e*64] from its slot.
struct AdjLayout {
    long laneD, adjD;   // forward and adjoint scratch doubles per cell
    long perScratch;    // laneD + adjD + K*n
    long nck;           // checkpoint slots per cell
    long maxSteps;      // recorded step sizes per cell
    int K;              // checkpoint interval
};

// Forward solve with recording (one lane per cell).
__global__ void __launch_bounds__(64)
adjoint_forward_kernel(SolveArgs A, AdjLayout W, double* scratch, double* ck, double* hs,
                       int* nAcc, double* yEnd, double* y0copy) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= A.nCells) return;
    const int nS = A.m.nS, n = nS + 1;
    constexpr long s = lane::kSlotStride;
    const lane::LaneBuf b{lane::wave_slot(scratch, cell, W.perScratch), s, n};
    double* y0c = lane::wave_slot(y0copy, cell, n + 1);
    const long ld = A.ldY;
    for (int k = 0; k < nS + 2; ++k) y0c[k * s] = A.Y0[k * ld + cell];
    for (int k = 0; k < n; ++k) b.y()[k] = A.Y0[k * ld + cell];
    const double p = A.Y0[(long)(nS + 1) * ld + cell];
    double dtChem = A.dtChem[cell];
    int attempts = 0, accepted = 0;
    adj::Recorder rec{lane::wave_slot(ck, cell, W.nck * n), lane::wave_slot(hs, cell, W.maxSteps),
                      s, n, W.K, 0};
    const int status = adj::cell_forward(A, p, b, dtChem, rec,
                                         lane::SV{lane::wave_slot(yEnd, cell, n), s},
                                         attempts, accepted);
    nAcc[cell] = rec.count;
    if (A.Ynew) {
        for (int k = 0; k < n; ++k) A.Ynew[k * ld + cell] = b.y()[k];
        A.Ynew[(long)(nS + 1) * ld + cell] = p;
    }
    A.dtChem[cell] = fmin(dtChem, A.deltaTChemMax);
    if (A.stats) {
        A.stats[cell] = status;
        A.stats[ld + cell] = attempts;
        A.stats[2 * ld + cell] = accepted;
    }
}

// Reverse sweep (one lane per cell).  Ybar/Y0bar: [n][ldb]; th: [6 nR][ldb].
__global__ void __launch_bounds__(64)
adjoint_reverse_kernel(SolveArgs A, AdjLayout W, double* scratch, double* ck, double* hs,
                       const int* nAcc, double* yEnd, double* y0copy, const double* Ybar,
                       long ldb, double* Y0bar, double* thBar) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= A.nCells) return;
    const int nS = A.m.nS, n = nS + 1;
    constexpr long s = lane::kSlotStride;
    double* slot = lane::wave_slot(scratch, cell, W.perScratch);
    lane::LaneBuf b{slot, s, n};
    adj::AdjBuf a{b, slot + W.laneD * s, s, n};
    a.b.raw = a.mat(1);
    a.b.jt = a.mat(2);
    a.b.f2 = a.f2().p;
    const double* y0c = lane::wave_slot(y0copy, cell, n + 1);
    const double p = y0c[(long)(nS + 1) * s];
    const double T0 = y0c[(long)nS * s];
    lane::SV yb = a.ybar();
    for (int e = 0; e < n; ++e) yb[e] = Ybar[e * ldb + cell];
    const adj::ThView th{thBar ? thBar + cell : nullptr, ldb};
    if (th.p)
        for (long q = 0; q < (long)adj::kParams * A.m.nR; ++q) th.p[q * ldb] = 0.0;
    adj::cell_reverse(A, p, T0, a, slot + (W.laneD + W.adjD) * s,
                      lane::wave_slot(ck, cell, W.nck * n), lane::wave_slot(hs, cell, W.maxSteps),
                      s, W.K, nAcc[cell], lane::SV{lane::wave_slot(yEnd, cell, n), s}, th);
    for (int e = 0; e < n; ++e) Y0bar[e * ldb + cell] = yb[e];
}

// Re-integrates the recorded accepted steps from Y0 (one lane per cell).
__global__ void __launch_bounds__(64)
adjoint_replay_kernel(SolveArgs A, AdjLayout W, double* scratch, double* hs, const int* nAcc,
                      double* y0copy) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= A.nCells) return;
    const int nS = A.m.nS, n = nS + 1;
    constexpr long s = lane::kSlotStride;
    const lane::LaneBuf b{lane::wave_slot(scratch, cell, W.perScratch), s, n};
    const double* hc = lane::wave_slot(hs, cell, W.maxSteps);
    const long ld = A.ldY;
    for (int k = 0; k < n; ++k) b.y()[k] = A.Y0[k * ld + cell];
    const double p = A.Y0[(long)(nS + 1) * ld + cell];
    const double T0 = lane::wave_slot(y0copy, cell, n + 1)[(long)nS * s];
    if (T0 > A.Treact && A.deltaT > kSmall) {
        const int N = nAcc[cell];
        for (int t = 0; t < N; ++t) {
            lane::derivatives(A.m, p, b.y(), b.dydx0(), b);
            (void)lane::ros_step(A, p, b.y(), hc[(long)t * s], b.yTemp(), b);
            for (int k = 0; k < n; ++k) b.y()[k] = b.yTemp()[k];
        }
        for (int i = 0; i < nS; i++) b.y()[i] = fmax(0.0, b.y()[i]);
    }
    for (int k = 0; k < n; ++k) A.Ynew[k * ld + cell] = b.y()[k];
    A.Ynew[(long)(nS + 1) * ld + cell] = p;
}

void adjoint_free(gfc_evaluator* ev) {
    auto* st = static_cast<AdjState*>(ev->adjoint);
    if (!st) return;
    st->release();
    delete st;
    ev->adjoint = nullptr;
}

static AdjLayout layout(gfc_evaluator* ev) {
    AdjState* st = state(ev);
    const long n = ev->nS + 1;
    AdjLayout W;
    W.laneD = st->laneD;
    W.adjD = st->adjD;
    W.K = st->K;
    W.perScratch = st->laneD + st->adjD + (long)st->K * n;
    W.nck = (st->maxSteps + st->K - 1) / st->K;
    W.maxSteps = st->maxSteps;
    return W;
}

// Bytes of the adjoint workspace for `cap` cells at checkpoint interval K:
// per cell the forward and adjoint scratch plus the K-state segment buffer,
// ceil(maxSteps/K) checkpoints of n, maxSteps step sizes, the initial state
// (n+1), the final state (n) and the accepted-step count.
double adjoint_workspace_bytes(long cap, long n, long maxSteps, long K, long laneD,
                               long adjD) {
    const long nck = (maxSteps + K - 1) / K;
    const double perCell = 8.0 * (double)(laneD + adjD + K * n + nck * n + maxSteps +
                                          (n + 1) + n) + 4.0;
    return perCell * (double)cap + sizeof(int);
}

// Automatic checkpoint interval: the smallest K whose whole workspace fits
// in `budget` bytes.  The total falls with K while the checkpoints dominate
// and rises again once the segment buffer does; its minimum lies at
// K = sqrt(maxSteps), so the search stops at ceil(sqrt(maxSteps)) and
// returns that K when nothing fits (the allocation then reports the
// shortage).
long adjoint_plan_interval(double budget, long cap, long n, long maxSteps, long laneD,
                           long adjD) {
    maxSteps = std::max(1L, maxSteps);
    const long kMax = std::min(maxSteps, (long)std::ceil(std::sqrt((double)maxSteps)));
    for (long K = 1; K <= kMax; ++K)
        if (adjoint_workspace_bytes(cap, n, maxSteps, K, laneD, adjD) <= budget) return K;
    return kMax;
}

// Sizes (and allocates) the workspace for nCells cells.
static int adjoint_reserve(gfc_evaluator* ev, int nCells) {
    AdjState* st = state(ev);
    const long n = ev->nS + 1;
    const long cap = ((long)nCells + 63) / 64 * 64;
    const long maxSteps = std::max(1, ev->ode.maxSteps);
    int K = st->Kreq;
    const long laneD = lane::lane_scratch_doubles(ev->nS);
    const long adjD = adj::adj_scratch_doubles(ev->nS);
    if (K <= 0) {
        size_t freeB = 0, totalB = 0;
        if (hipFail(hipMemGetInfo(&freeB, &totalB), "hipMemGetInfo")) return GFC_ERR_HIP;
        if (st->scratch) freeB += adjoint_workspace_bytes(st->cap, n, st->maxSteps, st->K,
                                                          st->laneD, st->adjD);
        K = (int)adjoint_plan_interval(0.8 * (double)freeB, cap, n, maxSteps, laneD, adjD);
    }
    if (st->cap >= cap && st->K == K && st->scratch) return GFC_OK;
    st->release();
    st->laneD = laneD;
    st->adjD = adjD;
    st->K = K;
    st->cap = cap;
    st->maxSteps = maxSteps;
    const long nck = (maxSteps + K - 1) / K;
    const size_t scratchB = (size_t)8 * (st->laneD + st->adjD + (long)K * n) * cap;
    if (hipFail(hipMalloc(&st->scratch, scratchB), "hipMalloc(adjoint scratch)") ||
        hipFail(hipMalloc(&st->ck, (size_t)8 * nck * n * cap), "hipMalloc(checkpoints)") ||
        hipFail(hipMalloc(&st->hs, (size_t)8 * maxSteps * cap), "hipMalloc(step sizes)") ||
        hipFail(hipMalloc(&st->y0, (size_t)8 * (n + 1) * cap), "hipMalloc(adjoint y0)") ||
        hipFail(hipMalloc(&st->yEnd, (size_t)8 * n * cap), "hipMalloc(adjoint yEnd)") ||
        hipFail(hipMalloc(&st->nAcc, (size_t)4 * cap), "hipMalloc(adjoint counts)") ||
        hipFail(hipMalloc(&st->queue, sizeof(int)), "hipMalloc(adjoint queue)")) {
        st->release();
        return GFC_ERR_HIP;
    }
    return GFC_OK;
}

static t16::T16Record record_of(const AdjState* st, const AdjLayout& W) {
    return t16::T16Record{st->ck, st->hs, st->nAcc, st->yEnd, st->y0,
                          W.nck, W.maxSteps, W.K};
}

// The wavefront-per-cell form (row capacity N) when the evaluator's shape
// leaves the choice to the mechanism (automatic, one or two wavefronts per
// cell) and 16 < nSpecie+1 <= 64; 0 otherwise.
static int wave_adjoint_rows(const gfc_evaluator* ev) {
    if (ev->cellsPerWave != 0 && ev->cellsPerWave != 1 && ev->cellsPerWave != 2) return 0;
    if (ev->nS + 1 <= 16) return 0;
    // the reverse keeps a cell's parameter adjoints in the cell's share of
    // the one-lane scratch, which it does not otherwise use
    if ((long)adj::kParams * 64 * ev->dmTeam.nPass >
        lane::lane_scratch_doubles(ev->nS) + adj::adj_scratch_doubles(ev->nS))
        return 0;
    return wave_kernel_rows(ev);
}

// Persistent grid of a kernel of one wavefront per `perWave` cells: the
// wavefronts that stay resident, at most one per perWave cells.
template <class K>
static int resident_grid(AdjState* st, int device, K kernel, int nCells, int perWave) {
    if (st->numCUs == 0) {
        if (hipDeviceGetAttribute(&st->numCUs, hipDeviceAttributeMultiprocessorCount, device) !=
                hipSuccess ||
            st->numCUs <= 0)
            st->numCUs = 1;
    }
    int perCU = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, (const void*)kernel, 64, 0) !=
            hipSuccess ||
        perCU <= 0)
        perCU = 1;
    const long waves = ((long)nCells + perWave - 1) / perWave;
    return (int)std::max(1L, std::min<long>((long)st->numCUs * perCU, waves));
}

// Persistent grid of a four-cells-per-wavefront kernel: the wavefronts that
// stay resident, at most one per four cells.
template <class K>
static int team16_grid(AdjState* st, int device, K kernel, int nCells) {
    if (st->numCUs == 0) {
        if (hipDeviceGetAttribute(&st->numCUs, hipDeviceAttributeMultiprocessorCount, device) !=
                hipSuccess ||
            st->numCUs <= 0)
            st->numCUs = 1;
    }
    int perCU = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, (const void*)kernel, 64, 0) !=
            hipSuccess ||
        perCU <= 0)
        perCU = 1;
    const long waves = ((long)nCells + 3) / 4;
    return (int)std::max(1L, std::min<long>((long)st->numCUs * perCU, waves));
}

} // namespace gfc

using namespace gfc;

extern "C" {

int gfc_evaluator_param_count(const gfc_evaluator* ev) {
    return ev ? adj::kParams * ev->nR : 0;
}

int gfc_evaluator_get_rate_parameters(const gfc_evaluator* ev, double* params) {
    if (!ev || !params) {
        setLastError("gfc_evaluator_get_rate_parameters: bad argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    const gfc_mechanism& m = ev->mech;
    for (int r = 0; r < m.nR; ++r) {
        double* q = params + adj::kParams * r;
        q[0] = std::log(std::fabs(m.k[3 * r]));
        q[1] = m.k[3 * r + 1];
        q[2] = m.k[3 * r + 2];
        const bool fo = m.type[r] == GFC_RATE_LINDEMANN || m.type[r] == GFC_RATE_TROE;
        q[3] = fo ? std::log(std::fabs(m.kInf[3 * r])) : 0.0;
        q[4] = fo ? m.kInf[3 * r + 1] : 0.0;
        q[5] = fo ? m.kInf[3 * r + 2] : 0.0;
    }
    return GFC_OK;
}

int gfc_evaluator_set_rate_parameters(gfc_evaluator* ev, const double* params) {
    if (!ev || !params) {
        setLastError("gfc_evaluator_set_rate_parameters: bad argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    gfc_mechanism& m = ev->mech;
    for (int r = 0; r < m.nR; ++r) {
        const double* q = params + adj::kParams * r;
        const double s0 = m.k[3 * r] < 0 ? -1.0 : 1.0;
        m.k[3 * r] = s0 * std::exp(q[0]);
        m.k[3 * r + 1] = q[1];
        m.k[3 * r + 2] = q[2];
        if (m.type[r] == GFC_RATE_LINDEMANN || m.type[r] == GFC_RATE_TROE) {
            const double s1 = m.kInf[3 * r] < 0 ? -1.0 : 1.0;
            m.kInf[3 * r] = s1 * std::exp(q[3]);
            m.kInf[3 * r + 1] = q[4];
            m.kInf[3 * r + 2] = q[5];
        }
    }
    if (hipFail(hipDeviceSynchronize(), "hipDeviceSynchronize")) return GFC_ERR_HIP;
    return evaluator_upload(ev);
}

int gfc_adjoint_set_checkpoint_interval(gfc_evaluator* ev, int K) {
    if (!ev || K < 0) {
        setLastError("gfc_adjoint_set_checkpoint_interval: bad argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    state(ev)->Kreq = K;
    return GFC_OK;
}

int gfc_adjoint_checkpoint_interval(const gfc_evaluator* ev) {
    if (!ev || !ev->adjoint) return 0;
    return static_cast<const AdjState*>(ev->adjoint)->K;
}

double gfc_adjoint_workspace_bytes(int nCells, int nSpecie, int maxSteps, int K) {
    if (nCells < 0 || nSpecie <= 0 || K <= 0) return -1.0;
    const long cap = ((long)nCells + 63) / 64 * 64;
    return adjoint_workspace_bytes(cap, nSpecie + 1, std::max(1, maxSteps), K,
                                   lane::lane_scratch_doubles(nSpecie),
                                   adj::adj_scratch_doubles(nSpecie));
}

int gfc_adjoint_plan_interval(double budgetBytes, int nCells, int nSpecie, int maxSteps) {
    if (nCells < 0 || nSpecie <= 0) return 0;
    const long cap = ((long)nCells + 63) / 64 * 64;
    return (int)adjoint_plan_interval(budgetBytes, cap, nSpecie + 1, maxSteps,
                                      lane::lane_scratch_doubles(nSpecie),
                                      adj::adj_scratch_doubles(nSpecie));
}

int gfc_adjoint_forward(gfc_evaluator* ev, double deltaT, int nCells, int ldY,
                        const double* Y0, double* deltaTChem, double* Ynew,
                        int* cellStats, void* stream) {
    if (!ev || nCells < 0 || ldY < nCells || !(deltaT > 0) ||
        (nCells > 0 && (!Y0 || !deltaTChem))) {
        setLastError("gfc_adjoint_forward: bad argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    if (nCells > ev->maxCells) {
        setLastError("gfc_adjoint_forward: nCells exceeds the evaluator's maxCells");
        return GFC_ERR_CAPACITY;
    }
    if (ev->dm.extended) {
        setLastError("gfc_adjoint_forward: the discrete adjoint covers gpuFoam's rate types "
                     "(Arrhenius, third body, Lindemann and Troe fall-off) only");
        return GFC_ERR_UNSUPPORTED;
    }
    AdjState* st = state(ev);
    st->valid = false;
    if (nCells == 0) {
        st->nCells = 0;
        st->deltaT = deltaT;
        st->valid = true;
        return GFC_OK;
    }
    int rc = adjoint_reserve(ev, nCells);
    if (rc != GFC_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    SolveArgs A = solve_args(ev, deltaT, 1e300, nCells);
    A.ldY = ldY;
    A.Y0 = Y0;
    A.Ynew = Ynew;
    A.dtChem = deltaTChem;
    A.stats = cellStats;
    int N = 0;
    // the reverse keeps a cell's parameter adjoints in the cell's share of
    // the one-lane scratch, which the team form does not otherwise use
    const bool teamAccFits = (long)adj::kParams * 16 * ((ev->nR + 15) / 16) <=
                             lane::lane_scratch_doubles(ev->nS) + adj::adj_scratch_doubles(ev->nS);
    st->teamN = (teamAccFits && team16_selected(ev, nCells, N)) ? N : 0;
    st->waveN = st->teamN ? 0 : wave_adjoint_rows(ev);
    if (st->waveN) {
        SolveArgs At = A;
        At.m = ev->dmTeam;
        const t16::T16Record R = record_of(st, layout(ev));
        auto k = st->waveN == 32 ? team_record_kernel<32>
               : st->waveN == 54 ? team_record_kernel<54> : team_record_kernel<64>;
        const int grid = resident_grid(st, ev->device, k, nCells, 1);
        queue_chunks_for(ev, At, grid); // the batched solve's tickets
        hipLaunchKernelGGL(adj_reset_queue_kernel, dim3(1), dim3(1), 0, s, st->queue);
        hipLaunchKernelGGL(k, dim3(grid), dim3(64), 0, s, At, st->queue, R);
    } else if (st->teamN) {
        SolveArgs At = A;
        At.m = ev->dmTeam;
        const t16::T16Record R = record_of(st, layout(ev));
        auto k = (N == 11) ? team16_record_kernel<11> : team16_record_kernel<16>;
        hipLaunchKernelGGL(adj_reset_queue_kernel, dim3(1), dim3(1), 0, s, st->queue);
        hipLaunchKernelGGL(k, dim3(team16_grid(st, ev->device, k, nCells)), dim3(64), 0, s, At, st->queue,
                           R);
    } else {
        const int blocks = (nCells + 63) / 64;
        hipLaunchKernelGGL(adjoint_forward_kernel, dim3(blocks), dim3(64), 0, s, A, layout(ev),
                           st->scratch, st->ck, st->hs, st->nAcc, st->yEnd, st->y0);
    }
    if (hipFail(hipGetLastError(), "adjoint forward launch")) return GFC_ERR_HIP;
    st->nCells = nCells;
    st->deltaT = deltaT;
    st->valid = true;
    return bounds_report(ev, s);
}

int gfc_adjoint_backward(gfc_evaluator* ev, const double* Ybar, int ldb, double* Y0bar,
                         double* paramBar, void* stream) {
    if (!ev || !ev->adjoint || !static_cast<AdjState*>(ev->adjoint)->valid) {
        setLastError("gfc_adjoint_backward: no recorded forward solve");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    AdjState* st = state(ev);
    if (st->nCells == 0) return GFC_OK;
    if (!Ybar || !Y0bar || ldb < st->nCells) {
        setLastError("gfc_adjoint_backward: bad argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    SolveArgs A = solve_args(ev, st->deltaT, 1e300, st->nCells);
    const AdjLayout W = layout(ev);
    if (st->waveN) {
        SolveArgs At = A;
        At.m = ev->dmTeam;
        auto k = st->waveN == 32 ? wave_adjoint_reverse_kernel<32>
               : st->waveN == 54 ? wave_adjoint_reverse_kernel<54>
                                 : wave_adjoint_reverse_kernel<64>;
        double* seg = st->scratch + (W.laneD + W.adjD) * lane::kSlotStride;
        hipLaunchKernelGGL(adj_reset_queue_kernel, dim3(1), dim3(1), 0, s, st->queue);
        hipLaunchKernelGGL(k, dim3(resident_grid(st, ev->device, k, st->nCells, 1)), dim3(64), 0,
                           s, At, record_of(st, W), seg, W.perScratch, Ybar, (long)ldb, Y0bar,
                           paramBar, st->scratch, W.perScratch * lane::kSlotStride,
                           W.laneD + W.adjD, st->queue);
    } else if (st->teamN) {
        SolveArgs At = A;
        At.m = ev->dmTeam;
        auto k = (st->teamN == 11) ? team16_adjoint_reverse_kernel<11>
                                   : team16_adjoint_reverse_kernel<16>;
        // the segment buffers sit after the forward and adjoint scratch
        double* seg = st->scratch + (W.laneD + W.adjD) * lane::kSlotStride;
        hipLaunchKernelGGL(adj_reset_queue_kernel, dim3(1), dim3(1), 0, s, st->queue);
        hipLaunchKernelGGL(k, dim3(team16_grid(st, ev->device, k, st->nCells)), dim3(64), 0, s, At,
                           record_of(st, W), seg, W.perScratch, Ybar, (long)ldb, Y0bar,
                           paramBar, st->scratch, W.perScratch * lane::kSlotStride,
                           W.laneD + W.adjD, st->queue);
    } else {
        const int blocks = (st->nCells + 63) / 64;
        hipLaunchKernelGGL(adjoint_reverse_kernel, dim3(blocks), dim3(64), 0, s, A, W,
                           st->scratch, st->ck, st->hs, st->nAcc, st->yEnd, st->y0, Ybar,
                           (long)ldb, Y0bar, paramBar);
    }
    if (hipFail(hipGetLastError(), "adjoint reverse launch")) return GFC_ERR_HIP;
    return bounds_report(ev, s);
}

int gfc_adjoint_replay(gfc_evaluator* ev, const double* Y0, int ldY, double* Ynew,
                       void* stream) {
    if (!ev || !ev->adjoint || !static_cast<AdjState*>(ev->adjoint)->valid) {
        setLastError("gfc_adjoint_replay: no recorded forward solve");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    AdjState* st = state(ev);
    if (st->nCells == 0) return GFC_OK;
    if (!Y0 || !Ynew || ldY < st->nCells) {
        setLastError("gfc_adjoint_replay: bad argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    SolveArgs A = solve_args(ev, st->deltaT, 1e300, st->nCells);
    A.ldY = ldY;
    A.Y0 = Y0;
    A.Ynew = Ynew;
    const int blocks = (st->nCells + 63) / 64;
    hipLaunchKernelGGL(adjoint_replay_kernel, dim3(blocks), dim3(64), 0, s, A, layout(ev),
                       st->scratch, st->hs, st->nAcc, st->y0);
    if (hipFail(hipGetLastError(), "adjoint replay launch")) return GFC_ERR_HIP;
    return bounds_report(ev, s);
}

} // extern "C"
