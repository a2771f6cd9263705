// Persistent work-stealing batched launcher (filled in below).
#include <hip/hip_runtime.h>

#include "evaluator.hpp"
#include "mechanism.hpp"

namespace gfc {

int batched_init(gfc_evaluator* ev) {
    (void)ev;
    return GFC_OK;
}

void batched_free(gfc_evaluator* ev) { (void)ev; }

int batched_launch(gfc_evaluator* ev, const SolveArgs& A, hipStream_t s,
                   bool graph) {
    (void)ev; (void)A; (void)s; (void)graph;
    setLastError("batched launch not built");
    return GFC_ERR_NOT_BUILT;
}

} // namespace gfc

extern "C" {
int gfc_evaluator_set_batched_shape(gfc_evaluator* ev, int cellsPerWave,
                                    int wavesPerBlock, int chunk) {
    if (!ev) return GFC_ERR_INVALID_ARGUMENT;
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
