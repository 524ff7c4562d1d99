// engine_inst.hip — instantiation of Engine<L> and the standalone GEMM for one limb count.
// Compiled once per L of SDPG_FOR_EACH_L with -DSDPG_INST_L=<L>.
#include "engine_impl.h"

#ifndef SDPG_INST_L
#error "compile with -DSDPG_INST_L=<limb count>"
#endif

#define SDPG_CAT2_(a, b) a##b
#define SDPG_CAT_(a, b) SDPG_CAT2_(a, b)

namespace sdpg {

EngineBase *SDPG_CAT_(new_engine_, SDPG_INST_L)() { return new Engine<SDPG_INST_L>(); }

int SDPG_CAT_(init_engine_, SDPG_INST_L)(EngineBase *e) { return static_cast<Engine<SDPG_INST_L> *>(e)->init(); }

int SDPG_CAT_(rgemm_, SDPG_INST_L)(int device, int transa, int transb, int m, int n, int k, const uint32_t *A,
                                   int lda, const uint32_t *B, int ldb, uint32_t *C, int ldc, int op, int mode,
                                   int slices, int lower, double *ms, uint8_t *scratch, size_t scratch_bytes) {
    return rgemm_L<SDPG_INST_L>(device, transa, transb, m, n, k, A, lda, B, ldb, C, ldc, op, mode, slices, lower,
                                ms, scratch, scratch_bytes);
}

int SDPG_CAT_(tune_time_, SDPG_INST_L)(int device, int S, const uint32_t *host, double *ms_long, double *ms_short) {
    return tune_time_L<SDPG_INST_L>(device, S, host, ms_long, ms_short);
}

int SDPG_CAT_(batch_create_, SDPG_INST_L)(EngineBase *const *es, int n, BatchBase **out) {
    return batch_create_L<SDPG_INST_L>(es, n, out);
}

}  // namespace sdpg
