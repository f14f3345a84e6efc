// kNN candidate tiles with lists of 32 or 64 per lane half (k + 1 in 17..64), FP16 GEMM.
#include "knn_tiles.h"

namespace ncvis_hip {

void knn_tiles_q1_fp16(int KL, int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn, float beta,
                        int32_t *cand, hipStream_t st, int split) {
    launch_tiles_q1<PREC_FP16, 0>(KL, DK, N, Npad, Xh, cn, beta, cand, st, split);
}

}  // namespace ncvis_hip
