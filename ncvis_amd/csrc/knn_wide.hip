// kNN candidate tiles for rows of 257..512 elements in the 16-bit GEMMs.
#include "knn_tiles.h"

namespace ncvis_hip {

void knn_tiles_wide(int prec, int KL, int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn,
                    float beta, int32_t *cand, hipStream_t st, int split) {
    if (prec == PREC_BF16)
        launch_tiles_wide<PREC_BF16>(KL, DK, N, Npad, Xh, cn, beta, cand, st, split);
    else
        launch_tiles_wide<PREC_FP16>(KL, DK, N, Npad, Xh, cn, beta, cand, st, split);
}

}  // namespace ncvis_hip
