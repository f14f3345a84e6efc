// kNN candidate tiles with lists of 16 per lane half (k + 1 <= 16): all three GEMM precisions.
#include "knn_tiles.h"

namespace ncvis_hip {

void knn_tiles16(int prec, int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn, float beta,
                 int32_t *cand, hipStream_t st, const TileVariant &tv) {
    switch (prec) {
        case PREC_FP32: launch_tiles16<PREC_FP32>(DK, N, Npad, Xh, cn, beta, cand, st, tv); break;
        case PREC_BF16: launch_tiles16<PREC_BF16>(DK, N, Npad, Xh, cn, beta, cand, st, tv); break;
        default: launch_tiles16<PREC_FP16>(DK, N, Npad, Xh, cn, beta, cand, st, tv); break;
    }
}

}  // namespace ncvis_hip
