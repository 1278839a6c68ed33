// bh_gbmv_slab_s.hip -- the FP32 'N' gbmv kernels for contiguous band storage (in
// bh_gbmv_slab.hpp), a unit of their own so the precisions compile in parallel.
#include "bh_gbmv_slab.hpp"

namespace bh {
template int slab_gbmv_n<float>(i64, i64, i64, i64, i64, float, const float *, i64, const float *, i64, i64, float, float *,
                             i64, i64, hipStream_t);
}  // namespace bh
