// bh_gbmv_slab_d.hip -- the FP64 'N' gbmv kernels for contiguous band storage (in
// bh_gbmv_slab.hpp), a unit of their own so the precisions compile in parallel.
#include "bh_gbmv_slab.hpp"

namespace bh {
template int slab_gbmv_n<double>(i64, i64, i64, i64, i64, double, const double *, i64, const double *, i64, i64, double, double *,
                             i64, i64, hipStream_t);
}  // namespace bh
