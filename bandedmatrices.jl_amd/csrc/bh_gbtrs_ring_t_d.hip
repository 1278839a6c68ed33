// bh_gbtrs_ring_t_d.hip -- the FP64 'T' lane-ring band solve (kernels in
// bh_gbtrs_ring.hpp), a unit of its own so the ring kernels compile in parallel.
#include "bh_gbtrs_ring.hpp"

namespace bh {
template int ring_gbtrs_t_launch<double>(i64, i64, i64, i64, const double *, i64, const i64 *, double *, i64, hipStream_t);
}  // namespace bh
