// bh_gbtrs_ring_t_s.hip -- the FP32 'T' lane-ring band solve (kernels in
// bh_gbtrs_ring.hpp), a unit of its own so the ring kernels compile in parallel.
#include "bh_gbtrs_ring.hpp"

namespace bh {
template int ring_gbtrs_t_launch<float>(i64, i64, i64, i64, const float *, i64, const i64 *, float *, i64, hipStream_t);
}  // namespace bh
