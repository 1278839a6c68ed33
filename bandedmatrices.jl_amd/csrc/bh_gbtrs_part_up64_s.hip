// bh_gbtrs_part_up64_s.hip -- the partitioned band solve's FP32 U and U^T sweeps with
// 64-wide windows (kernels in bh_gbtrs_part.hpp), a unit of their own so the sweeps compile
// in parallel.
#define BH_PART_INSTANTIATE
#define BH_PART_KW64
#include "bh_gbtrs_part.hpp"

namespace bh {
namespace part {
template int part_sweep_kw<float, 64, kUN>(PartArgs<float>, bool, hipStream_t);
template int part_sweep_kw<float, 64, kUT>(PartArgs<float>, bool, hipStream_t);
}  // namespace part
}  // namespace bh
