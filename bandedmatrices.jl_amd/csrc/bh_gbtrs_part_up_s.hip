// bh_gbtrs_part_up_s.hip -- the partitioned band solve's FP32 U and U^T sweeps (kernels in
// bh_gbtrs_part.hpp), a unit of their own so the sweeps compile in parallel.
#define BH_PART_INSTANTIATE
#include "bh_gbtrs_part.hpp"

namespace bh {
namespace part {
template int part_sweep<float, kUN>(const PartPlan &, PartArgs<float>, bool, hipStream_t);
template int part_sweep<float, kUT>(const PartPlan &, PartArgs<float>, bool, hipStream_t);
}  // namespace part
}  // namespace bh
