// bh_gbtrs_part_lo_d.hip -- the partitioned band solve's FP64 L and L^T sweeps (kernels in
// bh_gbtrs_part.hpp), a unit of their own so the sweeps compile in parallel.
#define BH_PART_INSTANTIATE
#include "bh_gbtrs_part.hpp"

namespace bh {
namespace part {
template int part_sweep<double, kLN>(const PartPlan &, PartArgs<double>, bool, hipStream_t);
template int part_sweep<double, kLT>(const PartPlan &, PartArgs<double>, bool, hipStream_t);
}  // namespace part
}  // namespace bh
