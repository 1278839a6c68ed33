// bh_lu_lane_s.hip -- the float instantiation of the narrow-band LU kernels
// (single-system gbtrf and the persistent batched gbsv) of bh_lu_lane.hpp.
#define BH_LANE_T float
#include "bh_lu_lane.hpp"
