// bh_lu_lane_d.hip -- the double instantiation of the narrow-band LU kernels
// (single-system gbtrf and the persistent batched gbsv) of bh_lu_lane.hpp.
#define BH_LANE_T double
#include "bh_lu_lane.hpp"
