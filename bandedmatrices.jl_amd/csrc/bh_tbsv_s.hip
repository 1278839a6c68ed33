// bh_tbsv_s.hip -- the triangular band kernels for float.
#define BH_TB_T float
#include "bh_tbsv.hpp"
