// bh_tbsv_d.hip -- the triangular band kernels for double.
#define BH_TB_T double
#include "bh_tbsv.hpp"
