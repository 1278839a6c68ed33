// bh_gbtrs_part_s.hip -- FP32 entry point of the partitioned band LU solve
// (kernels in bh_gbtrs_part.hpp).
#include "bh_gbtrs_part.hpp"

extern "C" {

int bh_sgbtrs_part(char trans, bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const float *AB, bh_int ldab,
                   const bh_int *ipiv, float *B, bh_int ldb, void *work, size_t lwork) {
    return gbtrs_part_impl<float>(trans, n, kl, ku, nrhs, AB, ldab, ipiv, B, ldb, work, lwork);
}

}  // extern "C"
