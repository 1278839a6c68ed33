// bh_gbtrs_part_d.hip -- FP64 entry points of the partitioned band LU solve
// (kernels in bh_gbtrs_part.hpp), and its workspace query.
#include "bh_gbtrs_part.hpp"

extern "C" {

int bh_gbtrs_part_workspace(char prec, bh_int n, bh_int kl, bh_int ku, bh_int nrhs, size_t *bytes) {
    BH_ARG(prec == 'd' || prec == 's', 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(kl >= 0, 3);
    BH_ARG(ku >= 0, 4);
    BH_ARG(nrhs >= 0, 5);
    BH_ARG(bytes != nullptr, 6);
    *bytes = prec == 'd' ? part_bytes<double>(n, kl, ku, nrhs) : part_bytes<float>(n, kl, ku, nrhs);
    return BH_OK;
}

int bh_dgbtrs_part(char trans, bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const double *AB, bh_int ldab,
                   const bh_int *ipiv, double *B, bh_int ldb, void *work, size_t lwork) {
    return gbtrs_part_impl<double>(trans, n, kl, ku, nrhs, AB, ldab, ipiv, B, ldb, work, lwork);
}

}  // extern "C"
