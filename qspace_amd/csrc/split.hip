// split.hip -- split-precision grouped GEMM (placeholder until the slice
// kernels land; the modes are rejected at run time).
#include "qsx_internal.hpp"

namespace qsx {

void launch_grouped_gemm_split(int mode, int nslices, const TileDesc* dtiles, int ntiles,
                               const PairDesc* dpairs, const double* A, const double* B, double* C, hipStream_t s) {
  (void)nslices; (void)dtiles; (void)ntiles; (void)dpairs; (void)A; (void)B; (void)C; (void)s;
  fail(QSX_ERR_UNSUPPORTED, "split-precision mode %d not available", mode);
}

}  // namespace qsx
