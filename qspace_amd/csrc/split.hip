// split.hip -- split-precision grouped GEMM (placeholder until the slice
// kernels land; the modes are rejected at run time).
#include "qsx_internal.hpp"

namespace qsx {

void launch_work_split(int mode, int nslices, const WorkQueue* dqueues, int nqueues, const WorkItem* ditems,
                       const ReduceItem* dreds, int nreds, const PairDesc* dpairs, const double* A, const double* B,
                       double* C, double* scratch, hipStream_t s) {
  (void)nslices; (void)dqueues; (void)nqueues; (void)ditems; (void)dreds; (void)nreds; (void)dpairs; (void)A; (void)B; (void)C;
  (void)scratch; (void)s;
  fail(QSX_ERR_UNSUPPORTED, "split-precision mode %d not available", mode);
}

}  // namespace qsx
