// vec.hpp -- FP64 device vector operations of the eigensolvers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace qsx {

constexpr int DOT_BLOCKS = 256;  // fixed partition of every reduction (one workgroup per CU)

// out[j] (+)= <V_j, w> for j < m; partial needs m*DOT_BLOCKS doubles
void vec_mdot(const double* V, int64_t ldv, int m, const double* w, int64_t n, double* partial, double* out,
              int accumulate, hipStream_t s);
// the same steps with the reduction's final sum folded into the consumer:
// partial[j*DOT_BLOCKS + b] from vec_mdot_partial
void vec_mdot_partial(const double* V, int64_t ldv, int m, const double* w, int64_t n, double* partial,
                      hipStream_t s);
void vec_normalize_partial(const double* src, double* dst, const double* partial, double* beta, int64_t n,
                           hipStream_t s);
// w -= V c(pin) (block 0 writes c[cj] to cout if given), then pout = the
// partials of <V2_j, w> (j < jn), then of <w, w> if wnorm; pin and pout must
// not overlap
void vec_gs_update_mdot(double* w, const double* V, int64_t ldv, const double* pin, int m, int64_t n, double* cout,
                        int cj, const double* V2, int jn, int wnorm, double* pout, hipStream_t s);
// dst = (w - V c(pin)) / beta, beta^2 = |w|^2 - |c|^2 from the partials pin
// (m coefficient sets, then |w|^2), written to *beta if given; beta = 0 and
// dst = 0 when |c|^2 > |w|^2 / 2 (w is rounding noise inside span V)
void vec_gs_update_normalize(const double* w, const double* V, int64_t ldv, const double* pin, int m, int64_t n,
                             double* dst, double* beta, hipStream_t s);
// dst = src .* w and the DOT_BLOCKS partials of |dst|^2
void vec_scale_norm_partial(const double* src, const double* w, double* dst, int64_t n, double* partial,
                            hipStream_t s);
// y = V c, divided elementwise by div if given
void vec_combine(double* y, const double* V, int64_t ldv, const double* c, int m, int64_t n, hipStream_t s,
                 const double* div = nullptr);
// u = V c, r = HV c - c[m] u (c holds m coefficients and then theta)
void vec_combine_res(double* u, double* r, const double* V, const double* HV, int64_t ldv, const double* c, int m,
                     int64_t n, hipStream_t s);

}  // namespace qsx
