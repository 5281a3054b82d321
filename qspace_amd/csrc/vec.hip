// vec.hip -- FP64 vector kernels of the Lanczos / Davidson loops
// (RunEigs Source/eigs2SiteQS.cc:98-334, update_psi_2site.m:223-391).
// Reductions are two-stage with a fixed partition (DOT_BLOCKS partials,
// each summed in one fixed order) so every dot product is deterministic; scalars
// stay in device memory so an iteration needs no host round trip and can
// be captured in a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "qsx_internal.hpp"
#include "vec.hpp"

namespace qsx {
namespace {

constexpr int RT = 256;

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[RT / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < RT / 64; ++i) s += red[i];
  __syncthreads();
  return s;  // valid in thread 0
}

// the DOT_BLOCKS partials of one reduction summed by one wave (four per
// lane in a fixed pairing, then a fixed shuffle tree): every block that
// needs the value sums it in the same order, so all of them hold the same
// bits
static_assert(DOT_BLOCKS == 256, "the partial sums read four partials per lane");
__device__ __forceinline__ double lane_partials(const double* __restrict__ p, int lane) {
  return (p[lane] + p[lane + 64]) + (p[lane + 128] + p[lane + 192]);
}
__device__ __forceinline__ double wave_reduce(double s) {
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  return s;  // valid in lane 0
}
// the result is in every lane
__device__ __forceinline__ double wave_sum_partials(const double* __restrict__ p) {
  const double s = wave_reduce(lane_partials(p, threadIdx.x & 63));
  return __shfl(s, 0, 64);
}

// c[j] = sum_b partial[j][b] for j < m, one wave per j; a wave takes up to
// kSumGroup coefficients per round with all their loads issued together
// (one L2 round trip for up to RT/64 * kSumGroup coefficients)
constexpr int kSumGroup = 8;
__device__ __forceinline__ void sum_coefficients(const double* __restrict__ partial, int m, double* c) {
  constexpr int NW = RT / 64;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int j0 = wv; j0 < m; j0 += NW * kSumGroup) {
    double v[kSumGroup];
#pragma unroll
    for (int q = 0; q < kSumGroup; ++q) {
      const int j = j0 + q * NW;
      v[q] = j < m ? lane_partials(partial + (int64_t)j * DOT_BLOCKS, lane) : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kSumGroup; ++q) {
      const double s = wave_reduce(v[q]);
      const int j = j0 + q * NW;
      if (lane == 0 && j < m) c[j] = s;
    }
  }
  __syncthreads();
}

// partial[j*DOT_BLOCKS + b] = sum over the b-th slice of V_j . w
__global__ __launch_bounds__(RT) void k_mdot_partial(const double* __restrict__ V, int64_t ldv,
                                                      const double* __restrict__ w, int64_t n,
                                                      double* __restrict__ partial) {
  const int j = blockIdx.y, b = blockIdx.x;
  const double* v = V + (int64_t)j * ldv;
  double s = 0;
  for (int64_t i = (int64_t)b * RT + threadIdx.x; i < n; i += (int64_t)DOT_BLOCKS * RT) s += v[i] * w[i];
  s = block_sum(s);
  if (threadIdx.x == 0) partial[(int64_t)j * DOT_BLOCKS + b] = s;
}

// out[j] = (accumulate ? out[j] : 0) + sum_b partial[j][b]
__global__ __launch_bounds__(RT) void k_mdot_final(const double* __restrict__ partial, int m, double* __restrict__ out,
                                                    int accumulate) {
  const int j = blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < DOT_BLOCKS; b += RT) s += partial[(int64_t)j * DOT_BLOCKS + b];
  s = block_sum(s);
  if (threadIdx.x == 0) out[j] = accumulate ? out[j] + s : s;
}

// dst = src .* w, partial[b] = sum over the b-th slice of dst^2 (the slice
// layout of k_mdot_partial)
__global__ __launch_bounds__(RT) void k_scale_norm_partial(const double* __restrict__ src,
                                                            const double* __restrict__ w, double* __restrict__ dst,
                                                            int64_t n, double* __restrict__ partial) {
  const int b = blockIdx.x;
  double s = 0;
  for (int64_t i = (int64_t)b * RT + threadIdx.x; i < n; i += (int64_t)DOT_BLOCKS * RT) {
    const double x = src[i] * w[i];
    dst[i] = x;
    s += x * x;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partial[b] = s;
}

// y = sum_j c[j] V_j (divided elementwise by div if given)
__global__ __launch_bounds__(RT) void k_combine(double* __restrict__ y, const double* __restrict__ V, int64_t ldv,
                                                 const double* __restrict__ c, int m, int64_t n,
                                                 const double* __restrict__ div) {
  for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) {
    double s = 0;
    for (int j = 0; j < m; ++j) s += c[j] * V[(int64_t)j * ldv + i];
    y[i] = div ? s / div[i] : s;
  }
}

// u = V c, r = HV c - theta u with theta = c[m] (the Davidson Ritz vector
// and its residual in one pass over V and HV)
__global__ __launch_bounds__(RT) void k_combine_res(double* __restrict__ u, double* __restrict__ r,
                                                     const double* __restrict__ V, const double* __restrict__ HV,
                                                     int64_t ldv, const double* __restrict__ c, int m, int64_t n) {
  const double th = c[m];
  for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) {
    double su = 0, sh = 0;
    for (int j = 0; j < m; ++j) {
      su += c[j] * V[(int64_t)j * ldv + i];
      sh += c[j] * HV[(int64_t)j * ldv + i];
    }
    u[i] = su;
    r[i] = sh - th * su;
  }
}

// One Gram-Schmidt step fused with the next reduction: w -= sum_j c[j] V_j
// (c from the partials pin[j][b], every block summing them in the same
// order), then pout[j][b] = <V2_j, w> over block b's slice for j < jn,
// followed by <w, w> at j = jn if wnorm is set.  The grid is exactly DOT_BLOCKS blocks with
// the slice layout of k_mdot_partial; each partial is its slice's sum in a
// fixed order (elements in index order per thread, a fixed shuffle tree,
// the waves in order): deterministic.  The dots of kDotGroup vectors are
// accumulated in one pass over the slice (their loads in flight together)
// and reduced with one barrier.
constexpr int kMaxM = 512;  // largest subspace (coefficients held in LDS)
constexpr int kDotGroup = 8;
__global__ __launch_bounds__(RT) void k_gs_update_mdot(double* __restrict__ w, const double* __restrict__ V,
                                                        int64_t ldv, const double* __restrict__ pin, int m,
                                                        int64_t n, double* __restrict__ cout, int cj,
                                                        const double* __restrict__ V2, int jn, int wnorm,
                                                        double* __restrict__ pout) {
  __shared__ double c[kMaxM];
  __shared__ double red[RT / 64][kDotGroup];
  sum_coefficients(pin, m, c);
  if (cout && blockIdx.x == 0 && threadIdx.x == 0) *cout = c[cj];
  const int b = blockIdx.x;
  for (int64_t i = (int64_t)b * RT + threadIdx.x; i < n; i += (int64_t)DOT_BLOCKS * RT) {
    double s = 0;
#pragma unroll 8
    for (int j = 0; j < m; ++j) s += c[j] * V[(int64_t)j * ldv + i];
    w[i] -= s;
  }
  // the slice's updated w is this thread's own elements: no barrier needed
  // before the dots
  const int nv = jn + (wnorm ? 1 : 0);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j0 = 0; j0 < nv; j0 += kDotGroup) {
    const int ng = min(kDotGroup, nv - j0);
    double acc[kDotGroup];
#pragma unroll
    for (int q = 0; q < kDotGroup; ++q) acc[q] = 0.0;
    for (int64_t i = (int64_t)b * RT + threadIdx.x; i < n; i += (int64_t)DOT_BLOCKS * RT) {
      const double wi = w[i];
#pragma unroll
      for (int q = 0; q < kDotGroup; ++q)
        if (q < ng) acc[q] += (j0 + q < jn ? V2[(int64_t)(j0 + q) * ldv + i] : wi) * wi;
    }
#pragma unroll
    for (int q = 0; q < kDotGroup; ++q) {
      double s = acc[q];
      for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
      if (lane == 0) red[wv][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < ng) {
      double t = 0;
      for (int q = 0; q < RT / 64; ++q) t += red[q][threadIdx.x];
      pout[(int64_t)(j0 + threadIdx.x) * DOT_BLOCKS + b] = t;
    }
    __syncthreads();
  }
}

// The second Gram-Schmidt step fused with the normalisation: c[j] (j < m)
// and |w|^2 (at j = m) from the partials pin, then dst = (w - sum_j c[j] V_j)
// / beta with beta^2 = |w|^2 - sum_j c[j]^2, the squared norm of the update
// for orthonormal V_j.  After a first pass the c[j] are at rounding level
// relative to |w| and the difference is exact to rounding; when they are
// not (|c|^2 > |w|^2 / 2: w is rounding noise inside span V, a breakdown)
// beta = 0 and dst = 0.  Every block sums the partials and the squares in
// the same order.
__global__ __launch_bounds__(RT) void k_gs_update_normalize(const double* __restrict__ w,
                                                             const double* __restrict__ V, int64_t ldv,
                                                             const double* __restrict__ pin, int m, int64_t n,
                                                             double* __restrict__ dst, double* __restrict__ beta) {
  __shared__ double c[kMaxM + 1];
  __shared__ double inv;
  sum_coefficients(pin, m + 1, c);
  if (threadIdx.x == 0) {
    double cs = 0;
    for (int j = 0; j < m; ++j) cs += c[j] * c[j];
    const double b = cs <= 0.5 * c[m] ? sqrt(c[m] - cs) : 0.0;
    inv = b > 0 ? 1.0 / b : 0.0;
    if (blockIdx.x == 0 && beta) *beta = b;
  }
  __syncthreads();
  const double f = inv;
  for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) {
    double s = 0;
#pragma unroll 8
    for (int j = 0; j < m; ++j) s += c[j] * V[(int64_t)j * ldv + i];
    dst[i] = (w[i] - s) * f;
  }
}

// k_gs_update_mdot for the first Gram-Schmidt pass of the eigensolvers (the
// dots against the same m <= MC vectors as the update): one pass over V, a
// thread's V_j[i] held in registers between the update and the dots.  Each
// partial is its slice's sum in a fixed order (the thread's elements in
// index order, the shuffle tree, the waves in order): deterministic.
template <int MC>
__global__ __launch_bounds__(RT) void k_gs_update_mdot_reg(double* __restrict__ w, const double* __restrict__ V,
                                                            int64_t ldv, const double* __restrict__ pin, int m,
                                                            int64_t n, double* __restrict__ cout, int cj, int wnorm,
                                                            double* __restrict__ pout) {
  __shared__ double c[MC];
  __shared__ double red[RT / 64][MC + 1];
  sum_coefficients(pin, m, c);
  if (cout && blockIdx.x == 0 && threadIdx.x == 0) *cout = c[cj];
  const int b = blockIdx.x;
  double acc[MC + 1];
#pragma unroll
  for (int q = 0; q <= MC; ++q) acc[q] = 0.0;
  for (int64_t i = (int64_t)b * RT + threadIdx.x; i < n; i += (int64_t)DOT_BLOCKS * RT) {
    double v[MC];
#pragma unroll
    for (int j = 0; j < MC; ++j) v[j] = j < m ? V[(int64_t)j * ldv + i] : 0.0;
    double s = 0;
#pragma unroll
    for (int j = 0; j < MC; ++j)
      if (j < m) s += c[j] * v[j];
    const double wi = w[i] - s;
    w[i] = wi;
#pragma unroll
    for (int j = 0; j < MC; ++j) acc[j] += v[j] * wi;
    acc[MC] += wi * wi;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q <= MC; ++q) {
    if (q < m || (q == MC && wnorm)) {
      const double t = wave_reduce(acc[q]);
      if (lane == 0) red[wv][q] = t;
    }
  }
  __syncthreads();
  const int nv = m + (wnorm ? 1 : 0);
  if (threadIdx.x < nv) {
    const int q = threadIdx.x < m ? threadIdx.x : MC;
    double t = 0;
    for (int k = 0; k < RT / 64; ++k) t += red[k][q];
    pout[(int64_t)threadIdx.x * DOT_BLOCKS + b] = t;
  }
}

// beta = sqrt(sum_b partial[b]); dst = src / beta (every block sums the
// partials in the same order)
__global__ __launch_bounds__(RT) void k_normalize_p(const double* __restrict__ src, double* __restrict__ dst,
                                                     const double* __restrict__ partial, double* __restrict__ beta,
                                                     int64_t n) {
  __shared__ double nb;
  if (threadIdx.x < 64) {
    const double s = wave_sum_partials(partial);
    if (threadIdx.x == 0) nb = s;
  }
  __syncthreads();
  const double b = sqrt(nb);
  const double inv = b > 0 ? 1.0 / b : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0 && beta) *beta = b;
  for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) dst[i] = src[i] * inv;
}

int grid_for(int64_t n) { return (int)std::min<int64_t>(2048, std::max<int64_t>(1, (n + RT - 1) / RT)); }

[[noreturn]] void fail_vec(int m) { fail(QSX_ERR_ARG, "eigensolver subspace of %d vectors exceeds %d", m, kMaxM); }

}  // namespace

void vec_mdot(const double* V, int64_t ldv, int m, const double* w, int64_t n, double* partial, double* out,
              int accumulate, hipStream_t s) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_mdot_partial, dim3(DOT_BLOCKS, m), dim3(RT), 0, s, V, ldv, w, n, partial);
  hipLaunchKernelGGL(k_mdot_final, dim3(m), dim3(RT), 0, s, partial, m, out, accumulate);
}

void vec_mdot_partial(const double* V, int64_t ldv, int m, const double* w, int64_t n, double* partial,
                      hipStream_t s) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_mdot_partial, dim3(DOT_BLOCKS, m), dim3(RT), 0, s, V, ldv, w, n, partial);
}

void vec_gs_update_mdot(double* w, const double* V, int64_t ldv, const double* pin, int m, int64_t n, double* cout,
                        int cj, const double* V2, int jn, int wnorm, double* pout, hipStream_t s) {
  if (m <= 0) return;
  if (m > kMaxM) fail_vec(m);
  if (jn > 0 && !V2) fail(QSX_ERR_ARG, "vec_gs_update_mdot: %d dots without vectors", jn);
  if (V2 == V && jn == m && m <= 32) {  // dots against the update's own vectors: one pass over V
    if (m <= 8)
      hipLaunchKernelGGL(k_gs_update_mdot_reg<8>, dim3(DOT_BLOCKS), dim3(RT), 0, s, w, V, ldv, pin, m, n, cout, cj,
                         wnorm, pout);
    else if (m <= 16)
      hipLaunchKernelGGL(k_gs_update_mdot_reg<16>, dim3(DOT_BLOCKS), dim3(RT), 0, s, w, V, ldv, pin, m, n, cout, cj,
                         wnorm, pout);
    else
      hipLaunchKernelGGL(k_gs_update_mdot_reg<32>, dim3(DOT_BLOCKS), dim3(RT), 0, s, w, V, ldv, pin, m, n, cout, cj,
                         wnorm, pout);
    return;
  }
  hipLaunchKernelGGL(k_gs_update_mdot, dim3(DOT_BLOCKS), dim3(RT), 0, s, w, V, ldv, pin, m, n, cout, cj, V2, jn,
                     wnorm, pout);
}

void vec_gs_update_normalize(const double* w, const double* V, int64_t ldv, const double* pin, int m, int64_t n,
                             double* dst, double* beta, hipStream_t s) {
  if (m <= 0) return;
  if (m > kMaxM) fail_vec(m);
  hipLaunchKernelGGL(k_gs_update_normalize, dim3(grid_for(n)), dim3(RT), 0, s, w, V, ldv, pin, m, n, dst, beta);
}

void vec_scale_norm_partial(const double* src, const double* w, double* dst, int64_t n, double* partial,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_scale_norm_partial, dim3(DOT_BLOCKS), dim3(RT), 0, s, src, w, dst, n, partial);
}

void vec_normalize_partial(const double* src, double* dst, const double* partial, double* beta, int64_t n,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_normalize_p, dim3(grid_for(n)), dim3(RT), 0, s, src, dst, partial, beta, n);
}

void vec_combine(double* y, const double* V, int64_t ldv, const double* c, int m, int64_t n, hipStream_t s,
                 const double* div) {
  hipLaunchKernelGGL(k_combine, dim3(grid_for(n)), dim3(RT), 0, s, y, V, ldv, c, m, n, div);
}

void vec_combine_res(double* u, double* r, const double* V, const double* HV, int64_t ldv, const double* c, int m,
                     int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_combine_res, dim3(grid_for(n)), dim3(RT), 0, s, u, r, V, HV, ldv, c, m, n);
}

}  // namespace qsx
