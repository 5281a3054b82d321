// svd.hip -- the block SVD of the truncation (orthoQS): the group
// matrices are assembled from the blocks on the device, each is decomposed
// by a one-sided Jacobi SVD in one workgroup, and the kept singular vectors
// are scattered into the blocks of the two result tensors.
//
// One-sided Jacobi (Hestenes, J. SIAM 6 (1958); the rotation and stopping
// rule of Demmel and Veselic, SIAM J. Matrix Anal. Appl. 13 (1992)): the
// columns of an m x n matrix (m >= n) are orthogonalised by plane rotations
// of column pairs; a sweep visits every pair once in the round-robin order,
// whose n/2 pairs per step are disjoint and run in parallel.  At the end
// a_j = sigma_j u_j and the accumulated rotations are V.
#include <cfloat>

#include "qsx_internal.hpp"

namespace qsx {
namespace {

constexpr int SVD_NT = 1024;          // threads per workgroup
constexpr int SVD_SEG = 16;           // lanes per column pair
constexpr int SVD_MAX_SWEEPS = 60;

// sum over the 16 lanes of a segment
__device__ inline double seg_sum(double v) {
  v += __shfl_xor(v, 8, SVD_SEG);
  v += __shfl_xor(v, 4, SVD_SEG);
  v += __shfl_xor(v, 2, SVD_SEG);
  v += __shfl_xor(v, 1, SVD_SEG);
  return v;
}

// zero-padded group matrices: block e (d1 x d2, column-major) goes to
// row i0, column j0 of its matrix (leading dim ld), transposed when trans
__global__ __launch_bounds__(256) void k_svd_gather(const SvdGather* __restrict__ items, const double* __restrict__ src,
                                                    double* __restrict__ work) {
  const SvdGather it = items[blockIdx.x];
  const int64_t n = (int64_t)it.d1 * it.d2;
  for (int64_t e = threadIdx.x; e < n; e += 256) {
    const int64_t i = e % it.d1, j = e / it.d1;
    const int64_t d = it.trans ? it.dst_off + j + i * it.ld : it.dst_off + i + j * it.ld;
    work[d] = src[it.src_off + e];
  }
}

__global__ __launch_bounds__(SVD_NT) void k_jacobi_svd(const SvdItem* __restrict__ items, double* __restrict__ work,
                                                       double* __restrict__ sig, int32_t* __restrict__ sweeps) {
  __shared__ double lds[SVD_LDS_DOUBLES];
  __shared__ int rotated;
  const SvdItem it = items[blockIdx.x];
  const int m = it.m, n = it.n;  // m >= n >= 1
  const int tid = threadIdx.x;
  double* A = it.lds ? lds : work + it.a_off;
  double* V = it.lds ? lds + (int64_t)m * n : work + it.v_off;
  if (it.lds)
    for (int64_t e = tid; e < (int64_t)m * n; e += SVD_NT) A[e] = work[it.a_off + e];
  for (int64_t e = tid; e < (int64_t)n * n; e += SVD_NT) V[e] = (e % (n + 1)) == 0 ? 1.0 : 0.0;
  __syncthreads();

  const int np = n + (n & 1);  // column n is a dummy when n is odd
  const int seg = tid / SVD_SEG, lane = tid % SVD_SEG;
  constexpr int nseg = SVD_NT / SVD_SEG;
  const double tol = (double)m * DBL_EPSILON;
  int sw = 0;
  for (; sw < SVD_MAX_SWEEPS; ++sw) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int step = 0; step < np - 1; ++step) {
      for (int k = seg; k < np / 2; k += nseg) {
        int p, q;
        if (k == 0) {
          p = np - 1;
          q = step;
        } else {
          p = (step + k) % (np - 1);
          q = (step - k + np - 1) % (np - 1);
        }
        if (p >= n || q >= n) continue;  // the dummy column (whole segment)
        double* ap = A + (int64_t)p * m;
        double* aq = A + (int64_t)q * m;
        double al = 0.0, be = 0.0, ga = 0.0;
        for (int i = lane; i < m; i += SVD_SEG) {
          const double x = ap[i], y = aq[i];
          al += x * x;
          be += y * y;
          ga += x * y;
        }
        al = seg_sum(al);
        be = seg_sum(be);
        ga = seg_sum(ga);
        if (!(fabs(ga) > tol * sqrt(al) * sqrt(be))) continue;
        // the rotation that zeroes the off-diagonal entry of the 2x2 Gram
        // matrix [al ga; ga be], the smaller angle
        const double zeta = (be - al) / (2.0 * ga);
        const double az = fabs(zeta);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (az + (az > 1e150 ? az : sqrt(1.0 + zeta * zeta)));
        const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
        for (int i = lane; i < m; i += SVD_SEG) {
          const double x = ap[i], y = aq[i];
          ap[i] = c * x - s * y;
          aq[i] = s * x + c * y;
        }
        double* vp = V + (int64_t)p * n;
        double* vq = V + (int64_t)q * n;
        for (int i = lane; i < n; i += SVD_SEG) {
          const double x = vp[i], y = vq[i];
          vp[i] = c * x - s * y;
          vq[i] = s * x + c * y;
        }
        if (lane == 0) rotated = 1;
      }
      __syncthreads();
    }
    const int any = rotated;
    __syncthreads();  // every thread has read the flag before it is reset
    if (!any) break;
  }
  // sigma_j = |a_j|, u_j = a_j / sigma_j (a zero column stays zero)
  for (int j = seg; j < n; j += nseg) {
    double* aj = A + (int64_t)j * m;
    double s2 = 0.0;
    for (int i = lane; i < m; i += SVD_SEG) s2 += aj[i] * aj[i];
    s2 = seg_sum(s2);
    const double sg = sqrt(s2);
    const double inv = sg > 0.0 ? 1.0 / sg : 0.0;
    for (int i = lane; i < m; i += SVD_SEG) aj[i] *= inv;
    if (lane == 0) sig[it.s_off + j] = sg;
  }
  __syncthreads();
  if (it.lds) {
    for (int64_t e = tid; e < (int64_t)m * n; e += SVD_NT) work[it.a_off + e] = A[e];
    for (int64_t e = tid; e < (int64_t)n * n; e += SVD_NT) work[it.v_off + e] = V[e];
  }
  if (tid == 0) sweeps[blockIdx.x] = sw;
}

// result block: dst[r + c * rows] = src[src_off + r + col_c * ld] * scale_c
// over the kept columns (col_c, scale_c) = cols[col0 + c]
__global__ __launch_bounds__(256) void k_svd_scatter(const SvdScatter* __restrict__ items,
                                                     const SvdCol* __restrict__ cols, const double* __restrict__ work,
                                                     double* __restrict__ out) {
  const SvdScatter it = items[blockIdx.x];
  const int64_t n = (int64_t)it.rows * it.ncols;
  for (int64_t e = threadIdx.x; e < n; e += 256) {
    const int64_t r = e % it.rows, c = e / it.rows;
    const SvdCol col = cols[it.col0 + c];
    out[it.dst_off + e] = work[it.src_off + r + (int64_t)col.col * it.ld] * col.scale;
  }
}

}  // namespace

void launch_svd_gather(const SvdGather* ditems, int nitems, const double* src, double* work, hipStream_t s) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL(k_svd_gather, dim3(nitems), dim3(256), 0, s, ditems, src, work);
  QSX_HIP(hipGetLastError());
}

void launch_jacobi_svd(const SvdItem* ditems, int nitems, double* work, double* sig, int32_t* sweeps,
                       hipStream_t s) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL(k_jacobi_svd, dim3(nitems), dim3(SVD_NT), 0, s, ditems, work, sig, sweeps);
  QSX_HIP(hipGetLastError());
}

int jacobi_max_sweeps() { return SVD_MAX_SWEEPS; }

void launch_svd_scatter(const SvdScatter* ditems, int nitems, const SvdCol* dcols, const double* work, double* out,
                        hipStream_t s) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL(k_svd_scatter, dim3(nitems), dim3(256), 0, s, ditems, dcols, work, out);
  QSX_HIP(hipGetLastError());
}

}  // namespace qsx
