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
constexpr int SVD_NT_BLOCK = 512;     // threads per workgroup of the block steps
constexpr int SVD_SEG = 16;           // lanes per column pair
constexpr int SVD_MAX_SWEEPS = 60;
// Orthogonality target of a column pair: sqrt(m) eps (as LAPACK's dgesvj)
// for the first kTightSweeps sweeps, then m eps, the rounding bound of the
// length-m dot products that decide a rotation -- a pair whose computed
// cosine stays above the tight target only through that rounding (e.g.
// columns of nearly equal norm) would otherwise rotate forever.
constexpr int kTightSweeps = 20;
__host__ __device__ inline double jacobi_tol(int m, int sweep) {
  return (sweep < kTightSweeps ? sqrt((double)m) : (double)m) * DBL_EPSILON;
}

// v + (v of the lane that the DPP row pattern CTRL names); a segment is one
// DPP row of 16 lanes, active or inactive as a whole
template <int CTRL>
__device__ __forceinline__ double dpp_add(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return v + __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// sum over the 16 lanes of a segment, the same bits in every lane: pairs
// (quad_perm 1,0,3,2), quads (quad_perm 2,3,0,1), half rows (row_half_mirror),
// rows (row_mirror); each step adds two equal-bit partial sums in either
// order, so all lanes agree
__device__ __forceinline__ double seg_sum(double v) {
  v = dpp_add<0xB1>(v);
  v = dpp_add<0x4E>(v);
  v = dpp_add<0x141>(v);
  v = dpp_add<0x140>(v);
  return v;
}

// the rotation of the Gram matrix [al ga; ga be] that zeroes ga (smaller
// angle); false when |ga| <= tol sqrt(al be): the pair is orthogonal.  A
// pair whose smaller column is below tol times the larger (norms) is left
// alone as well: such a column is rounding noise of a rank-deficient matrix
// (its singular value is below what FP64 resolves against the larger one),
// and rotating it against the larger column only makes new noise of the
// same size, so the sweeps would never end.
__device__ __forceinline__ bool jacobi_cs(double al, double be, double ga, double tol, double& c, double& s) {
  if (!(fabs(ga) > tol * sqrt(al) * sqrt(be))) return false;
  if (fmin(al, be) <= tol * tol * fmax(al, be)) return false;
  const double zeta = (be - al) / (2.0 * ga);
  const double az = fabs(zeta);
  const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (az + (az > 1e150 ? az : sqrt(1.0 + zeta * zeta)));
  c = 1.0 / sqrt(1.0 + t * t);
  s = c * t;
  return true;
}

// rotation of the column pair (ap, aq) of A (m rows) and (vp, vq) of V
// (n rows) by the 16 lanes of a segment; with R > 0 a lane holds its R
// elements of each column in registers (m, n <= 16 R), R = 0 runs over the
// rows in chunks of 4 per lane.  True if the pair was rotated.
template <int R>
__device__ __forceinline__ bool rotate_pair(double* ap, double* aq, double* vp, double* vq, int m, int n, int lane,
                                            double tol) {
  constexpr int L = SVD_SEG;
  double c, s;
  if constexpr (R > 0) {
    // V's elements are fetched with A's when registers allow (R <= 4)
    constexpr bool pre = R <= 4;
    double x[R], y[R], u[R], z[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = lane + L * r;
      x[r] = i < m ? ap[i] : 0.0;
      y[r] = i < m ? aq[i] : 0.0;
      if (pre) {
        u[r] = i < n ? vp[i] : 0.0;
        z[r] = i < n ? vq[i] : 0.0;
      }
    }
    double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      al += x[r] * x[r];
      be += y[r] * y[r];
      ga += x[r] * y[r];
    }
    al = seg_sum(al);
    be = seg_sum(be);
    ga = seg_sum(ga);
    if (!jacobi_cs(al, be, ga, tol, c, s)) return false;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = lane + L * r;
      if (i < m) {
        ap[i] = c * x[r] - s * y[r];
        aq[i] = s * x[r] + c * y[r];
      }
    }
    if (!pre) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = lane + L * r;
        u[r] = i < n ? vp[i] : 0.0;
        z[r] = i < n ? vq[i] : 0.0;
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = lane + L * r;
      if (i < n) {
        vp[i] = c * u[r] - s * z[r];
        vq[i] = s * u[r] + c * z[r];
      }
    }
    return true;
  } else {
    double al = 0.0, be = 0.0, ga = 0.0;
    for (int i0 = lane; i0 < m; i0 += 4 * L) {
      double x[4], y[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + L * r;
        x[r] = i < m ? ap[i] : 0.0;
        y[r] = i < m ? aq[i] : 0.0;
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        al += x[r] * x[r];
        be += y[r] * y[r];
        ga += x[r] * y[r];
      }
    }
    al = seg_sum(al);
    be = seg_sum(be);
    ga = seg_sum(ga);
    if (!jacobi_cs(al, be, ga, tol, c, s)) return false;
    for (int pass = 0; pass < 2; ++pass) {
      double* xp = pass ? vp : ap;
      double* yp = pass ? vq : aq;
      const int len = pass ? n : m;
      for (int i0 = lane; i0 < len; i0 += 4 * L) {
        double x[4], y[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = i0 + L * r;
          x[r] = i < len ? xp[i] : 0.0;
          y[r] = i < len ? yp[i] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = i0 + L * r;
          if (i < len) {
            xp[i] = c * x[r] - s * y[r];
            yp[i] = s * x[r] + c * y[r];
          }
        }
      }
    }
    return true;
  }
}

// Pair schedules over columns of A (leading dim m) and V (leading dim n),
// run by NT threads in segments of 16 lanes, one pair per segment at a
// time; every step's pairs are disjoint, steps are separated by barriers.
// Return true (in every thread) if any pair rotated; flag is LDS.
//  * sweep: all pairs of w columns in the round-robin order (w-1 steps of
//    w/2 pairs; column w is a dummy when w is odd);
//  * cross: all pairs (i, np + j), i < np, j < nq: step t pairs i with
//    (i + t) mod max(np, nq).
template <int NT, int R, bool CROSS>
__device__ __forceinline__ bool jacobi_pairs(double* A, double* V, int np, int nq, int m, int n, double tol,
                                             int* flag) {
  constexpr int L = SVD_SEG;
  constexpr int nseg = NT / L;
  const int tid = threadIdx.x;
  const int seg = tid / L, lane = tid % L;
  const int w = np + nq;
  const int wp = w + (w & 1);
  const int cmax = np > nq ? np : nq;
  const int nsteps = CROSS ? cmax : wp - 1;
  const int npairs = CROSS ? cmax : wp / 2;
  if (tid == 0) *flag = 0;
  __syncthreads();
  for (int step = 0; step < nsteps; ++step) {
    for (int k = seg; k < npairs; k += nseg) {
      int p, q;
      if (CROSS) {
        const int j = (k + step) % cmax;
        if (k >= np || j >= nq) continue;
        p = k;
        q = np + j;
      } else {
        if (k == 0) {
          p = wp - 1;
          q = step;
        } else {
          p = (step + k) % (wp - 1);
          q = (step - k + wp - 1) % (wp - 1);
        }
        if (p >= w || q >= w) continue;  // the dummy column (whole segment)
      }
      if (rotate_pair<R>(A + (int64_t)p * m, A + (int64_t)q * m, V + (int64_t)p * n, V + (int64_t)q * n, m, n,
                         lane, tol) &&
          lane == 0)
        *flag = 1;
    }
    __syncthreads();
  }
  const bool any = *flag != 0;
  __syncthreads();  // every thread has read the flag before it is reused
  return any;
}

// the cross pairs with each segment's own column of P held in registers for
// the whole schedule: segment k owns column k of P (A and V) and meets the
// columns of Q one step at a time, so only Q's columns move through LDS
// (half the LDS traffic of jacobi_pairs).  m, n <= 16 R.
template <int NT, int R>
__device__ __forceinline__ bool jacobi_cross_reg(double* A, double* V, int np, int nq, int m, int n, double tol,
                                                 int* flag) {
  constexpr int L = SVD_SEG;
  static_assert(R > 0, "register form");
  const int tid = threadIdx.x;
  const int seg = tid / L, lane = tid % L;
  const int cmax = np > nq ? np : nq;
  if (tid == 0) *flag = 0;
  __syncthreads();
  // one pass of segments over P's columns (np <= NT / 16, the caller's check)
  const int k = seg;
  const bool own = k < np;
  double x[R], u[R];
  double* ap = A + (int64_t)k * m;
  double* vp = V + (int64_t)k * n;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = lane + L * r;
    x[r] = own && i < m ? ap[i] : 0.0;
    u[r] = own && i < n ? vp[i] : 0.0;
  }
  for (int step = 0; step < cmax; ++step) {
    const int j = (k + step) % cmax;
    if (own && j < nq) {
      double* aq = A + (int64_t)(np + j) * m;
      double* vq = V + (int64_t)(np + j) * n;
      double y[R], z[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = lane + L * r;
        y[r] = i < m ? aq[i] : 0.0;
      }
      double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        al += x[r] * x[r];
        be += y[r] * y[r];
        ga += x[r] * y[r];
      }
      al = seg_sum(al);
      be = seg_sum(be);
      ga = seg_sum(ga);
      double c, s;
      if (jacobi_cs(al, be, ga, tol, c, s)) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = lane + L * r;
          const double xn = c * x[r] - s * y[r];
          if (i < m) aq[i] = s * x[r] + c * y[r];
          x[r] = xn;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = lane + L * r;
          z[r] = i < n ? vq[i] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = lane + L * r;
          const double un = c * u[r] - s * z[r];
          if (i < n) vq[i] = s * u[r] + c * z[r];
          u[r] = un;
        }
        if (lane == 0) *flag = 1;
      }
    }
    __syncthreads();
  }
  if (own) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = lane + L * r;
      if (i < m) ap[i] = x[r];
      if (i < n) vp[i] = u[r];
    }
  }
  __syncthreads();
  const bool any = *flag != 0;
  __syncthreads();  // every thread has read the flag before it is reused
  return any;
}

// the variant for column length m (NT threads; 512 leave the registers for
// 32 elements per lane)
template <int NT, bool CROSS>
__device__ __forceinline__ bool jacobi_pairs_m(double* A, double* V, int np, int nq, int m, int n, double tol,
                                               int* flag) {
  if constexpr (CROSS && NT <= 512) {
    if (np <= NT / SVD_SEG) {
      if (m <= 64) return jacobi_cross_reg<NT, 4>(A, V, np, nq, m, n, tol, flag);
      if (m <= 128) return jacobi_cross_reg<NT, 8>(A, V, np, nq, m, n, tol, flag);
      if (m <= 256) return jacobi_cross_reg<NT, 16>(A, V, np, nq, m, n, tol, flag);
    }
  }
  if (m <= 16) return jacobi_pairs<NT, 1, CROSS>(A, V, np, nq, m, n, tol, flag);
  if (m <= 32) return jacobi_pairs<NT, 2, CROSS>(A, V, np, nq, m, n, tol, flag);
  if (m <= 64) return jacobi_pairs<NT, 4, CROSS>(A, V, np, nq, m, n, tol, flag);
  if (m <= 128) return jacobi_pairs<NT, 8, CROSS>(A, V, np, nq, m, n, tol, flag);
  if (m <= 256) return jacobi_pairs<NT, 16, CROSS>(A, V, np, nq, m, n, tol, flag);
  if constexpr (NT <= 512)
    if (m <= 512) return jacobi_pairs<NT, 32, CROSS>(A, V, np, nq, m, n, tol, flag);
  return jacobi_pairs<NT, 0, CROSS>(A, V, np, nq, m, n, tol, flag);
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
    work[d] = src[it.src_off + e] * it.scale;
  }
}

__global__ __launch_bounds__(SVD_NT) void k_jacobi_svd(const SvdItem* __restrict__ items, double* __restrict__ work,
                                                       double* __restrict__ sig, int32_t* __restrict__ sweeps) {
  __shared__ double lds[SVD_LDS_DOUBLES];
  __shared__ int rotated;
  const SvdItem it = items[blockIdx.x];
  const int m = it.m, n = it.n;  // m >= n >= 1
  const int tid = threadIdx.x;
  // LDS only (m*n + n*n <= SVD_LDS_DOUBLES, checked by the host), so that
  // every access is an LDS instruction
  double* A = lds;
  double* V = lds + (int64_t)m * n;
  for (int64_t e = tid; e < (int64_t)m * n; e += SVD_NT) A[e] = work[it.a_off + e];
  for (int64_t e = tid; e < (int64_t)n * n; e += SVD_NT) V[e] = (e % (n + 1)) == 0 ? 1.0 : 0.0;
  __syncthreads();

  const int seg = tid / SVD_SEG, lane = tid % SVD_SEG;
  constexpr int nseg = SVD_NT / SVD_SEG;
  int sw = 0;
  for (; sw < SVD_MAX_SWEEPS; ++sw)
    if (!jacobi_pairs_m<SVD_NT, false>(A, V, n, 0, m, n, jacobi_tol(m, sw), &rotated)) break;
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
  for (int64_t e = tid; e < (int64_t)m * n; e += SVD_NT) work[it.a_off + e] = A[e];
  for (int64_t e = tid; e < (int64_t)n * n; e += SVD_NT) work[it.v_off + e] = V[e];
  if (tid == 0) sweeps[blockIdx.x] = sw;
}

// one step of the block one-sided Jacobi of a group matrix too large for
// LDS: the columns of block P = [p0,p1) (and of block Q = [q0,q1)) and
// their columns of V are staged in LDS; the pairs within P (q0 == q1) or
// the pairs across P and Q are rotated, the columns go back; rot[group] = 1
// if anything rotated
__global__ __launch_bounds__(SVD_NT_BLOCK) void k_jacobi_block(const SvdItem* __restrict__ groups,
                                                         const SvdBlockStep* __restrict__ steps,
                                                         double* __restrict__ work, int32_t* __restrict__ rot,
                                                         int sweep) {
  __shared__ double lds[SVD_LDS_DOUBLES];
  __shared__ int rotated;
  const SvdBlockStep st = steps[blockIdx.x];
  const SvdItem g = groups[st.grp];
  const int m = g.m, n = g.n;
  const int np_ = st.p1 - st.p0, w = np_ + (st.q1 - st.q0);
  const int tid = threadIdx.x;
  double* A = lds;                       // w columns of m
  double* V = lds + (int64_t)w * m;      // w columns of n
  for (int64_t e = tid; e < (int64_t)w * m; e += SVD_NT_BLOCK) {
    const int c = (int)(e / m), i = (int)(e % m);
    const int gc = c < np_ ? st.p0 + c : st.q0 + (c - np_);
    A[e] = work[g.a_off + (int64_t)gc * m + i];
  }
  for (int64_t e = tid; e < (int64_t)w * n; e += SVD_NT_BLOCK) {
    const int c = (int)(e / n), i = (int)(e % n);
    const int gc = c < np_ ? st.p0 + c : st.q0 + (c - np_);
    V[e] = work[g.v_off + (int64_t)gc * n + i];
  }
  __syncthreads();
  // one block (q0 == q1): all its pairs; two blocks: the pairs across them
  const double tol = jacobi_tol(m, sweep);
  const bool any = w == np_ ? jacobi_pairs_m<SVD_NT_BLOCK, false>(A, V, w, 0, m, n, tol, &rotated)
                            : jacobi_pairs_m<SVD_NT_BLOCK, true>(A, V, np_, w - np_, m, n, tol, &rotated);
  for (int64_t e = tid; e < (int64_t)w * m; e += SVD_NT_BLOCK) {
    const int c = (int)(e / m), i = (int)(e % m);
    const int gc = c < np_ ? st.p0 + c : st.q0 + (c - np_);
    work[g.a_off + (int64_t)gc * m + i] = A[e];
  }
  for (int64_t e = tid; e < (int64_t)w * n; e += SVD_NT_BLOCK) {
    const int c = (int)(e / n), i = (int)(e % n);
    const int gc = c < np_ ? st.p0 + c : st.q0 + (c - np_);
    work[g.v_off + (int64_t)gc * n + i] = V[e];
  }
  if (tid == 0 && any) rot[st.grp] = 1;
}

// sigma_j = |a_j| and u_j = a_j / sigma_j of a group matrix in the work buffer
__global__ __launch_bounds__(SVD_NT) void k_svd_finalize(const SvdItem* __restrict__ groups,
                                                         const int32_t* __restrict__ which,
                                                         double* __restrict__ work, double* __restrict__ sig) {
  const SvdItem g = groups[which[blockIdx.x]];
  const int m = g.m, n = g.n;
  const int seg = threadIdx.x / SVD_SEG, lane = threadIdx.x % SVD_SEG;
  constexpr int nseg = SVD_NT / SVD_SEG;
  for (int j = seg; j < n; j += nseg) {
    double* aj = work + g.a_off + (int64_t)j * m;
    double s2 = 0.0;
    for (int i = lane; i < m; i += SVD_SEG) s2 += aj[i] * aj[i];
    s2 = seg_sum(s2);
    const double sg = sqrt(s2);
    const double inv = sg > 0.0 ? 1.0 / sg : 0.0;
    for (int i = lane; i < m; i += SVD_SEG) aj[i] *= inv;
    if (lane == 0) sig[g.s_off + j] = sg;
  }
}

// V = identity for the groups listed
__global__ __launch_bounds__(256) void k_svd_init_v(const SvdItem* __restrict__ groups,
                                                    const int32_t* __restrict__ which, double* __restrict__ work) {
  const SvdItem g = groups[which[blockIdx.x]];
  const int64_t nn = (int64_t)g.n * g.n;
  for (int64_t e = threadIdx.x; e < nn; e += 256) work[g.v_off + e] = (e % (g.n + 1)) == 0 ? 1.0 : 0.0;
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
    out[it.dst_off + e] = work[it.src_off + r + (int64_t)col.col * it.ld] * col.scale * it.scale;
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

void launch_jacobi_block(const SvdItem* dgroups, const SvdBlockStep* dsteps, int nsteps, double* work, int32_t* rot,
                         int sweep, hipStream_t s) {
  if (nsteps <= 0) return;
  hipLaunchKernelGGL(k_jacobi_block, dim3(nsteps), dim3(SVD_NT_BLOCK), 0, s, dgroups, dsteps, work, rot, sweep);
  QSX_HIP(hipGetLastError());
}

void launch_svd_finalize(const SvdItem* dgroups, const int32_t* dwhich, int n, double* work, double* sig,
                         hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_svd_finalize, dim3(n), dim3(SVD_NT), 0, s, dgroups, dwhich, work, sig);
  QSX_HIP(hipGetLastError());
}

void launch_svd_init_v(const SvdItem* dgroups, const int32_t* dwhich, int n, double* work, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_svd_init_v, dim3(n), dim3(256), 0, s, dgroups, dwhich, work);
  QSX_HIP(hipGetLastError());
}

void launch_svd_scatter(const SvdScatter* ditems, int nitems, const SvdCol* dcols, const double* work, double* out,
                        hipStream_t s) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL(k_svd_scatter, dim3(nitems), dim3(256), 0, s, ditems, dcols, work, out);
  QSX_HIP(hipGetLastError());
}

}  // namespace qsx
