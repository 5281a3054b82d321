// split.hip -- split-precision H*psi data half: FP64 operands as S BF16
// slices, products on the BF16 matrix cores with FP32 accumulation that is
// exact by construction, recombined in FP64.
//
// Method (Ozaki-type error-free splitting, restated from the published
// scheme; no reference counterpart -- moondg/qspace has no GPU path):
//  * each contraction line of an operand block (a row of A over its k-range,
//    a column of B over its k-range) is scaled by a power of two 2^-e with
//    max |x| < 2^e, so every scaled element lies in (-1, 1);
//  * the scaled value r_0 = x 2^-e is cut into slices of 7 bits:
//    q_s = rint(r_s 2^{7(s+1)}) (|q_s| <= 128, exact in BF16),
//    sigma_s = q_s 2^{-7(s+1)}, r_{s+1} = r_s - sigma_s (exact in FP64),
//    so x = 2^e (sum_s sigma_s + r_S) with |r_S| <= 2^{-7S-1};
//  * products sigma^A_s sigma^B_t with s + t = u are integers times
//    2^{-7(u+2)} of at most 14 bits; a sum of at most (u+1) * 128 of them
//    fits the 24-bit FP32 significand, so accumulating one u-group over at
//    most 128 k (8 MFMAs of k = 16) in FP32 is exact (the slices are stored
//    with their scale, so acc_u holds the group's value itself);
//  * groups u = 0..S-1 are kept (S(S+1)/2 MFMAs per k-step of 16); every
//    128 k and at each pair end the FP32 accumulators are flushed into FP64:
//    C += afac 2^{e_row + f_col} sum_u acc_u.
// The truncation of the slice expansion and of the dropped groups u >= S
// leaves a relative error of about 2^{-7S} of |row| |col| per term: S = 6
// gives ~1e-13 relative to the operand scale; S <= 5 does not reach 1e-12.
//
// Layout: an operand's slices are S planes of BF16 (uint16) parallel to its
// FP64 buffer (plane s, element i at planes[s * plane + i]); exponents are
// int32 per contraction line.  One wave computes one 16x16 output tile
// (v_mfma_f32_16x16x16bf16_1k: A lane l holds row l%16, k 4(l/16)..+3;
// B lane l holds k 4(l/16)..+3, col l%16; result element v of lane l is
// row 4(l/16)+v, col l%16 -- checked on the device with exact integers).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "qsx_internal.hpp"

namespace qsx {

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;
constexpr int kBits = 7;       // bits per slice
constexpr int kFlushK = 128;   // k per exact FP32 accumulation window
constexpr int kSlot = 64 * 64; // scratch doubles per partial tile (the wave engine's reduce reads them)

// the 16 x 16 tile of a wave (lane l holds rows 4(l/16)+v, column l%16):
// into C, or as a k-chunk's partial tile (column-major 16 x 16) into its
// scratch slot for the wave engine's reduction
__device__ __forceinline__ void store_tile(const WorkItem& w, const double (&c64)[4], double* __restrict__ C,
                                           double* __restrict__ scratch) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
  if (w.slot >= 0) {
    double* sp = scratch + (int64_t)w.slot * kSlot;
#pragma unroll
    for (int v = 0; v < 4; ++v) sp[(4 * lq + v) + 16 * lr] = c64[v];
    return;
  }
  const int col = w.n0 + lr;
  if (col >= w.N) return;
  double* c = C + w.c_off;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = w.m0 + 4 * lq + v;
    if (r < w.M) c[r + (int64_t)w.M * col] = c64[v];
  }
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// one wave per contraction line: exponent, then the S slices of each element
__global__ __launch_bounds__(NT) void k_split_lines(const SplitLine* __restrict__ lines, int nlines,
                                                    const double* __restrict__ X, uint16_t* __restrict__ planes,
                                                    int64_t plane, int S, int32_t* __restrict__ exps) {
  const int li = uniform(blockIdx.x * (NT / 64) + (threadIdx.x >> 6));
  if (li >= nlines) return;
  const int lane = threadIdx.x & 63;
  const SplitLine L = lines[li];
  double mx = 0.0;
  for (int k = lane; k < L.K; k += 64) mx = fmax(mx, fabs(X[L.off + (int64_t)k * L.sk]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  // max |x| < 2^e
  const int e = mx > 0.0 ? ilogb(mx) + 1 : 0;
  if (lane == 0) exps[L.eidx] = e;
  for (int k = lane; k < L.K; k += 64) {
    const int64_t i = L.off + (int64_t)k * L.sk;
    double r = ldexp(X[i], -e);
    for (int s = 0; s < S; ++s) {
      const double q = rint(ldexp(r, kBits * (s + 1)));
      const double sig = ldexp(q, -kBits * (s + 1));
      r -= sig;
      // sig has at most 8 significant bits: exact in FP32 and in BF16
      planes[(int64_t)s * plane + i] = (uint16_t)(__float_as_uint((float)sig) >> 16);
    }
  }
}

// 16 contraction lines per wave, 4 lanes per line (lines whose elements are
// strided while neighbouring lines are adjacent, e.g. the rows of a
// column-major block): lane l takes line l%16 of the wave's group at
// k = l/16 (mod 4), so each load instruction reads 16 adjacent elements at
// 4 k; the line maximum is reduced over the 4 lanes of a line
__device__ __forceinline__ int line_exponent(const SplitLine& L, bool ok, int kq, const double* __restrict__ X) {
  double mx = 0.0;
  if (ok)
    for (int k = kq; k < L.K; k += 4) mx = fmax(mx, fabs(X[L.off + (int64_t)k * L.sk]));
  mx = fmax(mx, __shfl_xor(mx, 16, 64));
  mx = fmax(mx, __shfl_xor(mx, 32, 64));
  return mx > 0.0 ? ilogb(mx) + 1 : 0;
}

__global__ __launch_bounds__(NT) void k_split_lines_grp(const SplitLine* __restrict__ lines, int nlines,
                                                        const double* __restrict__ X, uint16_t* __restrict__ planes,
                                                        int64_t plane, int S, int32_t* __restrict__ exps) {
  const int lane = threadIdx.x & 63;
  const int li = (blockIdx.x * (NT / 64) + (threadIdx.x >> 6)) * 16 + (lane & 15), kq = lane >> 4;
  const bool ok = li < nlines;
  SplitLine L{};
  if (ok) L = lines[li];
  const int e = line_exponent(L, ok, kq, X);
  if (!ok) return;
  if (kq == 0) exps[L.eidx] = e;
  for (int k = kq; k < L.K; k += 4) {
    const int64_t i = L.off + (int64_t)k * L.sk;
    double r = ldexp(X[i], -e);
    for (int s = 0; s < S; ++s) {
      const double q = rint(ldexp(r, kBits * (s + 1)));
      const double sig = ldexp(q, -kBits * (s + 1));
      r -= sig;
      planes[(int64_t)s * plane + i] = (uint16_t)(__float_as_uint((float)sig) >> 16);
    }
  }
}

template <int S>
__device__ __forceinline__ void split_tile(const WorkItem& w, const SplitPair* __restrict__ pairs,
                                           const uint16_t* __restrict__ Ap, int64_t apl,
                                           const int32_t* __restrict__ expA, const uint16_t* __restrict__ Bp,
                                           int64_t bpl, const int32_t* __restrict__ expB, double* __restrict__ C,
                                           double* __restrict__ scratch) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
  const int row = w.m0 + lr, col = w.n0 + lr;
  const bool rowok = row < w.M, colok = col < w.N;
  double c64[4] = {0.0, 0.0, 0.0, 0.0};
  for (int p = w.p0; p <= w.p1; ++p) {
    const SplitPair pd = pairs[p];
    const int kb = p == w.p0 ? w.k0 : 0, ke = p == w.p1 ? w.k1 : pd.K;
    int er[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = w.m0 + 4 * lq + v;
      er[v] = r < w.M ? expA[pd.ea + r] : 0;
    }
    const int fc = colok ? expB[pd.eb + col] : 0;
    f4 acc[S];
#pragma unroll
    for (int u = 0; u < S; ++u) acc[u] = f4{0.f, 0.f, 0.f, 0.f};
    int win = 0;
    for (int k = kb; k < ke; k += 16) {
      s4 a[S], b[S];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int kk = k + 4 * lq + i;
        const bool ka = rowok && kk < ke, kbok = colok && kk < ke;
        const int64_t ia = pd.a_off + (int64_t)row * pd.sra + (int64_t)kk * pd.ska;
        const int64_t ib = pd.b_off + (int64_t)kk * pd.skb + (int64_t)col * pd.scb;
#pragma unroll
        for (int s = 0; s < S; ++s) {
          a[s][i] = ka ? (short)Ap[(int64_t)s * apl + ia] : (short)0;
          b[s][i] = kbok ? (short)Bp[(int64_t)s * bpl + ib] : (short)0;
        }
      }
#pragma unroll
      for (int u = 0; u < S; ++u)
#pragma unroll
        for (int s = 0; s <= u; ++s)
          acc[u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[s], b[u - s], acc[u], 0, 0, 0);
      win += 16;
      if (win == kFlushK || k + 16 >= ke) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          double t = 0.0;
#pragma unroll
          for (int u = S - 1; u >= 0; --u) t += (double)acc[u][v];
          c64[v] += pd.afac * ldexp(t, er[v] + fc);
        }
#pragma unroll
        for (int u = 0; u < S; ++u) acc[u] = f4{0.f, 0.f, 0.f, 0.f};
        win = 0;
      }
    }
  }
  store_tile(w, c64, C, scratch);
}

template <int S>
__global__ __launch_bounds__(NT) void k_split_gemm(const WorkItem* __restrict__ items, int nitems,
                                                   const SplitPair* __restrict__ pairs,
                                                   const uint16_t* __restrict__ Ap, int64_t apl,
                                                   const int32_t* __restrict__ expA, const uint16_t* __restrict__ Bp,
                                                   int64_t bpl, const int32_t* __restrict__ expB,
                                                   double* __restrict__ C, double* __restrict__ scratch) {
  const int wid = uniform(blockIdx.x * (NT / 64) + (threadIdx.x >> 6));
  if (wid >= nitems) return;
  const WorkItem w = items[wid];
  if (w.slot == -2) return;
  split_tile<S>(w, pairs, Ap, apl, expA, Bp, bpl, expB, C, scratch);
}

// ---------------------------------------------------------------- FP32 x 3
// Three FP32 slices per element, products on the FP32 matrix cores
// (v_mfma_f32_16x16x4f32: A lane l holds row l%16 at k l/16, B lane l holds k
// l/16 at column l%16; result element v of lane l is row 4(l/16)+v, column
// l%16), FP32 accumulation, FP64 recombination:
//  * a line's scaled value r = x 2^-e in (-1, 1) is cut as sigma_0 =
//    rint(r 2^11) 2^-11 (at most 12 significant bits), sigma_1 =
//    fl32(r - sigma_0), sigma_2 = fl32(r - sigma_0 - sigma_1); the remainder
//    is below 2^-60;
//  * sigma_0 products are integers times 2^-22 below 2^22 in magnitude, so one
//    MFMA's sum of four is exact in FP32 and is added to FP64 after every
//    MFMA (u = 0);
//  * the groups u = 1 (sigma_0 sigma_1 + sigma_1 sigma_0, |.| < 2^-11) and
//    u = 2 (sigma_1 sigma_1 + sigma_0 sigma_2 + sigma_2 sigma_0) accumulate in
//    FP32 over a window of 32 k and are rounded at 2^-24 of their size: about
//    2^-35 of |row| |col| per term, the bound of this variant (the sigma_1
//    sigma_2 and sigma_2 sigma_2 products, below 2^-58, are dropped).
constexpr int kF32Win = 32;  // k per FP32 accumulation window of u = 1, 2

__global__ __launch_bounds__(NT) void k_split_lines_f32(const SplitLine* __restrict__ lines, int nlines,
                                                        const double* __restrict__ X, float* __restrict__ planes,
                                                        int64_t plane, int32_t* __restrict__ exps) {
  const int li = uniform(blockIdx.x * (NT / 64) + (threadIdx.x >> 6));
  if (li >= nlines) return;
  const int lane = threadIdx.x & 63;
  const SplitLine L = lines[li];
  double mx = 0.0;
  for (int k = lane; k < L.K; k += 64) mx = fmax(mx, fabs(X[L.off + (int64_t)k * L.sk]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  const int e = mx > 0.0 ? ilogb(mx) + 1 : 0;
  if (lane == 0) exps[L.eidx] = e;
  for (int k = lane; k < L.K; k += 64) {
    const int64_t i = L.off + (int64_t)k * L.sk;
    double r = ldexp(X[i], -e);
    const double s0 = ldexp(rint(ldexp(r, 11)), -11);
    r -= s0;
    const float s1 = (float)r;
    r -= (double)s1;
    planes[i] = (float)s0;
    planes[plane + i] = s1;
    planes[2 * plane + i] = (float)r;
  }
}

__global__ __launch_bounds__(NT) void k_split_lines_f32_grp(const SplitLine* __restrict__ lines, int nlines,
                                                            const double* __restrict__ X, float* __restrict__ planes,
                                                            int64_t plane, int32_t* __restrict__ exps) {
  const int lane = threadIdx.x & 63;
  const int li = (blockIdx.x * (NT / 64) + (threadIdx.x >> 6)) * 16 + (lane & 15), kq = lane >> 4;
  const bool ok = li < nlines;
  SplitLine L{};
  if (ok) L = lines[li];
  const int e = line_exponent(L, ok, kq, X);
  if (!ok) return;
  if (kq == 0) exps[L.eidx] = e;
  for (int k = kq; k < L.K; k += 4) {
    const int64_t i = L.off + (int64_t)k * L.sk;
    double r = ldexp(X[i], -e);
    const double s0 = ldexp(rint(ldexp(r, 11)), -11);
    r -= s0;
    const float s1 = (float)r;
    r -= (double)s1;
    planes[i] = (float)s0;
    planes[plane + i] = s1;
    planes[2 * plane + i] = (float)r;
  }
}

__global__ __launch_bounds__(NT) void k_split_gemm_f32(const WorkItem* __restrict__ items, int nitems,
                                                       const SplitPair* __restrict__ pairs,
                                                       const float* __restrict__ Ap, int64_t apl,
                                                       const int32_t* __restrict__ expA,
                                                       const float* __restrict__ Bp, int64_t bpl,
                                                       const int32_t* __restrict__ expB, double* __restrict__ C,
                                                       double* __restrict__ scratch) {
  const int wid = uniform(blockIdx.x * (NT / 64) + (threadIdx.x >> 6));
  if (wid >= nitems) return;
  const WorkItem w = items[wid];
  if (w.slot == -2) return;
  const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
  const int row = w.m0 + lr, col = w.n0 + lr;
  const bool rowok = row < w.M, colok = col < w.N;
  double c64[4] = {0.0, 0.0, 0.0, 0.0};
  for (int p = w.p0; p <= w.p1; ++p) {
    const SplitPair pd = pairs[p];
    const int kb = p == w.p0 ? w.k0 : 0, ke = p == w.p1 ? w.k1 : pd.K;
    double sc[4];
    const int fc = colok ? expB[pd.eb + col] : 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = w.m0 + 4 * lq + v;
      sc[v] = pd.afac * ldexp(1.0, (r < w.M ? expA[pd.ea + r] : 0) + fc);
    }
    f4 acc1 = f4{0.f, 0.f, 0.f, 0.f}, acc2 = f4{0.f, 0.f, 0.f, 0.f};
    int win = 0;
    for (int k = kb; k < ke; k += 4) {
      const int kk = k + lq;
      const bool ka = rowok && kk < ke, kbok = colok && kk < ke;
      const int64_t ia = pd.a_off + (int64_t)row * pd.sra + (int64_t)kk * pd.ska;
      const int64_t ib = pd.b_off + (int64_t)kk * pd.skb + (int64_t)col * pd.scb;
      const float a0 = ka ? Ap[ia] : 0.f, a1 = ka ? Ap[apl + ia] : 0.f, a2 = ka ? Ap[2 * apl + ia] : 0.f;
      const float b0 = kbok ? Bp[ib] : 0.f, b1 = kbok ? Bp[bpl + ib] : 0.f, b2 = kbok ? Bp[2 * bpl + ib] : 0.f;
      const f4 acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc1, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc1, 0, 0, 0);
      acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc2, 0, 0, 0);
      acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b2, acc2, 0, 0, 0);
      acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, b0, acc2, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; ++v) c64[v] += sc[v] * (double)acc0[v];
      win += 4;
      if (win == kF32Win || k + 4 >= ke) {
#pragma unroll
        for (int v = 0; v < 4; ++v) c64[v] += sc[v] * ((double)acc2[v] + (double)acc1[v]);
        acc1 = f4{0.f, 0.f, 0.f, 0.f};
        acc2 = f4{0.f, 0.f, 0.f, 0.f};
        win = 0;
      }
    }
  }
  store_tile(w, c64, C, scratch);
}

}  // namespace

void launch_split_lines_f32(const SplitLine* dlines, int nlines, bool per_lane, const double* X, float* planes,
                            int64_t plane, int32_t* exps, hipStream_t s) {
  if (nlines <= 0) return;
  if (per_lane) {
    const int nb = (nlines + 16 * (NT / 64) - 1) / (16 * (NT / 64));
    hipLaunchKernelGGL(k_split_lines_f32_grp, dim3(nb), dim3(NT), 0, s, dlines, nlines, X, planes, plane, exps);
  } else {
    const int nb = (nlines + NT / 64 - 1) / (NT / 64);
    hipLaunchKernelGGL(k_split_lines_f32, dim3(nb), dim3(NT), 0, s, dlines, nlines, X, planes, plane, exps);
  }
  QSX_HIP(hipGetLastError());
}

void launch_split_gemm_f32(const WorkItem* ditems, int nitems, const SplitPair* dpairs, const float* Ap, int64_t apl,
                           const int32_t* expA, const float* Bp, int64_t bpl, const int32_t* expB, double* C,
                           double* scratch, hipStream_t s) {
  if (nitems <= 0) return;
  const int nb = (nitems + NT / 64 - 1) / (NT / 64);
  hipLaunchKernelGGL(k_split_gemm_f32, dim3(nb), dim3(NT), 0, s, ditems, nitems, dpairs, Ap, apl, expA, Bp, bpl,
                     expB, C, scratch);
  QSX_HIP(hipGetLastError());
}

void launch_split_lines(const SplitLine* dlines, int nlines, bool per_lane, const double* X, uint16_t* planes,
                        int64_t plane, int S, int32_t* exps, hipStream_t s) {
  if (nlines <= 0) return;
  if (per_lane) {
    const int nb = (nlines + 16 * (NT / 64) - 1) / (16 * (NT / 64));
    hipLaunchKernelGGL(k_split_lines_grp, dim3(nb), dim3(NT), 0, s, dlines, nlines, X, planes, plane, S, exps);
  } else {
    const int nb = (nlines + NT / 64 - 1) / (NT / 64);
    hipLaunchKernelGGL(k_split_lines, dim3(nb), dim3(NT), 0, s, dlines, nlines, X, planes, plane, S, exps);
  }
  QSX_HIP(hipGetLastError());
}

void launch_split_gemm(int S, const WorkItem* ditems, int nitems, const SplitPair* dpairs, const uint16_t* Ap,
                       int64_t apl, const int32_t* expA, const uint16_t* Bp, int64_t bpl, const int32_t* expB,
                       double* C, double* scratch, hipStream_t s) {
  if (nitems <= 0) return;
  const int nb = (nitems + NT / 64 - 1) / (NT / 64);
  switch (S) {
#define QSX_SPLIT_CASE(n)                                                                                    \
  case n:                                                                                                    \
    hipLaunchKernelGGL(k_split_gemm<n>, dim3(nb), dim3(NT), 0, s, ditems, nitems, dpairs, Ap, apl, expA, Bp, \
                       bpl, expB, C, scratch);                                                               \
    break;
    QSX_SPLIT_CASE(3) QSX_SPLIT_CASE(4) QSX_SPLIT_CASE(5) QSX_SPLIT_CASE(6) QSX_SPLIT_CASE(7) QSX_SPLIT_CASE(8)
#undef QSX_SPLIT_CASE
    default: fail(QSX_ERR_ARG, "split precision: %d slices (3..8 supported)", S);
  }
  QSX_HIP(hipGetLastError());
}

}  // namespace qsx
