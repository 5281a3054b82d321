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
// Layout: each contraction line is stored packed, quads of 4 consecutive k
// with the S slices of each quad together (slice s of element k at
// pbase + (k/4) 4S + 4s + k%4; the tail of the last quad is zero), so a lane
// loads the slices of its row (column) and quad of k as S 8-byte runs;
// exponents are int32 per line.  One wave computes one 16x16 output tile
// (v_mfma_f32_16x16x16bf16_1k: A lane l holds row l%16, k 4(l/16)..+3;
// B lane l holds k 4(l/16)..+3, col l%16; result element v of lane l is
// row 4(l/16)+v, col l%16 -- checked on the device with exact integers), or
// a k-chunk of it whose partial tile the wave engine's reduction sums.
#include <hip/hip_runtime.h>

#include <algorithm>
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

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// position of slice s of element k of a line in its packed storage: quads of
// 4 consecutive k, each quad holding its S slices of 4 values
__device__ __forceinline__ int64_t packed_at(const SplitLine& L, int k, int s, int S) {
  return L.pbase + (int64_t)(k >> 2) * (4 * S) + 4 * s + (k & 3);
}

// BF16 slices of one element (7 bits each, exact in BF16)
__device__ __forceinline__ void slice_bf16(double x, int e, int S, const SplitLine& L, int k,
                                           uint16_t* __restrict__ P) {
  double r = ldexp(x, -e);
  for (int s = 0; s < S; ++s) {
    const double q = rint(ldexp(r, kBits * (s + 1)));
    const double sig = ldexp(q, -kBits * (s + 1));
    r -= sig;
    // sig has at most 8 significant bits: exact in FP32 and in BF16
    P[packed_at(L, k, s, S)] = (uint16_t)(__float_as_uint((float)sig) >> 16);
  }
}

// FP32 x 3 slices of one element: sigma_0 on the 2^-11 grid, sigma_1 and
// sigma_2 the FP32 roundings of the remainders
__device__ __forceinline__ void slice_f32(double x, int e, const SplitLine& L, int k, float* __restrict__ P) {
  double r = ldexp(x, -e);
  const double s0 = ldexp(rint(ldexp(r, 11)), -11);
  r -= s0;
  const float s1 = (float)r;
  r -= (double)s1;
  P[packed_at(L, k, 0, 3)] = (float)s0;
  P[packed_at(L, k, 1, 3)] = s1;
  P[packed_at(L, k, 2, 3)] = (float)r;
}

// one wave per contraction line (lines contiguous along k): exponent, then
// the slices of each element
template <typename T, bool F32>
__global__ __launch_bounds__(NT) void k_split_lines(const SplitLine* __restrict__ lines, int nlines,
                                                    const double* __restrict__ X, T* __restrict__ P, int S,
                                                    int32_t* __restrict__ exps) {
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
    const double x = X[L.off + (int64_t)k * L.sk];
    if constexpr (F32) slice_f32(x, e, L, k, (float*)P);
    else slice_bf16(x, e, S, L, k, (uint16_t*)P);
  }
}

// Lines whose elements are strided while neighbouring lines are adjacent
// (e.g. the rows of a column-major block) run in two passes over a 2-D grid
// of (16 lines, 64 k) waves, 4 lanes per line, so that each load instruction
// reads 16 adjacent elements: the exponents (chunk maxima merged by an
// integer maximum), then the slices, each lane writing a quad's slices as
// one run.
constexpr int kQuadsPerWave = 16;  // quads of 4 k per wave of both passes
constexpr int kNoExp = -100000;    // exponent of a line not yet seen non-zero

__global__ __launch_bounds__(NT) void k_exps_fill(int32_t* __restrict__ exps, int n) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < n) exps[i] = kNoExp;
}

// per (16 lines, chunk of k) wave: the chunk's maximum of each line, its
// exponent merged into the line's by an integer maximum (the exponent is
// monotonic in the maximum)
__global__ __launch_bounds__(NT) void k_split_exps_grp(const SplitLine* __restrict__ lines, int nlines,
                                                       const double* __restrict__ X, int32_t* __restrict__ exps) {
  const int lane = threadIdx.x & 63;
  const int li = (blockIdx.x * (NT / 64) + (threadIdx.x >> 6)) * 16 + (lane & 15), kq = lane >> 4;
  const bool ok = li < nlines;
  SplitLine L{};
  if (ok) L = lines[li];
  double mx = 0.0;
  if (ok) {
    const int nq = (L.K + 3) >> 2, q1 = min(nq, (int)(blockIdx.y + 1) * kQuadsPerWave);
    for (int q = (int)blockIdx.y * kQuadsPerWave + kq; q < q1; q += 4) {
      double x[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = 4 * q + j;
        x[j] = k < L.K ? fabs(X[L.off + (int64_t)k * L.sk]) : 0.0;
      }
      mx = fmax(mx, fmax(fmax(x[0], x[1]), fmax(x[2], x[3])));
    }
  }
  mx = fmax(mx, __shfl_xor(mx, 16, 64));
  mx = fmax(mx, __shfl_xor(mx, 32, 64));
  if (ok && kq == 0 && mx > 0.0) atomicMax(exps + L.eidx, ilogb(mx) + 1);
}

template <typename T, bool F32>
__global__ __launch_bounds__(NT) void k_split_quads_grp(const SplitLine* __restrict__ lines, int nlines,
                                                        const double* __restrict__ X, T* __restrict__ P, int S,
                                                        const int32_t* __restrict__ exps) {
  const int lane = threadIdx.x & 63;
  const int li = (blockIdx.x * (NT / 64) + (threadIdx.x >> 6)) * 16 + (lane & 15), kq = lane >> 4;
  if (li >= nlines) return;
  const SplitLine L = lines[li];
  const int e = exps[L.eidx] == kNoExp ? 0 : exps[L.eidx];
  const int nq = (L.K + 3) >> 2, q1 = min(nq, (int)(blockIdx.y + 1) * kQuadsPerWave);
  for (int q = (int)blockIdx.y * kQuadsPerWave + kq; q < q1; q += 4) {
    double x[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * q + j;
      x[j] = k < L.K ? X[L.off + (int64_t)k * L.sk] : 0.0;
    }
    if constexpr (F32) {
      f4 o0, o1, o2;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        double r = ldexp(x[j], -e);
        const double s0 = ldexp(rint(ldexp(r, 11)), -11);
        r -= s0;
        const float s1 = (float)r;
        r -= (double)s1;
        o0[j] = (float)s0;
        o1[j] = s1;
        o2[j] = (float)r;
      }
      f4* d = (f4*)((float*)P + L.pbase + (int64_t)q * 12);
      d[0] = o0;
      d[1] = o1;
      d[2] = o2;
    } else {
      double r[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) r[j] = ldexp(x[j], -e);
      s4* d = (s4*)((uint16_t*)P + L.pbase + (int64_t)q * 4 * S);
      for (int s = 0; s < S; ++s) {
        s4 o;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double sig = ldexp(rint(ldexp(r[j], kBits * (s + 1))), -kBits * (s + 1));
          r[j] -= sig;
          o[j] = (short)(__float_as_uint((float)sig) >> 16);
        }
        d[s] = o;
      }
    }
  }
}

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

// BF16 slices: per k-step of 16 each lane loads, for its row of A (its
// column of B) and its quad of 4 k, the S slices as S 8-byte runs of the
// packed line (v_mfma_f32_16x16x16bf16_1k: A lane l holds row l%16, k
// 4(l/16)..+3; B lane l holds k 4(l/16)..+3, column l%16)
template <int S>
__global__ __launch_bounds__(NT) void k_split_gemm(const WorkItem* __restrict__ items, int nitems,
                                                   const SplitPair* __restrict__ pairs,
                                                   const uint16_t* __restrict__ Ap, const int32_t* __restrict__ expA,
                                                   const uint16_t* __restrict__ Bp, const int32_t* __restrict__ expB,
                                                   double* __restrict__ C, double* __restrict__ scratch) {
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
    const int64_t lenq = (int64_t)((pd.K + 3) >> 2) * (4 * S);  // packed length of a line
    const s4* pa = (const s4*)(Ap + pd.pa + (int64_t)(rowok ? row : 0) * lenq);
    const s4* pb = (const s4*)(Bp + pd.pb + (int64_t)(colok ? col : 0) * lenq);
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
      const int kq = k + 4 * lq;
      const bool in = kq < ke;
      const int q = kq >> 2;
      s4 a[S], b[S];
#pragma unroll
      for (int s = 0; s < S; ++s) {
        a[s] = in && rowok ? pa[(int64_t)q * S + s] : s4{0, 0, 0, 0};
        b[s] = in && colok ? pb[(int64_t)q * S + s] : s4{0, 0, 0, 0};
      }
#pragma unroll
      for (int u = 0; u < S; ++u)
#pragma unroll
        for (int s = 0; s <= u; ++s)
          acc[u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[s], b[u - s], acc[u], 0, 0, 0);
      win += 16;
      if (win >= kFlushK || k + 16 >= ke) {
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

// ---------------------------------------------------------------- FP32 x 3
// Three FP32 slices per element, products on the FP32 matrix cores
// (v_mfma_f32_16x16x4f32: A lane l holds row l%16 at one k, B lane l holds
// that k at column l%16, the four lane groups the four k of the MFMA;
// result element v of lane l is row 4(l/16)+v, column l%16), FP32
// accumulation, FP64 recombination:
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
// Per k-step of 16 lane group g takes the quad k + 4g..+3 for the 4 MFMAs
// (MFMA j uses element j of every group's quad: the k order inside the step
// is permuted, which the sum does not see), so each lane loads its 3 slices
// as 3 16-byte runs of the packed line.
constexpr int kF32Win = 32;  // k per FP32 accumulation window of u = 1, 2

__global__ __launch_bounds__(NT) void k_split_gemm_f32(const WorkItem* __restrict__ items, int nitems,
                                                       const SplitPair* __restrict__ pairs,
                                                       const float* __restrict__ Ap, const int32_t* __restrict__ expA,
                                                       const float* __restrict__ Bp, const int32_t* __restrict__ expB,
                                                       double* __restrict__ C, double* __restrict__ scratch) {
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
    const int64_t lenq = (int64_t)((pd.K + 3) >> 2) * 12;
    const f4* pa = (const f4*)(Ap + pd.pa + (int64_t)(rowok ? row : 0) * lenq);
    const f4* pb = (const f4*)(Bp + pd.pb + (int64_t)(colok ? col : 0) * lenq);
    double sc[4];
    const int fc = colok ? expB[pd.eb + col] : 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = w.m0 + 4 * lq + v;
      sc[v] = pd.afac * ldexp(1.0, (r < w.M ? expA[pd.ea + r] : 0) + fc);
    }
    f4 acc1 = f4{0.f, 0.f, 0.f, 0.f}, acc2 = f4{0.f, 0.f, 0.f, 0.f};
    int win = 0;
    for (int k = kb; k < ke; k += 16) {
      const int kq = k + 4 * lq;
      const bool in = kq < ke;
      const int64_t q = kq >> 2;
      const f4 z = f4{0.f, 0.f, 0.f, 0.f};
      const f4 a0 = in && rowok ? pa[q * 3] : z, a1 = in && rowok ? pa[q * 3 + 1] : z,
               a2 = in && rowok ? pa[q * 3 + 2] : z;
      const f4 b0 = in && colok ? pb[q * 3] : z, b1 = in && colok ? pb[q * 3 + 1] : z,
               b2 = in && colok ? pb[q * 3 + 2] : z;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const f4 acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[j], b0[j], z, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[j], b1[j], acc1, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[j], b0[j], acc1, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[j], b1[j], acc2, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[j], b2[j], acc2, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a2[j], b0[j], acc2, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) c64[v] += sc[v] * (double)acc0[v];
      }
      win += 16;
      if (win >= kF32Win || k + 16 >= ke) {
#pragma unroll
        for (int v = 0; v < 4; ++v) c64[v] += sc[v] * ((double)acc2[v] + (double)acc1[v]);
        acc1 = z;
        acc2 = z;
        win = 0;
      }
    }
  }
  store_tile(w, c64, C, scratch);
}

template <typename T, bool F32>
void lines_launch(const SplitLine* dlines, int nlines, int maxk, bool grouped, const double* X, T* P, int S,
                  int32_t* exps, hipStream_t s) {
  if (nlines <= 0) return;
  if (grouped) {
    const int nb = (nlines + 16 * (NT / 64) - 1) / (16 * (NT / 64));
    const int ny = std::max(1, ((maxk + 3) / 4 + kQuadsPerWave - 1) / kQuadsPerWave);
    // line indices are the exponent indices 0 .. nlines-1
    hipLaunchKernelGGL(k_exps_fill, dim3((nlines + NT - 1) / NT), dim3(NT), 0, s, exps, nlines);
    hipLaunchKernelGGL(k_split_exps_grp, dim3(nb, ny), dim3(NT), 0, s, dlines, nlines, X, exps);
    QSX_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_split_quads_grp<T, F32>), dim3(nb, ny), dim3(NT), 0, s, dlines, nlines, X, P, S, exps);
  } else {
    const int nb = (nlines + NT / 64 - 1) / (NT / 64);
    hipLaunchKernelGGL((k_split_lines<T, F32>), dim3(nb), dim3(NT), 0, s, dlines, nlines, X, P, S, exps);
  }
  QSX_HIP(hipGetLastError());
}

}  // namespace

void launch_split_lines(const SplitLine* dlines, int nlines, int maxk, bool per_lane, const double* X,
                        uint16_t* planes, int S, int32_t* exps, hipStream_t s) {
  lines_launch<uint16_t, false>(dlines, nlines, maxk, per_lane, X, planes, S, exps, s);
}

void launch_split_lines_f32(const SplitLine* dlines, int nlines, int maxk, bool per_lane, const double* X,
                            float* planes, int32_t* exps, hipStream_t s) {
  lines_launch<float, true>(dlines, nlines, maxk, per_lane, X, planes, 3, exps, s);
}

void launch_split_gemm(int S, const WorkItem* ditems, int nitems, const SplitPair* dpairs, const uint16_t* Ap,
                       const int32_t* expA, const uint16_t* Bp, const int32_t* expB, double* C, double* scratch,
                       hipStream_t s) {
  if (nitems <= 0) return;
  const int nb = (nitems + NT / 64 - 1) / (NT / 64);
  switch (S) {
#define QSX_SPLIT_CASE(n)                                                                                     \
  case n:                                                                                                     \
    hipLaunchKernelGGL(k_split_gemm<n>, dim3(nb), dim3(NT), 0, s, ditems, nitems, dpairs, Ap, expA, Bp, expB, C, \
                       scratch);                                                                              \
    break;
    QSX_SPLIT_CASE(3) QSX_SPLIT_CASE(4) QSX_SPLIT_CASE(5) QSX_SPLIT_CASE(6) QSX_SPLIT_CASE(7) QSX_SPLIT_CASE(8)
#undef QSX_SPLIT_CASE
    default: fail(QSX_ERR_ARG, "split precision: %d slices (3..8 supported)", S);
  }
  QSX_HIP(hipGetLastError());
}

void launch_split_gemm_f32(const WorkItem* ditems, int nitems, const SplitPair* dpairs, const float* Ap,
                           const int32_t* expA, const float* Bp, const int32_t* expB, double* C, double* scratch,
                           hipStream_t s) {
  if (nitems <= 0) return;
  const int nb = (nitems + NT / 64 - 1) / (NT / 64);
  hipLaunchKernelGGL(k_split_gemm_f32, dim3(nb), dim3(NT), 0, s, ditems, nitems, dpairs, Ap, expA, Bp, expB, C,
                     scratch);
  QSX_HIP(hipGetLastError());
}

}  // namespace qsx
