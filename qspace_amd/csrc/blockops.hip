// blockops.hip -- whole-QSpace block operations on the device: the block
// permutation behind permute_to / QSpace::permute / permuteQS and the block
// linear combination behind plusQS.
//
// Reference: QSpace::permute (Source/QSpace.hh:2721-2780) permutes every
// block with wbperm_helper (Source/wbarray.hh:1846-1935,
// Source/wbarray.cc:4454-4600): dimensions that stay adjacent are fused, small
// blocks take a gather path, large ones a cache-line tiled transpose.
// plusQS (Source/plusQS.cc:17-55) adds blocks with equal labels
// (QSpace::plus_plain, Source/QSpace.cc:3593-3705; Cat + MakeUnique,
// Source/QSpace.cc:3075-3140,3807-3905).
//
// Device design.  The host fuses each block's dimensions (unit dims dropped,
// source-adjacent dims that stay adjacent merged), so most permutations
// become a copy of contiguous runs or a batch of 2-D transposes.  One launch
// covers all blocks: a list of work items, each a chunk of one block.
//  * runs (fused perm[0] == 0, or a block too narrow to tile):
//    destination-linear chunks; each element's source offset follows from its
//    run index (destination-order outer dims); writes are unit-stride, and
//    so are reads when the source's fast dim stays first;
//  * tiles (fused perm[0] != 0): 32 x 32 tiles of (source dim 0, the
//    destination dims in front of it flattened) staged through LDS, so the
//    read is unit-stride along the source's fast dim and the write
//    unit-stride in the destination; the dims behind it are a batch index
//    decoded once per tile.
// The linear combination is a chunked streaming kernel; its products and sum
// are rounded separately (no contraction into an FMA), so the result is the
// same double as a*x + b*y evaluated in that order on the host.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "qsx_internal.hpp"

namespace qsx {

namespace {

constexpr int NT = 256;
constexpr int TILE = 32;

__global__ __launch_bounds__(NT) void k_perm(const PermDesc* __restrict__ descs, const PermItem* __restrict__ items,
                                             const double* __restrict__ src, double* __restrict__ dst) {
  __shared__ double t[TILE][TILE + 1];
  const PermItem it = items[blockIdx.x];
  const PermDesc& d = descs[it.blk];
  const int rank = d.rank;
  const double sc = d.scale;
  const double* s = src + d.src_off;
  double* o = dst + d.dst_off;
  if (d.kind == 0) {
    // destination-linear chunk: runs of nd[0] elements, contiguous in the
    // destination, at source stride ss[0] (1 unless the block is too narrow
    // for the tiled transpose)
    const uint32_t L = (uint32_t)d.nd[0];
    const int64_t s0 = d.ss[0];
    const uint32_t e0 = (uint32_t)it.a, ne = (uint32_t)it.b;
    for (uint32_t e = threadIdx.x; e < ne; e += NT) {
      const uint32_t lin = e0 + e;
      uint32_t run = lin / L;
      const uint32_t x = lin - run * L;
      int64_t off = (int64_t)x * s0;
      for (int k = 1; k < rank; ++k) {
        const uint32_t n = (uint32_t)d.nd[k];
        const uint32_t q = run / n;
        off += (int64_t)(run - q * n) * d.ss[k];
        run = q;
      }
      o[lin] = s[off] * sc;
    }
    return;
  }
  // tile: rows i = source dim 0 (destination dim di), columns j = the
  // destination dims before di flattened (destination-contiguous); the
  // dims after di are a batch index
  const int di = d.kind - 1;
  const int ni = d.nd[di];
  const int nj = (int)d.ds[di];
  const int i0 = it.a * TILE, j0 = it.b * TILE;
  int64_t sb = 0, db = 0;
  {
    uint32_t c = (uint32_t)it.c;
    for (int k = di + 1; k < rank; ++k) {
      const uint32_t n = (uint32_t)d.nd[k];
      const uint32_t q = c / n;
      const int64_t idx = c - q * n;
      sb += idx * d.ss[k];
      db += idx * d.ds[k];
      c = q;
    }
  }
  const int64_t dsi = d.ds[di];
  const int tx = threadIdx.x & (TILE - 1), ty = threadIdx.x / TILE;
#pragma unroll
  for (int r = 0; r < TILE; r += NT / TILE) {
    const int i = i0 + tx, j = j0 + ty + r;
    if (i < ni && j < nj) {
      // source offset of column j (mixed-radix decode over dims 0..di-1)
      int64_t sj = 0;
      uint32_t rem = (uint32_t)j;
      for (int k = 0; k < di - 1; ++k) {
        const uint32_t n = (uint32_t)d.nd[k];
        const uint32_t q = rem / n;
        sj += (int64_t)(rem - q * n) * d.ss[k];
        rem = q;
      }
      sj += (int64_t)rem * d.ss[di - 1];
      t[ty + r][tx] = s[sb + i + sj];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < TILE; r += NT / TILE) {
    const int i = i0 + ty + r, j = j0 + tx;
    if (i < ni && j < nj) o[db + j + i * dsi] = t[tx][ty + r] * sc;
  }
}

#pragma clang fp contract(off)
__global__ __launch_bounds__(NT) void k_axpby(const AxpbyItem* __restrict__ items, const double* __restrict__ A,
                                              const double* __restrict__ B, double* __restrict__ C) {
  const AxpbyItem it = items[blockIdx.x];
  const bool ha = it.a_off >= 0, hb = it.b_off >= 0;
  for (int e = threadIdx.x; e < it.n; e += NT) {
    double v;
    if (ha && hb) {
      const double x = it.fa * A[it.a_off + e];
      const double y = it.fb * B[it.b_off + e];
      v = x + y;
    } else if (ha) {
      v = it.fa * A[it.a_off + e];
    } else {
      v = it.fb * B[it.b_off + e];
    }
    C[it.c_off + e] = v;
  }
}

}  // namespace

void launch_perm_plan(const PermPlan& pp, const double* src, double* dst, hipStream_t s) {
  if (pp.items.empty()) return;
  if (!pp.ddescs || !pp.ditems) fail(QSX_ERR_ARG, "block permutation plan not uploaded");
  hipLaunchKernelGGL(k_perm, dim3((unsigned)pp.items.size()), dim3(NT), 0, s, pp.ddescs, pp.ditems, src, dst);
  QSX_HIP(hipGetLastError());
}

void launch_axpby(const AxpbyItem* ditems, int nitems, const double* A, const double* B, double* C, hipStream_t s) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL(k_axpby, dim3(nitems), dim3(NT), 0, s, ditems, A, B, C);
  QSX_HIP(hipGetLastError());
}

}  // namespace qsx
