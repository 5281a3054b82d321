// wave.hip -- the batched data half with one wave per work item.
//
// Replaces the per-output-block loop of contractDATA_group
// (Source/QSpace_aux.cc:27-355: per pair a dgemm into a temporary and
// Ck += afac*Ci).  Block dims in symmetric tensor networks are small (a few
// tens), so the unit of work is sized to the v_mfma_f64_16x16x4f64 tile, not
// to a 64x64 workgroup tile:
//  * a work item is one (16*fm) x (16*fn) tile (fm, fn in 1..4) of one output
//    block over a k-range of its pairs (reference pair order), run by ONE
//    wave; the tile shape is chosen per block, so a 21x21 block pays 32x32,
//    and the k padding is to a multiple of 4 per pair;
//  * no LDS: each lane loads its MFMA operand elements straight from memory
//    (A: lane l holds row l%16, k l/16; B: k l/16, col l%16), with per-pair
//    row / column strides, so an operand stored M x K or K x M (N x K or
//    K x N) takes the same path; afac is applied to the smaller operand side;
//  * the k-loop runs in chunks of KS k-steps of 4 and double-buffers the
//    operand registers: the next chunk's loads (also across pair boundaries)
//    are in flight while the current chunk's MFMAs issue;
//  * operand loads are raw buffer loads; a lane whose element lies outside
//    the block (row >= M, col >= N, k beyond the pair) gets an offset past
//    the buffer range, for which the hardware returns 0 without a memory
//    access, so padded rows, columns and k-steps contribute exact zeros;
//  * items are issued longest first, G waves (independent items, or the
//    k-shares of one grouped tile) per workgroup; split tiles store partial
//    tiles to scratch and the reduce kernel sums them in chunk order
//    (deterministic, no atomics).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "qsx_internal.hpp"

namespace qsx {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;          // large-tile kernel and reduce: 4 waves per workgroup
constexpr int kMaxWaves = 16;    // small-tile kernel: up to 16 waves per workgroup
constexpr int kSlot = 64 * 64;   // scratch doubles per partial tile

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Operand loads are raw buffer loads: a lane whose element is outside the
// pair's block (row >= M, col >= N or k beyond the pair) gets an offset past
// the buffer's range, for which the hardware returns 0 without a memory
// access -- no garbage, no out-of-bounds read, no select after the load.
constexpr uint32_t kInv = 0x80000000u;   // lane offset of an element outside the block
constexpr int kRange = 0x40000000;       // buffer range (bytes); blocks are smaller
constexpr int kRsrcWord3 = 0x00020000;   // raw buffer, 32-bit dwords (gfx9 family)

__device__ __forceinline__ double bload(const char* base, uint32_t off) {
  __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, kRange, kRsrcWord3);
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// the k-cursor of one wave: pair p, first k of the next chunk to issue, the
// pair's addressing (uniform), the lanes' element offsets, and the next
// pair's descriptor (loaded one pair ahead, so a pair boundary does not wait
// for a descriptor load)
template <int FM, int FN>
struct Cursor {
  int p, k, kend;
  const char* pa;      // A element (0, k) (bytes)
  const char* pb;      // B element (k, 0)
  int64_t sa4, sb4;    // bytes between k-steps of 4
  double af;
  uint32_t oa[FM];     // lane byte offsets (row, l/16) from pa; kInv outside
  uint32_t ob[FN];
  PairDesc nx;         // descriptor of pair p+1 (if p < item's last pair)
};

template <int FM, int FN>
__device__ __forceinline__ void set_pair(Cursor<FM, FN>& c, const PairDesc& pd, const PairDesc* __restrict__ pairs,
                                         int p, int k, const WorkItem& w, const double* A, const double* B,
                                         const double* X) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
  c.p = p;
  c.k = k;
  c.kend = p == w.p1 ? w.k1 : pd.K;
  c.af = pd.afac;
  // A: element (i, k) at i*sra + k*ska ; B: element (k, j) at k*skb + j*scb
  const int64_t sra = (pd.flags & 1) ? pd.lda : 1, ska = (pd.flags & 1) ? 1 : pd.lda;
  const int64_t skb = (pd.flags & 2) ? pd.ldb : 1, scb = (pd.flags & 2) ? 1 : pd.ldb;
  c.sa4 = 4 * ska * 8;
  c.sb4 = 4 * skb * 8;
  // flags bit 2 / 3: the A / B operand block lies in X (the input vector)
  // instead of the launch's A / B buffer
  const double* Ab = (pd.flags & kPairAinX) ? X : A;
  const double* Bb = (pd.flags & kPairBinX) ? X : B;
  c.pa = reinterpret_cast<const char*>(Ab + pd.a_off + (int64_t)k * ska);
  c.pb = reinterpret_cast<const char*>(Bb + pd.b_off + (int64_t)k * skb);
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int r = w.m0 + 16 * i + lr;
    c.oa[i] = r < w.M ? (uint32_t)(((int64_t)r * sra + lq * ska) * 8) : kInv;
  }
#pragma unroll
  for (int j = 0; j < FN; ++j) {
    const int col = w.n0 + 16 * j + lr;
    c.ob[j] = col < w.N ? (uint32_t)(((int64_t)col * scb + lq * skb) * 8) : kInv;
  }
  if (p < w.p1) c.nx = pairs[p + 1];
}

// move the cursor to the next chunk of the item; false at its end
template <int FM, int FN, int KS>
__device__ __forceinline__ bool advance(Cursor<FM, FN>& c, const PairDesc* __restrict__ pairs, const WorkItem& w,
                                        const double* A, const double* B, const double* X) {
  if (c.k + 4 * KS < c.kend) {
    c.k += 4 * KS;
    c.pa += KS * c.sa4;
    c.pb += KS * c.sb4;
    return true;
  }
  if (c.p >= w.p1) return false;
  const PairDesc pd = c.nx;
  set_pair<FM, FN>(c, pd, pairs, c.p + 1, 0, w, A, B, X);
  return true;
}

// issue the loads of one chunk with kv valid k (kv <= 0: all zeros)
template <int FM, int FN, int KS>
__device__ __forceinline__ void issue(const Cursor<FM, FN>& c, int kv, double (&ra)[KS][FM], double (&rb)[KS][FN]) {
  const int lq = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const bool ok = 4 * s + lq < kv;
    const char* pa = c.pa + s * c.sa4;
    const char* pb = c.pb + s * c.sb4;
#pragma unroll
    for (int i = 0; i < FM; ++i) ra[s][i] = bload(pa, ok ? c.oa[i] : kInv);
#pragma unroll
    for (int j = 0; j < FN; ++j) rb[s][j] = bload(pb, ok ? c.ob[j] : kInv);
  }
}

template <int FM, int FN, int KS>
__device__ __forceinline__ void mma(int kv, double af, const double (&ra)[KS][FM], const double (&rb)[KS][FN],
                                    d4 (&acc)[FM][FN]) {
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    if (4 * s < kv) {  // uniform: skip k-steps past the pair's end
      double a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) a[i] = FM <= FN ? ra[s][i] * af : ra[s][i];
#pragma unroll
      for (int j = 0; j < FN; ++j) b[j] = FM <= FN ? rb[s][j] : rb[s][j] * af;
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
}

// D layout of v_mfma_f64_16x16x4f64: element v of lane l is (row l/16 + 4v,
// col l%16).  A column-major C takes it through a per-wave LDS tile
// (stage, 16 x 17 doubles): written in the D layout, read back so that each
// store covers 4 columns x 16 consecutive rows (4 x 128 B instead of 16 x
// 32 B per instruction).  LDS operations of one wave complete in order, so
// the wave needs no barrier around its private tile.
constexpr int kStageLd = 17;
constexpr int kStageElems = 16 * kStageLd;
template <int FM, int FN>
__device__ __forceinline__ void store_acc(const WorkItem& w, const d4 (&acc)[FM][FN], double* __restrict__ C,
                                          double* __restrict__ scratch, double* stage) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
  if (w.slot < 0) {
    double* c = C + w.c_off;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
#pragma unroll
        for (int v = 0; v < 4; ++v) stage[(lq + 4 * v) * kStageLd + lr] = acc[i][j][v];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int row = w.m0 + 16 * i + lr;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int cc = lq + 4 * v, col = w.n0 + 16 * j + cc;
          const double x = stage[lr * kStageLd + cc];
          if (row < w.M && col < w.N) c[row + (int64_t)w.M * col] = x;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
  } else {
    double* c = scratch + (int64_t)w.slot * kSlot;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) c[(16 * i + lq + 4 * v) + 16 * FM * (16 * j + lr)] = acc[i][j][v];
  }
}

// The k-loop of one wave over its item: a ring of D register sets; chunk
// i+D-1 is issued before chunk i's MFMAs, so D-1 chunks of loads are in
// flight behind the MFMAs (each chunk is KS k-steps of 4).  Only
// compile-time ring indices are used, so no loaded register is copied (a
// copy would wait for the loads in flight).  Once the item is exhausted the
// ring is fed empty chunks (kv = 0: out-of-range loads, no memory access)
// and the first empty chunk consumed ends the loop.
template <int FM, int FN>
struct Pipe {
  static constexpr int KS = (FM + FN) >= 6 ? 1 : (FM + FN) >= 3 ? 2 : 4;
  static_assert(KS == ((FM + FN) >= 6 ? 1 : (FM + FN) >= 3 ? 2 : 4), "chunk steps must match the work lists");
  static constexpr int D = (FM + FN) >= 6 ? 2 : (FM + FN) >= 4 ? 3 : (FM + FN) == 3 ? 4 : 6;
};

template <int FM, int FN, int D>
__device__ __forceinline__ void run_k(const WorkItem& w, const PairDesc* __restrict__ pairs,
                                      const double* __restrict__ A, const double* __restrict__ B,
                                      const double* __restrict__ X, d4 (&acc)[FM][FN]) {
  constexpr int KS = Pipe<FM, FN>::KS;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  if (w.p0 > w.p1) return;  // no pairs (zero tile or empty share of a group)
  Cursor<FM, FN> c;
  {
    // the first pair's addressing comes with the item; its afac (needed
    // only by the first MFMAs) is loaded beside the operands
    PairDesc pd;
    pd.a_off = w.fa_off;
    pd.b_off = w.fb_off;
    pd.K = w.fK;
    pd.lda = w.flda;
    pd.ldb = w.fldb;
    pd.flags = w.fflags;
    pd.afac = pairs[w.p0].afac;
    set_pair<FM, FN>(c, pd, pairs, w.p0, w.k0, w, A, B, X);
  }
  bool have = true;
  double ra[D][KS][FM], rb[D][KS][FN];
  int kv[D];
  double af[D];
#pragma unroll
  for (int d = 0; d < D - 1; ++d) {
    kv[d] = have ? min(4 * KS, c.kend - c.k) : 0;
    af[d] = c.af;
    issue<FM, FN, KS>(c, kv[d], ra[d], rb[d]);
    if (have) have = advance<FM, FN, KS>(c, pairs, w, A, B, X);
  }
  for (;;) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int nslot = (d + D - 1) % D;
      kv[nslot] = have ? min(4 * KS, c.kend - c.k) : 0;
      af[nslot] = c.af;
      issue<FM, FN, KS>(c, kv[nslot], ra[nslot], rb[nslot]);
      if (have) have = advance<FM, FN, KS>(c, pairs, w, A, B, X);
      if (kv[d] <= 0) return;
      mma<FM, FN, KS>(kv[d], af[d], ra[d], rb[d], acc);
    }
  }
}

// Partial tiles of waves 1..G-1 of a group (G = the workgroup's waves),
// lane-fastest (conflict-free).  Groups are formed only for tiles with
// G * fragments <= kGroupBudget.  The same memory holds each wave's store
// staging tile (wave w at w * kStageElems; after a group's reduction only
// wave 0 stores).
constexpr int kGroupBudget = 16;
constexpr int kLdsDoubles = (kGroupBudget - 1) * 4 * 64 > kMaxWaves * kStageElems ? (kGroupBudget - 1) * 4 * 64
                                                                                  : kMaxWaves * kStageElems;
struct GroupLds {
  double buf[kLdsDoubles];
};

template <int FM, int FN>
__device__ __forceinline__ void run_item(const WorkItem& w, const PairDesc* __restrict__ pairs,
                                         const double* __restrict__ A, const double* __restrict__ B,
                                         const double* __restrict__ X, double* __restrict__ C,
                                         double* __restrict__ scratch, GroupLds& lds) {
  d4 acc[FM][FN];
  // an item of at most two chunks (flag set by the work-list builder) runs a
  // two-deep ring: a deeper ring would issue empty chunks' loads (out of
  // range, no memory access, but address-unit work) for nothing
  if (w.pad & kShortItem)
    run_k<FM, FN, 2>(w, pairs, A, B, X, acc);
  else
    run_k<FM, FN, Pipe<FM, FN>::D>(w, pairs, A, B, X, acc);
  if constexpr (FM * FN <= 4) {
    if (w.group) {  // uniform over the workgroup: all its waves share this tile
      // the partials are summed in wave order 0, 1, ..., G-1: one fixed order
      // per element, deterministic
      const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, G = blockDim.x >> 6;
      constexpr int F = FM * FN * 4;
      if (wave > 0) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) lds.buf[((wave - 1) * F + (i * FN + j) * 4 + v) * 64 + lane] = acc[i][j][v];
      }
      __syncthreads();
      if (wave > 0) return;
      for (int q = 0; q < G - 1; ++q)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[i][j][v] += lds.buf[(q * F + (i * FN + j) * 4 + v) * 64 + lane];
      store_acc<FM, FN>(w, acc, C, scratch, lds.buf);
      return;
    }
  }
  store_acc<FM, FN>(w, acc, C, scratch, lds.buf + (threadIdx.x >> 6) * kStageElems);
}

// MAXF: the largest fragment count per tile side the kernel handles.  Tiles
// of at most 2x2 fragments run in a kernel whose register budget allows 4
// waves per SIMD (every workgroup of a stage resident at once); larger
// tiles need the 4x4 variant's accumulators.
template <int MAXF>
__device__ __forceinline__ void wave_dispatch(const WorkItem* __restrict__ items, int nitems,
                                              const PairDesc* __restrict__ pairs, const double* __restrict__ A,
                                              const double* __restrict__ B, const double* __restrict__ X,
                                              double* __restrict__ C, double* __restrict__ scratch) {
  __shared__ GroupLds lds;
  const int wid = uniform(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (wid >= nitems) return;
  const WorkItem w = items[wid];
  if (w.slot == -2) return;  // padding (never part of a group)
  switch (w.shape - SH_WAVE) {
#define QSX_WAVE_CASE(fm, fn) \
  case ((fm) - 1) * 4 + (fn) - 1: \
    if constexpr ((fm) <= MAXF && (fn) <= MAXF) run_item<fm, fn>(w, pairs, A, B, X, C, scratch, lds); \
    break;
    QSX_WAVE_CASE(1, 1) QSX_WAVE_CASE(1, 2) QSX_WAVE_CASE(1, 3) QSX_WAVE_CASE(1, 4)
    QSX_WAVE_CASE(2, 1) QSX_WAVE_CASE(2, 2) QSX_WAVE_CASE(2, 3) QSX_WAVE_CASE(2, 4)
    QSX_WAVE_CASE(3, 1) QSX_WAVE_CASE(3, 2) QSX_WAVE_CASE(3, 3) QSX_WAVE_CASE(3, 4)
    QSX_WAVE_CASE(4, 1) QSX_WAVE_CASE(4, 2) QSX_WAVE_CASE(4, 3) QSX_WAVE_CASE(4, 4)
#undef QSX_WAVE_CASE
    default: break;
  }
}

__global__ __launch_bounds__(64 * kMaxWaves) __attribute__((amdgpu_waves_per_eu(4)))
void k_wave_s(const WorkItem* __restrict__ items, int nitems, const PairDesc* __restrict__ pairs,
              const double* __restrict__ A, const double* __restrict__ B, const double* __restrict__ X,
              double* __restrict__ C, double* __restrict__ scratch) {
  wave_dispatch<2>(items, nitems, pairs, A, B, X, C, scratch);
}

__global__ __launch_bounds__(NT) void k_wave_l(const WorkItem* __restrict__ items, int nitems,
                                               const PairDesc* __restrict__ pairs, const double* __restrict__ A,
                                               const double* __restrict__ B, const double* __restrict__ X,
                                               double* __restrict__ C, double* __restrict__ scratch) {
  wave_dispatch<4>(items, nitems, pairs, A, B, X, C, scratch);
}

// sum the chunk partials of split tiles in chunk order (fixed order:
// deterministic); the kernel boundary orders the partial stores before
__global__ __launch_bounds__(NT) void k_wave_reduce(const ReduceItem* __restrict__ reds,
                                                    const double* __restrict__ scratch, double* __restrict__ C) {
  const ReduceItem r = reds[blockIdx.x];
  int tm, tn;
  shape_dims(r.shape, tm, tn);
  const int mr = min(tm, r.M - r.m0), nc = min(tn, r.N - r.n0);
  double* c = C + r.c_off;
  for (int e = threadIdx.x; e < mr * nc; e += NT) {
    const int row = e % mr, col = e / mr;
    const double* s = scratch + (int64_t)r.slot0 * kSlot + row + tm * col;
    double v = 0.0;
    for (int q0 = 0; q0 < r.nslot; q0 += 8) {
      double x[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) x[q] = (q0 + q < r.nslot) ? s[(int64_t)(q0 + q) * kSlot] : 0.0;
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (q0 + q < r.nslot) v += x[q];
    }
    c[(r.m0 + row) + (int64_t)r.M * (r.n0 + col)] = v;
  }
}

}  // namespace

void launch_wave_reduce(const ReduceItem* dreds, int nreds, const double* scratch, double* C, hipStream_t s) {
  if (nreds <= 0) return;
  hipLaunchKernelGGL(k_wave_reduce, dim3(nreds), dim3(NT), 0, s, dreds, scratch, C);
  QSX_HIP(hipGetLastError());
}

void launch_wave(const WorkItem* ditems, int nitems, int maxf, int wg_waves, const ReduceItem* dreds, int nreds,
                 const PairDesc* dpairs, const double* A, const double* B, const double* X, double* C,
                 double* scratch, hipStream_t s) {
  if (nitems > 0) {
    if (maxf <= 2) {
      if (wg_waves < 1 || wg_waves > kMaxWaves)
        fail(QSX_ERR_ARG, "wave engine: %d waves per workgroup", wg_waves);
      const int nb = (nitems + wg_waves - 1) / wg_waves;
      hipLaunchKernelGGL(k_wave_s, dim3(nb), dim3(64 * wg_waves), 0, s, ditems, nitems, dpairs, A, B, X, C, scratch);
    } else {
      const int nb = (nitems + NT / 64 - 1) / (NT / 64);
      hipLaunchKernelGGL(k_wave_l, dim3(nb), dim3(NT), 0, s, ditems, nitems, dpairs, A, B, X, C, scratch);
    }
    QSX_HIP(hipGetLastError());
  }
  if (nreds > 0) {
    hipLaunchKernelGGL(k_wave_reduce, dim3(nreds), dim3(NT), 0, s, dreds, scratch, C);
    QSX_HIP(hipGetLastError());
  }
}

}  // namespace qsx
