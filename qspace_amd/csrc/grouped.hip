// grouped.hip -- the batched data half: one launch runs every (output tile,
// k-chunk) work item of a contraction (or of both terms of an H*psi stage).
//
// Replaces the per-output-block loop of contractDATA_group
// (Source/QSpace_aux.cc:27-355: per pair a dgemm into a temporary and
// Ck += afac*Ci).  Design for gfx950:
//  * work item = one TMxTN tile of one output block (shape chosen per block
//    from six shapes, 64x64 ... 16x64, so thin blocks do not pad a 64x64
//    tile) over a k-range that walks the block's pairs in the reference pair
//    order, accumulating afac*A*B in registers with v_mfma_f64_16x16x4f64;
//  * the grid is persistent: one workgroup (4 waves) per resident slot walks
//    a host-built queue of items of one shape (queues balanced longest-first),
//    as one stream of k-steps: operand tiles go global -> registers -> LDS
//    (double buffered) and the next step's loads -- also across item
//    boundaries -- are in flight while the current MFMAs and the previous
//    tile's store run; one barrier per step; afac folded into the A store;
//  * long k-ranges are split into chunks that run in parallel workgroups;
//    each chunk writes a partial tile to scratch and a second kernel sums
//    the partials in chunk order -- every output element has one fixed
//    summation order, no floating-point atomics, deterministic results.
//    (Summing in the last-arriving chunk instead needs device-scope fences,
//    which on this multi-XCD part write back / invalidate whole L2s and
//    measured slower than the extra launch.)
#include <hip/hip_runtime.h>

#include <cstdint>

#include "qsx_internal.hpp"

namespace qsx {

typedef double d4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int TK = 16;
constexpr int NT = 256;  // threads per workgroup
constexpr int kSlot = 64 * 64;  // scratch doubles per partial tile (any shape)

template <int TM, int TN>
struct Lds {
  double A[2][TK][TM + 1];
  double B[2][TK][TN + 1];
};

struct StepPair {  // the pair a k-step reads from
  const double* a;
  const double* b;
  int K, lda, ldb;
  int ta, tb;
  double af;
};

__device__ __forceinline__ StepPair load_pair(const PairDesc* __restrict__ pairs, int p, const double* A,
                                              const double* B, const double* X) {
  const PairDesc pd = pairs[p];
  StepPair s;
  // flags bit 2 / 3: the operand block lies in X (the input vector)
  s.a = ((pd.flags & kPairAinX) ? X : A) + pd.a_off;
  s.b = ((pd.flags & kPairBinX) ? X : B) + pd.b_off;
  s.K = pd.K;
  s.lda = pd.lda;
  s.ldb = pd.ldb;
  s.ta = pd.flags & 1;
  s.tb = (pd.flags >> 1) & 1;
  s.af = pd.afac;
  return s;
}

template <int TM, int TN>
__device__ __forceinline__ void gload(const StepPair& s, int k0, int M, int N, int m0, int n0,
                                      double (&ra)[(TM * TK) / NT], double (&rb)[(TN * TK) / NT]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int j = 0; j < (TM * TK) / NT; ++j) {
    const int idx = tid + NT * j;
    int i, k;
    if (!s.ta) { i = idx % TM; k = idx / TM; }
    else { k = idx % TK; i = idx / TK; }
    const int gi = m0 + i, gk = k0 + k;
    double v = 0.0;
    if (gi < M && gk < s.K) v = s.ta ? s.a[gk + (int64_t)s.lda * gi] : s.a[gi + (int64_t)s.lda * gk];
    ra[j] = v;
  }
#pragma unroll
  for (int j = 0; j < (TN * TK) / NT; ++j) {
    const int idx = tid + NT * j;
    int c, k;
    if (!s.tb) { k = idx % TK; c = idx / TK; }
    else { c = idx % TN; k = idx / TN; }
    const int gc = n0 + c, gk = k0 + k;
    double v = 0.0;
    if (gc < N && gk < s.K) v = s.tb ? s.b[gc + (int64_t)s.ldb * gk] : s.b[gk + (int64_t)s.ldb * gc];
    rb[j] = v;
  }
}

template <int TM, int TN>
__device__ __forceinline__ void lstore(Lds<TM, TN>& sm, int buf, const StepPair& s,
                                       const double (&ra)[(TM * TK) / NT], const double (&rb)[(TN * TK) / NT]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int j = 0; j < (TM * TK) / NT; ++j) {
    const int idx = tid + NT * j;
    int i, k;
    if (!s.ta) { i = idx % TM; k = idx / TM; }
    else { k = idx % TK; i = idx / TK; }
    sm.A[buf][k][i] = ra[j] * s.af;
  }
#pragma unroll
  for (int j = 0; j < (TN * TK) / NT; ++j) {
    const int idx = tid + NT * j;
    int c, k;
    if (!s.tb) { k = idx % TK; c = idx / TK; }
    else { c = idx % TN; k = idx / TN; }
    sm.B[buf][k][c] = rb[j];
  }
}

// Store a finished tile: into C, or (split tiles) its valid region into the
// item's scratch slot.  D layout of v_mfma_f64_16x16x4f64: row = lane/16 + 4v,
// col = lane%16.
template <int TM, int TN, int WM, int WN>
__device__ __forceinline__ void store_tile(const WorkItem& w, d4 (&acc)[TM / (16 * WM)][TN / (16 * WN)],
                                           double* __restrict__ C, double* __restrict__ scratch) {
  constexpr int FM = TM / (16 * WM), FN = TN / (16 * WN);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM;
  double* c = w.slot < 0 ? C + w.c_off : scratch + (int64_t)w.slot * kSlot;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = wm * (TM / WM) + i * 16 + (lane >> 4) + 4 * v;
        const int col = wn * (TN / WN) + j * 16 + (lane & 15);
        if (w.m0 + row < w.M && w.n0 + col < w.N) {
          if (w.slot < 0) c[(w.m0 + row) + (int64_t)w.M * (w.n0 + col)] = acc[i][j][v];
          else c[row + TM * col] = acc[i][j][v];
        }
        acc[i][j][v] = 0.0;
      }
}

// Cursor over the k-steps of a queue of items: (item, pair, k).
struct Cursor {
  int item, p, k, kend;
};

// One persistent workgroup: every item of its queue (one tile shape), as one
// stream of k-steps.  The next step's operand loads -- possibly of the next
// item -- are in flight while the current step's MFMAs and the previous
// item's tile store run.
template <int TM, int TN, int WM, int WN>
__device__ __forceinline__ void run_queue(const WorkQueue& q, const WorkItem* __restrict__ items,
                                          const PairDesc* __restrict__ pairs, const double* __restrict__ A,
                                          const double* __restrict__ B, const double* __restrict__ X,
                                          double* __restrict__ C, double* __restrict__ scratch, void* smem) {
  static_assert(WM * WN == 4, "4 waves");
  constexpr int FM = TM / (16 * WM), FN = TN / (16 * WN);
  Lds<TM, TN>& sm = *reinterpret_cast<Lds<TM, TN>*>(smem);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int fr = lane & 15, fk = lane >> 4;
  d4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  double ra[(TM * TK) / NT], rb[(TN * TK) / NT];

  // skip leading empty items (tiles without pairs store zeros)
  int it = q.i0;
  WorkItem w = items[it];
  while (w.p0 > w.p1) {
    store_tile<TM, TN, WM, WN>(w, acc, C, scratch);
    if (++it >= q.i1) return;
    w = items[it];
  }
  Cursor cur{it, w.p0, w.k0, 0};
  StepPair sp = load_pair(pairs, cur.p, A, B, X);
  cur.kend = (cur.p == w.p1) ? w.k1 : sp.K;
  gload<TM, TN>(sp, cur.k, w.M, w.N, w.m0, w.n0, ra, rb);
  lstore<TM, TN>(sm, 0, sp, ra, rb);
  __syncthreads();
  int buf = 0;
  for (;;) {
    // next k-step: same pair, next pair of the item, or the next item
    Cursor nxt = cur;
    WorkItem nw = w;
    StepPair nsp = sp;
    bool more = true, newitem = false;
    nxt.k = cur.k + TK;
    if (nxt.k >= cur.kend) {
      if (cur.p < w.p1) {
        nxt.p = cur.p + 1;
        nxt.k = 0;
        nsp = load_pair(pairs, nxt.p, A, B, X);
        nxt.kend = (nxt.p == w.p1) ? w.k1 : nsp.K;
      } else {
        newitem = true;
        more = false;
        for (int ni = cur.item + 1; ni < q.i1; ++ni) {
          nw = items[ni];
          if (nw.p0 <= nw.p1) {
            nxt = Cursor{ni, nw.p0, nw.k0, 0};
            nsp = load_pair(pairs, nxt.p, A, B, X);
            nxt.kend = (nxt.p == nw.p1) ? nw.k1 : nsp.K;
            more = true;
            break;
          }
          // an empty item between: its zero tile is stored after the current one
        }
      }
    }
    if (more) gload<TM, TN>(nsp, nxt.k, nw.M, nw.N, nw.m0, nw.n0, ra, rb);
#pragma unroll
    for (int kk = 0; kk < TK; kk += 4) {
      double af[FM], bf[FN];
#pragma unroll
      for (int f = 0; f < FM; ++f) af[f] = sm.A[buf][kk + fk][wm * (TM / WM) + f * 16 + fr];
#pragma unroll
      for (int f = 0; f < FN; ++f) bf[f] = sm.B[buf][kk + fk][wn * (TN / WN) + f * 16 + fr];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    if (newitem) {
      store_tile<TM, TN, WM, WN>(w, acc, C, scratch);
      // zero tiles of empty items skipped over
      const int stop = more ? nxt.item : q.i1;
      for (int ei = cur.item + 1; ei < stop; ++ei) {
        WorkItem e = items[ei];
        store_tile<TM, TN, WM, WN>(e, acc, C, scratch);
      }
    }
    if (!more) break;
    lstore<TM, TN>(sm, buf ^ 1, nsp, ra, rb);
    __syncthreads();
    buf ^= 1;
    cur = nxt;
    w = nw;
    sp = nsp;
  }
}

constexpr size_t kLdsBytes = sizeof(Lds<64, 64>);

__global__ __launch_bounds__(NT) void k_grouped(const WorkQueue* __restrict__ queues, const WorkItem* __restrict__ items,
                                                const PairDesc* __restrict__ pairs, const double* __restrict__ A,
                                                const double* __restrict__ B, const double* __restrict__ X,
                                                double* __restrict__ C, double* __restrict__ scratch) {
  __shared__ __align__(16) unsigned char smem[kLdsBytes];
  const WorkQueue q = queues[blockIdx.x];
  if (q.i0 >= q.i1) return;
  switch (q.shape) {
    case SH_64x64: run_queue<64, 64, 2, 2>(q, items, pairs, A, B, X, C, scratch, smem); break;
    case SH_32x32: run_queue<32, 32, 2, 2>(q, items, pairs, A, B, X, C, scratch, smem); break;
    case SH_16x64: run_queue<16, 64, 1, 4>(q, items, pairs, A, B, X, C, scratch, smem); break;
    case SH_64x16: run_queue<64, 16, 4, 1>(q, items, pairs, A, B, X, C, scratch, smem); break;
    case SH_32x64: run_queue<32, 64, 2, 2>(q, items, pairs, A, B, X, C, scratch, smem); break;
    case SH_64x32: run_queue<64, 32, 2, 2>(q, items, pairs, A, B, X, C, scratch, smem); break;
    default: break;
  }
}

// sum the chunk partials of split tiles in chunk order (fixed order:
// deterministic); the kernel boundary orders the partial stores before
__global__ __launch_bounds__(NT) void k_reduce(const ReduceItem* __restrict__ reds, const double* __restrict__ scratch,
                                               double* __restrict__ C) {
  const ReduceItem r = reds[blockIdx.x];
  int tm, tn;
  shape_dims(r.shape, tm, tn);
  const int mr = min(tm, r.M - r.m0), nc = min(tn, r.N - r.n0);
  double* c = C + r.c_off;
  for (int e = threadIdx.x; e < mr * nc; e += NT) {
    const int row = e % mr, col = e / mr;
    const double* s = scratch + (int64_t)r.slot0 * kSlot + row + tm * col;
    // issue up to 8 partial loads before summing them in chunk order
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

void launch_work(const WorkQueue* dqueues, int nqueues, const WorkItem* ditems, const ReduceItem* dreds, int nreds,
                 const PairDesc* dpairs, const double* A, const double* B, const double* X, double* C, double* scratch,
                 hipStream_t s) {
  if (nqueues > 0) {
    hipLaunchKernelGGL(k_grouped, dim3(nqueues), dim3(NT), 0, s, dqueues, ditems, dpairs, A, B, X, C, scratch);
    QSX_HIP(hipGetLastError());
  }
  if (nreds > 0) {
    hipLaunchKernelGGL(k_reduce, dim3(nreds), dim3(NT), 0, s, dreds, scratch, C);
    QSX_HIP(hipGetLastError());
  }
}

}  // namespace qsx
