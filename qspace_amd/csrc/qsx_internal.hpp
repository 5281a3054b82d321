// qsx_internal.hpp -- internal types of the qsx library (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <string>
#include <memory>
#include <vector>

#include "../../include/qsx.h"

namespace qsx {

struct Error {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define QSX_HIP(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      ::qsx::fail(QSX_ERR_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

enum SymType { SYM_A = 0, SYM_SU2 = 1 };

// Host view of one tensor (structure; data pointers optional)
struct TensorDesc {
  int n = 0, r = 0, qdim = 0, nsym = 0;
  std::vector<int> symtype;
  std::vector<int> symcol;           // int offset of each symmetry inside a leg's qdim ints
  std::string qtype;
  std::vector<int32_t> Q;            // n x r*qdim
  std::vector<int64_t> dims;         // n x r
  std::vector<int64_t> offs;         // n+1 packed element offsets
  std::vector<uint64_t> itags;       // r or empty
  std::vector<double> cgw;           // n x nsym
  std::vector<int32_t> om;           // n outer multiplicities, or empty (all 1)
  bool cgw_given = false;
  const double* const* hdata = nullptr;  // host block pointers (may be null)
  const double* const* hdata_im = nullptr;  // imaginary parts (complex data) or null
  int64_t numel(int i) const { return offs[i + 1] - offs[i]; }
  int64_t total() const { return offs.empty() ? 0 : offs.back(); }
  bool has_cg() const {
    for (int t : symtype)
      if (t == SYM_SU2) return true;
    return false;
  }
  int label(int blk, int leg, int col) const { return Q[(size_t)blk * r * qdim + (size_t)leg * qdim + col]; }
  int om_of(int blk) const { return om.empty() ? 1 : om[blk]; }
  bool has_om() const { return !om.empty(); }
  // block size of one outer-multiplicity component
  int64_t slice(int blk) const { return numel(blk) / om_of(blk); }
};
// host wall-clock phases of a call (printed to stderr with verbose >= 2)
struct PhaseClock {
  int verbose;
  const char* what;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void lap(const char* phase) {
    auto n = std::chrono::steady_clock::now();
    if (verbose >= 2)
      fprintf(stderr, "qsx %s: %-28s %8.3f ms\n", what, phase,
              std::chrono::duration<double, std::milli>(n - t).count());
    t = n;
  }
};

// entry points without outer-multiplicity support refuse such tensors
void require_no_om(const TensorDesc& T, const char* what);

TensorDesc make_desc(const qsx_tensor* t, const char* name);
void parse_qtype(const char* s, int qdim, std::vector<int>& st);

// ---------------------------------------------------------------- labels
// Device label engine (labels.hip): block matching and output grouping.
struct LabelResult {
  std::vector<int32_t> Ia, Ib;   // pairs grouped by output block, reference order
  std::vector<int64_t> seg;      // nC+1
  std::vector<int32_t> QC;       // nC x mc records
  int mc = 0;                    // ints per output record
};
// QAc/QBc: contracted-label records (na x k, nb x k); QAk/QBk kept records.
LabelResult device_label_engine(const std::vector<int32_t>& QAc, const std::vector<int32_t>& QAk, int na,
                                const std::vector<int32_t>& QBc, const std::vector<int32_t>& QBk, int nb, int kc,
                                int ka, int kb, hipStream_t stream);

// ---------------------------------------------------------------- SU(2)
struct Canon;  // canonical CGT description (su2.cc)
double su2_cg_exact(int tj1, int tm1, int tj2, int tm2, int tj3, int tm3);
double su2_x3(const std::vector<int>& tja, const std::vector<int>& dira, const std::vector<int>& tjb,
              const std::vector<int>& dirb, const std::vector<int>& ia, const std::vector<int>& ib);
double su2_norm_ext(const std::vector<int>& tj, const std::vector<int>& dir);
// grow the exact-arithmetic tables for labels 2S <= max_tj up front, so that
// su2_x3 / su2_cg_exact on such labels only read them (concurrent callers)
void su2_reserve_labels(int max_tj);
// +-1: the CGT implied by the legs in order perm (new leg k = old leg
// perm[k]) relative to the transposed CGT of the block
double su2_perm_sign(const std::vector<int>& tj, const std::vector<int>& dir, const std::vector<int>& perm);
// outer-multiplicity basis of a block's SU(2) CGT space (su2.cc states the
// convention): om unit tensors, each a sparse list of (block-order
// column-major index over the legs' 2j+1 states, value); om = 0 when the
// labels admit no invariant
struct SU2Basis {
  int r = 0, om = 0;
  std::vector<int> tj, dims;
  int64_t numel = 1;
  std::vector<std::vector<std::pair<int64_t, double>>> ent;
};
std::shared_ptr<const SU2Basis> su2_om_basis(const std::vector<int>& tj, const std::vector<int>& dir);
int su2_om(const std::vector<int>& tj, const std::vector<int>& dir);
std::vector<int> su2_canon_order(const std::vector<int>& tj, const std::vector<int>& dir);
// contraction coefficients of basis tensors, x[(a*ob + b)*oc + c]
struct SU2XTensor {
  int oa = 0, ob = 0, oc = 0;
  std::vector<double> x;
};
std::shared_ptr<const SU2XTensor> su2_xom(const std::vector<int>& tja, const std::vector<int>& dira,
                                          const std::vector<int>& tjb, const std::vector<int>& dirb,
                                          const std::vector<int>& ia, const std::vector<int>& ib,
                                          const std::vector<int>& perm);

// ---------------------------------------------------------------- kernels
// Work descriptors of the grouped GEMM (gemm.hip)
struct PairDesc {
  int64_t a_off;   // element offset of the A operand block
  int64_t b_off;   // element offset of the B operand block
  int32_t K;       // contracted dimension
  int32_t lda;     // leading dim of A as stored (M if A is MxK, K if KxM)
  int32_t ldb;     // leading dim of B as stored (K if B is KxN, N if NxK)
  int32_t flags;   // bit0: A stored KxM; bit1: B stored NxK; bit2 / bit3: the A / B block lies in
                   // the launch's X buffer (the input vector) instead of its A / B buffer
  double afac;     // X-symbol x cgw x sign x output normalisation
};

constexpr int32_t kPairAinX = 4;
constexpr int32_t kPairBinX = 8;

// tile shapes of the grouped GEMM (TM x TN, 4 waves per workgroup)
enum TileShape { SH_64x64 = 0, SH_32x32 = 1, SH_16x64 = 2, SH_64x16 = 3, SH_32x64 = 4, SH_64x32 = 5, SH_NUM = 6 };
// wave tiles (one wave per item): shape SH_WAVE + (fm-1)*4 + (fn-1) is a
// (16*fm) x (16*fn) tile, fm, fn in 1..4
constexpr int SH_WAVE = 64;
__host__ __device__ inline int wave_shape(int fm, int fn) { return SH_WAVE + (fm - 1) * 4 + (fn - 1); }

__host__ __device__ inline void shape_dims(int sh, int& tm, int& tn) {
  if (sh >= SH_WAVE) {
    tm = 16 * ((sh - SH_WAVE) / 4 + 1);
    tn = 16 * ((sh - SH_WAVE) % 4 + 1);
    return;
  }
  switch (sh) {
    case SH_64x64: tm = 64; tn = 64; break;
    case SH_32x32: tm = 32; tn = 32; break;
    case SH_16x64: tm = 16; tn = 64; break;
    case SH_64x16: tm = 64; tn = 16; break;
    case SH_32x64: tm = 32; tn = 64; break;
    default: tm = 64; tn = 32; break;
  }
}

// one work item: a TMxTN tile of one output block over a k-range that runs
// from pair p0 at k0 to pair p1 (inclusive) ending at k1 (exclusive)
struct WorkItem {
  int64_t c_off;     // element offset of the output block
  int32_t M, N;      // output block as MxN column-major (ldc = M)
  int32_t m0, n0;    // tile origin
  int32_t p0, k0, p1, k1;
  int32_t shape;     // TileShape
  int32_t slot;      // -1: store into C; >= 0: partial tile into scratch slot;
                     // -2 (wave engine): no work (padding of a workgroup)
  int32_t group;     // wave engine: 1 if the workgroup's 4 waves share this
                     // tile, each over its own k-range, summed through LDS
  int32_t pad;       // wave engine flags: kShortItem (at most two k-chunks)
  // wave engine: the addressing of the item's first pair (pairs[p0] but its
  // afac, which may differ between pair arrays sharing the work list), so
  // the wave issues its first operand loads without a descriptor load
  int64_t fa_off, fb_off;
  int32_t fK, flda, fldb, fflags;
};
constexpr int32_t kShortItem = 2;
// k-steps of 4 per chunk of the wave kernel's ring for a tile of fm x fn
// fragments (the kernel's Pipe<fm, fn>::KS)
__host__ __device__ inline int wave_chunk_steps(int fm, int fn) {
  return (fm + fn) >= 6 ? 1 : (fm + fn) >= 3 ? 2 : 4;
}

// sum of the partial tiles slot0 .. slot0+nslot-1 into one tile of C
struct ReduceItem {
  int64_t c_off;
  int32_t M, N, m0, n0;
  int32_t shape, slot0, nslot, pad;
};

// the items [i0, i1) one persistent workgroup runs; all of one shape
struct WorkQueue {
  int32_t i0, i1, shape, pad;
};

enum Engine { ENG_GROUPED = 0, ENG_WAVE = 1 };

struct WorkPolicy {
  int engine = ENG_WAVE;         // workgroup tiles (grouped.hip) or wave tiles (wave.hip)
  int wave_max = 2;              // wave tiles: at most wave_max x wave_max fragments of 16x16
  int split_mfma = 1 << 30;      // wave tiles: MFMAs per workgroup item before the k-range is split
  int group_mfma = 64;           // wave tiles: MFMAs per tile above which a workgroup's waves share it
  int xcd_bins = 8;              // wave tiles: XCD-aware workgroup order over this many L2s (<= 1: off)
  int group_waves = 4;           // wave tiles of <= 2x2 fragments: waves per workgroup (4, 8 or 16)
  int split_steps = 16;          // k-steps (of 16) per chunk before a tile is split
  double tile_overhead = 1024;   // per-tile cost (in MACs per k) when choosing shapes
  int shape_mask = (1 << SH_NUM) - 1;
  int slots = 1024;              // resident workgroups of the persistent grid
};

// exchange form of a policy: int32[QSX_POLICY_INTS] (see qsx.h)
inline void policy_to_ints(const WorkPolicy& p, int32_t* v) {
  v[0] = p.engine;
  v[1] = p.split_steps;
  v[2] = (int32_t)p.tile_overhead;
  v[3] = p.slots;
  v[4] = p.wave_max;
  v[5] = p.group_mfma;
  v[6] = p.split_mfma;
  v[7] = p.xcd_bins;
  v[8] = p.group_waves;
  v[9] = v[10] = v[11] = 0;
}
inline WorkPolicy policy_from_ints(const int32_t* v) {
  WorkPolicy p;
  p.engine = v[0] == ENG_GROUPED ? ENG_GROUPED : ENG_WAVE;
  p.split_steps = v[1] > 0 ? v[1] : 1;
  p.tile_overhead = v[2] >= 0 ? v[2] : 0;
  p.slots = v[3] > 0 ? v[3] : 1;
  p.wave_max = v[4] >= 1 && v[4] <= 4 ? v[4] : 4;
  p.group_mfma = v[5] > 0 ? v[5] : 1;
  p.split_mfma = v[6] > 0 ? v[6] : 1;
  p.xcd_bins = v[7] >= 0 && v[7] <= 64 ? v[7] : 8;
  p.group_waves = v[8] == 8 || v[8] == 16 ? v[8] : 4;
  return p;
}

// an output block and the range of its pairs in a pair array
struct WorkTarget {
  int64_t c_off;
  int32_t M, N;
  int64_t p0, p1;  // [p0, p1)
};

struct WorkList {
  int engine = ENG_GROUPED;
  int maxf = 4;                  // wave engine: largest fragment count per tile side
  int wg_waves = 4;              // wave engine: waves per workgroup (the group size)
  std::vector<WorkItem> items;   // in queue order (wave engine: longest first)
  std::vector<WorkQueue> queues; // one per persistent workgroup
  std::vector<ReduceItem> reds;  // one per split tile
  int64_t scratch_elems = 0;     // partial-tile scratch (doubles)
};

WorkList build_work(const std::vector<WorkTarget>& targets, const std::vector<PairDesc>& pairs, const WorkPolicy& pol);
// v_mfma_f64_16x16x4f64 instructions one launch of a wave-engine work list
// issues (padding of rows, columns and k to the tile included; 0 for the
// workgroup-tile engine)
double wave_exec_mfma(const WorkList& W, const std::vector<PairDesc>& pairs);
// persistent grouped GEMM over the queues, then (if any tile is split) the reduction
void launch_work(const WorkQueue* dqueues, int nqueues, const WorkItem* ditems, const ReduceItem* dreds, int nreds,
                 const PairDesc* dpairs, const double* A, const double* B, const double* X, double* C, double* scratch,
                 hipStream_t s);
// one wave per item, then (if any tile is split) the reduction
void launch_wave(const WorkItem* ditems, int nitems, int maxf, int wg_waves, const ReduceItem* dreds, int nreds,
                 const PairDesc* dpairs, const double* A, const double* B, const double* X, double* C,
                 double* scratch, hipStream_t s);
// the launch that matches the work list's engine; X is the buffer of pairs
// flagged kPairAinX / kPairBinX (may be null when no pair is)
inline void launch_worklist(const WorkList& W, const WorkQueue* dqueues, const WorkItem* ditems,
                            const ReduceItem* dreds, const PairDesc* dpairs, const double* A, const double* B,
                            double* C, double* scratch, hipStream_t s, const double* X = nullptr) {
  if (W.engine == ENG_WAVE)
    launch_wave(ditems, (int)W.items.size(), W.maxf, W.wg_waves, dreds, (int)W.reds.size(), dpairs, A, B, X, C,
                scratch, s);
  else
    launch_work(dqueues, (int)W.queues.size(), ditems, dreds, (int)W.reds.size(), dpairs, A, B, X, C, scratch, s);
}
void launch_copy(double* d, const double* src, int64_t n, hipStream_t s);
// device scratch of the no-allocation entry points (gemm.hip): a file-scope
// __device__ array of kNoAllocScratchBytes per device; setup takes its
// address on the current device (hipGetSymbolAddress), noalloc_scratch(dev)
// returns it (nullptr before the setup on that device)
constexpr size_t kNoAllocScratchBytes = size_t(1) << 20;
constexpr int kMaxDevices = 64;
// SkipZeroData test without allocation: at most this many blocks (offsets
// and flags in the scratch array)
constexpr int kNoAllocMaxBlocks = 65536;
static_assert(sizeof(int64_t) * (kNoAllocMaxBlocks + 1) + sizeof(int32_t) * kNoAllocMaxBlocks <= kNoAllocScratchBytes,
              "zero-test scratch exceeds the array");
void noalloc_scratch_setup();
void* noalloc_scratch(int dev);
void launch_pair_gemm_f64(const PairDesc& p, int64_t c_off, int M, int N, int beta, const double* A,
                          const double* B, double* C, hipStream_t s);
// split precision (split.hip): a contraction line of an operand block (a row
// of A or a column of B over its k-range) and a pair in slice addressing
struct SplitLine {
  int64_t off;    // element offset of the line's first element
  int64_t pbase;  // offset of the line's packed slices (quads of 4 k, S slices each)
  int32_t K;      // elements on the line
  int32_t sk;     // element stride along the line
  int32_t eidx;   // index of the line's exponent (= its index in the line list)
  int32_t pad;
};
struct SplitPair {
  int64_t pa, pb;     // packed offset of A row 0 / B column 0 (lines of a block are consecutive)
  int32_t K;
  int32_t ea, eb;     // line (exponent) index of A row 0 / B column 0 of the pair's blocks
  int32_t pad;
  double afac;
};
// per_lane: 16 lines per wave, 4 lanes per line, in two passes (exponents,
// then slices over k-chunks; neighbouring lines adjacent in memory), else one
// wave per line (lines contiguous along k); maxk bounds the lines' lengths;
// the slices go to the lines' packed storage
void launch_split_lines(const SplitLine* dlines, int nlines, int maxk, bool per_lane, const double* X,
                        uint16_t* planes, int S, int32_t* exps, hipStream_t s);
// items with slot >= 0 store partial tiles into scratch (reduced by
// launch_wave_reduce)
void launch_split_gemm(int S, const WorkItem* ditems, int nitems, const SplitPair* dpairs, const uint16_t* Ap,
                       const int32_t* expA, const uint16_t* Bp, const int32_t* expB, double* C, double* scratch,
                       hipStream_t s);
// FP32 x 3 variant: three FP32 slices per element
void launch_split_lines_f32(const SplitLine* dlines, int nlines, int maxk, bool per_lane, const double* X,
                            float* planes, int32_t* exps, hipStream_t s);
void launch_split_gemm_f32(const WorkItem* ditems, int nitems, const SplitPair* dpairs, const float* Ap,
                           const int32_t* expA, const float* Bp, const int32_t* expB, double* C, double* scratch,
                           hipStream_t s);
// sum the partial tiles of the wave engine's split items (k_wave_reduce)
void launch_wave_reduce(const ReduceItem* dreds, int nreds, const double* scratch, double* C, hipStream_t s);

// ---------------------------------------------------------------- block ops
// Block permutation of a whole QSpace (blockops.hip): per block the fused
// dims in destination order (nd), the source stride (ss) and destination
// stride (ds) of each destination dim; kind 0 = contiguous runs (fused
// perm[0] == 0), kind 1 + di = LDS-tiled transpose of source dim 0 (which
// is destination dim di) against destination dim 0.
struct PermDesc {
  int64_t src_off, dst_off;
  double scale;
  int32_t rank, kind;
  int32_t nd[QSX_MAX_RANK];
  int64_t ss[QSX_MAX_RANK];
  int64_t ds[QSX_MAX_RANK];
};
// runs: a = first destination element of the chunk, b = elements;
// tiles: a = tile row (source dim 0), b = tile column (destination dim 0),
// c = batch index over the other dims
struct PermItem {
  int32_t blk, a, b, c;
};
struct PermPlan {
  std::vector<PermDesc> descs;
  std::vector<PermItem> items;
  PermDesc* ddescs = nullptr;
  PermItem* ditems = nullptr;
  int64_t elems = 0;  // elements moved per launch
  PermPlan() = default;
  PermPlan(const PermPlan&) = delete;
  PermPlan& operator=(const PermPlan&) = delete;
  ~PermPlan();
};
// dst block b = scale[b] * permute(src block b, perm), perm[k] = source dim
// of destination dim k; dims n x r (source dims)
void perm_plan_build(PermPlan& pp, int n, int r, const int64_t* dims, const int64_t* src_offs,
                     const int64_t* dst_offs, const int* perm, const double* scale);
void perm_plan_upload(PermPlan& pp, hipStream_t s);
void launch_perm_plan(const PermPlan& pp, const double* src, double* dst, hipStream_t s);
// C[c_off + e] = fa * A[a_off + e] + fb * B[b_off + e], e < n (a_off / b_off
// < 0: that term absent)
struct AxpbyItem {
  int64_t c_off, a_off, b_off;
  double fa, fb;
  int32_t n, pad;
};
void launch_axpby(const AxpbyItem* ditems, int nitems, const double* A, const double* B, double* C, hipStream_t s);
// stable ascending sort of n records of m ints on the device (ties by index,
// the order of Wb::hpsort / groupRecs): order[i] = index of the i-th record,
// seg = starts of the runs of equal records (+ n at the end)
void device_group_records(const std::vector<int32_t>& R, int n, int m, std::vector<int32_t>& order,
                          std::vector<int64_t>& seg, hipStream_t s);
// per-block "all entries exactly zero" flags
void launch_block_zero_flags(const int64_t* doffs, int nblk, const double* data, int32_t* flags, hipStream_t s);
void launch_fill_zero(double* p, int64_t n, hipStream_t s);

// ---------------------------------------------------------------- block SVD
// (svd.hip) group matrices of the truncation, one-sided Jacobi SVD, result
// blocks.  A group matrix is stored m x n column-major with m >= n (the
// transpose of a wide matrix); it is decomposed by one workgroup in LDS
// when m*n + n*n <= SVD_LDS_DOUBLES (k_jacobi_svd takes only such items),
// by column-block steps otherwise.
constexpr int SVD_LDS_DOUBLES = 20224;
// block -> its place in a group matrix (transposed for a wide group)
struct SvdGather {
  int64_t src_off, dst_off;
  int32_t d1, d2, ld, trans;
  double scale;  // block factor (1 for abelian symmetries)
};
// one group matrix: A (m x n) at a_off, V (n x n) at v_off, singular values
// (unsorted) at s_off; on return A holds the left vectors
struct SvdItem {
  int64_t a_off, v_off, s_off;
  int32_t m, n, lds, pad;
};
struct SvdCol {
  int32_t col, pad;
  double scale;
};
// result block (rows x ncols, column-major) from the kept columns
// cols[col0 .. col0+ncols-1] of the matrix at src_off (leading dim ld)
struct SvdScatter {
  int64_t dst_off, src_off;
  int32_t rows, ncols, ld, col0;
  double scale;  // block factor on top of the column scales
};
void launch_svd_gather(const SvdGather* ditems, int nitems, const double* src, double* work, hipStream_t s);
// sweeps[i] = sweeps that rotated (jacobi_max_sweeps(): not converged)
void launch_jacobi_svd(const SvdItem* ditems, int nitems, double* work, double* sig, int32_t* sweeps, hipStream_t s);
int jacobi_max_sweeps();
void launch_svd_scatter(const SvdScatter* ditems, int nitems, const SvdCol* dcols, const double* work, double* out,
                        hipStream_t s);
// block one-sided Jacobi of group matrices beyond the LDS: one step = the
// column-block pairs (P, Q) of a round-robin over a group's column blocks,
// one workgroup each; rot[grp] set when a pair rotated
struct SvdBlockStep {
  int32_t grp, p0, p1, q0, q1, pad;
};
void launch_jacobi_block(const SvdItem* dgroups, const SvdBlockStep* dsteps, int nsteps, double* work, int32_t* rot,
                         int sweep, hipStream_t s);
void launch_svd_finalize(const SvdItem* dgroups, const int32_t* dwhich, int n, double* work, double* sig,
                         hipStream_t s);
void launch_svd_init_v(const SvdItem* dgroups, const int32_t* dwhich, int n, double* work, hipStream_t s);

}  // namespace qsx
