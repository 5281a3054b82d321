// contract.hpp -- contraction plans (structure half on host+device, data
// half as grouped GEMM launches).
#pragma once

#include <functional>

#include <memory>
#include <vector>

#include "qsx_internal.hpp"

namespace qsx {

struct OperandLayout {
  bool permute = false;        // a permuted packed copy is used
  std::vector<int> perm;       // leg order of the copy (new leg k = old leg perm[k])
  bool kmajor = false;         // A: stored KxM ; B: stored NxK
};

struct CtrPlan {
  TensorDesc A, B;
  int force_layout = 0;  // bit 0: A as M x K copy, bit 1: B as N x K copy
  std::vector<int> ica, icb, ka, kb;  // contracted legs (pair order used for K), kept legs
  int conjA = 0, conjB = 0;
  int mode = QSX_MODE_BATCHED;
  int nslices = 0;
  int verbose = 1;
  std::vector<uint64_t> tagsC;        // output itags (after perm)
  std::vector<int> dirA, dirB, dirC;  // effective directions (+1 in, -1 out); dirC before perm
  // structure
  LabelResult L;                      // all matched pairs grouped by output block
  std::vector<double> afac;           // per pair of L (0 = skipped)
  // executed output
  int nC = 0, rc = 0;
  std::vector<int32_t> Qout;          // nC x rc*qdim (after perm)
  std::vector<int64_t> dimsOut;       // nC x rc (after perm)
  std::vector<int64_t> offsOut;       // nC+1
  std::vector<double> cgwOut;         // nC x nsym
  std::vector<int> blkL;              // output block -> index into L's blocks
  std::vector<int> perm;              // output permutation (identity if none)
  bool has_perm = false;
  // outer multiplicity (SU(2) CGTs of rank > 3, su2_om_basis): each pair
  // expands into one GEMM per non-zero X-symbol entry (a, b, c) on the
  // operands' components a, b into the output component c
  bool om_path = false;
  std::vector<int32_t> omOut;         // nC outer multiplicities (om path), else empty
  // GEMM work: targets are the output blocks, or (om path) every output
  // block's components
  int nT = 0;
  std::vector<int64_t> tOff;          // nT output offsets of the targets
  std::vector<PairDesc> pairs;        // executed pairs, grouped by target
  std::vector<int64_t> pseg;          // nT+1 ranges into pairs
  std::vector<int32_t> pIa, pIb;      // executed pairs' block indices
  std::vector<int32_t> M, N;          // per target matrix dims
  OperandLayout la, lb;
  bool a_prepermuted = false, b_prepermuted = false;
  double flops = 0;                   // useful flops (2MNK over executed pairs)
  // device resources
  PairDesc* dpairs = nullptr;
  WorkPolicy policy;
  WorkList work;
  WorkItem* ditems = nullptr;
  ReduceItem* dreds = nullptr;
  WorkQueue* dqueues = nullptr;
  double* dscratch = nullptr;
  double* dAperm = nullptr;
  double* dBperm = nullptr;
  double* dCtmp = nullptr;
  PermPlan ppA, ppB, ppC;            // operand copies and the output permutation
  ~CtrPlan();
};

// force: bit 0 / bit 1 -- give A / B a permuted copy in the orientation the
// GEMM loads best (A as M x K, B as N x K: 16 consecutive rows / columns per
// k) even when the native layout would do (plans of constant operands)
std::unique_ptr<CtrPlan> plan_create(const qsx_tensor* A, const int32_t* ia, int na, int conjA, const qsx_tensor* B,
                                     const int32_t* ib, int nb, int conjB, const int32_t* perm, int nperm,
                                     const qsx_opts* o, bool upload = true, int force = 0);
// build the work list for the plan's policy and upload pairs/items
void plan_upload(CtrPlan& P);
void plan_execute(CtrPlan& P, const double* dA, const double* dB, double* dC, hipStream_t s);
// permute an operand's packed device data into the plan's layout
void plan_permute_operand(const CtrPlan& P, int which, const double* src, double* dst, hipStream_t s);

// itag auto-matching (iTags::getCtr, Source/wbindex.cc:1512-1590); legs
// flagged in skipA / skipB (itag_::SetFlags of icFlags::apply) match nothing
void match_itags(const TensorDesc& A, const TensorDesc& B, int conjA, int conjB, std::vector<int32_t>& ia,
                 std::vector<int32_t>& ib, int& conjA_out, const std::vector<char>* skipA = nullptr,
                 const std::vector<char>* skipB = nullptr);
// host result of a plan: the output blocks keep[] (plan order) with the
// packed host data (and imaginary parts, or null)
qsx_tensor_out* make_out(const CtrPlan& P, const std::vector<int>& keep, const double* data,
                         const double* data_im = nullptr);
// contractQS usage #2 (cell.cc)
qsx_tensor_out* contract_cell(const qsx_cell* items, int n, const int32_t* perm, int nperm, const qsx_opts* o);
// contractQS usage #1 on host tensors (qsx_contract without the error wrapper)
qsx_tensor_out* contract_oneshot(const qsx_tensor* A, const int32_t* ia, int na, int conjA, const qsx_tensor* B,
                                 const int32_t* ib, int nb, int conjB, const int32_t* perm, int nperm,
                                 const qsx_opts* o);

hipStream_t default_stream();
// host blocks h[i] (T's dims/offsets) -> the packed device buffer d, through
// the library's pinned staging buffer; returns when the copy is complete
void upload_packed(const TensorDesc& T, const double* const* h, double* d, hipStream_t s);
// host arrays -> device arrays through the pinned staging buffer (each part
// in its own region, one DMA each, one synchronisation at the end); returns
// when the copies are complete
struct UploadPart {
  void* d;
  const void* h;
  size_t bytes;
};
void upload_many(const std::vector<UploadPart>& parts, hipStream_t s);
// n doubles of device memory -> host, through the pinned staging buffer:
// use(h) reads them (h = nullptr when n == 0) while the buffer is held, so
// use must not upload or download itself
void download_staged(const double* d, int64_t n, hipStream_t s, const std::function<void(const double*)>& use);
// n doubles of device memory -> host memory dst
void download_packed(const double* d, int64_t n, double* dst, hipStream_t s);
// a second stream of the library (default priority) for work that may
// overlap the default stream's; joined back with events
hipStream_t side_stream();
void ensure_device();
void* dev_alloc(size_t bytes);
int resident_slots();
void dev_free(void* p);

}  // namespace qsx
