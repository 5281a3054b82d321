// contract.cc -- contractQS on the device: structure half (itags, device
// label engine, X-symbols, output bookkeeping, operand layouts, GEMM work
// lists) and data half (grouped FP64 MFMA GEMM or the plain per-pair path).
//
// Reference path restated (moondg/qspace):
//   contract_plain / CONTRACT_QS      Source/contractQS.cc:505-632
//   QSpace::contract                  Source/QSpace.cc:4296-4420
//   contract_getIdxSet                Source/QSpace.cc:4131-4294
//   permute_to                        Source/QSpace.hh:2823-2865
//   contractDATA_group                Source/QSpace_aux.cc:27-355
//   SkipZeroData / SkipEmptyData      Source/QSpace.cc:3200-3305
//   QSpace::Permute                   Source/QSpace.hh:2669-2719
//   NormCGW                           Source/QSpace.cc:3727-3805

#include "contract.hpp"

#include <algorithm>
#include <cstdlib>
#include <functional>
#include <mutex>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <numeric>

namespace qsx {

void fail(int code, const char* fmt, ...) {
  char buf[768];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Error{code, buf};
}

void parse_qtype(const char* s, int qdim, std::vector<int>& st) {
  st.clear();
  std::string q = s ? s : "";
  if (q.empty()) {
    for (int i = 0; i < qdim; ++i) st.push_back(SYM_A);
    return;
  }
  size_t p = 0;
  while (p <= q.size()) {
    size_t e = q.find(',', p);
    if (e == std::string::npos) e = q.size();
    std::string t = q.substr(p, e - p);
    while (!t.empty() && t.front() == ' ') t.erase(t.begin());
    while (!t.empty() && t.back() == ' ') t.pop_back();
    if (t == "A" || t == "U1") st.push_back(SYM_A);
    else if (t == "SU2") st.push_back(SYM_SU2);
    else fail(QSX_ERR_UNSUPPORTED, "unsupported symmetry '%s' in qtype '%s' (supported: A, SU2)", t.c_str(), q.c_str());
    p = e + 1;
  }
  if ((int)st.size() != qdim)
    fail(QSX_ERR_QDIM, "qtype '%s' has %d symmetries but QDIM=%d", q.c_str(), (int)st.size(), qdim);
}

TensorDesc make_desc(const qsx_tensor* t, const char* name) {
  if (!t) fail(QSX_ERR_ARG, "null tensor %s", name);
  TensorDesc T;
  T.n = t->nblk;
  T.r = t->rank;
  T.qdim = t->qdim;
  if (T.n < 0 || T.r < 0 || T.r > QSX_MAX_RANK || T.qdim <= 0)
    fail(QSX_ERR_STRUCT, "invalid tensor %s (nblk=%d rank=%d qdim=%d)", name, T.n, T.r, T.qdim);
  if (T.n && (!t->Q || !t->dims)) fail(QSX_ERR_ARG, "tensor %s: null Q or dims", name);
  parse_qtype(t->qtype, T.qdim, T.symtype);
  T.qtype = t->qtype ? t->qtype : "";
  T.nsym = (int)T.symtype.size();
  const size_t rq = (size_t)T.r * T.qdim;
  if (T.n) {
    T.Q.assign(t->Q, t->Q + (size_t)T.n * rq);
    T.dims.assign(t->dims, t->dims + (size_t)T.n * T.r);
  }
  if (t->itags) T.itags.assign(t->itags, t->itags + T.r);
  if (t->om && T.n) {
    bool any = false;
    for (int i = 0; i < T.n; ++i) {
      if (t->om[i] < 1) fail(QSX_ERR_STRUCT, "tensor %s: outer multiplicity %d of block %d", name, t->om[i], i + 1);
      any |= t->om[i] > 1;
    }
    if (any) T.om.assign(t->om, t->om + T.n);
  }
  T.offs.assign(T.n + 1, 0);
  for (int i = 0; i < T.n; ++i) {
    int64_t s = T.om_of(i);
    for (int k = 0; k < T.r; ++k) {
      int64_t d = T.dims[(size_t)i * T.r + k];
      if (d < 0) fail(QSX_ERR_STRUCT, "tensor %s: negative block dim", name);
      s *= d;
    }
    T.offs[i + 1] = T.offs[i] + s;
  }
  T.cgw.assign((size_t)T.n * T.nsym, 1.0);
  if (t->cgw) {
    T.cgw.assign(t->cgw, t->cgw + (size_t)T.n * T.nsym);
    T.cgw_given = true;
  }
  T.hdata = t->data;
  T.hdata_im = t->data_im;
  // QIDX rows must be unique (QSpace.hh:3214-3215)
  if (T.n > 1 && rq) {
    std::vector<int> p(T.n);
    std::iota(p.begin(), p.end(), 0);
    std::sort(p.begin(), p.end(), [&](int a, int b) {
      return std::lexicographical_compare(&T.Q[a * rq], &T.Q[a * rq] + rq, &T.Q[b * rq], &T.Q[b * rq] + rq);
    });
    for (int i = 1; i < T.n; ++i)
      if (std::equal(&T.Q[p[i] * rq], &T.Q[p[i] * rq] + rq, &T.Q[p[i - 1] * rq]))
        fail(QSX_ERR_STRUCT, "tensor %s: non-unique Q records (blocks %d and %d)", name, p[i - 1] + 1, p[i] + 1);
  }
  return T;
}

void require_no_om(const TensorDesc& T, const char* what) {
  if (T.has_om()) fail(QSX_ERR_UNSUPPORTED, "%s: blocks with outer multiplicity > 1 are supported by contractions only", what);
}

// ---------------------------------------------------------------- device
// Kernel arguments in device memory (HIP_FORCE_DEV_KERNARG): the command
// processor then reads a launch's arguments from HBM instead of host memory
// over PCIe, which shortens every dispatch (the eigensolver loops are bound
// by their launches).  Set when the library is loaded, before its first HIP
// call, unless the host program chose a value itself; no effect when the
// process initialised the HIP runtime earlier.
__attribute__((constructor)) static void kernarg_in_device_memory() { setenv("HIP_FORCE_DEV_KERNARG", "1", 0); }

// the library's device and stream, set up once per process (the first call
// from any thread; the others wait)
static std::atomic<int> g_dev{-1};
static hipStream_t g_stream = nullptr;
static hipStream_t g_side = nullptr;  // independent work beside g_stream (same priority)
static std::mutex g_dev_mu;

void ensure_device() {
  if (g_dev.load(std::memory_order_acquire) >= 0) return;
  std::lock_guard<std::mutex> lk(g_dev_mu);
  if (g_dev.load(std::memory_order_relaxed) >= 0) return;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) fail(QSX_ERR_DEVICE, "no HIP device available");
  int dev = -1;
  QSX_HIP(hipGetDevice(&dev));
  QSX_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  QSX_HIP(hipStreamCreateWithFlags(&g_side, hipStreamNonBlocking));
  g_dev.store(dev, std::memory_order_release);
}

hipStream_t default_stream() {
  ensure_device();
  return g_stream;
}

hipStream_t side_stream() {
  ensure_device();
  return g_side;
}

// resident workgroups of the persistent grouped GEMM: 4 per compute unit
// (the kernel's LDS and register use allow 4 per CU)
int resident_slots() {
  static int slots = 0;
  if (!slots) {
    ensure_device();
    hipDeviceProp_t prop;
    QSX_HIP(hipGetDeviceProperties(&prop, g_dev));
    slots = 4 * std::max(1, prop.multiProcessorCount);
  }
  return slots;
}

void* dev_alloc(size_t bytes) {
  ensure_device();
  void* p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) fail(QSX_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  return p;
}

void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

CtrPlan::~CtrPlan() {
  dev_free(dpairs);
  dev_free(ditems);
  dev_free(dreds);
  dev_free(dqueues);
  dev_free(dscratch);
  dev_free(dAperm);
  dev_free(dBperm);
  dev_free(dCtmp);
}

// ---------------------------------------------------------------- itags
static inline uint64_t tag_name(uint64_t t) { return t & ~(uint64_t)0x80; }
static inline bool tag_out(uint64_t t) { return (t & 0x80) != 0; }

void match_itags(const TensorDesc& A, const TensorDesc& B, int conjA, int conjB, std::vector<int32_t>& ia,
                 std::vector<int32_t>& ib, int& conjA_out, const std::vector<char>* skipA,
                 const std::vector<char>* skipB) {
  if ((int)A.itags.size() != A.r || (int)B.itags.size() != B.r)
    fail(QSX_ERR_ITAG, "itag auto-contraction needs itags on both operands");
  conjA_out = conjA;
  int cflag = conjA ^ conjB;
  for (int it = 0; it < 2; ++it) {
    ia.clear();
    ib.clear();
    std::vector<char> mb(B.r, 0);
    int e = 0;
    const bool want_same = cflag != 0;  // sameAs sign +1 (same dir) if one side is conjugated
    for (int i = 0; i < A.r; ++i) {
      if (skipA && (*skipA)[i]) continue;
      int hit = -1;
      for (int j = 0; j < B.r; ++j) {
        if (skipB && (*skipB)[j]) continue;
        uint64_t a = A.itags[i], b = B.itags[j];
        if (tag_name(a) != tag_name(b)) continue;  // two empty names count as equal (itag_::sameAs)
        bool same = tag_out(a) == tag_out(b);
        if (same != want_same) continue;
        if (hit >= 0 || mb[j]) ++e;
        else hit = j;
      }
      if (hit >= 0) {
        ia.push_back(i);
        ib.push_back(hit);
        mb[hit] = 1;
      }
    }
    if (e) fail(QSX_ERR_ITAG, "non-unique itags in auto-contraction");
    if (!ia.empty()) {
      if (it) conjA_out = !conjA;  // "check missing conj flag" -> conj(A)
      return;
    }
    cflag = !cflag;
  }
  fail(QSX_ERR_ITAG, "empty set of matching itags");
}

// ---------------------------------------------------------------- structure
static std::vector<int> su2_col(const TensorDesc& T, int blk, int sym) {
  std::vector<int> tj(T.r);
  for (int l = 0; l < T.r; ++l) tj[l] = T.label(blk, l, sym);
  return tj;
}

// contract_getIdxSet checks + output itags (QSpace.cc:4131-4271)
static void check_and_tags(CtrPlan& P) {
  TensorDesc &A = P.A, &B = P.B;
  const int ra = A.r, rb = B.r, k = (int)P.ica.size();
  if (A.qdim != B.qdim) fail(QSX_ERR_QDIM, "incompatible objects (QDIM=%d/%d)", A.qdim, B.qdim);
  if (k != (int)P.icb.size() || !k) fail(QSX_ERR_ARG, "invalid contraction (%d <> %d legs)", k, (int)P.icb.size());
  if (A.symtype != B.symtype) {
    bool allab = true;
    for (int t : A.symtype) allab &= t == SYM_A;
    for (int t : B.symtype) allab &= t == SYM_A;
    if (!allab || (!A.qtype.empty() && !B.qtype.empty()))
      fail(QSX_ERR_QDIM, "qtype inconsistency '%s' / '%s'", A.qtype.c_str(), B.qtype.c_str());
  }
  std::vector<char> ua(ra, 0), ub(rb, 0);
  for (int i = 0; i < k; ++i) {
    if (P.ica[i] < 0 || P.ica[i] >= ra || ua[P.ica[i]]++)
      fail(QSX_ERR_ARG, "invalid contraction indices for A (rank %d)", ra);
    if (P.icb[i] < 0 || P.icb[i] >= rb || ub[P.icb[i]]++)
      fail(QSX_ERR_ARG, "invalid contraction indices for B (rank %d)", rb);
  }
  P.ka.clear();
  P.kb.clear();
  for (int i = 0; i < ra; ++i)
    if (!ua[i]) P.ka.push_back(i);
  for (int i = 0; i < rb; ++i)
    if (!ub[i]) P.kb.push_back(i);
  P.rc = (int)(P.ka.size() + P.kb.size());
  const bool cg = A.has_cg();
  if (cg && ((int)A.itags.size() != ra || (int)B.itags.size() != rb))
    fail(QSX_ERR_ITAG, "missing itags for non-abelian QSpace");
  std::vector<uint64_t> ta = A.itags, tb = B.itags;
  if (!ta.empty() && !tb.empty()) {
    int zflag = 0;
    for (int i = 0; i < k; ++i) {
      uint64_t a = ta[P.ica[i]], b = tb[P.icb[i]];
      int q = 0;
      if (tag_name(a) == tag_name(b)) q = 3;
      else if (!tag_name(b) || !tag_name(a)) q = 1;
      if (q && tag_out(a) != tag_out(b)) q = -q;
      if (q) {
        q = q > 0 ? 1 : -1;
        if (zflag && zflag != q) q = 0;
        else zflag = q;
      }
      if (!q)
        fail(QSX_ERR_ITAG, "itag mismatch at contracted leg %d (A leg %d, B leg %d)", i + 1, P.ica[i] + 1,
             P.icb[i] + 1);
    }
    if (zflag > 0 && !(P.conjA ^ P.conjB)) {
      if (P.verbose) fprintf(stderr, "WRN contract() check missing conj-flag!? applying conj(A) flag\n");
      P.conjA = 1;
    }
  }
  if (!ta.empty() && P.conjA)
    for (auto& t : ta) t ^= 0x80;
  if (!tb.empty() && P.conjB)
    for (auto& t : tb) t ^= 0x80;
  P.tagsC.clear();
  if (!ta.empty() || !tb.empty()) {
    for (int i : P.ka) P.tagsC.push_back(ta.empty() ? 0 : ta[i]);
    for (int i : P.kb) P.tagsC.push_back(tb.empty() ? 0 : tb[i]);
  }
  P.dirA.assign(ra, 1);
  P.dirB.assign(rb, 1);
  for (int i = 0; i < ra && !ta.empty(); ++i) P.dirA[i] = tag_out(ta[i]) ? -1 : 1;
  for (int i = 0; i < rb && !tb.empty(); ++i) P.dirB[i] = tag_out(tb[i]) ? -1 : 1;
  P.dirC.assign(P.rc, 1);
  for (int i = 0; i < P.rc && !P.tagsC.empty(); ++i) P.dirC[i] = tag_out(P.tagsC[i]) ? -1 : 1;
  // default cgw: external norm of the stored tensors (before conj flags)
  auto defcgw = [](TensorDesc& T) {
    if (T.cgw_given) return;
    std::vector<int> dir(T.r, 1);
    for (int i = 0; i < T.r && !T.itags.empty(); ++i) dir[i] = tag_out(T.itags[i]) ? -1 : 1;
    for (int i = 0; i < T.n; ++i)
      for (int j = 0; j < T.nsym; ++j)
        if (T.symtype[j] == SYM_SU2) T.cgw[(size_t)i * T.nsym + j] = su2_norm_ext(su2_col(T, i, j), dir);
  };
  defcgw(A);
  defcgw(B);
}

// joint ordering of the contracted legs and the operand layouts (permute_to,
// QSpace.hh:2823-2865): a zero-copy matricisation when the contracted legs
// are contiguous at the front or end of both operands in a common order,
// otherwise a permuted copy of the operand(s)
static void choose_layouts(CtrPlan& P) {
  const int ra = P.A.r, rb = P.B.r, k = (int)P.ica.size();
  auto is_front_end = [](const std::vector<int>& legs, int r, int& where) {
    std::vector<int> s(legs);
    std::sort(s.begin(), s.end());
    int k = (int)s.size();
    bool front = true, end = true;
    for (int i = 0; i < k; ++i) {
      front &= s[i] == i;
      end &= s[i] == r - k + i;
    }
    where = end ? 2 : (front ? 1 : 0);  // prefer 'end' (MxK / NxK with kept legs first)
    return front || end;
  };
  auto increasing = [](const std::vector<int>& v) {
    for (size_t i = 1; i < v.size(); ++i)
      if (v[i] <= v[i - 1]) return false;
    return true;
  };
  std::vector<int> ordA(k), ordB(k);
  std::iota(ordA.begin(), ordA.end(), 0);
  std::iota(ordB.begin(), ordB.end(), 0);
  std::stable_sort(ordA.begin(), ordA.end(), [&](int x, int y) { return P.ica[x] < P.ica[y]; });
  std::stable_sort(ordB.begin(), ordB.end(), [&](int x, int y) { return P.icb[x] < P.icb[y]; });
  auto apply = [&](const std::vector<int>& o) {
    std::vector<int> a(k), b(k);
    for (int i = 0; i < k; ++i) {
      a[i] = P.ica[o[i]];
      b[i] = P.icb[o[i]];
    }
    P.ica = a;
    P.icb = b;
  };
  int wa = 0, wb = 0;
  bool fa = is_front_end(P.ica, ra, wa), fb = is_front_end(P.icb, rb, wb);
  if (P.force_layout & 1) fa = false;
  if (P.force_layout & 2) fb = false;
  // native layouts: A with contracted legs at its end is MxK, at its front
  // KxM; B with contracted legs at its front is KxN, at its end NxK
  auto front_of = [&](const std::vector<int>& legs) {
    std::vector<int> s(legs);
    std::sort(s.begin(), s.end());
    bool front = true;
    for (int i = 0; i < (int)s.size(); ++i) front &= s[i] == i;
    return front;
  };
  auto set_native = [&](void) {
    P.la.permute = false;
    P.la.kmajor = (wa == 1);
    P.lb.permute = false;
    P.lb.kmajor = !front_of(P.icb);
  };
  if (fa && fb) {
    std::vector<int> ica0 = P.ica, icb0 = P.icb;
    apply(ordA);
    if (increasing(P.icb)) {
      set_native();
      return;
    }
    P.ica = ica0;
    P.icb = icb0;
    apply(ordB);
    if (increasing(P.ica)) {
      set_native();
      return;
    }
    P.ica = ica0;
    P.icb = icb0;
  }
  // permuted copies: A as [kept | contracted] (M x K), B as [kept |
  // contracted] (N x K), the orientations whose MFMA fragments are 16
  // consecutive elements per k
  if (fa) {  // keep A, permute B to [kept | contracted (A order)]
    apply(ordA);
    P.la.permute = false;
    P.la.kmajor = (wa == 1);
    P.lb.permute = true;
    P.lb.perm.assign(P.kb.begin(), P.kb.end());
    P.lb.perm.insert(P.lb.perm.end(), P.icb.begin(), P.icb.end());
    P.lb.kmajor = true;
    return;
  }
  if (fb) {  // keep B, permute A to [kept | contracted (B order)]
    apply(ordB);
    P.lb.permute = false;
    P.lb.kmajor = !front_of(P.icb);
    P.la.permute = true;
    P.la.perm.assign(P.ka.begin(), P.ka.end());
    P.la.perm.insert(P.la.perm.end(), P.ica.begin(), P.ica.end());
    P.la.kmajor = false;
    return;
  }
  apply(ordA);  // permute both
  P.la.permute = true;
  P.la.perm.assign(P.ka.begin(), P.ka.end());
  P.la.perm.insert(P.la.perm.end(), P.ica.begin(), P.ica.end());
  P.la.kmajor = false;
  P.lb.permute = true;
  P.lb.perm.assign(P.kb.begin(), P.kb.end());
  P.lb.perm.insert(P.lb.perm.end(), P.icb.begin(), P.icb.end());
  P.lb.kmajor = true;
}

static void build_structure(CtrPlan& P, const int32_t* perm, int nperm, const double* fA, const double* fB) {
  PhaseClock clk{P.verbose, "contract plan"};
  check_and_tags(P);
  TensorDesc &A = P.A, &B = P.B;
  const int q = A.qdim, k = (int)P.ica.size();
  const int kc = k * q, kma = (int)P.ka.size() * q, kmb = (int)P.kb.size() * q;
  // label records (getQsub, QSpace.hh:2315-2340)
  std::vector<int32_t> QAc((size_t)A.n * kc), QAk((size_t)A.n * kma), QBc((size_t)B.n * kc), QBk((size_t)B.n * kmb);
  for (int i = 0; i < A.n; ++i) {
    for (int l = 0; l < k; ++l)
      for (int j = 0; j < q; ++j) QAc[(size_t)i * kc + l * q + j] = A.label(i, P.ica[l], j);
    for (size_t l = 0; l < P.ka.size(); ++l)
      for (int j = 0; j < q; ++j) QAk[(size_t)i * kma + l * q + j] = A.label(i, P.ka[l], j);
  }
  for (int i = 0; i < B.n; ++i) {
    for (int l = 0; l < k; ++l)
      for (int j = 0; j < q; ++j) QBc[(size_t)i * kc + l * q + j] = B.label(i, P.icb[l], j);
    for (size_t l = 0; l < P.kb.size(); ++l)
      for (int j = 0; j < q; ++j) QBk[(size_t)i * kmb + l * q + j] = B.label(i, P.kb[l], j);
  }
  clk.lap("label records");
  P.L = device_label_engine(QAc, QAk, A.n, QBc, QBk, B.n, kc, kma, kmb, default_stream());
  clk.lap("device label engine");
  const int nL = (int)P.L.seg.size() - 1;
  const size_t np = P.L.Ia.size();
  // output permutation (completed with the identity, bin/contractQS.m:154-170)
  P.perm.resize(P.rc);
  std::iota(P.perm.begin(), P.perm.end(), 0);
  if (nperm > 0) {
    if (nperm > P.rc) fail(QSX_ERR_ARG, "invalid permutation (length %d > rank %d)", nperm, P.rc);
    std::vector<char> used(P.rc, 0);
    for (int i = 0; i < nperm; ++i) {
      if (perm[i] < 0 || perm[i] >= P.rc || used[perm[i]]) fail(QSX_ERR_ARG, "invalid permutation");
      used[perm[i]] = 1;
      P.perm[i] = perm[i];
    }
    int l = nperm;
    for (int i = 0; i < P.rc; ++i)
      if (!used[i]) P.perm[l++] = i;
  }
  P.has_perm = false;
  for (int i = 0; i < P.rc; ++i) P.has_perm |= P.perm[i] != i;
  // outer multiplicity: CGTs of rank > 3 on the operands or the result, or
  // blocks that carry components
  int nsu2 = 0, jsu2 = -1;
  for (int j = 0; j < A.nsym; ++j)
    if (A.symtype[j] == SYM_SU2) {
      ++nsu2;
      jsu2 = j;
    }
  if ((A.has_om() || B.has_om()) && !nsu2)
    fail(QSX_ERR_STRUCT, "outer multiplicity > 1 in a tensor without SU(2) symmetry");
  P.om_path = nsu2 && (A.r > 3 || B.r > 3 || P.rc > 3 || A.has_om() || B.has_om());
  if (P.om_path && nsu2 > 1) fail(QSX_ERR_UNSUPPORTED, "outer multiplicity with more than one SU(2) symmetry");
  std::vector<std::shared_ptr<const SU2XTensor>> xs(P.om_path ? np : 0);
  // afac per pair (contractDATA_group, QSpace_aux.cc:100-230).  The
  // X-symbols are exact and memoised; a structure's new label sets are
  // evaluated by OpenMP threads (the exact-arithmetic tables are grown
  // first, so the threads only read them; an error is rethrown after the
  // loop)
  P.afac.assign(np, 1.0);
  int max_tj = 0;
  bool su2 = false;
  for (int j = 0; j < A.nsym; ++j) {
    if (A.symtype[j] != SYM_SU2) continue;
    su2 = true;
    for (int i = 0; i < A.n; ++i)
      for (int l = 0; l < A.r; ++l) max_tj = std::max(max_tj, A.label(i, l, j));
    for (int i = 0; i < B.n; ++i)
      for (int l = 0; l < B.r; ++l) max_tj = std::max(max_tj, B.label(i, l, j));
  }
  if (su2) su2_reserve_labels(max_tj * std::max(3, std::max(std::max(A.r, B.r), P.rc)));
  std::atomic<bool> bad{false};
  Error first{0, ""};
  std::mutex err_mu;
#pragma omp parallel for schedule(dynamic, 32) if (su2 && np > 256)
  for (size_t p = 0; p < np; ++p) {
    if (bad.load(std::memory_order_relaxed)) continue;
    try {
      const int ia = P.L.Ia[p], ib = P.L.Ib[p];
      double af = (fA ? fA[ia] : 1.0) * (fB ? fB[ib] : 1.0);
      if (P.om_path) {
        auto X = su2_xom(su2_col(A, ia, jsu2), P.dirA, su2_col(B, ib, jsu2), P.dirB, P.ica, P.icb, P.perm);
        if (X->oa != A.om_of(ia))
          fail(QSX_ERR_STRUCT, "block %d of A has outer multiplicity %d, its SU(2) labels give %d", ia + 1,
               A.om_of(ia), X->oa);
        if (X->ob != B.om_of(ib))
          fail(QSX_ERR_STRUCT, "block %d of B has outer multiplicity %d, its SU(2) labels give %d", ib + 1,
               B.om_of(ib), X->ob);
        bool nz = false;
        for (double v : X->x) nz |= v != 0;
        af = nz ? af * A.cgw[(size_t)ia * A.nsym + jsu2] * B.cgw[(size_t)ib * B.nsym + jsu2] : 0.0;
        xs[p] = std::move(X);
        P.afac[p] = af;
        continue;
      }
      for (int j = 0; j < A.nsym && af != 0; ++j) {
        if (A.symtype[j] != SYM_SU2) continue;
        double x = su2_x3(su2_col(A, ia, j), P.dirA, su2_col(B, ib, j), P.dirB, P.ica, P.icb);
        af *= x * A.cgw[(size_t)ia * A.nsym + j] * B.cgw[(size_t)ib * B.nsym + j];
      }
      if (std::fabs(af) < 1e-18) af = 0;  // CG_SKIP_EPS2
      P.afac[p] = af;
    } catch (const Error& e) {
      std::lock_guard<std::mutex> lk(err_mu);
      if (!bad.exchange(true)) first = e;
    }
  }
  if (bad.load()) throw first;
  clk.lap("X-symbols");
  // layouts (may reorder ica/icb jointly; afac is independent of that order)
  choose_layouts(P);
  // executed output blocks, pair descriptors, output normalisation
  P.nC = 0;
  P.Qout.clear();
  P.dimsOut.clear();
  P.offsOut.assign(1, 0);
  P.cgwOut.clear();
  P.omOut.clear();
  P.blkL.clear();
  P.nT = 0;
  P.tOff.clear();
  P.pairs.clear();
  P.pIa.clear();
  P.pIb.clear();
  P.pseg.assign(1, 0);
  P.M.clear();
  P.N.clear();
  P.flops = 0;
  const int rc = P.rc, mc = rc * q;
  for (int c = 0; c < nL; ++c) {
    int64_t p0 = P.L.seg[c], p1 = P.L.seg[c + 1];
    bool any = false;
    for (int64_t p = p0; p < p1; ++p) any |= P.afac[p] != 0;
    if (!any) continue;  // SkipEmptyData
    // output block dims [kept A | kept B] from the first executed pair
    std::vector<int64_t> d;
    int64_t first = -1;
    for (int64_t p = p0; p < p1; ++p)
      if (P.afac[p] != 0) { first = p; break; }
    const int ia0 = P.L.Ia[first], ib0 = P.L.Ib[first];
    for (int l : P.ka) d.push_back(A.dims[(size_t)ia0 * A.r + l]);
    for (int l : P.kb) d.push_back(B.dims[(size_t)ib0 * B.r + l]);
    int64_t Mc = 1, Nc = 1;
    for (int l : P.ka) Mc *= A.dims[(size_t)ia0 * A.r + l];
    for (int l : P.kb) Nc *= B.dims[(size_t)ib0 * B.r + l];
    if (Mc > INT32_MAX || Nc > INT32_MAX) fail(QSX_ERR_UNSUPPORTED, "output block too large");
    // output normalisation of NormCGW folded into afac: cgw -> normExt, data /= normExt
    std::vector<int32_t> rec(P.L.QC.begin() + (size_t)c * mc, P.L.QC.begin() + (size_t)(c + 1) * mc);
    double cfac = 1.0;
    std::vector<double> cgw(A.nsym, 1.0);
    for (int j = 0; j < A.nsym; ++j) {
      if (A.symtype[j] != SYM_SU2) continue;
      std::vector<int> tj(rc);
      for (int l = 0; l < rc; ++l) tj[l] = rec[l * q + j];
      double ne = su2_norm_ext(tj, P.dirC);
      cgw[j] = ne;
      cfac /= ne;
      // after the user permutation the block's CGT is the canonical CGT of
      // the new leg order: the data take its sign relative to the
      // transposed one (on the om path the X-symbols are taken on the
      // permuted output's basis instead)
      if (P.has_perm && !P.om_path) cfac *= su2_perm_sign(tj, P.dirC, P.perm);
    }
    auto pair_desc = [&](int64_t p, int64_t aoff, int64_t boff, double af) {
      const int ia = P.L.Ia[p], ib = P.L.Ib[p];
      int64_t K = 1, Mp = 1, Np = 1;
      for (int l = 0; l < k; ++l) {
        int64_t da = A.dims[(size_t)ia * A.r + P.ica[l]], db = B.dims[(size_t)ib * B.r + P.icb[l]];
        if (da != db) fail(QSX_ERR_STRUCT, "contracted block dims differ (%ld/%ld)", (long)da, (long)db);
        K *= da;
      }
      for (int l : P.ka) Mp *= A.dims[(size_t)ia * A.r + l];
      for (int l : P.kb) Np *= B.dims[(size_t)ib * B.r + l];
      if (Mp != Mc || Np != Nc) fail(QSX_ERR_STRUCT, "inconsistent output block dims within block %d", c + 1);
      if (K > INT32_MAX) fail(QSX_ERR_UNSUPPORTED, "contracted dim too large");
      PairDesc pd{};
      pd.a_off = aoff;
      pd.b_off = boff;
      pd.K = (int32_t)K;
      pd.flags = (P.la.kmajor ? 1 : 0) | (P.lb.kmajor ? 2 : 0);
      pd.lda = (int32_t)std::max<int64_t>(1, P.la.kmajor ? K : Mc);
      pd.ldb = (int32_t)std::max<int64_t>(1, P.lb.kmajor ? Nc : K);
      pd.afac = af;
      P.pairs.push_back(pd);
      P.pIa.push_back(ia);
      P.pIb.push_back(ib);
      if (af != 0) P.flops += 2.0 * (double)Mc * (double)Nc * (double)K;
    };
    int omC = 1;
    if (P.om_path) {
      // one GEMM target per output component; a component that no X-symbol
      // reaches gets one zero-weight pair, so that it is written (zeros)
      omC = xs[first]->oc;
      const int64_t cs = Mc * Nc;
      for (int g = 0; g < omC; ++g) {
        const size_t before = P.pairs.size();
        for (int64_t p = p0; p < p1; ++p) {
          if (P.afac[p] == 0) continue;
          const SU2XTensor& X = *xs[p];
          const int ia = P.L.Ia[p], ib = P.L.Ib[p];
          const int64_t sa = A.slice(ia), sb = B.slice(ib);
          for (int a = 0; a < X.oa; ++a)
            for (int b = 0; b < X.ob; ++b) {
              const double x = X.x[((size_t)a * X.ob + b) * X.oc + g];
              if (x == 0) continue;
              pair_desc(p, A.offs[ia] + a * sa, B.offs[ib] + b * sb, P.afac[p] * x * cfac);
            }
        }
        if (P.pairs.size() == before) pair_desc(first, A.offs[P.L.Ia[first]], B.offs[P.L.Ib[first]], 0.0);
        P.pseg.push_back((int64_t)P.pairs.size());
        P.M.push_back((int32_t)Mc);
        P.N.push_back((int32_t)Nc);
        P.tOff.push_back(P.offsOut.back() + g * cs);
        ++P.nT;
      }
      P.omOut.push_back(omC);
    }
    for (int64_t p = p0; p < p1 && !P.om_path; ++p) {
      if (P.afac[p] == 0) continue;
      const int ia = P.L.Ia[p], ib = P.L.Ib[p];
      int64_t K = 1, Mp = 1, Np = 1;
      for (int l = 0; l < k; ++l) {
        int64_t da = A.dims[(size_t)ia * A.r + P.ica[l]], db = B.dims[(size_t)ib * B.r + P.icb[l]];
        if (da != db) fail(QSX_ERR_STRUCT, "contracted block dims differ (%ld/%ld)", (long)da, (long)db);
        K *= da;
      }
      for (int l : P.ka) Mp *= A.dims[(size_t)ia * A.r + l];
      for (int l : P.kb) Np *= B.dims[(size_t)ib * B.r + l];
      if (Mp != Mc || Np != Nc) fail(QSX_ERR_STRUCT, "inconsistent output block dims within block %d", c + 1);
      if (K > INT32_MAX) fail(QSX_ERR_UNSUPPORTED, "contracted dim too large");
      PairDesc pd{};
      pd.a_off = A.offs[ia];
      pd.b_off = B.offs[ib];
      pd.K = (int32_t)K;
      pd.flags = (P.la.kmajor ? 1 : 0) | (P.lb.kmajor ? 2 : 0);
      pd.lda = (int32_t)std::max<int64_t>(1, P.la.kmajor ? K : Mc);
      pd.ldb = (int32_t)std::max<int64_t>(1, P.lb.kmajor ? Nc : K);
      pd.afac = P.afac[p] * cfac;
      P.pairs.push_back(pd);
      P.pIa.push_back(ia);
      P.pIb.push_back(ib);
      P.flops += 2.0 * (double)Mc * (double)Nc * (double)K;
    }
    if (!P.om_path) {
      P.pseg.push_back((int64_t)P.pairs.size());
      P.M.push_back((int32_t)Mc);
      P.N.push_back((int32_t)Nc);
      P.tOff.push_back(P.offsOut.back());
      ++P.nT;
    }
    // output record / dims after the user permutation
    std::vector<int32_t> r2(mc);
    std::vector<int64_t> d2(rc);
    for (int l = 0; l < rc; ++l) {
      for (int j = 0; j < q; ++j) r2[l * q + j] = rec[P.perm[l] * q + j];
      d2[l] = d[P.perm[l]];
    }
    P.Qout.insert(P.Qout.end(), r2.begin(), r2.end());
    P.dimsOut.insert(P.dimsOut.end(), d2.begin(), d2.end());
    P.offsOut.push_back(P.offsOut.back() + Mc * Nc * omC);
    P.cgwOut.insert(P.cgwOut.end(), cgw.begin(), cgw.end());
    P.blkL.push_back(c);
    ++P.nC;
  }
  if (P.has_perm && !P.tagsC.empty()) {
    std::vector<uint64_t> t2(rc);
    for (int l = 0; l < rc; ++l) t2[l] = P.tagsC[P.perm[l]];
    P.tagsC = t2;
  }
  clk.lap("pairs and output blocks");
}

// block permutation plan of n blocks of rank r whose outer-multiplicity
// components (om, empty = 1) form a trailing dimension that stays last
static void perm_plan_om(PermPlan& pp, int n, int r, const int64_t* dims, const std::vector<int32_t>& om,
                         const int64_t* offs, const std::vector<int>& perm) {
  if (om.empty()) {
    perm_plan_build(pp, n, r, dims, offs, offs, perm.data(), nullptr);
    return;
  }
  std::vector<int64_t> d((size_t)n * (r + 1));
  for (int b = 0; b < n; ++b) {
    for (int k = 0; k < r; ++k) d[(size_t)b * (r + 1) + k] = dims[(size_t)b * r + k];
    d[(size_t)b * (r + 1) + r] = om[b];
  }
  std::vector<int> p2(perm.begin(), perm.end());
  p2.push_back(r);
  perm_plan_build(pp, n, r + 1, d.data(), offs, offs, p2.data(), nullptr);
}

void plan_upload(CtrPlan& P) {
  PhaseClock clk{P.verbose, "contract plan upload"};
  hipStream_t s = default_stream();
  std::vector<WorkTarget> targets(P.nT);
  for (int t = 0; t < P.nT; ++t) targets[t] = WorkTarget{P.tOff[t], P.M[t], P.N[t], P.pseg[t], P.pseg[t + 1]};
  P.work = build_work(targets, P.pairs, P.policy);
  clk.lap("build_work");
  for (void** q : {(void**)&P.dpairs, (void**)&P.ditems, (void**)&P.dreds, (void**)&P.dqueues, (void**)&P.dscratch}) {
    dev_free(*q);
    *q = nullptr;
  }
  // pairs and work lists in one staged upload (pinned buffer, one synchronisation)
  std::vector<UploadPart> parts;
  P.dpairs = (PairDesc*)dev_alloc(sizeof(PairDesc) * std::max<size_t>(1, P.pairs.size()));
  parts.push_back({P.dpairs, P.pairs.data(), sizeof(PairDesc) * P.pairs.size()});
  P.ditems = (WorkItem*)dev_alloc(sizeof(WorkItem) * std::max<size_t>(1, P.work.items.size()));
  parts.push_back({P.ditems, P.work.items.data(), sizeof(WorkItem) * P.work.items.size()});
  P.dreds = (ReduceItem*)dev_alloc(sizeof(ReduceItem) * std::max<size_t>(1, P.work.reds.size()));
  parts.push_back({P.dreds, P.work.reds.data(), sizeof(ReduceItem) * P.work.reds.size()});
  P.dqueues = (WorkQueue*)dev_alloc(sizeof(WorkQueue) * std::max<size_t>(1, P.work.queues.size()));
  parts.push_back({P.dqueues, P.work.queues.data(), sizeof(WorkQueue) * P.work.queues.size()});
  P.dscratch = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, P.work.scratch_elems));
  upload_many(parts, s);
  clk.lap("pairs and work list to device");
  if (P.la.permute && !P.a_prepermuted) {
    perm_plan_om(P.ppA, P.A.n, P.A.r, P.A.dims.data(), P.A.om, P.A.offs.data(), P.la.perm);
    perm_plan_upload(P.ppA, s);
    dev_free(P.dAperm);
    P.dAperm = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, P.A.total()));
  }
  if (P.lb.permute && !P.b_prepermuted) {
    perm_plan_om(P.ppB, P.B.n, P.B.r, P.B.dims.data(), P.B.om, P.B.offs.data(), P.lb.perm);
    perm_plan_upload(P.ppB, s);
    dev_free(P.dBperm);
    P.dBperm = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, P.B.total()));
  }
  if (P.has_perm) {
    // unpermuted output blocks have dims [kept A | kept B]: d[perm[l]] = dimsOut[l]
    std::vector<int64_t> d((size_t)P.nC * P.rc);
    for (int c = 0; c < P.nC; ++c)
      for (int l = 0; l < P.rc; ++l) d[(size_t)c * P.rc + P.perm[l]] = P.dimsOut[(size_t)c * P.rc + l];
    perm_plan_om(P.ppC, P.nC, P.rc, d.data(), P.omOut, P.offsOut.data(), P.perm);
    perm_plan_upload(P.ppC, s);
    dev_free(P.dCtmp);
    P.dCtmp = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, P.offsOut.back()));
  }
  clk.lap("permutation plans");
  QSX_HIP(hipStreamSynchronize(s));
  clk.lap("sync");
}

void plan_permute_operand(const CtrPlan& P, int which, const double* src, double* dst, hipStream_t s) {
  const TensorDesc& T = which == 0 ? P.A : P.B;
  const OperandLayout& L = which == 0 ? P.la : P.lb;
  if (!L.permute) fail(QSX_ERR_ARG, "operand needs no permute");
  PermPlan pp;
  perm_plan_om(pp, T.n, T.r, T.dims.data(), T.om, T.offs.data(), L.perm);
  perm_plan_upload(pp, s);
  launch_perm_plan(pp, src, dst, s);
  QSX_HIP(hipStreamSynchronize(s));
}

std::unique_ptr<CtrPlan> plan_create(const qsx_tensor* A, const int32_t* ia, int na, int conjA, const qsx_tensor* B,
                                     const int32_t* ib, int nb, int conjB, const int32_t* perm, int nperm,
                                     const qsx_opts* o, bool upload, int force) {
  ensure_device();
  auto P = std::make_unique<CtrPlan>();
  P->force_layout = force;
  P->A = make_desc(A, "A");
  P->B = make_desc(B, "B");
  if (!ia || !ib || na <= 0 || nb <= 0) fail(QSX_ERR_ARG, "empty contraction index");
  P->ica.assign(ia, ia + na);
  P->icb.assign(ib, ib + nb);
  P->conjA = conjA ? 1 : 0;
  P->conjB = conjB ? 1 : 0;
  P->policy.slots = resident_slots();
  if (o) {
    P->mode = o->mode;
    P->nslices = o->nslices;
    P->verbose = o->verbose;
  }
  if (P->mode < QSX_MODE_BATCHED || P->mode > QSX_MODE_SPLIT_F32) fail(QSX_ERR_ARG, "invalid mode %d", P->mode);
  if (P->mode == QSX_MODE_SPLIT_BF16 || P->mode == QSX_MODE_SPLIT_F32)
    fail(QSX_ERR_UNSUPPORTED, "split precision is available for H*psi plans only (QSX_MODE_SPLIT_BF16)");
  build_structure(*P, perm, nperm, o ? o->fermA : nullptr, o ? o->fermB : nullptr);
  PhaseClock clk{P->verbose, "contract plan"};
  if (upload) plan_upload(*P);
  clk.lap("work list + upload");
  return P;
}

void plan_execute(CtrPlan& P, const double* dA, const double* dB, double* dC, hipStream_t s) {
  const double* a = dA;
  const double* b = dB;
  if (P.la.permute && !P.a_prepermuted) {
    launch_perm_plan(P.ppA, dA, P.dAperm, s);
    a = P.dAperm;
  }
  if (P.lb.permute && !P.b_prepermuted) {
    launch_perm_plan(P.ppB, dB, P.dBperm, s);
    b = P.dBperm;
  }
  double* c = P.has_perm ? P.dCtmp : dC;
  if (P.mode == QSX_MODE_PLAIN) {
    for (int t = 0; t < P.nT; ++t) {
      bool first = true;
      for (int64_t p = P.pseg[t]; p < P.pseg[t + 1]; ++p) {
        launch_pair_gemm_f64(P.pairs[p], P.tOff[t], P.M[t], P.N[t], first ? 0 : 1, a, b, c, s);
        first = false;
      }
      if (P.pseg[t] == P.pseg[t + 1]) launch_fill_zero(c + P.tOff[t], (int64_t)P.M[t] * P.N[t], s);
    }
  } else if (P.mode == QSX_MODE_BATCHED) {
    launch_worklist(P.work, P.dqueues, P.ditems, P.dreds, P.dpairs, a, b, c, P.dscratch, s);
  }
  if (P.has_perm) launch_perm_plan(P.ppC, P.dCtmp, dC, s);
}

// ---------------------------------------------------------------- one-shot
qsx_tensor_out* make_out(const CtrPlan& P, const std::vector<int>& keep, const double* data,
                         const double* data_im) {
  auto* o = (qsx_tensor_out*)calloc(1, sizeof(qsx_tensor_out));
  const int q = P.A.qdim, rc = P.rc;
  int n = (int)keep.size();
  o->nblk = n;
  o->rank = rc;
  o->qdim = q;
  o->nsym = P.A.nsym;
  snprintf(o->qtype, sizeof o->qtype, "%s", P.A.qtype.c_str());
  o->Q = (int32_t*)malloc(sizeof(int32_t) * std::max<size_t>(1, (size_t)n * rc * q));
  o->dims = (int64_t*)malloc(sizeof(int64_t) * std::max<size_t>(1, (size_t)n * rc));
  o->offs = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
  o->cgw = (double*)malloc(sizeof(double) * std::max<size_t>(1, (size_t)n * P.A.nsym));
  o->itags = (uint64_t*)calloc(std::max(1, rc), sizeof(uint64_t));
  o->offs[0] = 0;
  bool any_om = false;
  for (int c : keep) any_om |= !P.omOut.empty() && P.omOut[c] > 1;
  if (any_om) {
    o->om = (int32_t*)malloc(sizeof(int32_t) * std::max(1, n));
    for (int i = 0; i < n; ++i) o->om[i] = P.omOut[keep[i]];
  }
  for (int i = 0; i < n; ++i) {
    int c = keep[i];
    for (int l = 0; l < rc * q; ++l) o->Q[(size_t)i * rc * q + l] = P.Qout[(size_t)c * rc * q + l];
    for (int l = 0; l < rc; ++l) o->dims[(size_t)i * rc + l] = P.dimsOut[(size_t)c * rc + l];
    for (int j = 0; j < P.A.nsym; ++j) o->cgw[(size_t)i * P.A.nsym + j] = P.cgwOut[(size_t)c * P.A.nsym + j];
    o->offs[i + 1] = o->offs[i] + (P.offsOut[c + 1] - P.offsOut[c]);
  }
  for (size_t l = 0; l < P.tagsC.size(); ++l) o->itags[l] = P.tagsC[l];
  // the kept blocks, in parallel for large results
  auto gather = [&](const double* src, double*& dst) {
    dst = (double*)malloc(sizeof(double) * std::max<int64_t>(1, o->offs[n]));
#pragma omp parallel for schedule(dynamic, 16) if (o->offs[n] > (1 << 18))
    for (int i = 0; i < n; ++i) {
      const int c = keep[i];
      memcpy(dst + o->offs[i], src + P.offsOut[c], sizeof(double) * (P.offsOut[c + 1] - P.offsOut[c]));
    }
  };
  if (data) gather(data, o->data);
  if (data_im) gather(data_im, o->data_im);
  o->npairs = (int64_t)P.pairs.size();
  o->useful_flops = P.flops;
  return o;
}

// Host blocks -> one packed device buffer.  The blocks are gathered (OpenMP
// threads for large tensors) into a pinned staging buffer that the library
// keeps for the life of the process, then copied in one DMA per chunk: a
// pageable source would be staged by the driver at a fraction of the link
// rate, and fresh host vectors page-fault on first touch.
static std::mutex g_stage_mu;
static double* g_stage = nullptr;
static int64_t g_stage_n = 0;
constexpr int64_t kStageMax = (int64_t)1 << 26;  // 512 MB of doubles per chunk

// the staging buffer holds at least want doubles (caller holds g_stage_mu)
static void stage_reserve(int64_t want) {
  if (g_stage_n >= want) return;
  if (g_stage) (void)hipHostFree(g_stage);
  g_stage = nullptr;
  g_stage_n = 0;
  QSX_HIP(hipHostMalloc((void**)&g_stage, sizeof(double) * want, hipHostMallocDefault));
  g_stage_n = want;
}

// dst[0..n) = src[0..n), by OpenMP threads for large copies
static void par_copy(double* dst, const double* src, int64_t n) {
  constexpr int64_t kChunk = 1 << 17;
  const int64_t nch = (n + kChunk - 1) / kChunk;
#pragma omp parallel for schedule(static) if (nch > 2)
  for (int64_t c = 0; c < nch; ++c) {
    const int64_t lo = c * kChunk, hi = std::min(n, lo + kChunk);
    memcpy(dst + lo, src + lo, sizeof(double) * (hi - lo));
  }
}

void download_staged(const double* d, int64_t n, hipStream_t s, const std::function<void(const double*)>& use) {
  if (n <= 0) {
    use(nullptr);
    return;
  }
  if (n > kStageMax) {  // beyond the staging buffer: a plain host vector
    std::vector<double> v(n);
    QSX_HIP(hipMemcpyAsync(v.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    QSX_HIP(hipStreamSynchronize(s));
    use(v.data());
    return;
  }
  std::lock_guard<std::mutex> lk(g_stage_mu);
  stage_reserve(n);
  QSX_HIP(hipMemcpyAsync(g_stage, d, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  QSX_HIP(hipStreamSynchronize(s));
  use(g_stage);
}

void download_packed(const double* d, int64_t n, double* dst, hipStream_t s) {
  download_staged(d, n, s, [&](const double* h) {
    if (h) par_copy(dst, h, n);
  });
}

void upload_packed(const TensorDesc& T, const double* const* h, double* d, hipStream_t s) {
  const int64_t tot = T.total();
  if (!tot) return;
  if (!h) fail(QSX_ERR_ARG, "tensor without data");
  for (int i = 0; i < T.n; ++i)
    if (!h[i] && T.numel(i)) fail(QSX_ERR_ARG, "null data block %d", i + 1);
  std::lock_guard<std::mutex> lk(g_stage_mu);
  stage_reserve(std::min(tot, kStageMax));
  for (int64_t e0 = 0; e0 < tot; e0 += g_stage_n) {
    const int64_t e1 = std::min(tot, e0 + g_stage_n);
    // blocks overlapping [e0, e1)
    const int b0 = (int)(std::upper_bound(T.offs.begin(), T.offs.end(), e0) - T.offs.begin()) - 1;
    const int b1 = (int)(std::lower_bound(T.offs.begin(), T.offs.end(), e1) - T.offs.begin());
    double* st = g_stage;
#pragma omp parallel for schedule(dynamic, 8) if (e1 - e0 > (1 << 18))
    for (int i = b0; i < b1; ++i) {
      const int64_t lo = std::max(e0, T.offs[i]), hi = std::min(e1, T.offs[i + 1]);
      if (hi > lo) memcpy(st + (lo - e0), h[i] + (lo - T.offs[i]), sizeof(double) * (hi - lo));
    }
    QSX_HIP(hipMemcpyAsync(d + e0, st, sizeof(double) * (e1 - e0), hipMemcpyHostToDevice, s));
    QSX_HIP(hipStreamSynchronize(s));  // the staging buffer is reused by the next chunk / call
  }
}

void upload_many(const std::vector<UploadPart>& parts, hipStream_t s) {
  int64_t tot = 0;  // doubles, each part starting on a 256-byte boundary
  for (const UploadPart& p : parts) tot += ((int64_t)p.bytes + 255) / 256 * 32;
  if (!tot) return;
  if (tot > kStageMax) {  // beyond the staging buffer: the runtime's own path
    for (const UploadPart& p : parts)
      if (p.bytes) QSX_HIP(hipMemcpyAsync(p.d, p.h, p.bytes, hipMemcpyHostToDevice, s));
    QSX_HIP(hipStreamSynchronize(s));
    return;
  }
  std::lock_guard<std::mutex> lk(g_stage_mu);
  stage_reserve(tot);
  char* st = (char*)g_stage;
  for (const UploadPart& p : parts) {
    if (!p.bytes) continue;
    memcpy(st, p.h, p.bytes);
    QSX_HIP(hipMemcpyAsync(p.d, st, p.bytes, hipMemcpyHostToDevice, s));
    st += (p.bytes + 255) / 256 * 256;
  }
  QSX_HIP(hipStreamSynchronize(s));  // the staging buffer is reused by the next call
}

static void upload_tensor(const TensorDesc& T, double* d, hipStream_t s, bool imag = false) {
  upload_packed(T, imag ? T.hdata_im : T.hdata, d, s);
}

// device buffers freed on scope exit
struct DevBufs {
  std::vector<void*> p;
  double* add(int64_t n) {
    p.push_back(dev_alloc(sizeof(double) * std::max<int64_t>(1, n)));
    return (double*)p.back();
  }
  ~DevBufs() {
    for (void* q : p) dev_free(q);
  }
};

// complex operands (the zgemm path of contractDATA_group): the structure and
// its plan do not depend on the data, so one plan runs on the real and
// imaginary planes and a linear combination forms the result,
//   Re C = C(Ar,Br) - sA sB C(Ai,Bi),  Im C = sB C(Ar,Bi) + sA C(Ai,Br),
// with sX = -1 when operand X carries a conj flag (its data are conjugated)
static qsx_tensor_out* contract_oneshot_complex(const qsx_tensor* A, const int32_t* ia, int na, int conjA,
                                                const qsx_tensor* B, const int32_t* ib, int nb, int conjB,
                                                const int32_t* perm, int nperm, const qsx_opts* o) {
  hipStream_t s = default_stream();
  auto P = plan_create(A, ia, na, conjA, B, ib, nb, conjB, perm, nperm, o, true);
  const bool ai = P->A.hdata_im != nullptr, bi = P->B.hdata_im != nullptr;
  // the plan may have applied the missing-conj rule (contract_getIdxSet)
  const double sA = P->conjA ? -1.0 : 1.0, sB = P->conjB ? -1.0 : 1.0;
  const int64_t nC = P->offsOut.back();
  DevBufs m;
  double* dAr = m.add(P->A.total());
  double* dBr = m.add(P->B.total());
  double* dAi = ai ? m.add(P->A.total()) : nullptr;
  double* dBi = bi ? m.add(P->B.total()) : nullptr;
  upload_tensor(P->A, dAr, s);
  upload_tensor(P->B, dBr, s);
  if (ai) upload_tensor(P->A, dAi, s, true);
  if (bi) upload_tensor(P->B, dBi, s, true);
  double* C1 = m.add(nC);                          // Ar Br
  double* C2 = ai && bi ? m.add(nC) : nullptr;     // Ai Bi
  double* C3 = bi ? m.add(nC) : nullptr;           // Ar Bi
  double* C4 = ai ? m.add(nC) : nullptr;           // Ai Br
  double* Cr = m.add(nC);
  double* Ci = m.add(nC);
  hipEvent_t e0, e1;
  QSX_HIP(hipEventCreate(&e0));
  QSX_HIP(hipEventCreate(&e1));
  QSX_HIP(hipEventRecord(e0, s));
  plan_execute(*P, dAr, dBr, C1, s);
  if (C2) plan_execute(*P, dAi, dBi, C2, s);
  if (C3) plan_execute(*P, dAr, dBi, C3, s);
  if (C4) plan_execute(*P, dAi, dBr, C4, s);
  // Re C and Im C in one streaming launch each (offsets relative to the
  // two source buffers); a missing term is the other source with factor 0
  auto combine = [&](double* dst, double* x, double fx, double* y, double fy) {
    std::vector<AxpbyItem> items;
    for (int64_t e0i = 0; e0i < nC; e0i += 4096) {
      AxpbyItem it{};
      it.c_off = e0i;
      it.a_off = x ? e0i : -1;
      it.b_off = y ? e0i : -1;
      it.fa = fx;
      it.fb = fy;
      it.n = (int32_t)std::min<int64_t>(4096, nC - e0i);
      items.push_back(it);
    }
    if (items.empty()) return;
    AxpbyItem* d = (AxpbyItem*)dev_alloc(sizeof(AxpbyItem) * items.size());
    m.p.push_back(d);
    QSX_HIP(hipMemcpyAsync(d, items.data(), sizeof(AxpbyItem) * items.size(), hipMemcpyHostToDevice, s));
    launch_axpby(d, (int)items.size(), x ? x : y, y ? y : x, dst, s);
  };
  combine(Cr, C1, 1.0, C2, -sA * sB);
  if (C3 || C4) combine(Ci, C3, sB, C4, sA);
  else launch_fill_zero(Ci, nC, s);
  QSX_HIP(hipEventRecord(e1, s));
  QSX_HIP(hipEventSynchronize(e1));
  float ms = 0;
  QSX_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // SkipZeroData: a block is zero when both planes are
  std::vector<int> keep(P->nC);
  std::iota(keep.begin(), keep.end(), 0);
  const int skip_zero = o ? o->skip_zero : 1;
  if (skip_zero && P->nC) {
    bool blockdiag = false;
    const int q = P->A.qdim;
    if (P->rc == 2) {
      blockdiag = true;
      for (int c = 0; c < P->nC && blockdiag; ++c)
        for (int j = 0; j < q; ++j)
          if (P->Qout[(size_t)c * 2 * q + j] != P->Qout[(size_t)c * 2 * q + q + j]) blockdiag = false;
    }
    if (!(P->rc == 2 && blockdiag)) {
      int32_t* dz = (int32_t*)dev_alloc(sizeof(int32_t) * 2 * P->nC);
      m.p.push_back(dz);
      int64_t* doffs = (int64_t*)dev_alloc(sizeof(int64_t) * (P->nC + 1));
      m.p.push_back(doffs);
      QSX_HIP(hipMemcpyAsync(doffs, P->offsOut.data(), sizeof(int64_t) * (P->nC + 1), hipMemcpyHostToDevice, s));
      launch_block_zero_flags(doffs, P->nC, Cr, dz, s);
      launch_block_zero_flags(doffs, P->nC, Ci, dz + P->nC, s);
      std::vector<int32_t> z(2 * P->nC);
      QSX_HIP(hipMemcpyAsync(z.data(), dz, sizeof(int32_t) * 2 * P->nC, hipMemcpyDeviceToHost, s));
      QSX_HIP(hipStreamSynchronize(s));
      keep.clear();
      for (int c = 0; c < P->nC; ++c)
        if (!(z[c] && z[P->nC + c])) keep.push_back(c);
    }
  }
  std::vector<double> hr(std::max<int64_t>(1, nC)), hi(std::max<int64_t>(1, nC));
  QSX_HIP(hipMemcpyAsync(hr.data(), Cr, sizeof(double) * nC, hipMemcpyDeviceToHost, s));
  QSX_HIP(hipMemcpyAsync(hi.data(), Ci, sizeof(double) * nC, hipMemcpyDeviceToHost, s));
  QSX_HIP(hipStreamSynchronize(s));
  qsx_tensor_out* out = make_out(*P, keep, hr.data(), hi.data());
  out->gpu_ms = ms;
  out->useful_flops = P->flops * ((C2 ? 1 : 0) + (C3 ? 1 : 0) + (C4 ? 1 : 0) + 1);
  return out;
}

qsx_tensor_out* contract_oneshot(const qsx_tensor* A, const int32_t* ia, int na, int conjA, const qsx_tensor* B,
                                 const int32_t* ib, int nb, int conjB, const int32_t* perm, int nperm,
                                 const qsx_opts* o) {
  ensure_device();
  hipStream_t s = default_stream();
  if (A && B && (A->nblk == 0 || B->nblk == 0)) {  // empty input -> empty output
    auto* out = (qsx_tensor_out*)calloc(1, sizeof(qsx_tensor_out));
    out->offs = (int64_t*)calloc(1, sizeof(int64_t));
    out->qdim = A->qdim;
    return out;
  }
  if (A && B && (A->data_im || B->data_im))
    return contract_oneshot_complex(A, ia, na, conjA, B, ib, nb, conjB, perm, nperm, o);
  PhaseClock clk{o ? o->verbose : 0, "contract"};
  auto P = plan_create(A, ia, na, conjA, B, ib, nb, conjB, perm, nperm, o, true);
  clk.lap("plan (structure + upload)");
  double *dA = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, P->A.total()));
  double *dB = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, P->B.total()));
  double *dC = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, P->offsOut.back()));
  int32_t* dz = nullptr;
  int64_t* doffs = nullptr;
  qsx_tensor_out* out = nullptr;
  try {
    clk.lap("device buffers");
    upload_tensor(P->A, dA, s);
    upload_tensor(P->B, dB, s);
    clk.lap("upload A, B");
    hipEvent_t e0, e1;
    QSX_HIP(hipEventCreate(&e0));
    QSX_HIP(hipEventCreate(&e1));
    QSX_HIP(hipEventRecord(e0, s));
    plan_execute(*P, dA, dB, dC, s);
    QSX_HIP(hipEventRecord(e1, s));
    QSX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    QSX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    clk.lap("execute");
    std::vector<int> keep(P->nC);
    std::iota(keep.begin(), keep.end(), 0);
    const int skip_zero = o ? o->skip_zero : 1;
    if (skip_zero && P->nC) {
      // SkipZeroData (eps=0): exact-zero blocks are dropped unless the result
      // is a rank-2 block-diagonal matrix (QSpace.cc:3200-3270)
      bool blockdiag = false;
      const int q = P->A.qdim;
      if (P->rc == 2) {
        blockdiag = true;
        for (int c = 0; c < P->nC && blockdiag; ++c)
          for (int j = 0; j < q; ++j)
            if (P->Qout[(size_t)c * 2 * q + j] != P->Qout[(size_t)c * 2 * q + q + j]) blockdiag = false;
      }
      if (!(P->rc == 2 && blockdiag)) {
        dz = (int32_t*)dev_alloc(sizeof(int32_t) * P->nC);
        doffs = (int64_t*)dev_alloc(sizeof(int64_t) * (P->nC + 1));
        QSX_HIP(hipMemcpyAsync(doffs, P->offsOut.data(), sizeof(int64_t) * (P->nC + 1), hipMemcpyHostToDevice, s));
        launch_block_zero_flags(doffs, P->nC, dC, dz, s);
        std::vector<int32_t> z(P->nC);
        QSX_HIP(hipMemcpyAsync(z.data(), dz, sizeof(int32_t) * P->nC, hipMemcpyDeviceToHost, s));
        QSX_HIP(hipStreamSynchronize(s));
        keep.clear();
        for (int c = 0; c < P->nC; ++c)
          if (!z[c]) keep.push_back(c);
      }
    }
    clk.lap("SkipZeroData");
    static const double kNone = 0.0;
    download_staged(dC, P->offsOut.back(), s,
                    [&](const double* h) { out = make_out(*P, keep, h ? h : &kNone); });
    out->gpu_ms = ms;
    clk.lap("download + result");
  } catch (...) {
    dev_free(dA);
    dev_free(dB);
    dev_free(dC);
    dev_free(dz);
    dev_free(doffs);
    throw;
  }
  dev_free(dA);
  dev_free(dB);
  dev_free(dC);
  dev_free(dz);
  dev_free(doffs);
  return out;
}

}  // namespace qsx

// ====================================================================== C-ABI
using namespace qsx;

struct qsx_ctr_plan {
  std::unique_ptr<CtrPlan> p;
};

static void set_err(char* err, size_t errlen, const std::string& m) {
  if (err && errlen) snprintf(err, errlen, "%s", m.c_str());
}

template <class F>
static int guarded(char* err, size_t errlen, F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    set_err(err, errlen, e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_err(err, errlen, "out of host memory");
    return QSX_ERR_NOMEM;
  }
}

extern "C" {

const char* qsx_version(void) { return "qsx 0.1 (gfx950)"; }

int qsx_device_init(int device, char* err, size_t errlen) {
  return guarded(err, errlen, [&]() -> int {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) fail(QSX_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) fail(QSX_ERR_ARG, "device %d out of range (%d devices)", device, n);
    QSX_HIP(hipSetDevice(device));
    ensure_device();
    noalloc_scratch_setup();  // the no-allocation entry points' array on this device
    return 0;
  });
}

int qsx_malloc_device(void** ptr, size_t bytes, char* err, size_t errlen) {
  return guarded(err, errlen, [&]() -> int {
    *ptr = dev_alloc(bytes);
    return 0;
  });
}

int qsx_free_device(void* ptr) {
  dev_free(ptr);
  return 0;
}

int qsx_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  hipStream_t s = stream ? (hipStream_t)stream : default_stream();
  if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return QSX_ERR_DEVICE;
  return hipStreamSynchronize(s) == hipSuccess ? 0 : QSX_ERR_DEVICE;
}

int qsx_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  hipStream_t s = stream ? (hipStream_t)stream : default_stream();
  if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s) != hipSuccess) return QSX_ERR_DEVICE;
  return hipStreamSynchronize(s) == hipSuccess ? 0 : QSX_ERR_DEVICE;
}

int qsx_memset_device(void* dst, int value, size_t bytes, void* stream) {
  hipStream_t s = stream ? (hipStream_t)stream : default_stream();
  return hipMemsetAsync(dst, value, bytes, s) == hipSuccess ? 0 : QSX_ERR_DEVICE;
}

int qsx_copy_device(double* dst, const double* src, int64_t n, void* stream) {
  try {
    launch_copy(dst, src, n, stream ? (hipStream_t)stream : default_stream());
    return 0;
  } catch (const Error& e) {
    return e.code;
  }
}

int qsx_synchronize(void) {
  if (g_stream && hipStreamSynchronize(g_stream) != hipSuccess) return QSX_ERR_DEVICE;
  if (g_side && hipStreamSynchronize(g_side) != hipSuccess) return QSX_ERR_DEVICE;
  return hipDeviceSynchronize() == hipSuccess ? 0 : QSX_ERR_DEVICE;
}

int qsx_contract(const qsx_tensor* A, const int32_t* ia, int32_t na, int32_t conjA, const qsx_tensor* B,
                 const int32_t* ib, int32_t nb, int32_t conjB, const int32_t* perm, int32_t nperm,
                 const qsx_opts* opts, qsx_tensor_out** C, char* err, size_t errlen) {
  return guarded(err, errlen, [&]() -> int {
    if (!C) fail(QSX_ERR_ARG, "null output pointer");
    *C = contract_oneshot(A, ia, na, conjA, B, ib, nb, conjB, perm, nperm, opts);
    return 0;
  });
}

void qsx_tensor_free(qsx_tensor_out* t) {
  if (!t) return;
  free(t->Q);
  free(t->dims);
  free(t->offs);
  free(t->data);
  free(t->itags);
  free(t->cgw);
  free(t->data_im);
  free(t->om);
  free(t);
}

int qsx_parse_ctridx(const char* s, int32_t* idx, int32_t maxlen, int32_t* conj) {
  // ctrIdx::init (Source/wbindex.cc:482-530), Wb::Str2Idx (Source/wbio.c:182-297)
  if (!s || !s[0]) return -1;
  std::string t(s);
  int c = 0;
  while (!t.empty() && (t.back() == '*' || isspace((unsigned char)t.back()))) {
    if (t.back() == '*') c ^= 1;
    t.pop_back();
  }
  while (!t.empty() && (t.back() == ',' || t.back() == ';')) t.pop_back();
  std::vector<std::string> tok;
  std::string cur;
  for (char ch : t) {
    if (isspace((unsigned char)ch) || ch == ',' || ch == ';' || ch == '[' || ch == ']') {
      if (!cur.empty()) tok.push_back(cur);
      cur.clear();
    } else if (isalnum((unsigned char)ch)) {
      cur.push_back(ch);
    } else {
      return -2;
    }
  }
  if (!cur.empty()) tok.push_back(cur);
  std::vector<int> v;
  if (tok.size() == 1) {
    for (char ch : tok[0]) {
      if (ch >= '0' && ch <= '9') v.push_back(ch - '0');
      else if (ch >= 'A' && ch <= 'Z') v.push_back(ch - 'A' + 10);
      else v.push_back(ch - 'a' + 36);
    }
  } else {
    for (auto& x : tok) {
      char* end = nullptr;
      long y = strtol(x.c_str(), &end, 0);
      if (*end) return -2;
      v.push_back((int)y);
    }
  }
  if (v.empty()) return -1;
  std::vector<int> u(v);
  std::sort(u.begin(), u.end());
  if (std::adjacent_find(u.begin(), u.end()) != u.end()) return -3;
  if ((int)v.size() > maxlen) return -4;
  for (size_t i = 0; i < v.size(); ++i) {
    if (v[i] < 1) return -5;
    idx[i] = v[i] - 1;
  }
  if (conj) *conj = c;
  return (int)v.size();
}

int qsx_packed_offsets(const qsx_tensor* T, int64_t* offs) {
  try {
    TensorDesc d = make_desc(T, "T");
    std::copy(d.offs.begin(), d.offs.end(), offs);
    return 0;
  } catch (const Error& e) {
    return e.code;
  }
}

int qsx_ctr_plan_create(const qsx_tensor* A, const int32_t* ia, int32_t na, int32_t conjA, const qsx_tensor* B,
                        const int32_t* ib, int32_t nb, int32_t conjB, const int32_t* perm, int32_t nperm,
                        const qsx_opts* opts, qsx_ctr_plan** plan, char* err, size_t errlen) {
  return guarded(err, errlen, [&]() -> int {
    if (!plan) fail(QSX_ERR_ARG, "null plan pointer");
    auto* h = new qsx_ctr_plan;
    try {
      h->p = plan_create(A, ia, na, conjA, B, ib, nb, conjB, perm, nperm, opts, true);
    } catch (...) {
      delete h;
      throw;
    }
    *plan = h;
    return 0;
  });
}

int qsx_ctr_plan_output(const qsx_ctr_plan* plan, qsx_tensor_out** C) {
  if (!plan || !C) return QSX_ERR_ARG;
  std::vector<int> keep(plan->p->nC);
  std::iota(keep.begin(), keep.end(), 0);
  *C = make_out(*plan->p, keep, nullptr);
  return 0;
}

int qsx_ctr_plan_pairs(const qsx_ctr_plan* plan, int64_t* npairs, int64_t* nout, int32_t* ia, int32_t* ib,
                       double* afac, int64_t* seg) {
  if (!plan) return QSX_ERR_ARG;
  const CtrPlan& P = *plan->p;
  if (npairs) *npairs = (int64_t)P.pairs.size();
  if (nout) *nout = P.nT;
  if (ia) std::copy(P.pIa.begin(), P.pIa.end(), ia);
  if (ib) std::copy(P.pIb.begin(), P.pIb.end(), ib);
  if (afac)
    for (size_t i = 0; i < P.pairs.size(); ++i) afac[i] = P.pairs[i].afac;
  if (seg) std::copy(P.pseg.begin(), P.pseg.end(), seg);
  return 0;
}

int qsx_ctr_plan_idxset(const qsx_ctr_plan* plan, int64_t* npairs, int64_t* nout, int32_t* ia, int32_t* ib,
                        double* afac, int64_t* seg, int32_t* qc) {
  if (!plan) return QSX_ERR_ARG;
  const CtrPlan& P = *plan->p;
  const int64_t np = (int64_t)P.L.Ia.size(), nc = (int64_t)P.L.seg.size() - 1;
  if (npairs) *npairs = np;
  if (nout) *nout = nc;
  if (ia) std::copy(P.L.Ia.begin(), P.L.Ia.end(), ia);
  if (ib) std::copy(P.L.Ib.begin(), P.L.Ib.end(), ib);
  if (afac) std::copy(P.afac.begin(), P.afac.end(), afac);
  if (seg) std::copy(P.L.seg.begin(), P.L.seg.end(), seg);
  if (qc) std::copy(P.L.QC.begin(), P.L.QC.end(), qc);
  return 0;
}

int qsx_block_zero_flags_noalloc(const int64_t* offs, int32_t nblk, const double* ddata, int32_t* flags, char* err,
                                 size_t errlen) {
  return guarded(err, errlen, [&] {
    if (nblk < 0 || (nblk && (!offs || !flags))) fail(QSX_ERR_ARG, "invalid block list");
    if (nblk > kNoAllocMaxBlocks)
      fail(QSX_ERR_ARG, "qsx_block_zero_flags_noalloc: %d blocks exceed its maximum of %d (use qsx_block_zero_flags)",
           nblk, kNoAllocMaxBlocks);
    const int dev = g_dev.load(std::memory_order_acquire);
    if (dev < 0) fail(QSX_ERR_DEVICE, "qsx_block_zero_flags_noalloc: call qsx_device_init first");
    char* base = (char*)noalloc_scratch(dev);
    if (!base) fail(QSX_ERR_DEVICE, "qsx_block_zero_flags_noalloc: no scratch on device %d (qsx_device_init)", dev);
    if (!nblk) return 0;
    // scratch layout: nblk+1 offsets, then nblk flags (every flag is written
    // by the kernel, so nothing needs clearing)
    int64_t* doffs = (int64_t*)base;
    int32_t* dz = (int32_t*)(base + sizeof(int64_t) * (kNoAllocMaxBlocks + 1));
    hipStream_t s = g_stream;
    QSX_HIP(hipMemcpyAsync(doffs, offs, sizeof(int64_t) * (nblk + 1), hipMemcpyHostToDevice, s));
    launch_block_zero_flags(doffs, nblk, ddata, dz, s);
    QSX_HIP(hipMemcpyAsync(flags, dz, sizeof(int32_t) * nblk, hipMemcpyDeviceToHost, s));
    QSX_HIP(hipStreamSynchronize(s));
    return 0;
  });
}

int qsx_ctr_execute(qsx_ctr_plan* plan, const double* dA, const double* dB, double* dC, void* stream) {
  if (!plan) return QSX_ERR_ARG;
  try {
    plan_execute(*plan->p, dA, dB, dC, stream ? (hipStream_t)stream : default_stream());
    return 0;
  } catch (const Error& e) {
    fprintf(stderr, "qsx_ctr_execute: %s\n", e.msg.c_str());
    return e.code;
  }
}

int qsx_block_zero_flags(const int64_t* offs, int32_t nblk, const double* ddata, int32_t* flags, char* err,
                         size_t errlen) {
  return guarded(err, errlen, [&] {
    if (nblk < 0 || (nblk && (!offs || !flags))) fail(QSX_ERR_ARG, "invalid block list");
    if (!nblk) return 0;
    hipStream_t s = default_stream();
    int64_t* doffs = (int64_t*)dev_alloc(sizeof(int64_t) * (nblk + 1));
    int32_t* dz = (int32_t*)dev_alloc(sizeof(int32_t) * nblk);
    try {
      QSX_HIP(hipMemcpyAsync(doffs, offs, sizeof(int64_t) * (nblk + 1), hipMemcpyHostToDevice, s));
      launch_block_zero_flags(doffs, nblk, ddata, dz, s);
      QSX_HIP(hipMemcpyAsync(flags, dz, sizeof(int32_t) * nblk, hipMemcpyDeviceToHost, s));
      QSX_HIP(hipStreamSynchronize(s));
    } catch (...) {
      dev_free(doffs);
      dev_free(dz);
      throw;
    }
    dev_free(doffs);
    dev_free(dz);
    return 0;
  });
}

double qsx_ctr_plan_flops(const qsx_ctr_plan* plan) { return plan ? plan->p->flops : 0; }

int qsx_ctr_plan_set_policy(qsx_ctr_plan* plan, const int32_t* policy) {
  if (!plan || !policy) return QSX_ERR_ARG;
  try {
    plan->p->policy = policy_from_ints(policy);
    plan_upload(*plan->p);
    QSX_HIP(hipStreamSynchronize(default_stream()));
    return 0;
  } catch (const Error& e) {
    fprintf(stderr, "qsx_ctr_plan_set_policy: %s\n", e.msg.c_str());
    return e.code;
  }
}

void qsx_ctr_plan_destroy(qsx_ctr_plan* plan) { delete plan; }

double qsx_su2_cg(int32_t tj1, int32_t tm1, int32_t tj2, int32_t tm2, int32_t tj3, int32_t tm3) {
  return su2_cg_exact(tj1, tm1, tj2, tm2, tj3, tm3);
}

int qsx_su2_x3(int32_t ra, const int32_t* tja, const int32_t* dira, int32_t rb, const int32_t* tjb,
               const int32_t* dirb, const int32_t* ia, const int32_t* ib, int32_t nctr, double* x3, char* err,
               size_t errlen) {
  return guarded(err, errlen, [&]() -> int {
    std::vector<int> a(tja, tja + ra), da(dira, dira + ra), b(tjb, tjb + rb), db(dirb, dirb + rb);
    std::vector<int> ca(ia, ia + nctr), cb(ib, ib + nctr);
    *x3 = su2_x3(a, da, b, db, ca, cb);
    return 0;
  });
}

int qsx_su2_om_basis(int32_t r, const int32_t* tj, const int32_t* dir, double* out, int64_t outlen, char* err,
                     size_t errlen) {
  int om = 0;
  const int rc = guarded(err, errlen, [&]() -> int {
    if (r < 0 || r > QSX_MAX_RANK || (r && (!tj || !dir))) fail(QSX_ERR_ARG, "invalid SU(2) label set");
    std::vector<int> t(tj, tj + r), d(dir, dir + r);
    auto B = su2_om_basis(t, d);
    om = B->om;
    if (out) {
      if (outlen < (int64_t)om * B->numel) fail(QSX_ERR_ARG, "basis buffer too small (%lld < %lld)",
                                                (long long)outlen, (long long)om * B->numel);
      std::fill(out, out + (int64_t)om * B->numel, 0.0);
      for (int a = 0; a < om; ++a)
        for (const auto& e : B->ent[a]) out[(int64_t)a * B->numel + e.first] = e.second;
    }
    return 0;
  });
  return rc ? -rc : om;
}

int qsx_su2_canon_order(int32_t r, const int32_t* tj, const int32_t* dir, int32_t* ord) {
  if (r < 0 || r > QSX_MAX_RANK || (r && (!tj || !dir || !ord))) return QSX_ERR_ARG;
  std::vector<int> o = su2_canon_order(std::vector<int>(tj, tj + r), std::vector<int>(dir, dir + r));
  std::copy(o.begin(), o.end(), ord);
  return 0;
}

int qsx_su2_xom(int32_t ra, const int32_t* tja, const int32_t* dira, int32_t rb, const int32_t* tjb,
                const int32_t* dirb, const int32_t* ia, const int32_t* ib, int32_t nctr, const int32_t* perm,
                int32_t* dims3, double* x, int64_t xlen, char* err, size_t errlen) {
  return guarded(err, errlen, [&]() -> int {
    std::vector<int> a(tja, tja + ra), da(dira, dira + ra), b(tjb, tjb + rb), db(dirb, dirb + rb);
    std::vector<int> ca(ia, ia + nctr), cb(ib, ib + nctr);
    const int rc = ra + rb - 2 * nctr;
    if (rc < 0) fail(QSX_ERR_ARG, "invalid contraction");
    std::vector<int> p(rc);
    std::iota(p.begin(), p.end(), 0);
    if (perm) p.assign(perm, perm + rc);
    auto X = su2_xom(a, da, b, db, ca, cb, p);
    dims3[0] = X->oa;
    dims3[1] = X->ob;
    dims3[2] = X->oc;
    if (x) std::copy(X->x.begin(), X->x.begin() + std::min<int64_t>(xlen, (int64_t)X->x.size()), x);
    return 0;
  });
}

}  // extern "C"
