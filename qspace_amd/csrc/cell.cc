// cell.cc -- contractQS usage #2, the cell form with itag auto-contraction.
//
// Reference: contractQS_itags (Source/contractQS.cc:634-808) with the
// per-operand options of icFlags (:159-293) and the itag matching of
// iTags::getCtr (Source/wbindex.cc:1512-1590):
//   contractQS({A, opts.., {B, opts.., C, opts..}, opts..}, D, opts..)
// contracts right to left at every level -- X = D, then X = op . X for the
// operands before it, a nested cell evaluated first -- each step over all
// legs whose itags match (equal names, opposite conj, or equal conj when
// one side carries a '*' option).  Options after an operand: '*' (conj),
// '[!]##[*]' or a numeric index vector (1-based legs that may / '!' may not
// take part in the match).  One top-level operand: its conj flag and index
// option give a conj / permutation of it.
//
// Design: real operands stay on the device for the whole fold -- every
// distinct leaf is uploaded once, each step is a contraction plan built from
// the structures alone and executed on device buffers, intermediates keep
// their exactly-zero blocks (they add nothing later), and only the result
// comes back, after SkipZeroData's zero test on the device.  Complex leaves
// take the one-shot complex contraction per step (host tensors between the
// steps, zero blocks dropped after each, as the reference does).
#include <map>
#include <memory>
#include <numeric>

#include "qsops.hpp"

namespace qsx {
namespace {

// options of one operand (icFlags::set / apply, Source/contractQS.cc:159-293)
struct CellOpt {
  int conj = 0;
  bool has_idx = false;
  bool xflag = false;       // '!': the listed legs may NOT take part
  std::vector<int> idx;     // 0-based
};

void cell_option_str(CellOpt& op, const char* a) {
  std::string s = a ? a : "";
  std::string low = s;
  size_t k0 = low.find_first_not_of('-');
  low = k0 == std::string::npos ? "" : low.substr(k0);
  for (char& c : low) c = (char)tolower((unsigned char)c);
  if (low.rfind("op:", 0) == 0 || s.rfind("--itag:s/", 0) == 0)
    fail(QSX_ERR_UNSUPPORTED, "cell-contraction option '%s' (itag operations) not supported", s.c_str());
  const bool x = !s.empty() && s[0] == '!';
  if (x) s = s.substr(1);
  size_t n = 0;
  while (n < s.size() && isdigit((unsigned char)s[n])) ++n;
  const std::string rest = s.substr(n);
  if (rest == "*") {
    if (op.conj) fail(QSX_ERR_ARG, "cell-contraction option: conj already set");
    op.conj = 1;
  } else if (!rest.empty() || n == 0) {
    fail(QSX_ERR_ARG, "invalid cell-contraction option (expecting '[!##][*]': got '%s')", a ? a : "");
  }
  if (n) {
    if (op.has_idx) fail(QSX_ERR_ARG, "cell-contraction option: ctr-idx already set (%s)", a);
    op.has_idx = true;
    op.xflag = x;
    op.idx.clear();
    for (size_t i = 0; i < n; ++i) op.idx.push_back(s[i] - '1');
  }
}

void cell_option_idx(CellOpt& op, const int32_t* v, int n) {
  for (int i = 0; i < n; ++i)
    if (v[i] < 1 || v[i] > 9) fail(QSX_ERR_ARG, "invalid cell-contraction ctr-idx %d", v[i]);
  if (op.has_idx) fail(QSX_ERR_ARG, "cell-contraction option: ctr-idx already set");
  op.has_idx = true;
  op.xflag = false;
  op.idx.assign(v, v + n);
  for (int& i : op.idx) --i;
}

// legs excluded from the match: without '!' the legs not listed, with '!'
// the listed ones
std::vector<char> cell_skip(const CellOpt& op, int r) {
  std::vector<char> skip(r, 0);
  if (!op.has_idx) return skip;
  std::vector<char> listed(r, 0);
  for (int j : op.idx) {
    if (j < 0 || j >= r) fail(QSX_ERR_ARG, "cell-ctr index out of bounds (r=%d)", r);
    if (listed[j]) fail(QSX_ERR_ARG, "invalid cell-ctr index (non-unique)");
    listed[j] = 1;
  }
  for (int i = 0; i < r; ++i) skip[i] = (listed[i] != 0) == op.xflag;
  return skip;
}

// a qsx_tensor view of owned structure (and optional host data)
struct Held {
  std::vector<int32_t> Q;
  std::vector<int64_t> dims;
  std::vector<uint64_t> tags;
  std::vector<double> cgw;
  std::vector<int32_t> om;
  std::string qtype;
  std::vector<const double*> re, im;
  qsx_tensor t{};
  void bind(int n, int r, int q, bool with_cgw) {
    t.nblk = n;
    t.rank = r;
    t.qdim = q;
    t.Q = Q.empty() ? nullptr : Q.data();
    t.dims = dims.empty() ? nullptr : dims.data();
    t.data = re.empty() ? nullptr : re.data();
    t.data_im = im.empty() ? nullptr : im.data();
    t.itags = tags.empty() ? nullptr : tags.data();
    t.qtype = qtype.c_str();
    t.cgw = with_cgw && !cgw.empty() ? cgw.data() : nullptr;
    t.om = om.empty() ? nullptr : om.data();
  }
};

// a host result as an operand (views into the library-owned output)
std::unique_ptr<Held> view_out(const qsx_tensor_out* o) {
  auto h = std::make_unique<Held>();
  const int n = o->nblk, r = o->rank, q = o->qdim;
  h->Q.assign(o->Q, o->Q + (size_t)n * r * q);
  h->dims.assign(o->dims, o->dims + (size_t)n * r);
  if (o->itags) h->tags.assign(o->itags, o->itags + r);
  if (o->cgw && o->nsym) h->cgw.assign(o->cgw, o->cgw + (size_t)n * o->nsym);
  if (o->om) h->om.assign(o->om, o->om + n);
  h->qtype = o->qtype;
  for (int i = 0; i < n; ++i) {
    h->re.push_back(o->data ? o->data + o->offs[i] : nullptr);
    if (o->data_im) h->im.push_back(o->data_im + o->offs[i]);
  }
  h->bind(n, r, q, true);
  return h;
}

struct OutDel {
  void operator()(qsx_tensor_out* o) const { qsx_tensor_free(o); }
};
using OutPtr = std::unique_ptr<qsx_tensor_out, OutDel>;

OutPtr permute_out(const qsx_tensor* T, const std::vector<int>& perm, int conj) {
  qsx_tensor_out* out = nullptr;
  char err[512];
  std::vector<int32_t> p(perm.begin(), perm.end());
  const int rc = qsx_permute(T, p.empty() ? nullptr : p.data(), (int32_t)p.size(), conj, &out, err, sizeof err);
  if (rc) fail(rc, "%s", err);
  return OutPtr(out);
}

bool any_complex(const qsx_cell* items, int n) {
  for (int i = 0; i < n; ++i) {
    if (items[i].T && items[i].T->data_im) return true;
    if (!items[i].T && items[i].sub && any_complex(items[i].sub, items[i].nsub)) return true;
  }
  return false;
}

void read_options(const qsx_cell& c, CellOpt& op) {
  for (int k = 0; k < c.nopt; ++k) cell_option_str(op, c.opt[k]);
  if (c.nidx > 0) cell_option_idx(op, c.idx, c.nidx);
}

void check_item(const qsx_cell& c) {
  if (!c.T && (!c.sub || c.nsub <= 0)) fail(QSX_ERR_ARG, "invalid cell-contraction entry (neither QSpace nor cell)");
  if (!c.T && c.nsub < 2) fail(QSX_ERR_ARG, "invalid cell-contraction (got cell with fewer than two entries)");
}

// ------------------------------------------------------------ device fold
struct DevOperand {
  const qsx_tensor* t = nullptr;   // structure (and data for a leaf)
  std::shared_ptr<Held> held;      // owned structure of an intermediate
  std::shared_ptr<CtrPlan> plan;   // the plan that produced it
  double* d = nullptr;             // packed device data
  CellOpt op;
};

struct DevFold {
  const qsx_opts* o;
  hipStream_t s;
  std::map<const qsx_tensor*, double*> leaves;  // each distinct leaf uploaded once
  std::vector<void*> bufs;                      // every device buffer of the fold
  ~DevFold() {
    for (void* p : bufs) dev_free(p);
  }
  double* alloc(int64_t n) {
    bufs.push_back(dev_alloc(sizeof(double) * std::max<int64_t>(1, n)));
    return (double*)bufs.back();
  }
  double* leaf(const qsx_tensor* T) {
    auto it = leaves.find(T);
    if (it != leaves.end()) return it->second;
    TensorDesc D = make_desc(T, "cell operand");
    double* d = alloc(D.total());
    upload_packed(D, D.hdata, d, s);
    leaves[T] = d;
    return d;
  }

  DevOperand operand(const qsx_cell& c) {
    check_item(c);
    DevOperand x;
    if (c.T) {
      x.t = c.T;
      x.d = leaf(c.T);
    } else {
      x = fold(c.sub, c.nsub);
      x.op = CellOpt();
    }
    read_options(c, x.op);
    return x;
  }

  // right-to-left fold of one cell level
  DevOperand fold(const qsx_cell* items, int n) {
    std::vector<DevOperand> ops;
    for (int i = 0; i < n; ++i) ops.push_back(operand(items[i]));
    if (ops.size() < 2) fail(QSX_ERR_ARG, "invalid usage #2: at least two QSpaces required for each contraction");
    DevOperand cur = std::move(ops.back());
    int cflag = cur.op.conj;
    std::vector<char> skipB = cell_skip(cur.op, cur.t->rank);
    for (int k = (int)ops.size() - 2; k >= 0; --k) {
      DevOperand& a = ops[k];
      TensorDesc DA = make_desc(a.t, "A"), DB = make_desc(cur.t, "B");
      std::vector<char> skipA = cell_skip(a.op, a.t->rank);
      std::vector<int32_t> ia, ib;
      int cA = 0;
      match_itags(DA, DB, a.op.conj, cflag, ia, ib, cA, &skipA, &skipB);
      qsx_opts so = o ? *o : qsx_opts{};
      so.skip_zero = 0;
      so.hconj = 0;
      so.fermA = so.fermB = nullptr;
      std::shared_ptr<CtrPlan> P(plan_create(a.t, ia.data(), (int)ia.size(), cA, cur.t, ib.data(), (int)ib.size(),
                                             cflag, nullptr, 0, &so, true)
                                     .release());
      double* dC = alloc(P->offsOut.back());
      plan_execute(*P, a.d, cur.d, dC, s);
      auto h = std::make_shared<Held>();
      h->Q = P->Qout;
      h->dims = P->dimsOut;
      h->tags = P->tagsC;
      h->om = P->omOut;
      h->qtype = P->A.qtype;
      h->bind(P->nC, P->rc, P->A.qdim, false);  // results carry the external normalisation
      DevOperand r;
      r.t = &h->t;
      r.held = h;
      r.plan = P;
      r.d = dC;
      cur = std::move(r);
      cflag = 0;
      skipB.assign(cur.t->rank, 0);
    }
    return cur;
  }

  // the result back, SkipZeroData (eps = 0) on the device unless it is a
  // rank-2 block-diagonal matrix (Source/QSpace.cc:3200-3270)
  qsx_tensor_out* download(const DevOperand& R, int skip_zero) {
    const CtrPlan& P = *R.plan;
    std::vector<int> keep(P.nC);
    std::iota(keep.begin(), keep.end(), 0);
    if (skip_zero && P.nC) {
      const int q = P.A.qdim;
      bool blockdiag = P.rc == 2;
      for (int c = 0; c < P.nC && blockdiag; ++c)
        for (int j = 0; j < q; ++j)
          if (P.Qout[(size_t)c * 2 * q + j] != P.Qout[(size_t)c * 2 * q + q + j]) blockdiag = false;
      if (!blockdiag) {
        std::vector<int32_t> z(P.nC);
        char err[256];
        const int rc = qsx_block_zero_flags(P.offsOut.data(), P.nC, R.d, z.data(), err, sizeof err);
        if (rc) fail(rc, "%s", err);
        keep.clear();
        for (int c = 0; c < P.nC; ++c)
          if (!z[c]) keep.push_back(c);
      }
    }
    qsx_tensor_out* out = nullptr;
    static const double kNone = 0.0;
    download_staged(R.d, P.offsOut.back(), s, [&](const double* h) { out = make_out(P, keep, h ? h : &kNone); });
    return out;
  }
};

// ------------------------------------------------------------ host fold (complex leaves)
struct HostOperand {
  const qsx_tensor* t = nullptr;
  std::shared_ptr<qsx_tensor_out> out;  // owns the data of an intermediate
  std::shared_ptr<Held> held;
  CellOpt op;
};

struct HostFold {
  const qsx_opts* o;
  HostOperand operand(const qsx_cell& c) {
    check_item(c);
    HostOperand x;
    if (c.T) x.t = c.T;
    else x = fold(c.sub, c.nsub);
    x.op = CellOpt();
    read_options(c, x.op);
    return x;
  }
  HostOperand fold(const qsx_cell* items, int n) {
    std::vector<HostOperand> ops;
    for (int i = 0; i < n; ++i) ops.push_back(operand(items[i]));
    if (ops.size() < 2) fail(QSX_ERR_ARG, "invalid usage #2: at least two QSpaces required for each contraction");
    HostOperand cur = ops.back();
    int cflag = cur.op.conj;
    std::vector<char> skipB = cell_skip(cur.op, cur.t->rank);
    for (int k = (int)ops.size() - 2; k >= 0; --k) {
      HostOperand& a = ops[k];
      TensorDesc DA = make_desc(a.t, "A"), DB = make_desc(cur.t, "B");
      std::vector<char> skipA = cell_skip(a.op, a.t->rank);
      std::vector<int32_t> ia, ib;
      int cA = 0;
      match_itags(DA, DB, a.op.conj, cflag, ia, ib, cA, &skipA, &skipB);
      qsx_opts so = o ? *o : qsx_opts{};
      so.skip_zero = 1;
      so.hconj = 0;
      so.fermA = so.fermB = nullptr;
      qsx_tensor_out* r = contract_oneshot(a.t, ia.data(), (int)ia.size(), cA, cur.t, ib.data(), (int)ib.size(),
                                           cflag, nullptr, 0, &so);
      HostOperand x;
      x.out = std::shared_ptr<qsx_tensor_out>(r, OutDel());
      x.held = std::shared_ptr<Held>(view_out(r).release());
      x.t = &x.held->t;
      cur = x;
      cflag = 0;
      skipB.assign(cur.t->rank, 0);
    }
    return cur;
  }
};

qsx_tensor_out* copy_out(const qsx_tensor* T) {
  // a leaf returned as it is: the identity permutation of permuteQS copies it
  return permute_out(T, {}, 0).release();
}

}  // namespace

// contractQS usage #2 (contractQS_itags, Source/contractQS.cc:634-808)
qsx_tensor_out* contract_cell(const qsx_cell* items, int n, const int32_t* perm, int nperm, const qsx_opts* o) {
  if (!items || n <= 0) fail(QSX_ERR_ARG, "invalid usage #2: no QSpace");
  ensure_device();
  const int skip_zero = o ? o->skip_zero : 1;
  OutPtr res;
  if (n >= 2) {
    if (any_complex(items, n)) {
      HostFold f{o};
      HostOperand r = f.fold(items, n);
      res = permute_out(r.t, {}, 0);
    } else {
      DevFold f{o, default_stream(), {}, {}};
      DevOperand r = f.fold(items, n);
      res.reset(f.download(r, skip_zero));
    }
  } else {
    // one top-level operand: its conj flag, then its index option as a
    // permutation (contractQS_itags at level 0 with one entry, :676-704)
    const qsx_cell& c = items[0];
    check_item(c);
    CellOpt op;
    read_options(c, op);
    OutPtr base;
    std::unique_ptr<Held> hv;
    const qsx_tensor* T = c.T;
    if (!T) {
      qsx_tensor_out* r = contract_cell(c.sub, c.nsub, nullptr, 0, o);
      base.reset(r);
      hv = view_out(r);
      T = &hv->t;
    }
    OutPtr cur;
    std::unique_ptr<Held> cv;
    if (op.conj) {
      cur = permute_out(T, {}, 1);
      cv = view_out(cur.get());
      T = &cv->t;
    }
    if (op.has_idx) res = permute_out(T, op.idx, 0);
    else if (cur) res = std::move(cur);
    else if (base) res = std::move(base);
    else res.reset(copy_out(T));
  }
  if (nperm > 0) {
    if (nperm > res->rank) fail(QSX_ERR_ARG, "invalid input permutation (len=%d/%d)", nperm, res->rank);
    std::unique_ptr<Held> v = view_out(res.get());
    res = permute_out(&v->t, std::vector<int>(perm, perm + nperm), 0);
  }
  return res.release();
}

}  // namespace qsx

using namespace qsx;

int qsx_itag_match(const qsx_tensor* A, int32_t nopta, const char* const* opta, const int32_t* idxa, int32_t nidxa,
                   const qsx_tensor* B, int32_t noptb, const char* const* optb, const int32_t* idxb, int32_t nidxb,
                   int32_t* ia, int32_t* ib, int32_t* nctr, int32_t* conjA, int32_t* conjB, char* err,
                   size_t errlen) {
  return guarded(err, errlen, [&]() -> int {
    if (!A || !B || !ia || !ib || !nctr) fail(QSX_ERR_ARG, "null argument");
    qsx_cell ca{A, 0, nullptr, nopta, opta, nidxa, idxa}, cb{B, 0, nullptr, noptb, optb, nidxb, idxb};
    CellOpt oa, ob;
    read_options(ca, oa);
    read_options(cb, ob);
    TensorDesc DA = make_desc(A, "A"), DB = make_desc(B, "B");
    std::vector<char> sa = cell_skip(oa, A->rank), sb = cell_skip(ob, B->rank);
    std::vector<int32_t> va, vb;
    int cA = 0;
    match_itags(DA, DB, oa.conj, ob.conj, va, vb, cA, &sa, &sb);
    std::copy(va.begin(), va.end(), ia);
    std::copy(vb.begin(), vb.end(), ib);
    *nctr = (int32_t)va.size();
    if (conjA) *conjA = cA;
    if (conjB) *conjB = ob.conj;
    return 0;
  });
}

int qsx_contract_cell(const qsx_cell* items, int32_t nitems, const int32_t* perm, int32_t nperm,
                      const qsx_opts* opts, qsx_tensor_out** C, char* err, size_t errlen) {
  return guarded(err, errlen, [&]() -> int {
    if (!C) fail(QSX_ERR_ARG, "null result pointer");
    *C = contract_cell(items, nitems, perm, nperm, opts);
    return 0;
  });
}
