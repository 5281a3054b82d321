// ortho.cc -- host side of the truncation after the eigensolve: the C-ABI
// entry point qsx_ortho (the core of orthoQS / mpsOrthoQS).
//
// Reference semantics restated (Source/mpsortho.cc):
//  * SVD_Data::blockSVD (:545-613): the blocks of PSI are grouped by the
//    label sum (QSpace::getQsum, Source/QSpace.hh:2385-2420: plain sums for
//    U(1) labels) over legs [0] for K = 1, [r-1] for K = r-1 and [0..K-1]
//    otherwise, the groups in stable ascending order of that record
//    (groupRecs); for K other than 1 the first K legs must be incoming and
//    legs K+1..r-1 share leg K's direction.
//  * QBlock::toBlockMatrix (:30-256): a group's matrix has the sorted
//    distinct label records of legs 0..K-1 as row blocks and those of legs
//    K..r-1 as column blocks; each block enters in its column-major form
//    with its first K dims fused (cfac = 1 for abelian symmetries); a
//    repeated (row, column) record or a dims mismatch is an error.
//  * wbSVD (Source/wbarray_blas.cc:1090-1180): the economy SVD, singular
//    values descending; here a one-sided Jacobi SVD on the device (svd.hip)
//    whose values are sorted descending per group, ties by column.
//  * dmrgTruncate (:615-954): markSet 'desc' over all singular values
//    (Source/gathered.cc:1822-1884,1957-2079) with Nkmin, Nkeep and
//    stol; s2t = discarded sum of squares, sfac = |S| / sqrt(|S|^2 - s2t)
//    when something is discarded, applied to every singular value; the
//    flag count of its notes; the itags of the new leg (tU, tV, :786-818);
//    per group with kept states A1 gets one block per row record (legs
//    0..K-1 and the new leg, label = the group's label sum) and A2 one
//    block per column record (legs K..r-1 and the new leg) holding
//    V(:,kept) * S(kept) * sfac; A2's blocks of norm <= stol*1e-14 are
//    dropped unless A2 is a rank-2 block-diagonal matrix (SkipZeroData
//    with bflag 'b', Source/QSpace.cc:3201-3270).
//  * SU(2) (alone or with abelian factors) for K = 1 or r-1 splits of rank <= 3
//    (outer multiplicity 1): toBlockMatrix (:62-90, 192-210) scales each
//    block by prod cgw / sqrt(qdim_tot), qdim_tot the multiplet dimension
//    of the grouping leg; dmrgTruncate (:668-684, get_s2t in
//    Source/mpsortho.hh:98-106) weights each squared singular value by
//    qdim_tot and sorts the records (S, qdim_tot) descending; A1 keeps
//    PSI's CGT with weight sqrt(qdim_tot) renormalised to normExt
//    (QBlock::initFromBlockMatrix :447-452, CRef::init_wId and NormSignW,
//    Source/clebsch.hh:3287-3306, clebsch.cc:5035-5052), so its data carry
//    sqrt(qdim_tot) / normExt; A2 is the rank-2 identity, weight sqrt(d) =
//    normExt (CRef::initIdentityR).
//    For K = 1 of a rank-3 tensor A1 is the rank-2 identity and A2 carries
//    PSI's CGT permuted to (legs 1, 2, new) (cgp2 = [1,2,0]).
//  * Not restated: ranks above 3 (outer multiplicity), and the identity-CG
//    flip of a K = r-1
//    split with an incoming last leg (revert_Uidx, :786-797 and 887-916):
//    QSX_ERR_UNSUPPORTED.
#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <set>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "qsops.hpp"

namespace qsx {
namespace {

uint64_t pack_name(const char* s) {
  uint64_t v = 0;
  if (!s) return 0;
  std::string t(s);
  while (!t.empty() && t.back() == '*') t.pop_back();
  for (size_t i = 0; i < t.size() && i < 8; ++i) v |= (uint64_t)(unsigned char)t[i] << (8 * i);
  return v;
}

// one symmetry group: its blocks, row / column records and the matrix
struct Group {
  std::vector<int> blk;                 // PSI blocks in stable order
  std::vector<std::vector<int32_t>> rows, cols;
  std::vector<std::vector<int64_t>> S1, S2;  // dims of the row / column records
  std::vector<int64_t> cD1, cD2;        // row / column offsets (size + 1)
  std::vector<int32_t> key;             // the group's label sum
  int64_t dim1 = 0, dim2 = 0;
  bool trans = false;                   // stored as the transpose (dim1 < dim2)
  int64_t m = 0, n = 0;                 // stored m x n, m >= n
  int64_t a_off = 0, v_off = 0, s_off = 0;
  int item = -1;                        // SVD item (-1: no singular values)
  int64_t qd = 1;                       // multiplet dimension (SU(2) factors)
  std::vector<double> S;                // sorted descending
  std::vector<int32_t> col;             // column of the i-th sorted value
};

// stable grouping of records (ties by index): groups in ascending order
template <class Rec>
std::vector<std::vector<int>> sorted_groups(const std::vector<Rec>& R) {
  std::vector<int> o(R.size());
  std::iota(o.begin(), o.end(), 0);
  std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return R[a] < R[b]; });
  std::vector<std::vector<int>> g;
  for (int i : o) {
    if (g.empty() || R[g.back()[0]] != R[i]) g.emplace_back();
    g.back().push_back(i);
  }
  return g;
}

// markSet(F,L,E,...,'desc') on E ascending (Source/gathered.cc:1957-2079)
int64_t mark_desc(const std::vector<double>& E, int64_t Nkmin, int64_t Nkeep, double Etrunc,
                  std::vector<char>& mark) {
  const int64_t N = (int64_t)E.size();
  mark.assign(N, 0);
  if (Nkeep < 0 || Nkeep > N) Nkeep = N;
  if (Nkeep == 0) return 0;
  if (Nkeep >= N && Etrunc <= 0) {
    std::fill(mark.begin(), mark.end(), 1);
    return N;
  }
  int64_t l;
  if (Etrunc > 0) {
    const int64_t lx = Nkmin < N ? N - Nkmin : 0;
    l = -1;  // the reference's unsigned loop running past index 0
    for (int64_t i = N - 1; i >= 0; --i) {
      if (E[i] <= Etrunc && i < lx) {
        l = i;
        break;
      }
    }
    if ((l < 0 ? N : N - 1 - l) > Nkeep) l = N - 1 - Nkeep;
  } else {
    l = N - 1 - Nkeep;
  }
  if (l < 0) {
    std::fill(mark.begin(), mark.end(), 1);
    return N;
  }
  if (l == N - 1) return 0;
  for (int64_t i = l + 1; i < N; ++i) mark[i] = 1;
  return N - 1 - l;
}

// the CGT weight of block i for symmetry j (normExt when PSI carries none)
double block_cgw(const TensorDesc& P, int i, int j, const std::vector<int>& dir) {
  if (P.cgw_given) return P.cgw[(size_t)i * P.nsym + j];
  std::vector<int> tj(P.r);
  for (int l = 0; l < P.r; ++l) tj[l] = P.label(i, l, j);
  return su2_norm_ext(tj, dir);
}

struct OrthoResult {
  qsx_tensor_out* A1 = nullptr;
  qsx_tensor_out* A2 = nullptr;
  qsx_ortho_info* info = nullptr;
};

OrthoResult ortho_fused(const qsx_tensor* pt, const TensorDesc& P, int Nkmin, int64_t Nkeep, double stol,
                        const char* itag);

OrthoResult ortho_oneshot(const qsx_tensor* pt, int K, int Nkmin, int64_t Nkeep, double stol, const char* itag) {
  if (pt && pt->data_im) fail(QSX_ERR_UNSUPPORTED, "orthoQS: complex data not supported (real QSpaces only)");
  TensorDesc P = make_desc(pt, "PSI");
  const int r = P.r, q = P.qdim;
  bool su2 = false;
  for (int t : P.symtype) {
    if (t == SYM_SU2) su2 = true;
    else if (t != SYM_A) fail(QSX_ERR_UNSUPPORTED, "orthoQS: symmetry %s not supported", P.qtype.c_str());
  }
  if (P.n == 0) fail(QSX_ERR_ARG, "orthoQS: got empty PSI");
  if (K < 1 || K >= r) fail(QSX_ERR_ARG, "orthoQS: index specs out of bounds (%d/%d)", K, r);
  if ((int)P.itags.size() != r) fail(QSX_ERR_ITAG, "orthoQS: got invalid/empty itags (%d/%d)", (int)P.itags.size(), r);
  const std::vector<int> dir = leg_dirs(P.itags, r);
  if (K != 1) {
    for (int i = 0; i < K; ++i)
      if (dir[i] <= 0) fail(QSX_ERR_ITAG, "orthoQS: only ingoing indices can be fused into outgoing (K=%d)", K);
    for (int i = K + 1; i < r; ++i)
      if (dir[i] != dir[K]) fail(QSX_ERR_ITAG, "orthoQS: 2nd/right group is mixed in+out (K=%d)", K);
  }
  if (P.total() && !P.hdata) fail(QSX_ERR_ARG, "orthoQS: PSI without data");
  if (su2 && r == 4 && K == 2 && P.nsym == q) return ortho_fused(pt, P, Nkmin, Nkeep, stol, itag);
  require_no_om(P, "orthoQS");
  if (su2 && (r > 3 || (K != 1 && K != r - 1)))
    fail(QSX_ERR_UNSUPPORTED,
         "orthoQS: non-abelian symmetries (%s) need K = 1 or r-1 and rank <= 3, or K = 2 of rank 4 (got K=%d/%d)",
         P.qtype.c_str(), K, r);
  if (su2 && P.nsym != q) fail(QSX_ERR_UNSUPPORTED, "orthoQS: symmetries with several labels (%s)", P.qtype.c_str());

  // ---- groups and their block matrices
  std::vector<int> Ib;
  if (K == 1) Ib = {0};
  else if (K == r - 1) Ib = {r - 1};
  else for (int j = 0; j < K; ++j) Ib.push_back(j);
  std::vector<std::vector<int32_t>> Qb(P.n, std::vector<int32_t>(q, 0));
  for (int i = 0; i < P.n; ++i)
    for (int j : Ib)
      for (int c = 0; c < q; ++c) Qb[i][c] += P.label(i, j, c);
  std::vector<Group> G;
  for (auto& g : sorted_groups(Qb)) {
    Group b;
    b.blk = g;
    b.key = Qb[g[0]];
    std::vector<std::vector<int32_t>> rr(g.size()), cc(g.size());
    for (size_t t = 0; t < g.size(); ++t) {
      for (int j = 0; j < K; ++j)
        for (int c = 0; c < q; ++c) rr[t].push_back(P.label(g[t], j, c));
      for (int j = K; j < r; ++j)
        for (int c = 0; c < q; ++c) cc[t].push_back(P.label(g[t], j, c));
    }
    auto g1 = sorted_groups(rr), g2 = sorted_groups(cc);
    std::vector<int> ri(g.size()), ci(g.size());
    for (size_t a = 0; a < g1.size(); ++a)
      for (int t : g1[a]) ri[t] = (int)a;
    for (size_t a = 0; a < g2.size(); ++a)
      for (int t : g2[a]) ci[t] = (int)a;
    b.S1.assign(g1.size(), {});
    b.S2.assign(g2.size(), {});
    std::vector<char> seen(g1.size() * g2.size(), 0);
    for (size_t t = 0; t < g.size(); ++t) {
      const int a = ri[t], c = ci[t];
      if (seen[(size_t)a * g2.size() + c]++) fail(QSX_ERR_STRUCT, "orthoQS: block %d not unique", g[t] + 1);
      std::vector<int64_t> s1(P.dims.begin() + (size_t)g[t] * r, P.dims.begin() + (size_t)g[t] * r + K);
      std::vector<int64_t> s2(P.dims.begin() + (size_t)g[t] * r + K, P.dims.begin() + (size_t)(g[t] + 1) * r);
      if (b.S1[a].empty()) b.S1[a] = s1;
      else if (b.S1[a] != s1) fail(QSX_ERR_STRUCT, "orthoQS: block size mismatch (block %d)", g[t] + 1);
      if (b.S2[c].empty()) b.S2[c] = s2;
      else if (b.S2[c] != s2) fail(QSX_ERR_STRUCT, "orthoQS: block size mismatch (block %d)", g[t] + 1);
    }
    for (auto& x : g1) b.rows.push_back(rr[x[0]]);
    for (auto& x : g2) b.cols.push_back(cc[x[0]]);
    b.cD1.assign(1, 0);
    for (auto& s : b.S1) {
      int64_t d = 1;
      for (int64_t x : s) d *= x;
      b.cD1.push_back(b.cD1.back() + d);
    }
    b.cD2.assign(1, 0);
    for (auto& s : b.S2) {
      int64_t d = 1;
      for (int64_t x : s) d *= x;
      b.cD2.push_back(b.cD2.back() + d);
    }
    b.dim1 = b.cD1.back();
    b.dim2 = b.cD2.back();
    // multiplet dimension of the grouping leg (leg 0 for K = 1, leg K else)
    for (int j = 0; j < P.nsym && su2; ++j)
      if (P.symtype[j] == SYM_SU2) b.qd *= (int64_t)P.label(g[0], K == 1 ? 0 : K, j) + 1;
    b.trans = b.dim1 < b.dim2;
    b.m = std::max(b.dim1, b.dim2);
    b.n = std::min(b.dim1, b.dim2);
    G.push_back(std::move(b));
  }

  // ---- device: assemble, decompose
  ensure_device();
  hipStream_t s = default_stream();
  std::vector<SvdGather> gath;
  std::vector<SvdItem> items;
  int64_t work = 0, nsig = 0;
  for (auto& b : G) {
    if (b.n == 0) continue;
    if (b.n > INT32_MAX || b.m > INT32_MAX) fail(QSX_ERR_UNSUPPORTED, "orthoQS: group matrix too large");
    b.a_off = work;
    work += b.m * b.n;
    b.v_off = work;
    work += b.n * b.n;
    b.s_off = nsig;
    nsig += b.n;
    b.item = (int)items.size();
    SvdItem it{};
    it.a_off = b.a_off;
    it.v_off = b.v_off;
    it.s_off = b.s_off;
    it.m = (int32_t)b.m;
    it.n = (int32_t)b.n;
    it.lds = (b.m * b.n + b.n * b.n) <= SVD_LDS_DOUBLES ? 1 : 0;
    items.push_back(it);
    for (size_t t = 0; t < b.blk.size(); ++t) {
      const int i = b.blk[t];
      if (!P.numel(i)) continue;
      // the block's row / column block in its group
      std::vector<int32_t> rrec, crec;
      for (int j = 0; j < K; ++j)
        for (int c = 0; c < q; ++c) rrec.push_back(P.label(i, j, c));
      for (int j = K; j < r; ++j)
        for (int c = 0; c < q; ++c) crec.push_back(P.label(i, j, c));
      const size_t a = std::lower_bound(b.rows.begin(), b.rows.end(), rrec) - b.rows.begin();
      const size_t c = std::lower_bound(b.cols.begin(), b.cols.end(), crec) - b.cols.begin();
      SvdGather gg{};
      gg.src_off = P.offs[i];
      gg.d1 = (int32_t)(b.cD1[a + 1] - b.cD1[a]);
      gg.d2 = (int32_t)(b.cD2[c + 1] - b.cD2[c]);
      gg.ld = (int32_t)b.m;
      gg.trans = b.trans ? 1 : 0;
      gg.scale = 1.0;
      if (su2) {  // cfac = prod cgw / sqrt(qdim_tot)
        for (int j = 0; j < P.nsym; ++j)
          if (P.symtype[j] == SYM_SU2) gg.scale *= block_cgw(P, i, j, dir);
        gg.scale /= std::sqrt((double)b.qd);
      }
      gg.dst_off = b.trans ? b.a_off + b.cD2[c] + b.cD1[a] * b.m : b.a_off + b.cD1[a] + b.cD2[c] * b.m;
      if ((int64_t)gg.d1 * gg.d2 != P.numel(i)) fail(QSX_ERR_STRUCT, "orthoQS: block %d dims", i + 1);
      gath.push_back(gg);
    }
  }
  if (nsig == 0) fail(QSX_ERR_ARG, "orthoQS: got empty SVD data");
  // every access of the kernels inside its buffer (checked before launch)
  for (const SvdItem& it : items) {
    const int64_t mn = (int64_t)it.m * it.n, nn = (int64_t)it.n * it.n;
    if (it.m < it.n || it.n < 1 || it.a_off < 0 || it.a_off + mn > work || it.v_off + nn > work ||
        it.s_off + it.n > nsig || (it.lds && mn + nn > SVD_LDS_DOUBLES))
      fail(QSX_ERR_STRUCT, "orthoQS: internal SVD item out of range");
  }
  for (const SvdGather& g : gath) {
    const int64_t last = g.trans ? g.dst_off + (g.d2 - 1) + (int64_t)(g.d1 - 1) * g.ld
                                 : g.dst_off + (g.d1 - 1) + (int64_t)(g.d2 - 1) * g.ld;
    if (g.d1 < 1 || g.d2 < 1 || g.dst_off < 0 || last >= work || g.src_off < 0 ||
        g.src_off + (int64_t)g.d1 * g.d2 > P.total() || (g.trans ? g.d2 : g.d1) > g.ld)
      fail(QSX_ERR_STRUCT, "orthoQS: internal gather item out of range");
  }
  DevMem dpsi(P.total()), dwork(work), dsig(nsig);
  struct Free {
    std::vector<void*> p;
    void* add(size_t bytes) {
      p.push_back(dev_alloc(bytes));
      return p.back();
    }
    ~Free() {
      for (void* x : p) dev_free(x);
    }
  } fr;
  // groups that fit in LDS: one workgroup each (k_jacobi_svd); the others:
  // block one-sided Jacobi over column blocks of at most cb columns, two
  // blocks (and their columns of V) staged in LDS per workgroup
  std::vector<SvdItem> small;
  std::vector<int32_t> big, nbk, cbk;
  for (size_t i = 0; i < items.size(); ++i) {
    if (items[i].lds) {  // checked above: fits in LDS
      small.push_back(items[i]);
      continue;
    }
    const int64_t m = items[i].m, n = items[i].n;
    const int64_t cmax = SVD_LDS_DOUBLES / (2 * (m + n));
    if (cmax < 1) fail(QSX_ERR_UNSUPPORTED, "orthoQS: group matrix %lld x %lld too large", (long long)m, (long long)n);
    int64_t nb = (n + cmax - 1) / cmax;
    nb += nb & 1;
    if (nb < 2) nb = 2;
    const int64_t cb = (n + nb - 1) / nb;
    if (2 * cb * (m + n) > SVD_LDS_DOUBLES) fail(QSX_ERR_STRUCT, "orthoQS: internal column block too large");
    big.push_back((int32_t)i);
    nbk.push_back((int32_t)nb);
    cbk.push_back((int32_t)cb);
  }
  int32_t* dsw = (int32_t*)fr.add(sizeof(int32_t) * std::max<size_t>(1, small.size()));
  SvdItem* dsmall = (SvdItem*)fr.add(sizeof(SvdItem) * std::max<size_t>(1, small.size()));
  SvdItem* ditems = (SvdItem*)fr.add(sizeof(SvdItem) * items.size());
  SvdGather* dgath = (SvdGather*)fr.add(sizeof(SvdGather) * std::max<size_t>(1, gath.size()));
  int32_t* drot = (int32_t*)fr.add(sizeof(int32_t) * items.size());
  int32_t* dbig = (int32_t*)fr.add(sizeof(int32_t) * std::max<size_t>(1, big.size()));
  upload_packed(P, P.hdata, dpsi.p, s);
  QSX_HIP(hipMemcpyAsync(ditems, items.data(), sizeof(SvdItem) * items.size(), hipMemcpyHostToDevice, s));
  if (!small.empty())
    QSX_HIP(hipMemcpyAsync(dsmall, small.data(), sizeof(SvdItem) * small.size(), hipMemcpyHostToDevice, s));
  if (!big.empty())
    QSX_HIP(hipMemcpyAsync(dbig, big.data(), sizeof(int32_t) * big.size(), hipMemcpyHostToDevice, s));
  if (!gath.empty())
    QSX_HIP(hipMemcpyAsync(dgath, gath.data(), sizeof(SvdGather) * gath.size(), hipMemcpyHostToDevice, s));
  Timer tm(s);
  launch_fill_zero(dwork.p, work, s);
  launch_svd_gather(dgath, (int)gath.size(), dpsi.p, dwork.p, s);
  // the groups in LDS run on the side stream, beside the block steps
  hipStream_t s2 = side_stream();
  struct Ev {
    hipEvent_t e = nullptr;
    Ev() { QSX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }
    ~Ev() { (void)hipEventDestroy(e); }
  } ready, done;
  // on an error below, both streams drain before the buffers above are
  // freed (declared after them, so destroyed first)
  struct Drain {
    hipStream_t a, b;
    ~Drain() {
      (void)hipStreamSynchronize(a);
      (void)hipStreamSynchronize(b);
    }
  } drain{s, s2};
  QSX_HIP(hipEventRecord(ready.e, s));
  QSX_HIP(hipStreamWaitEvent(s2, ready.e, 0));
  launch_jacobi_svd(dsmall, (int)small.size(), dwork.p, dsig.p, dsw, s2);
  QSX_HIP(hipEventRecord(done.e, s2));
  int maxsw = 0;
  if (!big.empty()) {
    launch_svd_init_v(ditems, dbig, (int)big.size(), dwork.p, s);
    std::vector<int> active(big.size());
    std::iota(active.begin(), active.end(), 0);
    SvdBlockStep* dsteps = nullptr;
    size_t cap = 0;
    int sw = 0;
    std::vector<int32_t> rot(items.size());
    for (; sw < jacobi_max_sweeps() && !active.empty(); ++sw) {
      // a sweep: the pairs within every column block (one launch), then
      // the pairs across blocks P, Q in the round-robin order of the blocks
      // (nb - 1 launches of nb / 2 disjoint block pairs per group)
      std::vector<SvdBlockStep> steps;
      std::vector<size_t> off(1, 0);
      int maxnb = 0;
      for (int a : active) maxnb = std::max(maxnb, nbk[a]);
      for (int a : active) {
        const int nb = nbk[a], cb = cbk[a], n = items[big[a]].n;
        for (int k = 0; k < nb; ++k) {
          SvdBlockStep b{};
          b.grp = big[a];
          b.p0 = std::min(n, k * cb);
          b.p1 = std::min(n, (k + 1) * cb);
          b.q0 = b.q1 = b.p1;
          if (b.p1 - b.p0 >= 2) steps.push_back(b);
        }
      }
      off.push_back(steps.size());
      for (int st = 0; st < maxnb - 1; ++st) {
        for (int a : active) {
          const int nb = nbk[a], cb = cbk[a], n = items[big[a]].n;
          if (st >= nb - 1) continue;
          for (int k = 0; k < nb / 2; ++k) {
            const int P0 = k == 0 ? nb - 1 : (st + k) % (nb - 1);
            const int Q0 = k == 0 ? st : (st - k + nb - 1) % (nb - 1);
            SvdBlockStep b{};
            b.grp = big[a];
            b.p0 = std::min(n, P0 * cb);
            b.p1 = std::min(n, (P0 + 1) * cb);
            b.q0 = std::min(n, Q0 * cb);
            b.q1 = std::min(n, (Q0 + 1) * cb);
            if (b.p1 == b.p0 || b.q1 == b.q0) continue;
            steps.push_back(b);
          }
        }
        off.push_back(steps.size());
      }
      if (steps.size() > cap) {
        cap = steps.size();
        dsteps = (SvdBlockStep*)fr.add(sizeof(SvdBlockStep) * cap);
      }
      if (!steps.empty())
        QSX_HIP(hipMemcpyAsync(dsteps, steps.data(), sizeof(SvdBlockStep) * steps.size(), hipMemcpyHostToDevice, s));
      QSX_HIP(hipMemsetAsync(drot, 0, sizeof(int32_t) * items.size(), s));
      for (size_t st = 0; st + 1 < off.size(); ++st)
        launch_jacobi_block(ditems, dsteps + off[st], (int)(off[st + 1] - off[st]), dwork.p, drot, sw, s);
      QSX_HIP(hipMemcpyAsync(rot.data(), drot, sizeof(int32_t) * items.size(), hipMemcpyDeviceToHost, s));
      QSX_HIP(hipStreamSynchronize(s));
      std::vector<int> next;
      for (int a : active)
        if (rot[big[a]]) next.push_back(a);
      active.swap(next);
    }
    if (!active.empty()) fail(QSX_ERR_ARG, "orthoQS: block Jacobi SVD did not converge in %d sweeps", sw);
    maxsw = sw;
    launch_svd_finalize(ditems, dbig, (int)big.size(), dwork.p, dsig.p, s);
  }
  QSX_HIP(hipStreamWaitEvent(s, done.e, 0));
  std::vector<double> sig(nsig);
  std::vector<int32_t> sweeps(small.size());
  QSX_HIP(hipMemcpyAsync(sig.data(), dsig.p, sizeof(double) * nsig, hipMemcpyDeviceToHost, s));
  if (!small.empty())
    QSX_HIP(hipMemcpyAsync(sweeps.data(), dsw, sizeof(int32_t) * small.size(), hipMemcpyDeviceToHost, s));
  QSX_HIP(hipStreamSynchronize(s));
  for (int x : sweeps) {
    if (x >= jacobi_max_sweeps()) fail(QSX_ERR_ARG, "orthoQS: Jacobi SVD did not converge in %d sweeps", x);
    maxsw = std::max(maxsw, x + 1);
  }

  // ---- truncation (dmrgTruncate)
  std::vector<double> E, Ed;  // singular values and their multiplet dimensions
  for (auto& b : G) {
    if (b.item < 0) continue;
    b.col.resize(b.n);
    std::iota(b.col.begin(), b.col.end(), 0);
    const double* sg = sig.data() + b.s_off;
    std::stable_sort(b.col.begin(), b.col.end(), [&](int x, int y) { return sg[x] > sg[y]; });
    b.S.resize(b.n);
    for (int64_t i = 0; i < b.n; ++i) b.S[i] = sg[b.col[i]];
    E.insert(E.end(), b.S.begin(), b.S.end());
    Ed.insert(Ed.end(), b.S.size(), (double)b.qd);
  }
  const int64_t N = (int64_t)E.size();
  double snorm2 = 0.0;
  for (int64_t i = 0; i < N; ++i) snorm2 += E[i] * E[i] * Ed[i];
  // the records (S, qdim) sorted descending (SortRecs(P,-1))
  std::vector<int64_t> Ps(N);
  std::iota(Ps.begin(), Ps.end(), 0);
  std::sort(Ps.begin(), Ps.end(), [&](int64_t a, int64_t b) {
    return E[a] != E[b] ? E[a] > E[b] : Ed[a] > Ed[b];
  });
  std::vector<double> SM(N), SMd(N);
  for (int64_t i = 0; i < N; ++i) {
    SM[i] = E[Ps[i]];
    SMd[i] = Ed[Ps[i]];
  }
  // markSet over all values (stable ascending sort, marks mapped back)
  std::vector<char> markAll(N, 0);
  int64_t NK = Nkeep < 0 ? N : Nkeep;
  const int64_t nkmin = std::max(0, Nkmin);
  if (NK >= N && stol <= 0) {
    std::fill(markAll.begin(), markAll.end(), 1);
    NK = N;
  } else if (NK > 0) {
    std::vector<int64_t> Pm(N);
    std::iota(Pm.begin(), Pm.end(), 0);
    std::stable_sort(Pm.begin(), Pm.end(), [&](int64_t a, int64_t b) { return E[a] < E[b]; });
    std::vector<double> Es(N);
    for (int64_t i = 0; i < N; ++i) Es[i] = E[Pm[i]];
    std::vector<char> mm;
    NK = mark_desc(Es, nkmin, NK, stol, mm);
    for (int64_t i = 0; i < N; ++i) markAll[Pm[i]] = mm[i];
  }
  double s2t = 0.0;
  int64_t nk = 0;
  for (int64_t i = 0; i < N; ++i) {
    if (markAll[i]) ++nk;
    else s2t += E[i] * E[i] * Ed[i];
  }
  int flag = 0;
  double xtol2 = stol > 0 ? stol * 1e-14 : 0.0;
  const bool unlimited = Nkeep < 0;
  if (SM.back() > stol) {
    if (!unlimited && N > Nkeep) {
      if (stol) ++flag;
    } else {
      xtol2 = 0.0;
    }
  } else if (!(!unlimited && NK < Nkeep) && NK < N) {
    ++flag;
  }
  double sfac = 1.0;
  if (s2t) {
    if (snorm2 > s2t) {
      sfac = std::sqrt(snorm2) / std::sqrt(snorm2 - s2t);
      if (std::fabs(sfac - 1) > 0.01) ++flag;
    } else if (snorm2 != s2t) {
      fail(QSX_ERR_ARG, "orthoQS: discarded weight %g exceeds the total %g", s2t, snorm2);
    }
  }

  // ---- itags of the new leg (tU, tV)
  const uint64_t tx = pack_name(itag);
  uint64_t tu;
  if (r - K == 1) {
    const uint64_t last = P.itags[r - 1];
    tu = tx ? (tx | (last & 0x80)) : last;
    if (!tag_out(tu) && nk) {
      const bool is_op = r == 2 && (tag_out(P.itags[0]) != tag_out(P.itags[1]));
      if (!is_op)
        fail(QSX_ERR_UNSUPPORTED,
             "orthoQS: K = r-1 with an incoming last leg (the reference's identity-CG flip of the new leg, "
             "revert_Uidx, is not supported)");
    }
  } else if (K == 1) {
    const uint64_t first = P.itags[0];
    tu = (tx ? (tx | (first & 0x80)) : first) ^ 0x80;
  } else {
    tu = tag_out(P.itags[0]) ? tx : (tx | 0x80);
  }
  std::vector<uint64_t> tU(P.itags.begin(), P.itags.begin() + K), tV(P.itags.begin() + K, P.itags.end());
  tU.push_back(tu);
  tV.push_back(tu ^ 0x80);
  const std::vector<int> dirU = leg_dirs(tU, K + 1), dirV = leg_dirs(tV, r - K + 1);

  // ---- result blocks
  struct OutBlk {
    std::vector<int32_t> Q;
    std::vector<int64_t> dims;
    int64_t off;
    int grp;
  };
  std::vector<OutBlk> U, X;
  std::vector<SvdScatter> sc;
  std::vector<SvdCol> cols;
  int64_t totU = 0, totX = 0;
  std::vector<std::vector<int>> keptOf(G.size());
  {
    int64_t o = 0;
    for (size_t gi = 0; gi < G.size(); ++gi) {
      auto& b = G[gi];
      if (b.item < 0) continue;
      for (int64_t i = 0; i < b.n; ++i)
        if (markAll[o + i]) keptOf[gi].push_back((int)i);
      o += b.n;
    }
  }
  for (size_t gi = 0; gi < G.size(); ++gi) {
    auto& b = G[gi];
    const auto& kept = keptOf[gi];
    if (kept.empty()) continue;
    const int nkg = (int)kept.size();
    // new-leg label: the right record (K = r-1) or the left labels' sum
    std::vector<int32_t> nl(q, 0);
    if (r - K == 1) {
      for (int c = 0; c < q; ++c) nl[c] = b.cols[0][c];
    } else {
      for (int j = 0; j < K; ++j)
        for (int c = 0; c < q; ++c) nl[c] += b.rows[0][(size_t)j * q + c];
    }
    // A1 keeps PSI's CGT: the data carry sqrt(qdim_tot) / prod normExt
    auto u_scale = [&](const std::vector<int32_t>& Qrec) {
      if (!su2 || K != r - 1) return 1.0;  // K = 1 of rank 3: A1 is the identity
      double ne = 1.0;
      for (int j = 0; j < P.nsym; ++j) {
        if (P.symtype[j] != SYM_SU2) continue;
        std::vector<int> tj(K + 1);
        for (int l = 0; l <= K; ++l) tj[l] = Qrec[(size_t)l * q + j];
        ne *= su2_norm_ext(tj, dirU);
      }
      return std::sqrt((double)b.qd) / ne;
    };
    // K = 1 of a rank-3 tensor: A2 carries PSI's CGT (weight sqrt(qdim_tot)
    // renormalised to normExt) permuted from (new, 1, 2) to (1, 2, new), the
    // canonical CGT of that order times su2_perm_sign (initFromBlockMatrix
    // with cgp2 = [1,2,0], Source/mpsortho.cc:440-452)
    auto v_scale = [&](const std::vector<int32_t>& crec) {
      if (!su2 || K != 1 || r != 3) return 1.0;
      double f = std::sqrt((double)b.qd);
      for (int j = 0; j < P.nsym; ++j) {
        if (P.symtype[j] != SYM_SU2) continue;
        const std::vector<int> tj = {nl[j], crec[j], crec[(size_t)q + j]};
        f *= su2_perm_sign(tj, dir, {1, 2, 0}) / su2_norm_ext(tj, dir);
      }
      return f;
    };
    // U columns: the left vectors of M (A part, or V part when transposed)
    const int64_t u_base = b.trans ? b.v_off : b.a_off, u_ld = b.trans ? b.n : b.m;
    const int64_t v_base = b.trans ? b.a_off : b.v_off, v_ld = b.trans ? b.m : b.n;
    const int32_t c0u = (int32_t)cols.size();
    for (int k : kept) cols.push_back(SvdCol{b.col[k], 0, 1.0});
    const int32_t c0v = (int32_t)cols.size();
    for (int k : kept) cols.push_back(SvdCol{b.col[k], 0, b.S[k] * sfac});
    for (size_t a = 0; a < b.rows.size(); ++a) {
      if (b.cD1[a + 1] == b.cD1[a]) continue;  // empty sub-block: no block
      OutBlk ob;
      ob.Q = b.rows[a];
      ob.Q.insert(ob.Q.end(), nl.begin(), nl.end());
      ob.dims = b.S1[a];
      ob.dims.push_back(nkg);
      ob.off = totU;
      ob.grp = (int)gi;
      const int64_t rows = b.cD1[a + 1] - b.cD1[a];
      sc.push_back(SvdScatter{totU, u_base + b.cD1[a], (int32_t)rows, nkg, (int32_t)u_ld, c0u, u_scale(ob.Q)});
      totU += rows * nkg;
      U.push_back(std::move(ob));
    }
    for (size_t c = 0; c < b.cols.size(); ++c) {
      if (b.cD2[c + 1] == b.cD2[c]) continue;
      OutBlk ob;
      ob.Q = b.cols[c];
      ob.Q.insert(ob.Q.end(), nl.begin(), nl.end());
      ob.dims = b.S2[c];
      ob.dims.push_back(nkg);
      ob.off = totX;
      ob.grp = (int)gi;
      const int64_t rows = b.cD2[c + 1] - b.cD2[c];
      sc.push_back(SvdScatter{-1 - totX, v_base + b.cD2[c], (int32_t)rows, nkg, (int32_t)v_ld, c0v,
                              v_scale(b.cols[c])});
      totX += rows * nkg;
      X.push_back(std::move(ob));
    }
  }
  // A1's blocks first, then A2's, in one output buffer
  for (auto& x : sc)
    if (x.dst_off < 0) x.dst_off = totU + (-1 - x.dst_off);
  for (const SvdScatter& x : sc) {
    int32_t maxc = 0;
    if (x.col0 < 0 || x.ncols < 1 || x.col0 + x.ncols > (int64_t)cols.size())
      fail(QSX_ERR_STRUCT, "orthoQS: internal scatter columns out of range");
    for (int c = 0; c < x.ncols; ++c) maxc = std::max(maxc, cols[x.col0 + c].col);
    if (x.rows < 1 || x.dst_off < 0 || x.dst_off + (int64_t)x.rows * x.ncols > totU + totX || x.src_off < 0 ||
        x.rows > x.ld || x.src_off + (x.rows - 1) + (int64_t)maxc * x.ld >= work)
      fail(QSX_ERR_STRUCT, "orthoQS: internal scatter item out of range");
  }
  std::vector<double> out(std::max<int64_t>(1, totU + totX));
  if (!sc.empty()) {
    double* dout = (double*)fr.add(sizeof(double) * (totU + totX));  // freed after the drain
    SvdScatter* dsc = (SvdScatter*)fr.add(sizeof(SvdScatter) * sc.size());
    SvdCol* dcol = (SvdCol*)fr.add(sizeof(SvdCol) * cols.size());
    QSX_HIP(hipMemcpyAsync(dsc, sc.data(), sizeof(SvdScatter) * sc.size(), hipMemcpyHostToDevice, s));
    QSX_HIP(hipMemcpyAsync(dcol, cols.data(), sizeof(SvdCol) * cols.size(), hipMemcpyHostToDevice, s));
    launch_svd_scatter(dsc, (int)sc.size(), dcol, dwork.p, dout, s);
    download_packed(dout, totU + totX, out.data(), s);
  }
  const float ms = tm.stop(s);
  // SkipZeroData(xtol2, 'b') of each group's A2 part: blocks with
  // |block|^2 <= xtol2^2 dropped unless the part is rank-2 block diagonal
  std::vector<char> keepX(X.size(), 1);
  {
    size_t i0 = 0;
    while (i0 < X.size()) {
      size_t i1 = i0;
      while (i1 < X.size() && X[i1].grp == X[i0].grp) ++i1;
      bool bd = (r - K + 1) == 2;
      for (size_t i = i0; i < i1 && bd; ++i)
        for (int c = 0; c < q; ++c)
          if (X[i].Q[c] != X[i].Q[q + c]) bd = false;
      if (!bd) {
        for (size_t i = i0; i < i1; ++i) {
          int64_t ne = 1;
          for (int64_t d : X[i].dims) ne *= d;
          double n2 = 0.0;
          const double* x = out.data() + totU + X[i].off;
          for (int64_t e = 0; e < ne; ++e) n2 += x[e] * x[e];
          if (n2 <= xtol2 * xtol2) keepX[i] = 0;
        }
      }
      i0 = i1;
    }
  }
  auto make = [&](const std::vector<OutBlk>& B, const std::vector<char>* keep, int rank,
                  const std::vector<uint64_t>& tags, const std::vector<int>& dd, const double* base) {
    int n = 0;
    int64_t tot = 0;
    for (size_t i = 0; i < B.size(); ++i) {
      if (keep && !(*keep)[i]) continue;
      ++n;
      int64_t ne = 1;
      for (int64_t d : B[i].dims) ne *= d;
      tot += ne;
    }
    // abelian: no weights; SU(2): normExt per block (1 for abelian factors)
    const int nsym = su2 ? P.nsym : 0;
    qsx_tensor_out* o = alloc_out(n, rank, q, nsym, P.qtype, tot);
    int k = 0;
    for (size_t i = 0; i < B.size(); ++i) {
      if (keep && !(*keep)[i]) continue;
      int64_t ne = 1;
      for (int64_t d : B[i].dims) ne *= d;
      std::copy(B[i].Q.begin(), B[i].Q.end(), o->Q + (size_t)k * rank * q);
      std::copy(B[i].dims.begin(), B[i].dims.end(), o->dims + (size_t)k * rank);
      o->offs[k + 1] = o->offs[k] + ne;
      std::copy(base + B[i].off, base + B[i].off + ne, o->data + o->offs[k]);
      for (int j = 0; j < nsym; ++j) {
        double w = 1.0;
        if (P.symtype[j] == SYM_SU2) {
          std::vector<int> tj(rank);
          for (int l = 0; l < rank; ++l) tj[l] = B[i].Q[(size_t)l * q + j];
          w = su2_norm_ext(tj, dd);
        }
        o->cgw[(size_t)k * nsym + j] = w;
      }
      ++k;
    }
    for (int l = 0; l < rank; ++l) o->itags[l] = tags[l];
    o->gpu_ms = ms;
    return o;
  };
  OrthoResult R;
  R.A1 = make(U, nullptr, K + 1, tU, dirU, out.data());
  R.A2 = make(X, &keepX, r - K + 1, tV, dirV, out.data() + totU);
  auto* info = (qsx_ortho_info*)calloc(1, sizeof(qsx_ortho_info));
  info->ntot = N;
  info->nkeep = nk;
  info->nkmin = (int32_t)nkmin;
  info->flag = flag;
  info->stol = stol;
  info->svd2tr = snorm2 > 0 ? s2t / snorm2 : 0.0;
  info->sv2tot = snorm2;
  info->sfac = sfac;
  info->svd = (double*)calloc(std::max<int64_t>(1, N), sizeof(double));
  std::copy(SM.begin(), SM.end(), info->svd);
  info->svd_qdim = (double*)calloc(std::max<int64_t>(1, N), sizeof(double));
  std::copy(SMd.begin(), SMd.end(), info->svd_qdim);
  info->gpu_ms = ms;
  info->sweeps = maxsw;
  info->ngroups = (int32_t)items.size();
  R.info = info;
  return R;
}

// a host tensor owned here, as a qsx_tensor view
struct HostQ {
  std::vector<int32_t> Q, om;
  std::vector<int64_t> dims;
  std::vector<uint64_t> tags;
  std::vector<double> cgw;
  std::vector<std::vector<double>> blocks;
  std::vector<const double*> ptr;
  std::string qtype;
  qsx_tensor t{};
  void bind(int n, int r, int q, bool with_cgw) {
    ptr.clear();
    for (auto& b : blocks) ptr.push_back(b.data());
    t.nblk = n;
    t.rank = r;
    t.qdim = q;
    t.Q = Q.empty() ? nullptr : Q.data();
    t.dims = dims.empty() ? nullptr : dims.data();
    t.data = ptr.empty() ? nullptr : ptr.data();
    t.data_im = nullptr;
    t.itags = tags.empty() ? nullptr : tags.data();
    t.qtype = qtype.c_str();
    t.cgw = with_cgw && !cgw.empty() ? cgw.data() : nullptr;
    t.om = om.empty() ? nullptr : om.data();
  }
};

// a library result as an input view (borrowed data)
struct OutView {
  std::vector<const double*> ptr;
  qsx_tensor t{};
  explicit OutView(const qsx_tensor_out* o) {
    for (int i = 0; i < o->nblk; ++i) ptr.push_back(o->data ? o->data + o->offs[i] : nullptr);
    t.nblk = o->nblk;
    t.rank = o->rank;
    t.qdim = o->qdim;
    t.Q = o->Q;
    t.dims = o->dims;
    t.data = ptr.empty() ? nullptr : ptr.data();
    t.data_im = nullptr;
    t.itags = o->itags;
    t.qtype = o->qtype;
    t.cgw = o->cgw;
    t.om = o->om;
  }
};

struct OutFree {
  void operator()(qsx_tensor_out* o) const { qsx_tensor_free(o); }
};
using OutPtr = std::unique_ptr<qsx_tensor_out, OutFree>;

// K = 2 of a rank-4 SU(2) PSI (the two-site wave function after the
// eigensolve, mpsOrthoQS at Source/mpsortho.cc:1031-1078 from
// eigs2SiteQS.cc:481-484): legs (0, 1) are fused into one leg X by an
// isometry F(0, 1, X*) (data: identity maps of each label pair into the
// multiplet space of each coupled label, CGT weight normExt, so that F is an
// isometry in the full M-space), PSI~(X, 2, 3) = conj(F).PSI (a rank-3
// tensor: outer multiplicity resolved by the X-symbols), the K = 1 split of
// PSI~ gives U(X, X'*) and A2(2, 3, X'), and A1 = F.U.  F being a complete
// isometry, the singular values, their multiplets and the truncation are
// those of PSI split (0, 1 | 2, 3) per total spin of legs (0, 1).
OrthoResult ortho_fused(const qsx_tensor* pt, const TensorDesc& P, int Nkmin, int64_t Nkeep, double stol,
                        const char* itag) {
  (void)pt;
  const int q = P.qdim;
  // label records of legs 0 and 1 with their dims, and the fused labels of
  // each pair that a block of PSI couples to (SU(2): J in (j0 x j1) and in
  // (j2 x j3); abelian: the sum of legs 0 and 1)
  using Rec = std::vector<int32_t>;
  std::map<std::pair<Rec, Rec>, std::pair<std::pair<int64_t, int64_t>, std::set<Rec>>> pairs;
  for (int i = 0; i < P.n; ++i) {
    Rec r0(q), r1(q), x(q);
    for (int c = 0; c < q; ++c) {
      r0[c] = P.label(i, 0, c);
      r1[c] = P.label(i, 1, c);
    }
    auto& e = pairs[{r0, r1}];
    e.first = {P.dims[(size_t)i * 4 + 0], P.dims[(size_t)i * 4 + 1]};
    std::vector<Rec> xs(1, Rec(q, 0));
    for (int c = 0; c < q; ++c) {
      std::vector<Rec> nx;
      for (const Rec& base : xs) {
        if (P.symtype[c] == SYM_A) {
          Rec b = base;
          b[c] = r0[c] + r1[c];
          nx.push_back(b);
          continue;
        }
        const int a = r0[c], b1 = r1[c], c2 = P.label(i, 2, c), c3 = P.label(i, 3, c);
        for (int J = std::abs(a - b1); J <= a + b1; J += 2) {
          if (J < std::abs(c2 - c3) || J > c2 + c3 || ((J + c2 + c3) & 1)) continue;
          Rec b = base;
          b[c] = J;
          nx.push_back(b);
        }
      }
      xs.swap(nx);
    }
    for (auto& x2 : xs) e.second.insert(x2);
  }
  // multiplet space of each fused label: the pairs in sorted order
  std::map<Rec, int64_t> dimX;
  HostQ F;
  struct FB {
    Rec r0, r1, x;
    int64_t d0, d1, off;
  };
  std::vector<FB> fb;
  for (auto& [k, v] : pairs)
    for (const Rec& x : v.second) {
      int64_t& D = dimX[x];
      fb.push_back(FB{k.first, k.second, x, v.first.first, v.first.second, D});
      D += v.first.first * v.first.second;
    }
  for (const FB& b : fb) {
    F.Q.insert(F.Q.end(), b.r0.begin(), b.r0.end());
    F.Q.insert(F.Q.end(), b.r1.begin(), b.r1.end());
    F.Q.insert(F.Q.end(), b.x.begin(), b.x.end());
    const int64_t DX = dimX[b.x];
    F.dims.insert(F.dims.end(), {b.d0, b.d1, DX});
    std::vector<double> d((size_t)(b.d0 * b.d1 * DX), 0.0);
    for (int64_t i1 = 0; i1 < b.d1; ++i1)
      for (int64_t i0 = 0; i0 < b.d0; ++i0) {
        const int64_t row = i0 + b.d0 * i1;
        d[(size_t)(row + b.d0 * b.d1 * (b.off + row))] = 1.0;
      }
    F.blocks.push_back(std::move(d));
  }
  const uint64_t tagF = pack_name("qsxF") | 0x80;
  F.tags = {P.itags[0], P.itags[1], tagF};
  F.qtype = P.qtype;
  F.bind((int)fb.size(), 3, q, false);
  // PSI~ = conj(F) . PSI over legs (0, 1): legs (X, 2, 3)
  qsx_opts o{};
  o.mode = QSX_MODE_BATCHED;
  o.verbose = 0;
  o.skip_zero = 1;
  const int32_t i01[2] = {0, 1}, i2[1] = {2}, i0[1] = {0};
  OutPtr Pt(contract_oneshot(&F.t, i01, 2, 1, pt, i01, 2, 0, nullptr, 0, &o));
  OutView Pv(Pt.get());
  // the new leg's itag: the K = 2 rule (named after itag, outgoing on A1
  // when leg 0 is incoming)
  const uint64_t tx = pack_name(itag);
  const uint64_t tu = tag_out(P.itags[0]) ? tx : (tx | 0x80);
  OrthoResult R1 = ortho_oneshot(&Pv.t, 1, Nkmin, Nkeep, stol, "qsxX");
  OutPtr U(R1.A1);
  OutView Uv(U.get());
  OrthoResult R;
  R.A1 = contract_oneshot(&F.t, i2, 1, 0, &Uv.t, i0, 1, 0, nullptr, 0, &o);
  R.A2 = R1.A2;
  R.info = R1.info;
  R.A1->itags[2] = tu;
  R.A2->itags[R.A2->rank - 1] = tu ^ 0x80;
  return R;
}

}  // namespace
}  // namespace qsx

using namespace qsx;

extern "C" {

int qsx_ortho(const qsx_tensor* psi, int32_t K, int32_t nkmin, int64_t nkeep, double stol, const char* itag,
              qsx_tensor_out** A1, qsx_tensor_out** A2, qsx_ortho_info** info, char* err, size_t errlen) {
  return guarded(err, errlen, [&] {
    if (!A1 || !A2 || !info) fail(QSX_ERR_ARG, "null output");
    OrthoResult R = ortho_oneshot(psi, K, nkmin, nkeep, stol, itag);
    *A1 = R.A1;
    *A2 = R.A2;
    *info = R.info;
    return 0;
  });
}

void qsx_ortho_info_free(qsx_ortho_info* info) {
  if (!info) return;
  free(info->svd);
  free(info->svd_qdim);
  free(info);
}

}  // extern "C"
