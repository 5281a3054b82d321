// qsops.cc -- host side of the whole-QSpace block operations: the
// permutation plan (dimension fusion, work items) and the C-ABI entry
// points qsx_permute (permuteQS) and qsx_plus (plusQS).
//
// Reference semantics restated:
//  * permuteQS (Source/permuteQS.cc:25-48): a permutation shorter than the
//    rank is completed with the identity (wbperm::Extend), a longer one is
//    an error; an identity permutation without conj returns the input as it
//    is; otherwise every block is permuted (QSpace::permute,
//    Source/QSpace.hh:2721-2780), the itags follow the legs (conj flips
//    their conj bit), and NormCGW (Source/QSpace.cc:3727-3805) brings each
//    SU(2) weight to the external norm, the data taking the removed factor.
//    A block's CGT is the canonical CGT of its leg set in its leg order,
//    with label ties broken by leg position (the tensor format has no CGT
//    permutation, the reference's cgp), so the CGT implied by the new leg
//    order is the transposed one up to a sign (su2_perm_sign): the data
//    take that sign and the weights stay normalised.  conj flips the leg
//    directions exactly as the conj flag of a contraction does.
//  * plusQS (Source/plusQS.cc:17-55): C = A + bfac B.  Abelian tensors take
//    QSpace::plus_plain (Source/QSpace.cc:3593-3705): the union of the label
//    records, stably sorted and grouped (getQOverlapU, QSpace.cc:1266-1320),
//    a matched pair summed A + bfac B.  Tensors with SU(2) labels take Cat +
//    MakeUnique (QSpace.cc:3075-3140,3807-3905): without any common record
//    the result is A's blocks then B's blocks in their own order; with one,
//    the sorted unique records, each the sum of its group (zero blocks of a
//    group are not added).  Weights enter as NormCGW does: each operand's
//    block is scaled by cgw / normExt.  Then SkipZeroData (exact zeros
//    dropped unless the result is a rank-2 block-diagonal matrix,
//    QSpace.cc:3200-3270) and NormCGW.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "qsops.hpp"

namespace qsx {

PermPlan::~PermPlan() {
  dev_free(ddescs);
  dev_free(ditems);
}

namespace {

constexpr int kRunChunk = 4096;  // destination elements per run item
constexpr int kTile = 32;        // tile edge of the transpose kernel (blockops.hip)
constexpr int kTileMin = 8;      // both tile extents at least this long, else a gather

// fuse one block's dims: unit dims dropped, source-adjacent dims that stay
// adjacent in the destination merged.  Out: fused dims in destination order,
// the source position of each, and its source stride.
void fuse_block(int r, const int64_t* dims, const int* perm, std::vector<int64_t>& nd, std::vector<int>& spos,
                std::vector<int64_t>& ss) {
  nd.clear();
  spos.clear();
  ss.clear();
  std::vector<int64_t> sstride(r);
  int64_t st = 1;
  for (int k = 0; k < r; ++k) {
    sstride[k] = st;
    st *= dims[k];
  }
  int prev = -2;
  for (int k = 0; k < r; ++k) {
    const int s = perm[k];
    if (dims[s] == 1) continue;
    // extends the current group if it is the next non-unit source dim
    bool ext = false;
    if (prev >= 0) {
      int nxt = prev + 1;
      while (nxt < r && dims[nxt] == 1) ++nxt;
      ext = nxt == s;
    }
    if (ext) {
      nd.back() *= dims[s];
    } else {
      nd.push_back(dims[s]);
      spos.push_back(s);
      ss.push_back(sstride[s]);
    }
    prev = s;
  }
  if (nd.empty()) {  // a single element
    nd.push_back(1);
    spos.push_back(0);
    ss.push_back(1);
  }
}

}  // namespace

void perm_plan_build(PermPlan& pp, int n, int r, const int64_t* dims, const int64_t* src_offs,
                     const int64_t* dst_offs, const int* perm, const double* scale) {
  pp.descs.clear();
  pp.items.clear();
  pp.elems = 0;
  if (r > QSX_MAX_RANK) fail(QSX_ERR_UNSUPPORTED, "rank %d > %d", r, QSX_MAX_RANK);
  std::vector<int64_t> nd, ss;
  std::vector<int> spos;
  for (int b = 0; b < n; ++b) {
    const int64_t* d = dims + (size_t)b * r;
    int64_t N = 1;
    for (int k = 0; k < r; ++k) N *= d[k];
    if (N == 0) continue;
    if (N > INT32_MAX) fail(QSX_ERR_UNSUPPORTED, "block of %lld elements (permute limit 2^31-1)", (long long)N);
    fuse_block(r, d, perm, nd, spos, ss);
    PermDesc pd{};
    pd.src_off = src_offs[b];
    pd.dst_off = dst_offs[b];
    pd.scale = scale ? scale[b] : 1.0;
    pd.rank = (int)nd.size();
    int64_t dst = 1;
    int di = -1;  // destination position of source dim 0
    for (int k = 0; k < pd.rank; ++k) {
      pd.nd[k] = (int32_t)nd[k];
      pd.ss[k] = ss[k];
      pd.ds[k] = dst;
      dst *= nd[k];
      if (ss[k] == 1) di = k;
    }
    const int bi = (int)pp.descs.size();
    // tile columns: the destination dims in front of source dim 0, flattened
    const bool tiled = di > 0 && pd.ds[di] >= kTileMin && nd[di] >= kTileMin;
    if (!tiled) {
      pd.kind = 0;
      for (int64_t e0 = 0; e0 < N; e0 += kRunChunk)
        pp.items.push_back(PermItem{bi, (int32_t)e0, (int32_t)std::min<int64_t>(kRunChunk, N - e0), 0});
    } else {
      pd.kind = 1 + di;
      const int64_t ni = nd[di], nj = pd.ds[di], nb = N / (ni * nj);
      const int ti = (int)((ni + kTile - 1) / kTile), tj = (int)((nj + kTile - 1) / kTile);
      for (int64_t c = 0; c < nb; ++c)
        for (int a = 0; a < ti; ++a)
          for (int bb = 0; bb < tj; ++bb) pp.items.push_back(PermItem{bi, a, bb, (int32_t)c});
    }
    pp.descs.push_back(pd);
    pp.elems += N;
  }
}

void perm_plan_upload(PermPlan& pp, hipStream_t s) {
  dev_free(pp.ddescs);
  dev_free(pp.ditems);
  pp.ddescs = nullptr;
  pp.ditems = nullptr;
  if (pp.items.empty()) return;
  pp.ddescs = (PermDesc*)dev_alloc(sizeof(PermDesc) * pp.descs.size());
  pp.ditems = (PermItem*)dev_alloc(sizeof(PermItem) * pp.items.size());
  upload_many({{pp.ddescs, pp.descs.data(), sizeof(PermDesc) * pp.descs.size()},
               {pp.ditems, pp.items.data(), sizeof(PermItem) * pp.items.size()}},
              s);
}

namespace {

// product over SU(2) symmetries of cgw / normExt (the NormCGW data factor)
double norm_factor(const TensorDesc& T, int blk, const std::vector<int>& dir) {
  double f = 1.0;
  for (int j = 0; j < T.nsym; ++j) {
    if (T.symtype[j] != SYM_SU2) continue;
    std::vector<int> tj(T.r);
    for (int l = 0; l < T.r; ++l) tj[l] = T.label(blk, l, j);
    const double ne = su2_norm_ext(tj, dir);
    const double w = T.cgw_given ? T.cgw[(size_t)blk * T.nsym + j] : ne;
    f *= w / ne;
  }
  return f;
}

// the input returned unchanged (identity permutation, or an empty operand)
qsx_tensor_out* copy_out(const TensorDesc& T, const std::vector<double>& data, double fac,
                         const std::vector<double>* data_im = nullptr) {
  qsx_tensor_out* o = alloc_out(T.n, T.r, T.qdim, T.nsym, T.qtype, T.total());
  if (data_im) {
    o->data_im = (double*)calloc(std::max<int64_t>(1, T.total()), sizeof(double));
    for (int64_t e = 0; e < T.total(); ++e) o->data_im[e] = fac * (*data_im)[e];
  }
  std::copy(T.Q.begin(), T.Q.end(), o->Q);
  std::copy(T.dims.begin(), T.dims.end(), o->dims);
  std::copy(T.offs.begin(), T.offs.end(), o->offs);
  std::copy(T.cgw.begin(), T.cgw.end(), o->cgw);
  if (!T.cgw_given) {  // default weights are the external norms
    const std::vector<int> dir = leg_dirs(T.itags, T.r);
    for (int i = 0; i < T.n; ++i)
      for (int j = 0; j < T.nsym; ++j) {
        if (T.symtype[j] != SYM_SU2) continue;
        std::vector<int> tj(T.r);
        for (int l = 0; l < T.r; ++l) tj[l] = T.label(i, l, j);
        o->cgw[(size_t)i * T.nsym + j] = su2_norm_ext(tj, dir);
      }
  }
  for (size_t l = 0; l < T.itags.size(); ++l) o->itags[l] = T.itags[l];
  for (int64_t e = 0; e < T.total(); ++e) o->data[e] = fac == 1.0 ? data[e] : fac * data[e];
  return o;
}

qsx_tensor_out* permute_oneshot(const qsx_tensor* At, const int32_t* perm, int nperm, int conj) {
  ensure_device();
  hipStream_t s = default_stream();
  TensorDesc A = make_desc(At, "A");
  require_no_om(A, "permuteQS");
  const int r = A.r, q = A.qdim;
  if (nperm < 0 || nperm > r) fail(QSX_ERR_ARG, "invalid permutation (len=%d/%d)", nperm, r);
  if (nperm && !perm) fail(QSX_ERR_ARG, "null permutation");
  std::vector<int> p(r);
  {
    std::vector<char> used(r, 0);
    for (int i = 0; i < nperm; ++i) {
      if (perm[i] < 0 || perm[i] >= r || used[perm[i]]) fail(QSX_ERR_ARG, "invalid permutation");
      used[perm[i]] = 1;
      p[i] = perm[i];
    }
    int l = nperm;
    for (int i = 0; i < r; ++i)
      if (!used[i]) p[l++] = i;
  }
  bool ident = true;
  for (int i = 0; i < r; ++i) ident &= p[i] == i;
  const bool cplx = A.hdata_im != nullptr;
  if (!r || (ident && !conj)) {
    std::vector<double> host = packed_host(A), him;
    if (cplx) {
      TensorDesc Ai = A;
      Ai.hdata = A.hdata_im;
      him = packed_host(Ai);
    }
    return copy_out(A, host, 1.0, cplx ? &him : nullptr);
  }
  // output structure
  std::vector<uint64_t> tags(A.itags.size());
  for (size_t l = 0; l < tags.size(); ++l) tags[l] = A.itags[p[l]] ^ (conj ? 0x80u : 0u);
  const std::vector<int> dirIn = leg_dirs(A.itags, r), dirOut = leg_dirs(tags, r);
  qsx_tensor_out* o = alloc_out(A.n, r, q, A.nsym, A.qtype, A.total());
  std::vector<double> scale(A.n, 1.0);
  for (int i = 0; i < A.n; ++i) {
    for (int l = 0; l < r; ++l) {
      for (int j = 0; j < q; ++j) o->Q[(size_t)i * r * q + l * q + j] = A.label(i, p[l], j);
      o->dims[(size_t)i * r + l] = A.dims[(size_t)i * r + p[l]];
    }
    o->offs[i + 1] = A.offs[i + 1];
    scale[i] = norm_factor(A, i, dirIn);
    for (int j = 0; j < A.nsym; ++j) {
      double w = 1.0;
      if (A.symtype[j] == SYM_SU2) {
        std::vector<int> tj(r), tj0(r);
        for (int l = 0; l < r; ++l) tj[l] = o->Q[(size_t)i * r * q + l * q + j];
        for (int l = 0; l < r; ++l) tj0[l] = A.label(i, l, j);
        // the block's CGT is the canonical CGT of the new leg order (and
        // directions): the transposed CGT times su2_perm_sign, which the
        // data take
        std::vector<int> dflip(dirIn);
        for (auto& x : dflip) x = conj ? -x : x;
        w = su2_norm_ext(tj, dirOut);
        scale[i] *= su2_perm_sign(tj0, dflip, p);
      }
      o->cgw[(size_t)i * A.nsym + j] = w;
    }
  }
  for (size_t l = 0; l < tags.size(); ++l) o->itags[l] = tags[l];
  try {
    PermPlan pp;
    perm_plan_build(pp, A.n, r, A.dims.data(), A.offs.data(), A.offs.data(), p.data(), scale.data());
    perm_plan_upload(pp, s);
    DevMem dA(A.total()), dB(A.total());
    upload_packed(A, A.hdata, dA.p, s);
    Timer t(s);
    launch_perm_plan(pp, dA.p, dB.p, s);
    o->gpu_ms = t.stop(s);
    download_packed(dB.p, A.total(), o->data, s);
    if (cplx) {
      // the imaginary plane through the same permutation; conj conjugates
      // complex data (wbarray::permute with P.conj, Source/wbarray.cc:2738)
      PermPlan pi;
      std::vector<double> si(scale);
      if (conj)
        for (double& x : si) x = -x;
      perm_plan_build(pi, A.n, r, A.dims.data(), A.offs.data(), A.offs.data(), p.data(), si.data());
      perm_plan_upload(pi, s);
      upload_packed(A, A.hdata_im, dA.p, s);
      launch_perm_plan(pi, dA.p, dB.p, s);
      o->data_im = (double*)calloc(std::max<int64_t>(1, A.total()), sizeof(double));
      download_packed(dB.p, A.total(), o->data_im, s);
    }
  } catch (...) {
    qsx_tensor_free(o);
    throw;
  }
  return o;
}

// lenient itag equality (iTags ==, Source/wbindex.hh:983-989 with
// itag_::sameAs(.,'l'), Source/wbindex.cc:989-1007): conj bits equal, names
// equal or one of them empty
bool tags_equal(const std::vector<uint64_t>& a, const std::vector<uint64_t>& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i) {
    if (tag_out(a[i]) != tag_out(b[i])) return false;
    const uint64_t na = tag_name(a[i]), nb = tag_name(b[i]);
    if (na != nb && na && nb) return false;
  }
  return true;
}

qsx_tensor_out* plus_oneshot(const qsx_tensor* At, const qsx_tensor* Bt, double bfac, int skip_zero) {
  ensure_device();
  hipStream_t s = default_stream();
  TensorDesc A = make_desc(At, "A"), B = make_desc(Bt, "B");
  require_no_om(A, "plusQS");
  require_no_om(B, "plusQS");
  if (!A.n && !B.n) return copy_out(A, packed_host(A), 1.0);
  // an empty operand contributes no block (Cat / plus_plain return the
  // other one, scaled by bfac if it is B); the rest of the path is common
  const bool both = A.n && B.n;
  const TensorDesc& ref = A.n ? A : B;
  if (both && (A.qdim != B.qdim || A.r != B.r))
    fail(QSX_ERR_QDIM, "QSpace mismatch (QDIM=%d/%d; r=%d/%d)", A.qdim, B.qdim, A.r, B.r);
  if (both && A.symtype != B.symtype && !A.qtype.empty() && !B.qtype.empty())
    fail(QSX_ERR_QDIM, "got inconsistent qtype %s/%s", A.qtype.c_str(), B.qtype.c_str());
  const bool cg = ref.has_cg();
  const int r = ref.r, q = ref.qdim, rq = r * q;
  // output itags
  std::vector<uint64_t> tags = A.itags.empty() || !A.n ? B.itags : A.itags;
  if (both && !A.itags.empty() && !B.itags.empty()) {
    if (cg) {
      if (!tags_equal(A.itags, B.itags)) fail(QSX_ERR_ITAG, "itag inconsistency");
    } else {
      for (int l = 0; l < r; ++l)
        if (tag_out(A.itags[l]) != tag_out(B.itags[l])) fail(QSX_ERR_ITAG, "got inconsistent itags (conj)");
      if (!tags_equal(A.itags, B.itags))
        for (auto& t : tags) t &= 0x80;  // init_tags: names cleared, conj kept
    }
  }
  if (cg && ((A.n && A.itags.empty()) || (B.n && B.itags.empty())))
    fail(QSX_ERR_ITAG, "missing itags for non-abelian QSpace");
  // union of the records: stable sort of [A; B], grouped
  std::vector<int32_t> R(A.Q);
  R.insert(R.end(), B.Q.begin(), B.Q.end());
  std::vector<int32_t> order;
  std::vector<int64_t> seg;
  device_group_records(R, A.n + B.n, rq, order, seg, s);
  const int ng = (int)seg.size() - 1;
  std::vector<int> ia(ng, -1), ib(ng, -1);
  bool overlap = false;
  for (int g = 0; g < ng; ++g) {
    const int64_t d = seg[g + 1] - seg[g];
    if (d > 2) fail(QSX_ERR_STRUCT, "got non-unique QIDX (%lld)", (long long)d);
    for (int64_t k = seg[g]; k < seg[g + 1]; ++k) {
      const int i = order[k];
      if (i < A.n) {
        if (ia[g] >= 0) fail(QSX_ERR_STRUCT, "non-unique QIDX in A");
        ia[g] = i;
      } else {
        if (ib[g] >= 0) fail(QSX_ERR_STRUCT, "non-unique QIDX in B");
        ib[g] = i - A.n;
      }
    }
    if (ia[g] >= 0 && ib[g] >= 0) {
      overlap = true;
      for (int l = 0; l < r; ++l)
        if (A.dims[(size_t)ia[g] * r + l] != B.dims[(size_t)ib[g] * r + l])
          fail(QSX_ERR_STRUCT, "block size mismatch (A block %d, B block %d)", ia[g] + 1, ib[g] + 1);
    }
  }
  // output order: sorted groups; with SU(2) labels and no common record,
  // A's blocks then B's blocks (MakeUnique returns before reordering)
  std::vector<int> oa, ob;
  if (cg && !overlap) {
    for (int i = 0; i < A.n; ++i) oa.push_back(i), ob.push_back(-1);
    for (int i = 0; i < B.n; ++i) oa.push_back(-1), ob.push_back(i);
  } else {
    oa = ia;
    ob = ib;
  }
  const int nC = (int)oa.size();
  const std::vector<int> dA = leg_dirs(A.itags, r), dB = leg_dirs(B.itags, r), dC = leg_dirs(tags, r);
  std::vector<int64_t> offs(nC + 1, 0);
  for (int c = 0; c < nC; ++c) offs[c + 1] = offs[c] + (oa[c] >= 0 ? A.numel(oa[c]) : B.numel(ob[c]));
  // zero blocks of a group are not added (MakeUnique's mark): flags per operand block
  DevMem dAm(A.total()), dBm(B.total()), dCm(offs[nC]);
  upload_packed(A, A.hdata, dAm.p, s);
  upload_packed(B, B.hdata, dBm.p, s);
  // complex data: the imaginary planes take the same items (a real operand
  // has a zero plane); a block is zero when both planes are
  const bool cplx = A.hdata_im || B.hdata_im;
  DevMem dAi(cplx ? A.total() : 0), dBi(cplx ? B.total() : 0), dCi(cplx ? offs[nC] : 0);
  if (cplx) {
    if (A.hdata_im) upload_packed(A, A.hdata_im, dAi.p, s);
    else launch_fill_zero(dAi.p, A.total(), s);
    if (B.hdata_im) upload_packed(B, B.hdata_im, dBi.p, s);
    else launch_fill_zero(dBi.p, B.total(), s);
  }
  // per-block "all exact zeros" over both planes
  auto zero_flags = [&](const std::vector<int64_t>& of, int nb, const double* re, const double* im,
                        std::vector<int32_t>& z) {
    int32_t* dz = (int32_t*)dev_alloc(sizeof(int32_t) * std::max(1, 2 * nb));
    int64_t* doff = (int64_t*)dev_alloc(sizeof(int64_t) * (nb + 1));
    try {
      QSX_HIP(hipMemcpyAsync(doff, of.data(), sizeof(int64_t) * (nb + 1), hipMemcpyHostToDevice, s));
      launch_block_zero_flags(doff, nb, re, dz, s);
      if (im) launch_block_zero_flags(doff, nb, im, dz + nb, s);
      std::vector<int32_t> zz(2 * nb, 1);
      QSX_HIP(hipMemcpyAsync(zz.data(), dz, sizeof(int32_t) * (im ? 2 : 1) * nb, hipMemcpyDeviceToHost, s));
      QSX_HIP(hipStreamSynchronize(s));
      for (int i = 0; i < nb; ++i) z[i] = zz[i] && zz[nb + i];
    } catch (...) {
      dev_free(dz);
      dev_free(doff);
      throw;
    }
    dev_free(dz);
    dev_free(doff);
  };
  std::vector<int32_t> zA(A.n, 0), zB(B.n, 0);
  if (cg && overlap) {
    zero_flags(A.offs, A.n, dAm.p, cplx ? dAi.p : nullptr, zA);
    zero_flags(B.offs, B.n, dBm.p, cplx ? dBi.p : nullptr, zB);
  }
  std::vector<AxpbyItem> items;
  for (int c = 0; c < nC; ++c) {
    int a = oa[c], b = ob[c];
    if (cg && a >= 0 && b >= 0) {  // a group of a zero and a non-zero block keeps the non-zero one
      if (zA[a] && !zB[b]) a = -1;
      else if (zB[b] && !zA[a]) b = -1;
      else if (zA[a] && zB[b]) b = -1;
    }
    const double fa = a >= 0 ? (cg ? norm_factor(A, a, dA) : 1.0) : 0.0;
    const double fb = b >= 0 ? (cg ? bfac * norm_factor(B, b, dB) : bfac) : 0.0;
    const int64_t n = offs[c + 1] - offs[c];
    for (int64_t e0 = 0; e0 < n; e0 += 4096) {
      AxpbyItem it{};
      it.c_off = offs[c] + e0;
      it.a_off = a >= 0 ? A.offs[a] + e0 : -1;
      it.b_off = b >= 0 ? B.offs[b] + e0 : -1;
      it.fa = fa;
      it.fb = fb;
      it.n = (int32_t)std::min<int64_t>(4096, n - e0);
      items.push_back(it);
    }
  }
  AxpbyItem* ditems = (AxpbyItem*)dev_alloc(sizeof(AxpbyItem) * std::max<size_t>(1, items.size()));
  float ms = 0;
  std::vector<int32_t> zC(nC, 0);
  try {
    if (!items.empty())
      QSX_HIP(hipMemcpyAsync(ditems, items.data(), sizeof(AxpbyItem) * items.size(), hipMemcpyHostToDevice, s));
    Timer t(s);
    launch_axpby(ditems, (int)items.size(), dAm.p, dBm.p, dCm.p, s);
    if (cplx) launch_axpby(ditems, (int)items.size(), dAi.p, dBi.p, dCi.p, s);
    ms = t.stop(s);
    // SkipZeroData: exact-zero blocks dropped unless a rank-2 block-diagonal matrix
    bool blockdiag = false;
    if (r == 2) {
      blockdiag = true;
      for (int c = 0; c < nC && blockdiag; ++c) {
        const int32_t* rec = oa[c] >= 0 ? &A.Q[(size_t)oa[c] * rq] : &B.Q[(size_t)ob[c] * rq];
        for (int j = 0; j < q; ++j)
          if (rec[j] != rec[q + j]) blockdiag = false;
      }
    }
    if (skip_zero && nC && !(r == 2 && blockdiag)) zero_flags(offs, nC, dCm.p, cplx ? dCi.p : nullptr, zC);
    QSX_HIP(hipStreamSynchronize(s));
  } catch (...) {
    dev_free(ditems);
    throw;
  }
  dev_free(ditems);
  std::vector<int> keep;
  for (int c = 0; c < nC; ++c)
    if (!zC[c]) keep.push_back(c);
  int64_t tot = 0;
  for (int c : keep) tot += offs[c + 1] - offs[c];
  qsx_tensor_out* o = alloc_out((int)keep.size(), r, q, ref.nsym, A.qtype.empty() ? B.qtype : A.qtype, tot);
  for (size_t i = 0; i < keep.size(); ++i) {
    const int c = keep[i];
    const TensorDesc& S = oa[c] >= 0 ? A : B;
    const int sb = oa[c] >= 0 ? oa[c] : ob[c];
    std::copy(&S.Q[(size_t)sb * rq], &S.Q[(size_t)sb * rq] + rq, o->Q + i * rq);
    std::copy(&S.dims[(size_t)sb * r], &S.dims[(size_t)sb * r] + r, o->dims + i * r);
    o->offs[i + 1] = o->offs[i] + (offs[c + 1] - offs[c]);
    for (int j = 0; j < ref.nsym; ++j) {
      double w = 1.0;
      if (ref.symtype[j] == SYM_SU2) {
        std::vector<int> tj(r);
        for (int l = 0; l < r; ++l) tj[l] = S.label(sb, l, j);
        w = su2_norm_ext(tj, dC);
      }
      o->cgw[i * ref.nsym + j] = w;
    }
  }
  for (size_t l = 0; l < tags.size(); ++l) o->itags[l] = tags[l];
  // the kept blocks straight from the pinned staging copy of the result
  try {
    auto fetch = [&](const double* src, double* dst) {
      download_staged(src, offs[nC], s, [&](const double* h) {
        if (!h) return;
#pragma omp parallel for schedule(dynamic, 16) if (offs[nC] > (1 << 18))
        for (size_t i = 0; i < keep.size(); ++i) {
          const int c = keep[i];
          memcpy(dst + o->offs[i], h + offs[c], sizeof(double) * (offs[c + 1] - offs[c]));
        }
      });
    };
    fetch(dCm.p, o->data);
    if (cplx) {
      o->data_im = (double*)calloc(std::max<int64_t>(1, tot), sizeof(double));
      fetch(dCi.p, o->data_im);
    }
  } catch (...) {
    qsx_tensor_free(o);
    throw;
  }
  o->gpu_ms = ms;
  return o;
}

}  // namespace

}  // namespace qsx

using namespace qsx;

extern "C" {

int qsx_permute(const qsx_tensor* A, const int32_t* perm, int32_t nperm, int32_t conj, qsx_tensor_out** B,
                char* err, size_t errlen) {
  return guarded(err, errlen, [&] {
    if (!B) fail(QSX_ERR_ARG, "null output");
    *B = permute_oneshot(A, perm, nperm, conj);
    return 0;
  });
}

int qsx_plus(const qsx_tensor* A, const qsx_tensor* B, double bfac, const qsx_opts* opts, qsx_tensor_out** C,
             char* err, size_t errlen) {
  return guarded(err, errlen, [&] {
    if (!C) fail(QSX_ERR_ARG, "null output");
    *C = plus_oneshot(A, B, bfac, opts ? opts->skip_zero : 1);
    return 0;
  });
}

}  // extern "C"
