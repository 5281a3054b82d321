// hpsi.cc -- two-site H*psi as two batched-GEMM launches (or the
// split-precision BF16 path), replayed from a hipGraph, and
// the FP64 Lanczos / Davidson loops around it.
//
// Reference: get_HPsi.m:54-57 (Class/@Hamilton1D/private),
//   HPsi = contractQS({HK_L,Psi},HK_R) [+ contractQS({HK_L,'*',Psi},HK_R,'*')]
// with itag auto-matching (contractQS_itags, Source/contractQS.cc:634-808);
// the legacy C++ CPAT::HZTimes (Source/cpat.hh:487-527) reuses its pair
// lists across iterations (check_contraction, cpat.hh:572-593) -- here the
// whole plan (pairs, X-symbol factors, tiles, operand layouts) is built once
// and the per-iteration work is two kernel launches:
//   stage 1: T_t = HL_t . psi      for both terms t (one launch)
//   stage 2: Hpsi = sum_t T_t . HR_t  (one launch; the pairs of both terms
//            of an output block are summed in one fixed order)
// Environments are copied to the device once, in the layouts the GEMM
// consumes (permuted copies made once at plan time).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <numeric>

#include "contract.hpp"
#include "smalleig.hpp"
#include "vec.hpp"

namespace qsx {

// one two-stage term: contractQS({L [,'*'], psi}, R [,'*']) with L = Ls[li]
// (conj if cl) and R = Rs[ri] (conj if cr)
struct HpsiTerm {
  int li, cl, ri, cr;
};

struct HpsiPlan {
  std::vector<TensorDesc> Ls, Rs;  // left / right operators of the terms
  TensorDesc psi;
  std::vector<HpsiTerm> terms;
  TensorDesc HKL, HKR;             // local terms HK_L . psi and psi . HK_R (pseudo MPO), if present
  bool has_hkl = false, has_hkr = false;
  int nterm = 1;
  int verbose = 0;
  int mode = QSX_MODE_BATCHED;
  int nslices = 0;
  std::vector<std::unique_ptr<CtrPlan>> s1, s2;  // per term
  std::unique_ptr<CtrPlan> locL, locR;           // the local terms' contractions
  int64_t N = 0;
  double* dHL = nullptr;   // arena: [Ls as uploaded | permuted copies per term]
  double* dHR = nullptr;   // arena: [R copies per term in the stage-2 layouts | HK_R copy]
  double* dT = nullptr;    // arena: [T_t per term | HK_L copy]
  std::vector<int64_t> hl_base, hr_base, t_base;  // per term
  int64_t t_total = 0, hr_total = 0, hl_total = 0;  // T part of dT, whole dHR / dHL arenas
  int64_t hkl_base = 0, hkr_base = 0, dt_size = 0;
  bool xpairs = false;     // stage 2 has pairs that read psi itself (local terms)
  std::vector<PairDesc> pairs1, pairs2;
  std::vector<WorkTarget> targets1, targets2;
  WorkList work1, work2;
  WorkPolicy pol1, pol2;
  WorkItem *ditems1 = nullptr, *ditems2 = nullptr;
  ReduceItem *dreds1 = nullptr, *dreds2 = nullptr;
  WorkQueue *dq1 = nullptr, *dq2 = nullptr;
  double* dscratch = nullptr;
  PairDesc *dpairs1 = nullptr, *dpairs2 = nullptr;
  // the same work with afac scaled to the orthonormal reduced basis
  // x~ = sqrt(w) x, w = full-space norm of each psi block's CGT, in which
  // H is a symmetric matrix (used by the eigensolvers)
  PairDesc *dpairs1s = nullptr, *dpairs2s = nullptr;
  std::vector<double> welem;  // per-element weight w of psi's packed layout
  std::vector<double> wblk;
  std::vector<PairDesc> pairs1s, pairs2s;
  double flops = 0, bytes = 0;
  double stage_flops[2] = {0, 0}, stage_bytes[2] = {0, 0};
  // graph
  bool use_graph = true;
  hipGraphExec_t gexec = nullptr;
  const double* g_psi = nullptr;
  double* g_hpsi = nullptr;
  bool g_sym = false;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  float last_ms[2] = {0, 0};
  // split precision (QSX_MODE_SPLIT_BF16 / _F32): BF16 or FP32 slice planes
  // and line exponents of the four operand arenas, slice-addressed pairs,
  // 16x16 tiles
  struct SplitSide {
    uint16_t* planes = nullptr;
    float* fplanes = nullptr;
    int64_t plane = 0;
    int32_t* exps = nullptr;
    SplitLine* dlines = nullptr;
    int nlines = 0;
    bool lane = false;  // 16 lines per wave (lines adjacent, strided along k)
    int maxk = 0;       // longest line
  };
  int nsl = 6;
  SplitSide spHL, spPsi, spT, spHR;
  SplitPair *dsp1 = nullptr, *dsp2 = nullptr, *dsp1s = nullptr, *dsp2s = nullptr;
  WorkList swork1, swork2;
  WorkItem *dsitems1 = nullptr, *dsitems2 = nullptr;
  ReduceItem *dsreds1 = nullptr, *dsreds2 = nullptr;  // k-chunked tiles of the split work lists
  double* dsscratch = nullptr;
  double split_exec_flops[2] = {0, 0};
  // eigensolver state kept between calls: sqrt(welem) on the device and one
  // workspace (Krylov vectors, partials, scalars), grown on demand
  double* dsw = nullptr;
  double* dws = nullptr;
  size_t ws_doubles = 0;
  double* hstage = nullptr;  // pinned host staging of psi0 and the ground state (N doubles)
  double* hsmall = nullptr;  // pinned host scalars of the eigensolvers (coefficients, T, projected columns)
  double* dsmall = nullptr;  // the same memory as kernels address it (hipHostGetDevicePointer)
  size_t hsmall_n = 0;
  ~HpsiPlan() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    for (SplitSide* sp : {&spHL, &spPsi, &spT, &spHR}) {
      dev_free(sp->planes);
      dev_free(sp->fplanes);
      dev_free(sp->exps);
      dev_free(sp->dlines);
    }
    for (void* p : {(void*)dsp1, (void*)dsp2, (void*)dsp1s, (void*)dsp2s, (void*)dsitems1, (void*)dsitems2,
                    (void*)dsreds1, (void*)dsreds2, (void*)dsscratch})
      dev_free(p);
    dev_free(dHL);
    dev_free(dHR);
    dev_free(dT);
    dev_free(dpairs1);
    dev_free(dpairs2);
    dev_free(dpairs1s);
    dev_free(dpairs2s);
    dev_free(ditems1);
    dev_free(ditems2);
    dev_free(dreds1);
    dev_free(dreds2);
    dev_free(dq1);
    dev_free(dq2);
    dev_free(dscratch);
    dev_free(dsw);
    dev_free(dws);
    if (hstage) (void)hipHostFree(hstage);
    if (hsmall) (void)hipHostFree(hsmall);
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

namespace {

struct HostTensor {  // qsx_tensor view over owned vectors
  std::vector<int32_t> Q;
  std::vector<int64_t> dims;
  std::vector<uint64_t> tags;
  std::string qtype;
  qsx_tensor t{};
  void bind(int n, int r, int q) {
    t.nblk = n;
    t.rank = r;
    t.qdim = q;
    t.Q = Q.empty() ? nullptr : Q.data();
    t.dims = dims.empty() ? nullptr : dims.data();
    t.data = nullptr;
    t.itags = tags.empty() ? nullptr : tags.data();
    t.qtype = qtype.c_str();
    t.cgw = nullptr;  // results carry the external normalisation
  }
};

// multi-term work: output block c of term t's plan goes to target block
// target[t][c]; the pairs of all terms of one target are concatenated (term
// order, then the plan's reference pair order) into one pair range
void build_multi(const std::vector<CtrPlan*>& plans, const std::vector<int64_t>& a_base,
                 const std::vector<int64_t>& b_base, const std::vector<std::vector<int>>& target,
                 const std::vector<int64_t>& t_off, const std::vector<int>& tM, const std::vector<int>& tN,
                 std::vector<PairDesc>& pairs, std::vector<WorkTarget>& targets,
                 std::vector<int>* pair_target = nullptr, std::vector<std::pair<int, int64_t>>* pair_src = nullptr,
                 const std::vector<int32_t>* xflags = nullptr) {
  const int nt = (int)t_off.size();
  std::vector<std::vector<std::pair<int, int>>> src(nt);  // (term, block)
  for (size_t t = 0; t < plans.size(); ++t)
    for (int c = 0; c < plans[t]->nC; ++c) {
      int g = target[t][c];
      if (g < 0) continue;
      if (plans[t]->M[c] != tM[g] || plans[t]->N[c] != tN[g])
        fail(QSX_ERR_STRUCT, "H*psi term %zu block %d: dims %dx%d differ from the target %dx%d", t, c, plans[t]->M[c],
             plans[t]->N[c], tM[g], tN[g]);
      src[g].push_back({(int)t, c});
    }
  pairs.clear();
  targets.clear();
  if (pair_target) pair_target->clear();
  if (pair_src) pair_src->clear();
  for (int g = 0; g < nt; ++g) {
    int64_t p0 = (int64_t)pairs.size();
    for (auto [t, c] : src[g]) {
      const CtrPlan& P = *plans[t];
      for (int64_t p = P.pseg[c]; p < P.pseg[c + 1]; ++p) {
        PairDesc pd = P.pairs[p];
        pd.a_off += a_base[t];
        pd.b_off += b_base[t];
        if (xflags) pd.flags |= (*xflags)[t];
        pairs.push_back(pd);
        if (pair_target) pair_target->push_back(g);
        if (pair_src) pair_src->push_back({(int)t, p});
      }
    }
    targets.push_back(WorkTarget{t_off[g], tM[g], tN[g], p0, (int64_t)pairs.size()});
  }
}

// a device array for v, and its upload as a part of upload_many
template <class T>
T* alloc_for(const std::vector<T>& v, std::vector<UploadPart>& parts) {
  T* d = (T*)dev_alloc(sizeof(T) * std::max<size_t>(1, v.size()));
  parts.push_back({d, v.data(), sizeof(T) * v.size()});
  return d;
}

template <class T>
T* upload_vec(const std::vector<T>& v, hipStream_t s) {
  T* d = (T*)dev_alloc(sizeof(T) * std::max<size_t>(1, v.size()));
  if (!v.empty()) QSX_HIP(hipMemcpyAsync(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
  return d;
}

void upload_blocks(const TensorDesc& T, double* d, hipStream_t s) {
  if (!T.hdata) fail(QSX_ERR_ARG, "environment tensor without data");
  upload_packed(T, T.hdata, d, s);
}

std::string rec_key(const int32_t* r, int m) { return std::string((const char*)r, sizeof(int32_t) * m); }

}  // namespace

// (re)build and upload the two stages' work lists for the current policies
static void hpsi_build_work(HpsiPlan& H) {
  hipStream_t s = default_stream();
  H.work1 = build_work(H.targets1, H.pairs1, H.pol1);
  H.work2 = build_work(H.targets2, H.pairs2, H.pol2);
  for (void* p : {(void*)H.ditems1, (void*)H.ditems2, (void*)H.dreds1, (void*)H.dreds2, (void*)H.dq1, (void*)H.dq2,
                  (void*)H.dscratch})
    dev_free(p);
  for (auto* p : {(void**)&H.ditems1, (void**)&H.ditems2, (void**)&H.dreds1, (void**)&H.dreds2, (void**)&H.dq1,
                  (void**)&H.dq2, (void**)&H.dscratch})
    *p = nullptr;
  std::vector<UploadPart> parts;
  H.ditems1 = alloc_for(H.work1.items, parts);
  H.ditems2 = alloc_for(H.work2.items, parts);
  H.dreds1 = alloc_for(H.work1.reds, parts);
  H.dreds2 = alloc_for(H.work2.reds, parts);
  H.dq1 = alloc_for(H.work1.queues, parts);
  H.dq2 = alloc_for(H.work2.queues, parts);
  H.dscratch = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, std::max(H.work1.scratch_elems, H.work2.scratch_elems)));
  upload_many(parts, s);
  if (H.gexec) {
    (void)hipGraphExecDestroy(H.gexec);
    H.gexec = nullptr;
  }
}

// ---------------------------------------------------------------- autotuner
// Tile shapes and k-splits come from timing candidate work decompositions of
// each stage on the plan itself (HIP events, median of reps).  The result is
// cached per process under a key built from the stage's block-size
// histogram (log2 buckets of M, N and K, the STAT_DGEMM idea of
// Source/wbblas.hh:235-268), so plans with the same size profile reuse it.
// autotuned policies per block-size histogram, shared by all plans of the
// process (guarded: plans may be built and tuned from several host threads)
static std::map<std::string, std::pair<WorkPolicy, WorkPolicy>> g_tune_cache;
static std::mutex g_tune_mu;

static std::string hist_key(const std::vector<WorkTarget>& targets, const std::vector<PairDesc>& pairs) {
  std::map<int, int64_t> h;
  auto lg = [](int64_t x) {
    int b = 0;
    while ((1ll << b) < x) ++b;
    return b;
  };
  for (auto& t : targets)
    for (int64_t p = t.p0; p < t.p1; ++p) h[lg(t.M) * 400 + lg(t.N) * 20 + lg(pairs[p].K)]++;
  std::string k;
  for (auto& [b, n] : h) k += std::to_string(b) + ":" + std::to_string(lg(n)) + ";";
  return k;
}

static double time_stage(HpsiPlan& H, int stage, const WorkList& W, const double* dpsi, double* dout, int reps) {
  hipStream_t s = default_stream();
  WorkItem* di = upload_vec(W.items, s);
  ReduceItem* dr = upload_vec(W.reds, s);
  WorkQueue* dq = upload_vec(W.queues, s);
  double* sc = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, W.scratch_elems));
  const PairDesc* p = stage == 0 ? H.dpairs1 : H.dpairs2;
  std::vector<float> ms;
  hipEvent_t e0, e1;
  QSX_HIP(hipEventCreate(&e0));
  QSX_HIP(hipEventCreate(&e1));
  for (int r = 0; r < reps + 2; ++r) {
    QSX_HIP(hipEventRecord(e0, s));
    if (stage == 0)
      launch_worklist(W, dq, di, dr, p, H.dHL, dpsi, H.dT, sc, s);
    else
      launch_worklist(W, dq, di, dr, p, H.dT, H.dHR, dout, sc, s, dpsi);
    QSX_HIP(hipEventRecord(e1, s));
    QSX_HIP(hipEventSynchronize(e1));
    float t = 0;
    QSX_HIP(hipEventElapsedTime(&t, e0, e1));
    if (r >= 2) ms.push_back(t);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(di);
  dev_free(dr);
  dev_free(dq);
  dev_free(sc);
  std::sort(ms.begin(), ms.end());
  return ms[ms.size() / 2];
}

static void hpsi_launch(HpsiPlan& H, const double* dpsi, double* dhpsi, hipStream_t s, bool timed, bool sym);
static void hpsi_build_work(HpsiPlan& H);

// device ms of one whole H*psi (both stages, as uploaded), from runs of 8
// back-to-back applies: median over reps
static double time_whole(HpsiPlan& H, const double* dpsi, double* dout, int reps) {
  hipStream_t s = default_stream();
  constexpr int kRun = 8;
  std::vector<float> ms;
  hipEvent_t e0, e1;
  QSX_HIP(hipEventCreate(&e0));
  QSX_HIP(hipEventCreate(&e1));
  for (int r = 0; r < reps + 1; ++r) {
    QSX_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < kRun; ++i) hpsi_launch(H, dpsi, dout, s, false, false);
    QSX_HIP(hipEventRecord(e1, s));
    QSX_HIP(hipEventSynchronize(e1));
    float t = 0;
    QSX_HIP(hipEventElapsedTime(&t, e0, e1));
    if (r >= 1) ms.push_back(t / kRun);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::sort(ms.begin(), ms.end());
  return ms[ms.size() / 2];
}

static void hpsi_autotune(HpsiPlan& H, int reps, double* out) {
  if (H.mode != QSX_MODE_BATCHED) {  // the other modes have fixed work lists
    if (out) out[0] = out[1] = -1;
    return;
  }
  const std::string key = hist_key(H.targets1, H.pairs1) + "|" + hist_key(H.targets2, H.pairs2);
  std::pair<WorkPolicy, WorkPolicy> cached;
  bool hit = false;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      cached = it->second;
      hit = true;
    }
  }
  if (hit) {
    H.pol1 = cached.first;
    H.pol2 = cached.second;
    hpsi_build_work(H);
    if (out) out[0] = out[1] = -1;
    return;
  }
  const WorkPolicy in1 = H.pol1, in2 = H.pol2;
  double* dpsi = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, H.N));
  double* dout = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, H.N));
  launch_fill_zero(dpsi, H.N, default_stream());
  const int splits[] = {2, 4, 8, 16, 1 << 28};
  const double overheads[] = {0, 2048, 8192};
  const int slots[] = {resident_slots(), 2 * resident_slots()};
  for (int stage = 0; stage < 2; ++stage) {
    const auto& tg = stage == 0 ? H.targets1 : H.targets2;
    const auto& pr = stage == 0 ? H.pairs1 : H.pairs2;
    double best = 1e300;
    WorkPolicy bp;
    std::vector<WorkPolicy> cands;
    for (int sp : splits)
      for (double ov : overheads)
        for (int sl : slots) {
          WorkPolicy pol;
          pol.engine = ENG_GROUPED;
          pol.split_steps = sp;
          pol.tile_overhead = ov;
          pol.slots = sl;
          cands.push_back(pol);
        }
    for (int xb : {8, 0})
      for (int wm : {1, 2})
        for (int gw : {4, 8, 16}) {
          if (wm * wm * gw > 16) continue;
          for (int gm : {1 << 30, 256, 128, 64, 32})
            for (int sm : {1 << 30, 256, 128}) {
              WorkPolicy pol;
              pol.engine = ENG_WAVE;
              pol.wave_max = wm;
              pol.group_waves = gw;
              pol.group_mfma = gm;
              pol.split_mfma = sm;
              pol.xcd_bins = xb;
              cands.push_back(pol);
            }
        }
    for (const WorkPolicy& pol : cands) {
      WorkList W = build_work(tg, pr, pol);
      double t = time_stage(H, stage, W, dpsi, dout, reps);
      if (H.verbose)
        fprintf(stderr,
                "qsx autotune stage %d engine %d split %d/%d group %d/%d overhead %g slots %d wave_max %d xcd %d: "
                "%zu items %zu reds %.4f ms\n",
                stage + 1, W.engine, pol.split_steps, pol.split_mfma, pol.group_mfma, W.wg_waves, pol.tile_overhead,
                pol.slots, pol.wave_max, pol.xcd_bins, W.items.size(), W.reds.size(), t);
      if (t < best) {
        best = t;
        bp = pol;
      }
    }
    (stage == 0 ? H.pol1 : H.pol2) = bp;
    if (out) out[stage] = best;
  }
  // each stage was timed alone; a stage's cost inside an H*psi (after the
  // other stage, with the caches holding that stage's operands) can rank the
  // candidates differently, so the tuned pair is kept only if the whole
  // H*psi beats the plan's decomposition on entry (the default one of a
  // fresh plan, or the caller's)
  try {
    const WorkPolicy t1 = H.pol1, t2 = H.pol2;
    hpsi_build_work(H);
    const double t_tuned = time_whole(H, dpsi, dout, reps);
    H.pol1 = in1;
    H.pol2 = in2;
    hpsi_build_work(H);
    const double t_in = time_whole(H, dpsi, dout, reps);
    if (H.verbose)
      fprintf(stderr, "qsx autotune: whole H*psi %.4f ms tuned, %.4f ms with the entry decomposition\n", t_tuned,
              t_in);
    if (t_tuned < t_in) {
      H.pol1 = t1;
      H.pol2 = t2;
    }
  } catch (...) {
    dev_free(dpsi);
    dev_free(dout);
    throw;
  }
  dev_free(dpsi);
  dev_free(dout);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = {H.pol1, H.pol2};
  }
  hpsi_build_work(H);
}

// ---------------------------------------------------------------- split precision
// Lines of one operand side of a pair list: A side = rows of each A block view
// (over its K), B side = columns of each B block view.  Blocks shared by
// several pairs (same offset and strides) share their lines.
static std::vector<SplitLine> split_lines(const std::vector<PairDesc>& pairs, const std::vector<WorkTarget>& targets,
                                          bool aside, std::vector<int32_t>& ebase) {
  std::map<std::tuple<int64_t, int64_t, int64_t, int>, int32_t> seen;
  std::vector<SplitLine> lines;
  ebase.assign(pairs.size(), 0);
  for (const WorkTarget& t : targets)
    for (int64_t p = t.p0; p < t.p1; ++p) {
      const PairDesc& d = pairs[p];
      const int64_t off = aside ? d.a_off : d.b_off;
      const int64_t sl = aside ? ((d.flags & 1) ? d.lda : 1) : ((d.flags & 2) ? 1 : d.ldb);  // between lines
      const int64_t sk = aside ? ((d.flags & 1) ? 1 : d.lda) : ((d.flags & 2) ? d.ldb : 1);  // along a line
      const int n = aside ? t.M : t.N;
      auto key = std::make_tuple(off, sl, sk, n);
      auto it = seen.find(key);
      if (it != seen.end()) {
        ebase[p] = it->second;
        continue;
      }
      const int32_t base = (int32_t)lines.size();
      seen[key] = base;
      ebase[p] = base;
      for (int i = 0; i < n; ++i) lines.push_back(SplitLine{off + i * sl, 0, d.K, (int32_t)sk, base + i, base});
    }
  return lines;
}

static std::vector<SplitPair> split_pairs(const std::vector<PairDesc>& pairs, const std::vector<int32_t>& ea,
                                          const std::vector<int32_t>& eb, const std::vector<SplitLine>& la,
                                          const std::vector<SplitLine>& lb) {
  std::vector<SplitPair> v(pairs.size());
  for (size_t p = 0; p < pairs.size(); ++p) {
    const PairDesc& d = pairs[p];
    SplitPair& q = v[p];
    q.pa = la[ea[p]].pbase;
    q.pb = lb[eb[p]].pbase;
    q.K = d.K;
    q.ea = ea[p];
    q.eb = eb[p];
    q.pad = 0;
    q.afac = d.afac;
  }
  return v;
}

// packed storage of each line (quads of 4 k, S slices each, the last quad
// padded): offsets in line order; returns the total
static int64_t split_pack(std::vector<SplitLine>& lines, int S) {
  int64_t tot = 0;
  for (SplitLine& l : lines) {
    l.pbase = tot;
    tot += (int64_t)((l.K + 3) / 4) * 4 * S;
  }
  return tot;
}

// every line element lies inside its buffer (each line has its own packed
// slices, so one element read in two layouts is sliced twice)
static void split_check_lines(const std::vector<SplitLine>& lines, int64_t plane, const char* what) {
  for (const SplitLine& l : lines) {
    const int64_t last = l.off + (int64_t)(l.K - 1) * l.sk;
    if (l.K > 0 && (l.off < 0 || last < 0 || l.off >= plane || last >= plane))
      fail(QSX_ERR_STRUCT, "split precision: %s line outside its buffer", what);
  }
}

static void split_side_alloc(HpsiPlan::SplitSide& sp, const std::vector<SplitLine>& lines, int64_t packed, int S,
                             hipStream_t s, bool f32) {
  (void)S;
  sp.plane = packed;
  sp.nlines = (int)lines.size();
  int64_t strided = 0;
  for (const SplitLine& l : lines) {
    strided += l.sk != 1;
    sp.maxk = std::max(sp.maxk, l.K);
  }
  sp.lane = 2 * strided > (int64_t)lines.size();
  // zero once: the tails of the last quads are never written
  if (f32) {
    sp.fplanes = (float*)dev_alloc(sizeof(float) * std::max<int64_t>(1, packed));
    QSX_HIP(hipMemsetAsync(sp.fplanes, 0, sizeof(float) * std::max<int64_t>(1, packed), s));
  } else {
    sp.planes = (uint16_t*)dev_alloc(sizeof(uint16_t) * std::max<int64_t>(1, packed));
    QSX_HIP(hipMemsetAsync(sp.planes, 0, sizeof(uint16_t) * std::max<int64_t>(1, packed), s));
  }
  sp.exps = (int32_t*)dev_alloc(sizeof(int32_t) * std::max<size_t>(1, lines.size()));
  sp.dlines = upload_vec(lines, s);
}

// slice planes of the constant environments, lines of psi and T, the
// slice-addressed pairs and the 16x16-tile work lists
static void hpsi_split_setup(HpsiPlan& H) {
  hipStream_t s = default_stream();
  const int S = H.nsl;
  std::vector<int32_t> ea1, eb1, ea2, eb2;
  std::vector<SplitLine> lA1 = split_lines(H.pairs1, H.targets1, true, ea1);
  std::vector<SplitLine> lB1 = split_lines(H.pairs1, H.targets1, false, eb1);
  std::vector<SplitLine> lA2 = split_lines(H.pairs2, H.targets2, true, ea2);
  std::vector<SplitLine> lB2 = split_lines(H.pairs2, H.targets2, false, eb2);
  const int64_t hl_arena = H.hl_total;
  split_check_lines(lA1, hl_arena, "HL");
  split_check_lines(lB1, H.N, "psi");
  split_check_lines(lA2, H.t_total, "T");
  split_check_lines(lB2, H.hr_total, "HR");
  const bool f32 = H.mode == QSX_MODE_SPLIT_F32;
  split_side_alloc(H.spHL, lA1, split_pack(lA1, S), S, s, f32);
  split_side_alloc(H.spPsi, lB1, split_pack(lB1, S), S, s, f32);
  split_side_alloc(H.spT, lA2, split_pack(lA2, S), S, s, f32);
  split_side_alloc(H.spHR, lB2, split_pack(lB2, S), S, s, f32);
  if (f32) {
    launch_split_lines_f32(H.spHL.dlines, H.spHL.nlines, H.spHL.maxk, H.spHL.lane, H.dHL, H.spHL.fplanes, H.spHL.exps, s);
    launch_split_lines_f32(H.spHR.dlines, H.spHR.nlines, H.spHR.maxk, H.spHR.lane, H.dHR, H.spHR.fplanes, H.spHR.exps, s);
  } else {
    launch_split_lines(H.spHL.dlines, H.spHL.nlines, H.spHL.maxk, H.spHL.lane, H.dHL, H.spHL.planes, S, H.spHL.exps, s);
    launch_split_lines(H.spHR.dlines, H.spHR.nlines, H.spHR.maxk, H.spHR.lane, H.dHR, H.spHR.planes, S, H.spHR.exps, s);
  }
  H.dsp1 = upload_vec(split_pairs(H.pairs1, ea1, eb1, lA1, lB1), s);
  H.dsp2 = upload_vec(split_pairs(H.pairs2, ea2, eb2, lA2, lB2), s);
  H.dsp1s = upload_vec(split_pairs(H.pairs1s, ea1, eb1, lA1, lB1), s);
  H.dsp2s = upload_vec(split_pairs(H.pairs2s, ea2, eb2, lA2, lB2), s);
  // one 16 x 16 tile per wave; a tile longer than 32 k-steps of 4 is cut
  // into k-chunks whose partial tiles the wave engine's reduction sums
  WorkPolicy pol;
  pol.engine = ENG_WAVE;
  pol.wave_max = 1;
  pol.group_mfma = 1 << 30;
  pol.split_mfma = 32;
  H.swork1 = build_work(H.targets1, H.pairs1, pol);
  H.swork2 = build_work(H.targets2, H.pairs2, pol);
  // the split kernel knows only 16 x 16 tiles stored into C or into a
  // partial slot: a work list of another form (the workgroup-tile fallback
  // for blocks beyond the wave engine's 32-bit offsets, shared tiles)
  // would leave output elements unwritten
  for (const WorkList* W : {&H.swork1, &H.swork2}) {
    bool ok = W->engine == ENG_WAVE;
    for (const WorkItem& w : W->items) {
      if (w.slot == -2) continue;  // padding
      if (w.shape != wave_shape(1, 1) || w.group != 0) ok = false;
    }
    for (const ReduceItem& r : W->reds)
      if (r.shape != wave_shape(1, 1)) ok = false;
    if (!ok) fail(QSX_ERR_UNSUPPORTED, "split precision: blocks too large for the 16 x 16 tile kernel");
  }
  H.dsitems1 = upload_vec(H.swork1.items, s);
  H.dsitems2 = upload_vec(H.swork2.items, s);
  H.dsreds1 = upload_vec(H.swork1.reds, s);
  H.dsreds2 = upload_vec(H.swork2.reds, s);
  H.dsscratch = (double*)dev_alloc(sizeof(double) *
                                   std::max<int64_t>(1, std::max(H.swork1.scratch_elems, H.swork2.scratch_elems)));
  // executed slice-product flops: per 16x16 tile and k-step of a pair,
  // BF16: S(S+1)/2 MFMAs of 2*16*16*16 per 16 k; FP32 x 3: 6 MFMAs of
  // 2*16*16*4 per 4 k
  const WorkList* W[2] = {&H.swork1, &H.swork2};
  const std::vector<PairDesc>* P[2] = {&H.pairs1, &H.pairs2};
  const int ks = f32 ? 4 : 16;
  const double per_step = f32 ? 6 * 2.0 * 16 * 16 * 4 : (S * (S + 1) / 2) * 2.0 * 16 * 16 * 16;
  for (int st = 0; st < 2; ++st) {
    double steps = 0;
    for (const WorkItem& w : W[st]->items)
      for (int p = w.p0; p <= w.p1 && w.slot != -2; ++p) {
        const int kb = p == w.p0 ? w.k0 : 0, ke = p == w.p1 ? w.k1 : (*P[st])[p].K;
        steps += (ke - kb + ks - 1) / ks;
      }
    H.split_exec_flops[st] = steps * per_step;
  }
  QSX_HIP(hipStreamSynchronize(s));
}

static void hpsi_launch(HpsiPlan& H, const double* dpsi, double* dhpsi, hipStream_t s, bool timed, bool sym) {
  const PairDesc* p1 = sym ? H.dpairs1s : H.dpairs1;
  const PairDesc* p2 = sym ? H.dpairs2s : H.dpairs2;
  // events: [0] start, [1] stage 1 done, [2] stage 2 done
  if (timed) QSX_HIP(hipEventRecord(H.ev[0], s));
  if (H.mode == QSX_MODE_SPLIT_F32) {
    launch_split_lines_f32(H.spPsi.dlines, H.spPsi.nlines, H.spPsi.maxk, H.spPsi.lane, dpsi, H.spPsi.fplanes, H.spPsi.exps, s);
    launch_split_gemm_f32(H.dsitems1, (int)H.swork1.items.size(), sym ? H.dsp1s : H.dsp1, H.spHL.fplanes,
                          H.spHL.exps, H.spPsi.fplanes, H.spPsi.exps, H.dT, H.dsscratch, s);
    launch_wave_reduce(H.dsreds1, (int)H.swork1.reds.size(), H.dsscratch, H.dT, s);
    if (timed) QSX_HIP(hipEventRecord(H.ev[1], s));
    launch_split_lines_f32(H.spT.dlines, H.spT.nlines, H.spT.maxk, H.spT.lane, H.dT, H.spT.fplanes, H.spT.exps, s);
    launch_split_gemm_f32(H.dsitems2, (int)H.swork2.items.size(), sym ? H.dsp2s : H.dsp2, H.spT.fplanes,
                          H.spT.exps, H.spHR.fplanes, H.spHR.exps, dhpsi, H.dsscratch, s);
    launch_wave_reduce(H.dsreds2, (int)H.swork2.reds.size(), H.dsscratch, dhpsi, s);
  } else if (H.mode == QSX_MODE_SPLIT_BF16) {
    const int S = H.nsl;
    launch_split_lines(H.spPsi.dlines, H.spPsi.nlines, H.spPsi.maxk, H.spPsi.lane, dpsi, H.spPsi.planes, S, H.spPsi.exps, s);
    launch_split_gemm(S, H.dsitems1, (int)H.swork1.items.size(), sym ? H.dsp1s : H.dsp1, H.spHL.planes, H.spHL.exps,
                      H.spPsi.planes, H.spPsi.exps, H.dT, H.dsscratch, s);
    launch_wave_reduce(H.dsreds1, (int)H.swork1.reds.size(), H.dsscratch, H.dT, s);
    if (timed) QSX_HIP(hipEventRecord(H.ev[1], s));
    launch_split_lines(H.spT.dlines, H.spT.nlines, H.spT.maxk, H.spT.lane, H.dT, H.spT.planes, S, H.spT.exps, s);
    launch_split_gemm(S, H.dsitems2, (int)H.swork2.items.size(), sym ? H.dsp2s : H.dsp2, H.spT.planes, H.spT.exps,
                      H.spHR.planes, H.spHR.exps, dhpsi, H.dsscratch, s);
    launch_wave_reduce(H.dsreds2, (int)H.swork2.reds.size(), H.dsscratch, dhpsi, s);
  } else {
    launch_worklist(H.work1, H.dq1, H.ditems1, H.dreds1, p1, H.dHL, dpsi, H.dT, H.dscratch, s);
    if (timed) QSX_HIP(hipEventRecord(H.ev[1], s));
    launch_worklist(H.work2, H.dq2, H.ditems2, H.dreds2, p2, H.dT, H.dHR, dhpsi, H.dscratch, s, dpsi);
  }
  if (timed) QSX_HIP(hipEventRecord(H.ev[2], s));
}

// plain per-pair path (no batching, no graph): the baseline
static void hpsi_plain(HpsiPlan& H, const double* dpsi, double* dhpsi, hipStream_t s) {
  auto run = [&](const std::vector<PairDesc>& pairs, const std::vector<WorkTarget>& targets, const double* A,
                 const double* B, double* C) {
    for (const WorkTarget& t : targets) {
      if (t.p0 == t.p1) {
        launch_fill_zero(C + t.c_off, (int64_t)t.M * t.N, s);
        continue;
      }
      for (int64_t p = t.p0; p < t.p1; ++p) {
        const PairDesc& pd = pairs[p];
        launch_pair_gemm_f64(pd, t.c_off, t.M, t.N, p == t.p0 ? 0 : 1, (pd.flags & kPairAinX) ? dpsi : A,
                             (pd.flags & kPairBinX) ? dpsi : B, C, s);
      }
    }
  };
  run(H.pairs1, H.targets1, H.dHL, dpsi, H.dT);
  run(H.pairs2, H.targets2, H.dT, H.dHR, dhpsi);
}

// The plan of H*psi = [HK_L.psi] + [psi.HK_R] + sum_t contractQS({L_t[*], psi}, R_t[*]):
// stage 1 computes the intermediates T_t = L_t[*] . psi of every term, stage
// 2 every block of H*psi as one sum over the local terms' pairs (operand psi
// read directly) and the terms' T_t . R_t[*] pairs, in that order.
static void hpsi_create(HpsiPlan& H, const std::vector<const qsx_tensor*>& Ls, const std::vector<const qsx_tensor*>& Rs,
                        const std::vector<HpsiTerm>& terms, const qsx_tensor* psi, const qsx_tensor* hkl,
                        const qsx_tensor* hkr, const qsx_opts* o) {
  for (const qsx_tensor* t : Ls)
    if (t && t->data_im) fail(QSX_ERR_UNSUPPORTED, "H*psi plans take real tensors (complex data not supported)");
  for (const qsx_tensor* t : Rs)
    if (t && t->data_im) fail(QSX_ERR_UNSUPPORTED, "H*psi plans take real tensors (complex data not supported)");
  for (const qsx_tensor* t : {psi, hkl, hkr})
    if (t && t->data_im) fail(QSX_ERR_UNSUPPORTED, "H*psi plans take real tensors (complex data not supported)");
  if (!psi) fail(QSX_ERR_ARG, "H*psi: psi missing");
  for (const HpsiTerm& t : terms)
    if (t.li < 0 || t.li >= (int)Ls.size() || t.ri < 0 || t.ri >= (int)Rs.size() || !Ls[t.li] || !Rs[t.ri])
      fail(QSX_ERR_ARG, "H*psi: term operator missing");
  if (terms.empty() && !hkl && !hkr) fail(QSX_ERR_ARG, "H*psi: no terms");
  ensure_device();
  PhaseClock clk{o ? o->verbose : 0, "H*psi plan"};
  hipStream_t s = default_stream();
  for (size_t i = 0; i < Ls.size(); ++i) H.Ls.push_back(make_desc(Ls[i], "HL"));
  for (size_t i = 0; i < Rs.size(); ++i) H.Rs.push_back(make_desc(Rs[i], "HR"));
  H.psi = make_desc(psi, "psi");
  H.terms = terms;
  H.nterm = (int)terms.size();
  H.has_hkl = hkl != nullptr;
  H.has_hkr = hkr != nullptr;
  if (hkl) H.HKL = make_desc(hkl, "HK_L");
  if (hkr) H.HKR = make_desc(hkr, "HK_R");
  for (const auto& T : H.Ls) require_no_om(T, "H*psi");
  for (const auto& T : H.Rs) require_no_om(T, "H*psi");
  require_no_om(H.psi, "H*psi");
  if (hkl) require_no_om(H.HKL, "H*psi");
  if (hkr) require_no_om(H.HKR, "H*psi");
  H.xpairs = H.has_hkl || H.has_hkr;
  if (o) {
    H.mode = o->mode;
    H.nslices = o->nslices;
    H.verbose = o->verbose;
  }
  if (H.mode < QSX_MODE_BATCHED || H.mode > QSX_MODE_SPLIT_F32) fail(QSX_ERR_ARG, "invalid mode %d", H.mode);
  if (H.mode == QSX_MODE_SPLIT_BF16 || H.mode == QSX_MODE_SPLIT_F32) {
    if (H.xpairs || Ls.size() != 1 || Rs.size() != 1)
      fail(QSX_ERR_UNSUPPORTED, "split precision: plans of one environment pair (full MPO) only");
    if (H.mode == QSX_MODE_SPLIT_F32) {
      if (H.nslices != 0 && H.nslices != 3) fail(QSX_ERR_ARG, "FP32 split precision uses 3 slices (got %d)", H.nslices);
      H.nsl = 3;
    } else {
      H.nsl = H.nslices > 0 ? H.nslices : 7;
      if (H.nsl < 3 || H.nsl > 8) fail(QSX_ERR_ARG, "split precision: %d slices (3..8 supported)", H.nsl);
    }
  }
  const double* fermA = o ? o->fermA : nullptr;  // exact +-1 factors of the blocks of Ls[0] (parity operators)
  if (fermA && Ls.size() != 1) fail(QSX_ERR_ARG, "H*psi: fermion signs need a single left environment");
  H.N = H.psi.total();
  H.pol1.slots = H.pol2.slots = resident_slots();
  // default decomposition of an untuned plan: the k-range of a wave item
  // split beyond 256 MFMAs in both stages; stage 2 (T . R into psi's ragged
  // blocks) in single 16x16 wave tiles.  Measured on the Heisenberg (D* =
  // 2000, 4000, 8000) and Hubbard (D* = 4000) operators, this is within the
  // autotuner's timing noise of its choice, where the generic contraction
  // default (2x2-fragment tiles, no split) left stage 2 up to 1.5x slower.
  H.pol1.split_mfma = 256;
  H.pol2.split_mfma = 256;
  H.pol2.wave_max = 1;
  qsx_opts so{};
  so.mode = QSX_MODE_BATCHED;
  so.verbose = 0;
  so.skip_zero = 0;
  // stage 1 plans: ({L_t [,'*'], psi}) with itag matching
  H.s1.resize(H.nterm);
  H.s2.resize(H.nterm);
  for (int t = 0; t < H.nterm; ++t) {
    const HpsiTerm& tm = terms[t];
    std::vector<int32_t> ia, ib;
    int cA = 0;
    match_itags(H.Ls[tm.li], H.psi, tm.cl, 0, ia, ib, cA);
    qsx_opts s1o = so;
    s1o.fermA = fermA;
    H.s1[t] = plan_create(Ls[tm.li], ia.data(), (int)ia.size(), cA, psi, ib.data(), (int)ib.size(), 0, nullptr, 0,
                          &s1o, false, 1);
    if (H.s1[t]->lb.permute) fail(QSX_ERR_UNSUPPORTED, "H*psi: psi needs a permute in stage 1");
    if (H.s1[t]->om_path) fail(QSX_ERR_UNSUPPORTED, "H*psi: SU(2) intermediates of rank > 3 (outer multiplicity)");
  }
  clk.lap("stage-1 structure");
  // stage 2 plans: (T_t, R_t [,'*'])
  std::vector<HostTensor> Tt(H.nterm);
  for (int t = 0; t < H.nterm; ++t) {
    const HpsiTerm& tm = terms[t];
    CtrPlan& P = *H.s1[t];
    Tt[t].Q = P.Qout;
    Tt[t].dims = P.dimsOut;
    Tt[t].tags = P.tagsC;
    Tt[t].qtype = P.A.qtype;
    Tt[t].bind(P.nC, P.rc, P.A.qdim);
    TensorDesc Td = make_desc(&Tt[t].t, "T");
    std::vector<int32_t> ia, ib;
    int cA = 0;
    match_itags(Td, H.Rs[tm.ri], 0, tm.cr, ia, ib, cA);
    H.s2[t] = plan_create(&Tt[t].t, ia.data(), (int)ia.size(), cA, Rs[tm.ri], ib.data(), (int)ib.size(), tm.cr,
                          nullptr, 0, &so, false, 2);
    if (H.s2[t]->la.permute) fail(QSX_ERR_UNSUPPORTED, "H*psi: intermediate needs a permute in stage 2");
    if (H.s2[t]->om_path) fail(QSX_ERR_UNSUPPORTED, "H*psi: SU(2) intermediates of rank > 3 (outer multiplicity)");
  }
  // local terms: HK_L . psi (itag matching, psi as the B operand) and
  // psi . HK_R (psi as the A operand); psi is read where it lies
  if (hkl) {
    std::vector<int32_t> ia, ib;
    int cA = 0;
    match_itags(H.HKL, H.psi, 0, 0, ia, ib, cA);
    H.locL = plan_create(hkl, ia.data(), (int)ia.size(), cA, psi, ib.data(), (int)ib.size(), 0, nullptr, 0, &so,
                         false, 1);
    if (H.locL->lb.permute) fail(QSX_ERR_UNSUPPORTED, "H*psi: psi needs a permute in HK_L . psi");
  }
  if (hkr) {
    std::vector<int32_t> ia, ib;
    int cA = 0;
    match_itags(H.psi, H.HKR, 0, 0, ia, ib, cA);
    H.locR = plan_create(psi, ia.data(), (int)ia.size(), cA, hkr, ib.data(), (int)ib.size(), 0, nullptr, 0, &so,
                         false, 2);
    if (H.locR->la.permute) fail(QSX_ERR_UNSUPPORTED, "H*psi: psi needs a permute in psi . HK_R");
  }
  clk.lap("stage-2 structure");
  // device arenas.  dHL: every left operator as uploaded, then a permuted
  // copy per term whose stage-1 layout needs one
  std::vector<int64_t> l_base(Ls.size(), 0);
  int64_t off = 0;
  for (size_t i = 0; i < Ls.size(); ++i) {
    l_base[i] = off;
    off += H.Ls[i].total();
  }
  H.hl_base.assign(H.nterm, 0);
  for (int t = 0; t < H.nterm; ++t) {
    if (H.s1[t]->la.permute) {
      H.hl_base[t] = off;
      off += H.Ls[terms[t].li].total();
    } else H.hl_base[t] = l_base[terms[t].li];
  }
  H.hl_total = off;
  H.dHL = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, off));
  for (size_t i = 0; i < Ls.size(); ++i) upload_blocks(H.Ls[i], H.dHL + l_base[i], s);
  for (int t = 0; t < H.nterm; ++t)
    if (H.s1[t]->la.permute) plan_permute_operand(*H.s1[t], 0, H.dHL + l_base[terms[t].li], H.dHL + H.hl_base[t], s);
  // dHR: a copy of R_t per term in its stage-2 layout, then HK_R in its layout
  int64_t roff = 0;
  H.hr_base.assign(H.nterm, 0);
  for (int t = 0; t < H.nterm; ++t) {
    H.hr_base[t] = roff;
    roff += H.Rs[terms[t].ri].total();
  }
  H.hkr_base = roff;
  if (hkr) roff += H.HKR.total();
  H.hr_total = roff;
  H.dHR = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, roff));
  {
    int64_t tmp_n = hkr ? H.HKR.total() : 0;
    for (const TensorDesc& R : H.Rs) tmp_n = std::max(tmp_n, R.total());
    double* tmp = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, tmp_n));
    try {
      for (int t = 0; t < H.nterm; ++t) {
        const TensorDesc& R = H.Rs[terms[t].ri];
        upload_blocks(R, tmp, s);
        if (H.s2[t]->lb.permute) plan_permute_operand(*H.s2[t], 1, tmp, H.dHR + H.hr_base[t], s);
        else QSX_HIP(hipMemcpyAsync(H.dHR + H.hr_base[t], tmp, sizeof(double) * R.total(), hipMemcpyDeviceToDevice, s));
        QSX_HIP(hipStreamSynchronize(s));
      }
      if (hkr) {
        upload_blocks(H.HKR, tmp, s);
        if (H.locR->lb.permute) plan_permute_operand(*H.locR, 1, tmp, H.dHR + H.hkr_base, s);
        else QSX_HIP(hipMemcpyAsync(H.dHR + H.hkr_base, tmp, sizeof(double) * H.HKR.total(), hipMemcpyDeviceToDevice, s));
      }
      QSX_HIP(hipStreamSynchronize(s));
    } catch (...) {
      dev_free(tmp);
      throw;
    }
    dev_free(tmp);
  }
  // dT: the intermediates T_t (rewritten by every apply), then HK_L in its
  // layout (constant)
  int64_t toff = 0;
  H.t_base.assign(H.nterm, 0);
  for (int t = 0; t < H.nterm; ++t) {
    H.t_base[t] = toff;
    toff += H.s1[t]->offsOut.back();
  }
  H.t_total = toff;
  H.hkl_base = toff;
  if (hkl) toff += H.HKL.total();
  H.dt_size = toff;
  H.dT = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, toff));
  if (hkl) {
    if (H.locL->la.permute) {
      double* tmp = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, H.HKL.total()));
      try {
        upload_blocks(H.HKL, tmp, s);
        plan_permute_operand(*H.locL, 0, tmp, H.dT + H.hkl_base, s);
      } catch (...) {
        dev_free(tmp);
        throw;
      }
      dev_free(tmp);
    } else {
      upload_blocks(H.HKL, H.dT + H.hkl_base, s);
    }
  }
  clk.lap("environment upload + permute");
  // weights of psi's blocks: prod_j cgw_j^2 (|CData|_F = 1)
  std::vector<double> wblk(H.psi.n, 1.0);
  {
    std::vector<int> dir(H.psi.r, 1);
    for (int i = 0; i < H.psi.r && !H.psi.itags.empty(); ++i) dir[i] = (H.psi.itags[i] & 0x80) ? -1 : 1;
    for (int i = 0; i < H.psi.n; ++i)
      for (int j = 0; j < H.psi.nsym; ++j) {
        if (H.psi.symtype[j] != SYM_SU2) continue;
        double c;
        if (H.psi.cgw_given) c = H.psi.cgw[(size_t)i * H.psi.nsym + j];
        else {
          std::vector<int> tj(H.psi.r);
          for (int l = 0; l < H.psi.r; ++l) tj[l] = H.psi.label(i, l, j);
          c = su2_norm_ext(tj, dir);
        }
        wblk[i] *= c * c;
      }
  }
  H.welem.assign(H.N, 1.0);
  for (int i = 0; i < H.psi.n; ++i)
    for (int64_t e = H.psi.offs[i]; e < H.psi.offs[i + 1]; ++e) H.welem[e] = wblk[i];
  H.wblk = wblk;
  // stage 1 work: every T_t block is its own target
  {
    std::vector<CtrPlan*> plans;
    std::vector<int64_t> ab, bb, toffs;
    std::vector<std::vector<int>> target;
    std::vector<int> tM, tN;
    for (int t = 0; t < H.nterm; ++t) {
      CtrPlan& P = *H.s1[t];
      plans.push_back(&P);
      ab.push_back(H.hl_base[t]);
      bb.push_back(0);
      std::vector<int> tg(P.nC);
      for (int c = 0; c < P.nC; ++c) {
        tg[c] = (int)toffs.size();
        toffs.push_back(H.t_base[t] + P.offsOut[c]);
        tM.push_back(P.M[c]);
        tN.push_back(P.N[c]);
      }
      target.push_back(tg);
    }
    std::vector<std::pair<int, int64_t>> src1;
    build_multi(plans, ab, bb, target, toffs, tM, tN, H.pairs1, H.targets1, nullptr, &src1);
    H.pairs1s = H.pairs1;
    for (size_t p = 0; p < H.pairs1s.size(); ++p) {
      int ib = H.s1[src1[p].first]->pIb[src1[p].second];
      H.pairs1s[p].afac /= std::sqrt(wblk[ib]);
    }
  }
  // stage 2 work: targets are psi's blocks (H*psi in psi's packed layout);
  // per target the local terms' pairs first, then the terms' in term order
  {
    const int q = H.psi.qdim, r = H.psi.r;
    std::map<std::string, int> psi_idx;
    for (int i = 0; i < H.psi.n; ++i) psi_idx[rec_key(&H.psi.Q[(size_t)i * r * q], r * q)] = i;
    std::vector<CtrPlan*> plans;
    std::vector<int64_t> ab, bb, toffs(H.psi.n);
    std::vector<int32_t> xf;
    std::vector<std::vector<int>> target;
    std::vector<int> tM(H.psi.n), tN(H.psi.n);
    std::vector<int> kind;  // per plan: 0 term, 1 HK_L . psi (psi = B), 2 psi . HK_R (psi = A)
    for (int i = 0; i < H.psi.n; ++i) {
      toffs[i] = H.psi.offs[i];
      int64_t m = H.psi.dims[(size_t)i * r], n = 1;
      for (int k = 1; k < r; ++k) n *= H.psi.dims[(size_t)i * r + k];
      tM[i] = (int)m;
      tN[i] = (int)n;
    }
    auto add = [&](CtrPlan& P, int64_t a0, int64_t b0, int32_t flags, int kd) {
      if (P.rc != r) fail(QSX_ERR_STRUCT, "H*psi result has rank %d, psi rank %d", P.rc, r);
      for (int l = 0; l < r; ++l)
        if (P.tagsC.size() == (size_t)r && H.psi.itags.size() == (size_t)r && P.tagsC[l] != H.psi.itags[l])
          fail(QSX_ERR_ITAG, "H*psi result itags differ from psi's at leg %d", l + 1);
      plans.push_back(&P);
      ab.push_back(a0);
      bb.push_back(b0);
      xf.push_back(flags);
      kind.push_back(kd);
      std::vector<int> tg(P.nC);
      for (int c = 0; c < P.nC; ++c) {
        auto it = psi_idx.find(rec_key(&P.Qout[(size_t)c * r * q], r * q));
        if (it == psi_idx.end()) fail(QSX_ERR_STRUCT, "H*psi produces a block outside psi's block structure");
        tg[c] = it->second;
      }
      target.push_back(tg);
    };
    if (hkl) add(*H.locL, H.hkl_base, 0, kPairBinX, 1);
    if (hkr) add(*H.locR, 0, H.hkr_base, kPairAinX, 2);
    for (int t = 0; t < H.nterm; ++t) add(*H.s2[t], H.t_base[t], H.hr_base[t], 0, 0);
    std::vector<int> tgt2;
    std::vector<std::pair<int, int64_t>> src2;
    build_multi(plans, ab, bb, target, toffs, tM, tN, H.pairs2, H.targets2, &tgt2, &src2, &xf);
    H.pairs2s = H.pairs2;
    for (size_t p = 0; p < H.pairs2s.size(); ++p) {
      double f = std::sqrt(H.wblk[tgt2[p]]);
      const int pl = src2[p].first;
      if (kind[pl] == 1) f /= std::sqrt(wblk[plans[pl]->pIb[src2[p].second]]);
      if (kind[pl] == 2) f /= std::sqrt(wblk[plans[pl]->pIa[src2[p].second]]);
      H.pairs2s[p].afac *= f;
    }
  }
  clk.lap("pair lists");
  {
    std::vector<UploadPart> parts;
    H.dpairs1 = alloc_for(H.pairs1, parts);
    H.dpairs2 = alloc_for(H.pairs2, parts);
    H.dpairs1s = alloc_for(H.pairs1s, parts);
    H.dpairs2s = alloc_for(H.pairs2s, parts);
    upload_many(parts, s);
  }
  hpsi_build_work(H);
  clk.lap("work lists");
  if (H.mode == QSX_MODE_SPLIT_BF16 || H.mode == QSX_MODE_SPLIT_F32) hpsi_split_setup(H);
  QSX_HIP(hipStreamSynchronize(s));
  H.flops = 0;
  for (int t = 0; t < H.nterm; ++t) {
    H.stage_flops[0] += H.s1[t]->flops;
    H.stage_flops[1] += H.s2[t]->flops;
  }
  if (H.locL) H.stage_flops[1] += H.locL->flops;
  if (H.locR) H.stage_flops[1] += H.locR->flops;
  H.flops = H.stage_flops[0] + H.stage_flops[1];
  // algorithmic bytes per stage: unique operand and result elements
  int64_t nl = 0, nr = 0;
  for (const TensorDesc& L : H.Ls) nl += L.total();
  for (const TensorDesc& R : H.Rs) nr += R.total();
  const int64_t nloc = (hkl ? H.HKL.total() : 0) + (hkr ? H.HKR.total() : 0);
  H.stage_bytes[0] = 8.0 * ((double)nl + (double)H.N + (double)H.t_total);
  H.stage_bytes[1] = 8.0 * ((double)H.t_total + (double)nr + (double)nloc + (double)H.N * (H.xpairs ? 2 : 1));
  // algorithmic bytes: unique operators + psi + H*psi (+ the T_t write and re-read)
  H.bytes = 8.0 * ((double)nl + (double)nr + (double)nloc + 2.0 * (double)H.N + 2.0 * (double)H.t_total);
  for (auto& e : H.ev) QSX_HIP(hipEventCreate(&e));
}

static void hpsi_apply(HpsiPlan& H, const double* dpsi, double* dhpsi, hipStream_t s, bool sym = false) {
  if (H.mode == QSX_MODE_PLAIN) {
    if (sym) fail(QSX_ERR_ARG, "plain mode has no symmetric-basis variant");
    hpsi_plain(H, dpsi, dhpsi, s);
    return;
  }
  if (!H.use_graph) {
    hpsi_launch(H, dpsi, dhpsi, s, true, sym);
    return;
  }
  if (!H.gexec || H.g_psi != dpsi || H.g_hpsi != dhpsi || H.g_sym != sym) {
    if (H.gexec) {
      (void)hipGraphExecDestroy(H.gexec);
      H.gexec = nullptr;
    }
    hipGraph_t g;
    QSX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hpsi_launch(H, dpsi, dhpsi, s, false, sym);
    QSX_HIP(hipStreamEndCapture(s, &g));
    QSX_HIP(hipGraphInstantiate(&H.gexec, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    H.g_psi = dpsi;
    H.g_hpsi = dhpsi;
    H.g_sym = sym;
  }
  QSX_HIP(hipGraphLaunch(H.gexec, s));
}

// ---------------------------------------------------------------- eigensolvers
// symmetric eigen-decomposition by cyclic Jacobi (small dense matrices)

// eigensolver set-up shared by Lanczos and Davidson: sqrt(welem) on the
// device and a pinned host staging buffer (made once per plan), and a
// workspace of at least `doubles` elements (grown on demand, kept until the
// plan is destroyed)
static double* eig_workspace(HpsiPlan& H, size_t doubles, size_t small, hipStream_t s) {
  // nothing of an earlier call (one that failed part-way included) may
  // still read the pinned buffers the host is about to write
  QSX_HIP(hipStreamSynchronize(s));
  if (!H.dsw) {
    std::vector<double> sw(H.N);
    for (int64_t i = 0; i < H.N; ++i) sw[i] = std::sqrt(H.welem[i]);
    H.dsw = upload_vec(sw, s);
  }
  if (!H.hstage) {
    hipError_t e = hipHostMalloc((void**)&H.hstage, sizeof(double) * std::max<int64_t>(1, H.N), hipHostMallocDefault);
    if (e != hipSuccess) {
      H.hstage = nullptr;
      fail(QSX_ERR_NOMEM, "hipHostMalloc(%lld doubles) failed: %s", (long long)H.N, hipGetErrorString(e));
    }
  }
  if (H.hsmall_n < small) {
    if (H.hsmall) (void)hipHostFree(H.hsmall);
    H.hsmall = nullptr;
    H.hsmall_n = 0;
    H.dsmall = nullptr;
    // mapped and coherent: kernels store the Davidson scalars here and the
    // host reads them after the stream's synchronisation
    hipError_t e = hipHostMalloc((void**)&H.hsmall, sizeof(double) * small, hipHostMallocMapped | hipHostMallocCoherent);
    if (e != hipSuccess) {
      H.hsmall = nullptr;
      fail(QSX_ERR_NOMEM, "hipHostMalloc(%zu doubles) failed: %s", small, hipGetErrorString(e));
    }
    e = hipHostGetDevicePointer((void**)&H.dsmall, H.hsmall, 0);
    if (e != hipSuccess || !H.dsmall) {
      (void)hipHostFree(H.hsmall);
      H.hsmall = H.dsmall = nullptr;
      fail(QSX_ERR_DEVICE, "hipHostGetDevicePointer failed: %s", hipGetErrorString(e));
    }
    H.hsmall_n = small;
  }
  if (H.ws_doubles < doubles) {
    dev_free(H.dws);
    H.dws = nullptr;
    H.ws_doubles = 0;
    H.dws = (double*)dev_alloc(sizeof(double) * doubles);
    H.ws_doubles = doubles;
  }
  return H.dws;
}

// host <-> device copies of N-element vectors through the pinned staging
// buffer (a DMA each way instead of the runtime's pageable path); the
// stream orders the staging buffer's uses
static void eig_upload(HpsiPlan& H, double* d, const double* h, hipStream_t s) {
  std::memcpy(H.hstage, h, sizeof(double) * H.N);
  QSX_HIP(hipMemcpyAsync(d, H.hstage, sizeof(double) * H.N, hipMemcpyHostToDevice, s));
}
static void eig_download(HpsiPlan& H, double* h, const double* d, hipStream_t s) {
  QSX_HIP(hipMemcpyAsync(H.hstage, d, sizeof(double) * H.N, hipMemcpyDeviceToHost, s));
  QSX_HIP(hipStreamSynchronize(s));
  std::memcpy(h, H.hstage, sizeof(double) * H.N);
}

// carve consecutive regions out of the workspace, 256-byte aligned
struct Carve {
  double* base;
  size_t used = 0;
  double* take(size_t n) {
    double* p = base + used;
    used += (n + 31) / 32 * 32;
    return p;
  }
  static size_t size(std::initializer_list<size_t> parts) {
    size_t t = 0;
    for (size_t n : parts) t += (n + 31) / 32 * 32;
    return t;
  }
};

// H*psi in the symmetric basis inside the eigensolvers: the stage kernels
// launched directly (the input and output move through the Krylov vectors,
// so a graph would be re-instantiated on every step); plain-mode plans run
// their batched work lists here, the per-pair path having no symmetric-basis
// variant
static void eig_hpsi(HpsiPlan& H, const double* x, double* y, hipStream_t s) {
  hpsi_launch(H, x, y, s, false, true);
}

static void lanczos(HpsiPlan& H, const double* psi0, int niter, double* e0, double* ritz, double* gs) {
  hipStream_t s = default_stream();
  const int64_t N = H.N;
  if (niter < 1 || niter > 500) fail(QSX_ERR_ARG, "niter must be in 1..500");
  PhaseClock clk{H.verbose, "Lanczos"};
  const size_t nparts = (size_t)DOT_BLOCKS * (niter + 1);
  Carve ws{eig_workspace(H, Carve::size({(size_t)N * (niter + 1), (size_t)std::max<int64_t>(1, N), nparts, nparts,
                                         (size_t)niter}),
                         2 * (size_t)niter, s)};
  double* V = ws.take((size_t)N * (niter + 1));
  double* w = ws.take(std::max<int64_t>(1, N));
  double* pa = ws.take(nparts);
  double* pb = ws.take(nparts);
  double* dy = ws.take(niter);
  // alpha_j and beta_j are stored by the kernels into the mapped pinned
  // buffer [alpha | beta] and read by the host after the loop's
  // synchronisation
  double* alpha = H.dsmall;
  double* beta = H.dsmall + niter;
  // v_0 = sqrt(w) psi0 / |sqrt(w) psi0| (psi0 lands in the unused slot v_1)
  eig_upload(H, V + N, psi0, s);
  vec_scale_norm_partial(V + N, H.dsw, w, N, pa, s);
  vec_normalize_partial(w, V, pa, nullptr, N, s);
  clk.lap("upload psi0");
  // the iteration sequence is launched directly: the host enqueues all
  // steps while the device runs the earlier ones (5 launches per step)
  for (int j = 0; j < niter; ++j) {
    double* vj = V + (int64_t)j * N;
    eig_hpsi(H, vj, w, s);
    // full re-orthogonalisation, classical Gram-Schmidt twice; the first
    // pass's coefficient of v_j is alpha_j = <v_j, H v_j>; the second pass
    // is fused with the normalisation (beta_j from |w'|^2 - |c'|^2)
    vec_mdot_partial(V, N, j + 1, w, N, pa, s);
    vec_gs_update_mdot(w, V, N, pa, j + 1, N, alpha + j, j, V, j + 1, 1, pb, s);
    vec_gs_update_normalize(w, V, N, pb, j + 1, N, V + (int64_t)(j + 1) * N, beta + j, s);
  }
  clk.lap("enqueue the steps");
  QSX_HIP(hipStreamSynchronize(s));
  clk.lap("device drain");
  std::vector<double> al(H.hsmall, H.hsmall + niter), b(H.hsmall + niter, H.hsmall + 2 * niter);
  b.push_back(0.0);
  // breakdown: a beta_j at rounding level means the Krylov space of psi0
  // is invariant after j+1 vectors (psi0 inside an invariant subspace, or
  // niter beyond its dimension); the later vectors are then zero or
  // normalised rounding noise, not Lanczos vectors, so T is cut to the
  // first m = j+1 rows and columns.  Rounding level is relative to |H|: the
  // noise of H v is ~eps |H| even when |H v| is small (psi0 an eigenvector
  // of a small eigenvalue), so the scale is the largest entry of the whole
  // T (every alpha_j and beta_j is bounded by |H|, and the entries after a
  // breakdown, from noise directions, reveal it)
  int m = niter;
  double scale = 0.0;
  for (int j = 0; j < niter; ++j) scale = std::max(scale, std::max(std::fabs(al[j]), std::fabs(b[j])));
  for (int j = 0; j + 1 < niter; ++j)
    if (!(b[j] > 64.0 * 2.220446049250313e-16 * scale)) {
      m = j + 1;
      break;
    }
  // Ritz values of every T_k (T_m beyond a breakdown), lowest eigenpair of T_m
  if (ritz)
    for (int k = 1; k < niter; ++k) ritz[k - 1] = tri_min_eigenvalue(std::min(k, m), al.data(), b.data());
  double lam = 0.0;
  std::vector<double> y;
  tri_lowest(m, al.data(), b.data(), lam, y);
  if (ritz) ritz[niter - 1] = lam;
  if (e0) *e0 = lam;
  clk.lap("tridiagonal eigenproblems");
  if (gs) {
    // the Ritz vector back in the caller's basis: (V y) / sqrt(w)
    std::copy(y.begin(), y.begin() + m, H.hsmall);
    QSX_HIP(hipMemcpyAsync(dy, H.hsmall, sizeof(double) * m, hipMemcpyHostToDevice, s));
    vec_combine(w, V, N, dy, m, N, s, H.dsw);
    eig_download(H, gs, w, s);
  }
  clk.lap("Ritz vector");
}

static void davidson(HpsiPlan& H, const double* psi0, int niter, double* e0, double* ritz, double* gs) {
  hipStream_t s = default_stream();
  const int64_t N = H.N;
  if (niter < 1 || niter > 500) fail(QSX_ERR_ARG, "niter must be in 1..500");
  const int mmax = niter + 1;
  const size_t nparts = (size_t)DOT_BLOCKS * (mmax + 1);
  const size_t n1 = std::max<int64_t>(1, N);
  Carve ws{eig_workspace(
      H, Carve::size({(size_t)N * mmax, (size_t)N * mmax, n1, n1, nparts, nparts, (size_t)mmax + 1}),
      2 * ((size_t)mmax + 1), s)};
  // pinned scalars: [y and theta for the device (mmax + 1) | the projected
  // column and |r| (mmax + 1)]; the kernels that form the column and |r|
  // store them there directly (a few doubles over the link, read by the
  // host after the step's synchronisation); a step's host writes follow the
  // previous step's synchronisation, so the copies never overlap them
  double* hy = H.hsmall;
  double* hback = H.hsmall + mmax + 1;
  double* dback = H.dsmall + mmax + 1;
  double* V = ws.take((size_t)N * mmax);
  double* HV = ws.take((size_t)N * mmax);
  double* u = ws.take(n1);
  double* hu = ws.take(n1);
  double* pa = ws.take(nparts);
  double* pb = ws.take(nparts);
  double* dy = ws.take(mmax + 1);
  eig_upload(H, hu, psi0, s);
  vec_scale_norm_partial(hu, H.dsw, u, N, pa, s);
  vec_normalize_partial(u, V, pa, nullptr, N, s);
  eig_hpsi(H, V, HV, s);
  std::vector<double> Hk(1);
  vec_mdot(V, 0, 1, HV, N, pa, dback, 0, s);
  QSX_HIP(hipStreamSynchronize(s));
  Hk[0] = hback[0];
  int k = 1;
  for (int it = 0; it < niter; ++it) {
    double th = 0.0;
    std::vector<double> y;
    sym_lowest(k, Hk, th, y);
    if (ritz) ritz[it] = th;
    if (e0) *e0 = th;
    // u = V y and the residual r = H u - theta u in one pass (y and theta
    // in one host-to-device copy)
    std::copy(y.begin(), y.begin() + k, hy);
    hy[k] = th;
    QSX_HIP(hipMemcpyAsync(dy, hy, sizeof(double) * (k + 1), hipMemcpyHostToDevice, s));
    if (it + 1 == niter) break;
    vec_combine_res(u, hu, V, HV, N, dy, k, N, s);
    // r orthogonalised twice against V and normalised: classical
    // Gram-Schmidt, the reductions' final sums folded into their consumers,
    // the second pass fused with the normalisation (|r| to ret[mmax])
    vec_mdot_partial(V, N, k, hu, N, pa, s);
    vec_gs_update_mdot(hu, V, N, pa, k, N, nullptr, 0, V, k, 1, pb, s);
    vec_gs_update_normalize(hu, V, N, pb, k, N, V + (int64_t)k * N, dback + mmax, s);
    eig_hpsi(H, V + (int64_t)k * N, HV + (int64_t)k * N, s);
    // grow the projected matrix by one row/column; one host round trip
    // brings the new column and the residual norm
    vec_mdot(V, N, k + 1, HV + (int64_t)k * N, N, pa, dback, 0, s);
    QSX_HIP(hipStreamSynchronize(s));
    const double* back = hback;
    // converged when the orthogonalised residual is at rounding level
    // relative to the projected matrix (its entries are bounded by |H|):
    // the new direction would be noise; keep u = V y
    const double rn = back[mmax];
    double hs = 0.0;
    for (double x : Hk) hs = std::max(hs, std::fabs(x));
    for (int i = 0; i <= k; ++i) hs = std::max(hs, std::fabs(back[i]));
    if (!(rn > 64.0 * 2.220446049250313e-16 * hs)) {
      for (int j = it + 1; j < niter && ritz; ++j) ritz[j] = th;
      break;
    }
    std::vector<double> Hn((size_t)(k + 1) * (k + 1));
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) Hn[(size_t)i * (k + 1) + j] = Hk[(size_t)i * k + j];
    for (int i = 0; i <= k; ++i) {
      Hn[(size_t)i * (k + 1) + k] = back[i];
      Hn[(size_t)k * (k + 1) + i] = back[i];
    }
    Hk.swap(Hn);
    ++k;
  }
  if (gs) {
    // the Ritz vector V y (dy holds the last y) back in the caller's basis
    vec_combine(u, V, N, dy, k, N, s, H.dsw);
    eig_download(H, gs, u, s);
  }
}

}  // namespace qsx

using namespace qsx;

struct qsx_hpsi_plan {
  std::unique_ptr<HpsiPlan> h;
};

static void set_err2(char* err, size_t errlen, const std::string& m) {
  if (err && errlen) snprintf(err, errlen, "%s", m.c_str());
}

extern "C" {

int qsx_hpsi_plan_create(const qsx_tensor* HL, const qsx_tensor* psi, const qsx_tensor* HR, const qsx_opts* opts,
                         qsx_hpsi_plan** plan, char* err, size_t errlen) {
  try {
    if (!plan) fail(QSX_ERR_ARG, "null plan pointer");
    auto* p = new qsx_hpsi_plan;
    p->h = std::make_unique<HpsiPlan>();
    try {
      std::vector<HpsiTerm> terms{{0, 0, 0, 0}};
      if (opts && opts->hconj) terms.push_back({0, 1, 0, 1});  // get_HPsi.m:55-56
      hpsi_create(*p->h, {HL}, {HR}, terms, psi, nullptr, nullptr, opts);
    } catch (...) {
      delete p;
      throw;
    }
    *plan = p;
    return 0;
  } catch (const Error& e) {
    set_err2(err, errlen, e.msg);
    return e.code;
  }
}

int qsx_hpsi_plan_create_pseudo(const qsx_tensor* HKL, const qsx_tensor* HKR, int32_t nop, const qsx_tensor* OPL,
                                const qsx_tensor* OPR, const int32_t* hconj, const qsx_tensor* psi,
                                const qsx_opts* opts, qsx_hpsi_plan** plan, char* err, size_t errlen) {
  try {
    if (!plan) fail(QSX_ERR_ARG, "null plan pointer");
    if (nop < 0 || (nop > 0 && (!OPL || !OPR))) fail(QSX_ERR_ARG, "invalid operator list");
    if (opts && opts->fermA) fail(QSX_ERR_ARG, "pseudo-MPO H*psi: parity signs belong in the operators");
    // get_HPsi.m:62-83: HK_L.psi + psi.HK_R + sum_j {OP_L(j),'*',psi}.OP_R(j)
    // [+ {OP_L(j),psi}.OP_R(j)' when hconj(j)], in that order
    std::vector<const qsx_tensor*> Ls, Rs;
    std::vector<HpsiTerm> terms;
    for (int j = 0; j < nop; ++j) {
      Ls.push_back(OPL + j);
      Rs.push_back(OPR + j);
      terms.push_back({j, 1, j, 0});
      if (hconj && hconj[j]) terms.push_back({j, 0, j, 1});
    }
    auto* p = new qsx_hpsi_plan;
    p->h = std::make_unique<HpsiPlan>();
    try {
      hpsi_create(*p->h, Ls, Rs, terms, psi, HKL, HKR, opts);
    } catch (...) {
      delete p;
      throw;
    }
    *plan = p;
    return 0;
  } catch (const Error& e) {
    set_err2(err, errlen, e.msg);
    return e.code;
  }
}

int64_t qsx_hpsi_size(const qsx_hpsi_plan* plan) { return plan ? plan->h->N : 0; }
double qsx_hpsi_flops(const qsx_hpsi_plan* plan) { return plan ? plan->h->flops : 0; }
double qsx_hpsi_bytes(const qsx_hpsi_plan* plan) { return plan ? plan->h->bytes : 0; }

int qsx_hpsi_apply(qsx_hpsi_plan* plan, const double* dpsi, double* dhpsi, void* stream) {
  if (!plan) return QSX_ERR_ARG;
  try {
    hpsi_apply(*plan->h, dpsi, dhpsi, stream ? (hipStream_t)stream : default_stream());
    return 0;
  } catch (const Error& e) {
    fprintf(stderr, "qsx_hpsi_apply: %s\n", e.msg.c_str());
    return e.code;
  }
}

int qsx_hpsi_stage_info(const qsx_hpsi_plan* plan, double* info) {
  if (!plan || !info) return QSX_ERR_ARG;
  const HpsiPlan& H = *plan->h;
  info[0] = H.stage_flops[0];
  info[1] = H.stage_flops[1];
  info[2] = H.stage_bytes[0];
  info[3] = H.stage_bytes[1];
  info[4] = (double)H.work1.items.size();
  info[5] = (double)H.work2.items.size();
  info[6] = (double)H.pairs1.size();
  info[7] = (double)H.pairs2.size();
  const WorkList* W[2] = {&H.work1, &H.work2};
  for (int t = 0; t < 2; ++t) {
    info[8 + t] = W[t]->engine;
    info[10 + t] = W[t]->engine == ENG_WAVE ? (double)((W[t]->items.size() + W[t]->wg_waves - 1) / W[t]->wg_waves)
                                            : (double)W[t]->queues.size();
    info[12 + t] = W[t]->maxf;
    info[14 + t] = (double)W[t]->reds.size();
    info[16 + t] = H.split_exec_flops[t];
    info[18 + t] = 2.0 * 16 * 16 * 4 * wave_exec_mfma(*W[t], t == 0 ? H.pairs1 : H.pairs2);
    if (H.mode == QSX_MODE_SPLIT_BF16 || H.mode == QSX_MODE_SPLIT_F32) {
      info[10 + t] = (double)((H.swork1.items.size() * (t == 0) + H.swork2.items.size() * (t == 1) + 3) / 4);
      info[8 + t] = 2;
    }
  }
  return 0;
}

int qsx_hpsi_autotune(qsx_hpsi_plan* plan, int32_t reps, double* stage_ms, int32_t* policy) {
  if (!plan) return QSX_ERR_ARG;
  try {
    hpsi_autotune(*plan->h, reps > 0 ? reps : 5, stage_ms);
    if (policy) {
      policy_to_ints(plan->h->pol1, policy);
      policy_to_ints(plan->h->pol2, policy + QSX_POLICY_INTS);
    }
    return 0;
  } catch (const Error& e) {
    fprintf(stderr, "qsx_hpsi_autotune: %s\n", e.msg.c_str());
    return e.code;
  }
}

int qsx_hpsi_set_policy(qsx_hpsi_plan* plan, const int32_t* policy) {
  if (!plan || !policy) return QSX_ERR_ARG;
  try {
    HpsiPlan& H = *plan->h;
    H.pol1 = policy_from_ints(policy);
    H.pol2 = policy_from_ints(policy + QSX_POLICY_INTS);
    hpsi_build_work(H);
    return 0;
  } catch (const Error& e) {
    fprintf(stderr, "qsx_hpsi_set_policy: %s\n", e.msg.c_str());
    return e.code;
  }
}


int qsx_hpsi_weights(const qsx_hpsi_plan* plan, double* w) {
  if (!plan || !w) return QSX_ERR_ARG;
  std::copy(plan->h->welem.begin(), plan->h->welem.end(), w);
  return 0;
}

int qsx_hpsi_use_graph(qsx_hpsi_plan* plan, int32_t on) {
  if (!plan) return QSX_ERR_ARG;
  plan->h->use_graph = on != 0;
  return 0;
}

int qsx_hpsi_last_timing(const qsx_hpsi_plan* plan, double* ms2) {
  if (!plan || !ms2) return QSX_ERR_ARG;
  if (hipEventSynchronize(plan->h->ev[2]) != hipSuccess) return QSX_ERR_DEVICE;
  for (int i = 0; i < 2; ++i) {
    float t = 0;
    if (hipEventElapsedTime(&t, plan->h->ev[i], plan->h->ev[i + 1]) != hipSuccess) return QSX_ERR_DEVICE;
    ms2[i] = t;
  }
  return 0;
}

int qsx_hpsi_stage_ms(qsx_hpsi_plan* plan, int32_t stage, int32_t reps, const double* dpsi, double* dhpsi,
                      double* ms) {
  if (!plan || !ms || (stage != 1 && stage != 2) || reps < 1 || !dpsi || !dhpsi) return QSX_ERR_ARG;
  try {
    HpsiPlan& H = *plan->h;
    if (H.mode != QSX_MODE_BATCHED) fail(QSX_ERR_ARG, "stage timing is for the batched FP64 mode");
    hipStream_t s = default_stream();
    hpsi_launch(H, dpsi, dhpsi, s, false, false);  // T of this psi for stage 2; warm
    auto one = [&]() {
      if (stage == 1)
        launch_worklist(H.work1, H.dq1, H.ditems1, H.dreds1, H.dpairs1, H.dHL, dpsi, H.dT, H.dscratch, s);
      else
        launch_worklist(H.work2, H.dq2, H.ditems2, H.dreds2, H.dpairs2, H.dT, H.dHR, dhpsi, H.dscratch, s, dpsi);
    };
    one();
    QSX_HIP(hipEventRecord(H.ev[0], s));
    for (int r = 0; r < reps; ++r) one();
    QSX_HIP(hipEventRecord(H.ev[1], s));
    QSX_HIP(hipEventSynchronize(H.ev[1]));
    float t = 0;
    QSX_HIP(hipEventElapsedTime(&t, H.ev[0], H.ev[1]));
    *ms = (double)t / reps;
    return 0;
  } catch (const Error& e) {
    fprintf(stderr, "qsx_hpsi_stage_ms: %s\n", e.msg.c_str());
    return e.code;
  }
}

void qsx_hpsi_plan_destroy(qsx_hpsi_plan* plan) { delete plan; }

int qsx_lanczos(qsx_hpsi_plan* plan, const double* psi0, int32_t niter, double* e0, double* ritz, double* gs,
                char* err, size_t errlen) {
  try {
    if (!plan || !psi0) fail(QSX_ERR_ARG, "null argument");
    lanczos(*plan->h, psi0, niter, e0, ritz, gs);
    return 0;
  } catch (const Error& e) {
    set_err2(err, errlen, e.msg);
    return e.code;
  }
}

int qsx_davidson(qsx_hpsi_plan* plan, const double* psi0, int32_t niter, double* e0, double* ritz, double* gs,
                 char* err, size_t errlen) {
  try {
    if (!plan || !psi0) fail(QSX_ERR_ARG, "null argument");
    davidson(*plan->h, psi0, niter, e0, ritz, gs);
    return 0;
  } catch (const Error& e) {
    set_err2(err, errlen, e.msg);
    return e.code;
  }
}

}  // extern "C"
