// worklist.cc -- host builders of the work items of the batched data half.
//
// Two engines consume them:
//  * wave engine (build_wave_work, the default): one wave per tile of at
//    most 4x4 MFMA fragments, grouped or split for long k-ranges, workgroups
//    binned per XCD by operand locality and ordered longest first;
//  * workgroup engine (the rest of build_work): per output block, pick the
//    workgroup tile shape that minimises padded work plus a per-tile
//    overhead, cut the block into tiles and each tile's k-range (the
//    concatenated k-steps of the block's pairs, in the reference pair order)
//    into chunks of at most split_steps steps, then deal the items to the
//    queues of a persistent grid (one queue per resident workgroup, one tile
//    shape per queue, longest-processing-time first).
// The autotuner picks the engine and its parameters per stage.
#include <algorithm>
#include <cmath>
#include <map>

#include "qsx_internal.hpp"

namespace qsx {

static constexpr int TKS = 16;  // k-step of the workgroup-tile kernel
static constexpr int WKS = 4;   // k-step of the wave-tile kernel (one MFMA)

// Wave tiles: per output block, ceil(M/16) x ceil(N/16) fragments cut into
// tiles of at most wave_max x wave_max fragments (fragments spread evenly
// over the tiles).  A tile's k-range is the concatenation of its pairs'
// k-steps (of 4, reference pair order).  A tile with more than group_mfma
// MFMAs is shared by the G waves of a workgroup (G = group_waves: 4, 8 or 16,
// with G x fragments <= 16; contiguous k-ranges, partials
// summed through LDS); one with more than split_mfma MFMAs per workgroup is
// further cut into chunks whose partial tiles the reduction sums.  Other
// tiles run one per wave, G to a workgroup.  Workgroups are ordered
// longest first: the hardware issues them in order to free slots, which is
// list scheduling with the longest-processing-time rule.
static WorkList build_wave_work(const std::vector<WorkTarget>& targets, const std::vector<PairDesc>& pairs,
                                const WorkPolicy& pol) {
  WorkList out;
  out.engine = ENG_WAVE;
  const int maxf = std::max(1, std::min(4, pol.wave_max));
  out.maxf = maxf;
  // waves per workgroup: the group size; a group's partial tiles live in LDS,
  // at most 16 fragments per workgroup (the kernel's budget)
  int G = maxf <= 2 ? pol.group_waves : 4;
  while (G > 4 && G * maxf * maxf > 16) G /= 2;
  out.wg_waves = G;
  struct WG {
    double cost;
    int64_t key;  // locality key: the larger operand's block of the target's first pair
    std::vector<WorkItem> it;
  };
  struct Solo {
    double cost;
    int64_t key;
    WorkItem w;
  };
  std::vector<WG> groups;   // workgroups of 4 tasks of one tile
  std::vector<Solo> solo;   // one-wave items
  int64_t slot = 0;
  for (const WorkTarget& t : targets) {
    const int M = t.M, N = t.N;
    if (M <= 0 || N <= 0) continue;
    int64_t steps = 0;
    for (int64_t p = t.p0; p < t.p1; ++p) steps += (pairs[p].K + WKS - 1) / WKS;
    const int mf = (M + 15) / 16, nf = (N + 15) / 16;
    const int ntm = (mf + maxf - 1) / maxf, ntn = (nf + maxf - 1) / maxf;
    int64_t key = -1;
    if (t.p1 > t.p0) {
      const PairDesc& d = pairs[t.p0];
      const int64_t ea = (int64_t)M * d.K, eb = (int64_t)d.K * N;
      key = eb >= ea ? (d.b_off << 1) | 1 : (d.a_off << 1);
    }
    // item over the global k-steps [s0, s1) of the tile
    auto range = [&](WorkItem w, int64_t s0, int64_t s1) {
      if (s0 >= s1) {
        w.p0 = 0;
        w.p1 = -1;
        return w;
      }
      auto locate = [&](int64_t s, int& p, int& k) {
        for (int64_t q = t.p0; q < t.p1; ++q) {
          const int64_t n = (pairs[q].K + WKS - 1) / WKS;
          if (s < n) {
            p = (int)q;
            k = (int)(s * WKS);
            return;
          }
          s -= n;
        }
        p = (int)t.p1;
        k = 0;
      };
      int kb;
      locate(s0, w.p0, w.k0);
      locate(s1 - 1, w.p1, kb);
      w.k1 = std::min(kb + WKS, pairs[w.p1].K);
      return w;
    };
    for (int ti = 0, m0f = 0; ti < ntm; ++ti) {
      const int fm = mf / ntm + (ti < mf % ntm ? 1 : 0);
      for (int tj = 0, n0f = 0; tj < ntn; ++tj) {
        const int fn = nf / ntn + (tj < nf % ntn ? 1 : 0);
        WorkItem base{};
        base.c_off = t.c_off;
        base.M = M;
        base.N = N;
        base.m0 = 16 * m0f;
        base.n0 = 16 * n0f;
        base.shape = wave_shape(fm, fn);
        base.slot = -1;
        base.group = 0;
        const double per_step = (double)fm * fn;
        const double cost = per_step * (double)steps;
        const bool grp = fm * fn * G <= 16 && cost > pol.group_mfma && steps >= G;
        const int ways = grp ? G : 1;
        const double per_wg = std::max<double>(1.0, pol.split_mfma);
        const int nchunk = (int)std::max<double>(1.0, std::ceil(cost / ways / per_wg));
        const int64_t nch = std::min<int64_t>(nchunk, std::max<int64_t>(1, steps / ways));
        for (int64_t q = 0; q < nch; ++q) {
          const int64_t c0 = steps * q / nch, c1 = steps * (q + 1) / nch;
          WorkItem w = base;
          if (nch > 1) w.slot = (int32_t)(slot + q);
          if (grp) {
            WG g;
            g.key = key;
            g.cost = per_step * (double)((c1 - c0 + G - 1) / G) + 16.0;
            for (int v = 0; v < G; ++v) {
              WorkItem x = range(w, c0 + (c1 - c0) * v / G, c0 + (c1 - c0) * (v + 1) / G);
              x.group = 1;
              g.it.push_back(x);
            }
            groups.push_back(std::move(g));
          } else {
            solo.push_back({per_step * (double)(c1 - c0) + 8.0, key, range(w, c0, c1)});
          }
        }
        if (nch > 1) {
          ReduceItem r{};
          r.c_off = t.c_off;
          r.M = M;
          r.N = N;
          r.m0 = base.m0;
          r.n0 = base.n0;
          r.shape = base.shape;
          r.slot0 = (int32_t)slot;
          r.nslot = (int32_t)nch;
          out.reds.push_back(r);
          slot += nch;
        }
        n0f += fn;
      }
      m0f += fm;
    }
  }
  // items of at most two ring chunks run the kernel's two-deep ring
  auto mark_short = [&](WorkItem& w) {
    if (w.p0 > w.p1 || w.slot == -2) return;
    int64_t chunks = 0;
    const int fm = (w.shape - SH_WAVE) / 4 + 1, fn = (w.shape - SH_WAVE) % 4 + 1;
    const int ck = 4 * wave_chunk_steps(fm, fn);
    for (int p = w.p0; p <= w.p1; ++p) {
      const int kb = p == w.p0 ? w.k0 : 0, ke = p == w.p1 ? w.k1 : pairs[p].K;
      chunks += (ke - kb + ck - 1) / ck;
    }
    if (chunks <= 2) w.pad |= kShortItem;
  };
  for (Solo& x : solo) mark_short(x.w);
  for (WG& g : groups)
    for (WorkItem& w : g.it) mark_short(w);
  auto pad_item = []() {
    WorkItem pad{};
    pad.slot = -2;
    pad.p0 = 0;
    pad.p1 = -1;
    pad.shape = wave_shape(1, 1);
    return pad;
  };
  auto by_cost = [](const WG& x, const WG& y) { return x.cost > y.cost; };
  // one bin's workgroups: its group workgroups plus its solo items packed G
  // to a workgroup longest first (only the bin's last workgroup is padded;
  // locality lives in the L2 of the bin's XCD, so a workgroup may mix keys)
  auto pack = [&](std::vector<WG>& wgs, std::vector<Solo>& so) {
    std::stable_sort(so.begin(), so.end(), [](const Solo& x, const Solo& y) { return x.cost > y.cost; });
    for (size_t i = 0; i < so.size(); i += G) {
      WG g;
      g.cost = so[i].cost;
      g.key = so[i].key;
      for (size_t j = i; j < i + G; ++j) g.it.push_back(j < so.size() ? so[j].w : pad_item());
      wgs.push_back(std::move(g));
    }
    std::stable_sort(wgs.begin(), wgs.end(), by_cost);
  };
  const int nbins = pol.xcd_bins;
  if (nbins <= 1) {
    pack(groups, solo);
    for (const WG& g : groups)
      for (const WorkItem& w : g.it) out.items.push_back(w);
  } else {
    // XCD-aware order: workgroup i runs on XCD i % nbins (round-robin
    // dispatch), each XCD with its own L2.  The work of one locality key
    // goes to one XCD so its operand blocks are fetched into one L2; keys
    // are dealt to XCDs longest first to the least loaded (a key heavier
    // than an XCD's share is spread item by item); each XCD's list is
    // longest first and the lists are interleaved (short lists padded with
    // empty workgroups).
    struct Key {
      double cost = 0;
      std::vector<WG> wgs;
      std::vector<Solo> so;
    };
    std::map<int64_t, Key> bykey;
    double total = 0;
    for (WG& g : groups) {
      total += g.cost;
      Key& k = bykey[g.key];
      k.cost += g.cost;
      k.wgs.push_back(std::move(g));
    }
    for (Solo& x : solo) {
      const double c = x.cost / G;  // a solo item is one wave of a workgroup
      total += c;
      Key& k = bykey[x.key];
      k.cost += c;
      k.so.push_back(x);
    }
    std::vector<Key*> keys;
    for (auto& kv : bykey) keys.push_back(&kv.second);
    std::stable_sort(keys.begin(), keys.end(), [](const Key* x, const Key* y) { return x->cost > y->cost; });
    std::vector<std::vector<WG>> bins(nbins);
    std::vector<std::vector<Solo>> bsolo(nbins);
    std::vector<double> load(nbins, 0.0);
    const double share = total / nbins;
    auto least = [&]() { return (int)(std::min_element(load.begin(), load.end()) - load.begin()); };
    for (Key* k : keys) {
      if (k->cost > share) {
        std::stable_sort(k->wgs.begin(), k->wgs.end(), by_cost);
        for (WG& g : k->wgs) {
          const int b = least();
          load[b] += g.cost;
          bins[b].push_back(std::move(g));
        }
        std::stable_sort(k->so.begin(), k->so.end(), [](const Solo& x, const Solo& y) { return x.cost > y.cost; });
        for (Solo& x : k->so) {
          const int b = least();
          load[b] += x.cost / G;
          bsolo[b].push_back(x);
        }
      } else {
        const int b = least();
        load[b] += k->cost;
        for (WG& g : k->wgs) bins[b].push_back(std::move(g));
        for (Solo& x : k->so) bsolo[b].push_back(x);
      }
    }
    size_t depth = 0;
    for (int b = 0; b < nbins; ++b) {
      pack(bins[b], bsolo[b]);
      depth = std::max(depth, bins[b].size());
    }
    for (size_t r = 0; r < depth; ++r)
      for (int b = 0; b < nbins; ++b) {
        if (r < bins[b].size()) {
          for (const WorkItem& w : bins[b][r].it) out.items.push_back(w);
        } else {
          for (int v = 0; v < G; ++v) out.items.push_back(pad_item());
        }
      }
  }
  for (WorkItem& w : out.items) {
    if (w.slot == -2 || w.p0 > w.p1) continue;
    const PairDesc& d = pairs[w.p0];
    w.fa_off = d.a_off;
    w.fb_off = d.b_off;
    w.fK = d.K;
    w.flda = d.lda;
    w.fldb = d.ldb;
    w.fflags = d.flags;
  }
  out.scratch_elems = slot * 64 * 64;
  return out;
}

// the wave kernel addresses operand elements with 32-bit byte offsets inside
// a block
static bool wave_fits(const std::vector<WorkTarget>& targets, const std::vector<PairDesc>& pairs) {
  for (const WorkTarget& t : targets)
    for (int64_t p = t.p0; p < t.p1; ++p) {
      const PairDesc& d = pairs[p];
      const int64_t ea = (int64_t)d.lda * ((d.flags & 1) ? t.M : d.K);
      const int64_t eb = (int64_t)d.ldb * ((d.flags & 2) ? d.K : t.N);
      if (std::max(ea, eb) >= (int64_t(1) << 26)) return false;
    }
  return true;
}

WorkList build_work(const std::vector<WorkTarget>& targets, const std::vector<PairDesc>& pairs, const WorkPolicy& pol) {
  if (pol.engine == ENG_WAVE && wave_fits(targets, pairs)) return build_wave_work(targets, pairs, pol);
  struct Group {
    std::vector<WorkItem> items;
    double cost;  // per item
  };
  WorkList out;
  std::vector<Group> groups;
  int64_t slot = 0;
  for (const WorkTarget& t : targets) {
    const int M = t.M, N = t.N;
    if (M <= 0 || N <= 0) continue;
    int64_t steps = 0;
    for (int64_t p = t.p0; p < t.p1; ++p) steps += (pairs[p].K + TKS - 1) / TKS;
    int best = SH_64x64;
    double bestc = 1e300;
    for (int sh = 0; sh < SH_NUM; ++sh) {
      if (!(pol.shape_mask & (1 << sh))) continue;
      int tm, tn;
      shape_dims(sh, tm, tn);
      double nt = std::ceil((double)M / tm) * std::ceil((double)N / tn);
      double c = nt * ((double)tm * tn * std::max<int64_t>(1, steps) + pol.tile_overhead);
      if (c < bestc) {
        bestc = c;
        best = sh;
      }
    }
    int tm, tn;
    shape_dims(best, tm, tn);
    const int nchunk = steps <= pol.split_steps ? 1 : (int)((steps + pol.split_steps - 1) / pol.split_steps);
    std::vector<int64_t> bounds(nchunk + 1);
    for (int q = 0; q <= nchunk; ++q) bounds[q] = steps * q / nchunk;
    auto locate = [&](int64_t s, int& p, int& k) {  // global step -> (pair, k)
      for (int64_t q = t.p0; q < t.p1; ++q) {
        int64_t n = (pairs[q].K + TKS - 1) / TKS;
        if (s < n) {
          p = (int)q;
          k = (int)(s * TKS);
          return;
        }
        s -= n;
      }
      p = (int)t.p1;
      k = 0;
    };
    for (int m0 = 0; m0 < M; m0 += tm)
      for (int n0 = 0; n0 < N; n0 += tn) {
        WorkItem base{};
        base.c_off = t.c_off;
        base.M = M;
        base.N = N;
        base.m0 = m0;
        base.n0 = n0;
        base.shape = best;
        base.slot = -1;
        const double tile_cost = (double)tm * tn;
        Group g;
        if (steps == 0) {  // no pairs: the tile stores zeros
          base.p0 = 0;
          base.p1 = -1;
          g.items.push_back(base);
          g.cost = 1.0;
        } else if (nchunk == 1) {
          base.p0 = (int)t.p0;
          base.k0 = 0;
          base.p1 = (int)t.p1 - 1;
          base.k1 = pairs[t.p1 - 1].K;
          g.items.push_back(base);
          g.cost = tile_cost * steps;
        } else {
          for (int q = 0; q < nchunk; ++q) {
            WorkItem w = base;
            int pa, ka, pb, kb;
            locate(bounds[q], pa, ka);
            locate(bounds[q + 1] - 1, pb, kb);  // last step of the chunk
            w.p0 = pa;
            w.k0 = ka;
            w.p1 = pb;
            w.k1 = std::min(kb + TKS, pairs[pb].K);
            w.slot = (int32_t)(slot + q);
            g.items.push_back(w);
          }
          ReduceItem r{};
          r.c_off = t.c_off;
          r.M = M;
          r.N = N;
          r.m0 = m0;
          r.n0 = n0;
          r.shape = best;
          r.slot0 = (int32_t)slot;
          r.nslot = nchunk;
          out.reds.push_back(r);
          g.cost = tile_cost * (double)steps / nchunk;
          slot += nchunk;
        }
        groups.push_back(std::move(g));
      }
  }
  out.scratch_elems = slot * 64 * 64;  // uniform slot of 64x64 doubles (any shape fits)
  // persistent queues: each shape gets a share of the resident workgroups in
  // proportion to its work; within a shape, items are dealt longest first to
  // the least loaded queue (LPT), so the queues finish close together
  std::vector<std::vector<std::pair<WorkItem, double>>> byshape(SH_NUM);
  for (auto& g : groups)
    for (auto& it : g.items) byshape[it.shape].push_back({it, g.cost + pol.tile_overhead + 2.0 * 64 * 64});
  double total = 0;
  std::vector<double> scost(SH_NUM, 0.0);
  for (int s = 0; s < SH_NUM; ++s)
    for (auto& x : byshape[s]) scost[s] += x.second;
  for (double c : scost) total += c;
  const int slots = std::max(1, pol.slots);
  std::vector<int> nq(SH_NUM, 0);
  int used = 0;
  for (int s = 0; s < SH_NUM; ++s) {
    if (byshape[s].empty()) continue;
    int n = (int)std::floor(slots * scost[s] / std::max(total, 1e-300));
    n = std::max(1, std::min<int>(n, (int)byshape[s].size()));
    nq[s] = n;
    used += n;
  }
  // hand out the rounding remainder to the shapes with the longest queues
  while (used < slots) {
    int bs = -1;
    double bl = 0;
    for (int s = 0; s < SH_NUM; ++s)
      if (nq[s] && nq[s] < (int)byshape[s].size() && scost[s] / nq[s] > bl) {
        bl = scost[s] / nq[s];
        bs = s;
      }
    if (bs < 0) break;
    ++nq[bs];
    ++used;
  }
  for (int s = 0; s < SH_NUM; ++s) {
    if (!nq[s]) continue;
    auto& v = byshape[s];
    std::stable_sort(v.begin(), v.end(), [](const auto& x, const auto& y) { return x.second > y.second; });
    std::vector<std::vector<WorkItem>> qs(nq[s]);
    std::vector<double> load(nq[s], 0.0);
    for (auto& x : v) {
      int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      qs[b].push_back(x.first);
      load[b] += x.second;
    }
    for (auto& qv : qs) {
      if (qv.empty()) continue;
      WorkQueue q{};
      q.i0 = (int32_t)out.items.size();
      for (auto& it : qv) out.items.push_back(it);
      q.i1 = (int32_t)out.items.size();
      q.shape = s;
      out.queues.push_back(q);
    }
  }
  return out;
}

double wave_exec_mfma(const WorkList& W, const std::vector<PairDesc>& pairs) {
  if (W.engine != ENG_WAVE) return 0.0;
  double n = 0.0;
  for (const WorkItem& w : W.items) {
    if (w.slot == -2 || w.p0 > w.p1) continue;
    int tm, tn;
    shape_dims(w.shape, tm, tn);
    int64_t steps = 0;
    for (int p = w.p0; p <= w.p1; ++p) {
      const int kb = p == w.p0 ? w.k0 : 0, ke = p == w.p1 ? w.k1 : pairs[p].K;
      if (ke > kb) steps += (ke - kb + 3) / 4;
    }
    n += (double)steps * (tm / 16) * (tn / 16);
  }
  return n;
}

}  // namespace qsx
