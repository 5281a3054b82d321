// labels.hip -- device Q-label engine: block matching and output-block
// bookkeeping of QSpace::contract_getIdxSet (Source/QSpace.cc:4131-4294).
//
// Reference semantics restated (bit-exact by construction, integer work):
//  * records compare lexicographically, column 0 most significant, ties by
//    original row index (Wb::hpsort / wbRecs, Source/wbsort.cc:88-143,
//    Source/wbMatrix.hh:1377-1429) -- a stable LSD radix sort gives exactly
//    that order;
//  * matchIndex emits, for each equal contracted key in ascending order, the
//    cross product with A outer and B inner, both in original index order
//    (Source/wbsort.cc:153-271);
//  * QC = [A kept | B kept] per pair (QC.Cat, QSpace.cc:4278), then a stable
//    sort + grouping of the pairs by QC (groupRecs, Source/wbMatrix.cc:1682-1731).
//
// Device design: label columns are range-packed into 128-bit keys (two
// uint64) so each sort is ceil(bits/8) stable 8-bit LSD passes.  The record
// sets are small (1e2..1e6), so every sort runs in ONE workgroup of 1024
// threads (16 waves) over global ping-pong buffers: per pass each wave
// histograms a contiguous chunk into LDS, one scan over (digit, wave) gives
// stable destinations, and each wave scatters its chunk in order using
// ballot-based ranks among lanes with equal digits.  No atomics leave LDS.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include "qsx_internal.hpp"

namespace qsx {
namespace {

constexpr int SORT_THREADS = 1024;
constexpr int SORT_WAVES = SORT_THREADS / 64;

// ---- range packing of label columns into a 128-bit key ------------------
// col_lo[c], col_shift[c], col_in_hi[c]: value (v - lo) << shift placed in
// the hi or lo word.
struct PackSpec {
  int ncols;
  int32_t lo[32];
  int32_t shift[32];
  int32_t in_hi[32];
};

__global__ void k_pack_keys(const int32_t* __restrict__ rec, int n, int m, PackSpec spec, uint64_t* __restrict__ khi,
                            uint64_t* __restrict__ klo, uint32_t* __restrict__ vals) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t h = 0, l = 0;
  const int32_t* r = rec + (size_t)i * m;
  for (int c = 0; c < spec.ncols; ++c) {
    uint64_t v = (uint64_t)(uint32_t)(r[c] - spec.lo[c]) << spec.shift[c];
    if (spec.in_hi[c]) h |= v;
    else l |= v;
  }
  khi[i] = h;
  klo[i] = l;
  vals[i] = (uint32_t)i;
}

__device__ __forceinline__ uint32_t digit_of(uint64_t h, uint64_t l, int bit) {
  // 8-bit digit starting at 'bit' of the 128-bit key (bit < 128, multiple of 8)
  return bit >= 64 ? (uint32_t)((h >> (bit - 64)) & 0xff) : (uint32_t)((l >> bit) & 0xff);
}

// One workgroup, all passes.  Buffers: (h0,l0,v0) input/output, (h1,l1,v1) temp.
__global__ __launch_bounds__(SORT_THREADS) void k_radix_sort_1wg(uint64_t* h0, uint64_t* l0, uint32_t* v0, uint64_t* h1,
                                                                  uint64_t* l1, uint32_t* v1, int n, int npass) {
  __shared__ uint32_t hist[SORT_WAVES][256];
  __shared__ uint32_t wsum[SORT_WAVES];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int chunk = (n + SORT_WAVES - 1) / SORT_WAVES;
  const int c0 = min(n, w * chunk), c1 = min(n, c0 + chunk);
  uint64_t *sh = h0, *sl = l0, *dh = h1, *dl = l1;
  uint32_t *sv = v0, *dv = v1;
  for (int pass = 0; pass < npass; ++pass) {
    const int bit = pass * 8;
    for (int i = tid; i < SORT_WAVES * 256; i += SORT_THREADS) (&hist[0][0])[i] = 0;
    __syncthreads();
    for (int i = c0 + lane; i < c1; i += 64) atomicAdd(&hist[w][digit_of(sh[i], sl[i], bit)], 1u);
    __syncthreads();
    // exclusive scan over (digit-major, wave-minor): 4096 entries, 4 per thread
    {
      uint32_t loc[4];
      uint32_t s = 0;
      for (int j = 0; j < 4; ++j) {
        int e = tid * 4 + j;  // e = digit*16 + wave
        loc[j] = hist[e & 15][e >> 4];
        s += loc[j];
      }
      // block scan of s over 1024 threads: wave scan then wave sums
      uint32_t x = s;
      for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
      }
      if (lane == 63) wsum[w] = x;
      __syncthreads();
      if (w == 0) {
        uint32_t t = lane < SORT_WAVES ? wsum[lane] : 0;
        for (int o = 1; o < SORT_WAVES; o <<= 1) {
          uint32_t y = __shfl_up(t, o, 64);
          if (lane >= o) t += y;
        }
        if (lane < SORT_WAVES) wsum[lane] = t;  // inclusive
      }
      __syncthreads();
      uint32_t excl = x - s + (w ? wsum[w - 1] : 0);
      __syncthreads();
      for (int j = 0; j < 4; ++j) {
        int e = tid * 4 + j;
        hist[e & 15][e >> 4] = excl;
        excl += loc[j];
      }
    }
    __syncthreads();
    // stable scatter: each wave walks its chunk in order, 64 at a time
    for (int base = c0; base < c1; base += 64) {
      int i = base + lane;
      bool valid = i < c1;
      uint64_t kh = valid ? sh[i] : 0, kl = valid ? sl[i] : 0;
      uint32_t kv = valid ? sv[i] : 0;
      uint32_t d = valid ? digit_of(kh, kl, bit) : 0;
      uint64_t peers = __ballot(valid);
      for (int b = 0; b < 8; ++b) {
        uint64_t m = __ballot((d >> b) & 1);
        peers &= ((d >> b) & 1) ? m : ~m;
      }
      uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
      uint32_t rank = __popcll(peers & lt);
      uint32_t cnt = __popcll(peers);
      uint32_t start = valid ? hist[w][d] : 0;
      __builtin_amdgcn_wave_barrier();
      if (valid) {
        uint32_t dst = start + rank;
        dh[dst] = kh;
        dl[dst] = kl;
        dv[dst] = kv;
        if (rank == cnt - 1) hist[w][d] = start + cnt;
      }
      __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // swap roles
    uint64_t* t64 = sh; sh = dh; dh = t64;
    t64 = sl; sl = dl; dl = t64;
    uint32_t* t32 = sv; sv = dv; dv = t32;
    __threadfence_block();
    __syncthreads();
  }
  // result must end in (h0,l0,v0)
  if (npass & 1) {
    for (int i = tid; i < n; i += SORT_THREADS) {
      h0[i] = sh[i];
      l0[i] = sl[i];
      v0[i] = sv[i];
    }
  }
}

// equal-range of A key i in sorted B keys
__device__ __forceinline__ bool key_less(uint64_t ah, uint64_t al, uint64_t bh, uint64_t bl) {
  return ah < bh || (ah == bh && al < bl);
}

__global__ void k_match_count(const uint64_t* ah, const uint64_t* al, int na, const uint64_t* bh, const uint64_t* bl,
                              int nb, uint32_t* lo_out, uint32_t* cnt_out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na) return;
  uint64_t h = ah[i], l = al[i];
  int lo = 0, hi = nb;
  while (lo < hi) {  // lower bound
    int mid = (lo + hi) >> 1;
    if (key_less(bh[mid], bl[mid], h, l)) lo = mid + 1;
    else hi = mid;
  }
  int lb = lo;
  hi = nb;
  while (lo < hi) {  // upper bound
    int mid = (lo + hi) >> 1;
    if (key_less(h, l, bh[mid], bl[mid])) hi = mid;
    else lo = mid + 1;
  }
  lo_out[i] = (uint32_t)lb;
  cnt_out[i] = (uint32_t)(lo - lb);
}

// single-workgroup exclusive scan of n uint32 -> uint32 offsets + total
__global__ __launch_bounds__(1024) void k_scan_1wg(const uint32_t* in, uint32_t* out, int n, uint32_t* total) {
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t carry;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += 1024) {
    int i = base + tid;
    uint32_t v = i < n ? in[i] : 0;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
      uint32_t t = lane < 16 ? wsum[lane] : 0;
      for (int o = 1; o < 16; o <<= 1) {
        uint32_t y = __shfl_up(t, o, 64);
        if (lane >= o) t += y;
      }
      if (lane < 16) wsum[lane] = t;
    }
    __syncthreads();
    uint32_t c = carry;
    uint32_t excl = c + x - v + (w ? wsum[w - 1] : 0);
    if (i < n) out[i] = excl;
    __syncthreads();
    if (tid == 1023) carry = excl + v;
    __syncthreads();
  }
  if (tid == 0) *total = carry;
}

__global__ void k_match_emit(const uint32_t* av, int na, const uint32_t* bv, const uint32_t* lo, const uint32_t* cnt,
                             const uint32_t* off, int32_t* pa, int32_t* pb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na) return;
  uint32_t o = off[i], c = cnt[i], l = lo[i];
  int32_t a = (int32_t)av[i];
  for (uint32_t j = 0; j < c; ++j) {
    pa[o + j] = a;
    pb[o + j] = (int32_t)bv[l + j];
  }
}

// QC records of the pairs: [A kept | B kept]
__global__ void k_gather_qc(const int32_t* pa, const int32_t* pb, int np, const int32_t* QAk, int ka,
                            const int32_t* QBk, int kb, int32_t* qc) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  int32_t* d = qc + (size_t)p * (ka + kb);
  const int32_t* a = QAk + (size_t)pa[p] * ka;
  const int32_t* b = QBk + (size_t)pb[p] * kb;
  for (int c = 0; c < ka; ++c) d[c] = a[c];
  for (int c = 0; c < kb; ++c) d[ka + c] = b[c];
}

// after sorting pairs by QC key: head flags and permuted pair lists
__global__ void k_group_heads(const uint64_t* kh, const uint64_t* kl, const uint32_t* perm, int np, const int32_t* pa,
                              const int32_t* pb, uint32_t* head, int32_t* ga, int32_t* gb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  head[i] = (i == 0 || kh[i] != kh[i - 1] || kl[i] != kl[i - 1]) ? 1u : 0u;
  uint32_t p = perm[i];
  ga[i] = pa[p];
  gb[i] = pb[p];
}

__global__ void k_group_emit(const uint32_t* head, const uint32_t* hoff, const uint32_t* perm, int np,
                             const int32_t* qc, int mc, int64_t* seg, int32_t* qout) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  if (!head[i]) return;
  uint32_t g = hoff[i];
  seg[g] = i;
  const int32_t* s = qc + (size_t)perm[i] * mc;
  for (int c = 0; c < mc; ++c) qout[(size_t)g * mc + c] = s[c];
}

// head flag of each run of equal keys in a sorted key array
__global__ void k_key_heads(const uint64_t* kh, const uint64_t* kl, int n, uint32_t* head) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  head[i] = (i == 0 || kh[i] != kh[i - 1] || kl[i] != kl[i - 1]) ? 1u : 0u;
}

// ---- host helpers ----------------------------------------------------------

// The engine's temporaries (sort buffers, match and group arrays: about 25
// per contraction structure, ~150 per H*psi plan) come from a free list by
// power-of-two size class.  Every use of such a buffer is on the library
// stream inside one call and the host reads results only after a
// synchronisation, so a buffer returned here can be handed out again at
// once: stream order separates its old and new uses.  At most
// kTempPoolBytes stay cached; the rest go back to hipFree.
constexpr size_t kTempPoolBytes = size_t(256) << 20;
struct TempPool {
  std::mutex mu;
  std::multimap<size_t, void*> free;
  size_t held = 0;
};
TempPool& temp_pool() {
  static TempPool* pool = new TempPool;  // lives until the process ends
  return *pool;
}
size_t temp_class(size_t b) {
  size_t c = 4096;
  while (c < b) c <<= 1;
  return c;
}
void* temp_alloc(size_t b) {
  const size_t c = temp_class(b);
  {
    TempPool& P = temp_pool();
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.free.find(c);
    if (it != P.free.end()) {
      void* p = it->second;
      P.free.erase(it);
      P.held -= c;
      return p;
    }
  }
  void* p = nullptr;
  QSX_HIP(hipMalloc(&p, c));
  return p;
}
void temp_release(void* p, size_t b) {
  const size_t c = temp_class(b);
  TempPool& P = temp_pool();
  {
    std::lock_guard<std::mutex> lk(P.mu);
    if (P.held + c <= kTempPoolBytes) {
      P.free.emplace(c, p);
      P.held += c;
      return;
    }
  }
  (void)hipFree(p);
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  explicit DevBuf(size_t b) : bytes(b) {
    if (b) p = temp_alloc(b);
  }
  ~DevBuf() {
    if (p) temp_release(p, bytes);
  }
  DevBuf(const DevBuf&) = delete;
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

// bit layout for the columns of a record set (host side; the ranges are a
// property of the label values, computed once here)
PackSpec make_spec(const std::vector<int32_t>& R, int n, int m, int* nbits) {
  PackSpec s{};
  s.ncols = m;
  if (m > 32) fail(QSX_ERR_UNSUPPORTED, "label record of %d ints exceeds 32", m);
  std::vector<int> w(m, 0);
  for (int c = 0; c < m; ++c) {
    int32_t lo = 0, hi = 0;
    for (int i = 0; i < n; ++i) {
      int32_t v = R[(size_t)i * m + c];
      if (i == 0 || v < lo) lo = v;
      if (i == 0 || v > hi) hi = v;
    }
    s.lo[c] = lo;
    uint64_t span = (uint64_t)((int64_t)hi - (int64_t)lo);
    int b = 0;
    while (b < 63 && (span >> b)) ++b;
    w[c] = b;
  }
  int total = 0;
  for (int c = 0; c < m; ++c) total += w[c];
  if (total > 128) fail(QSX_ERR_UNSUPPORTED, "label records need %d bits (> 128)", total);
  // column 0 most significant: assign from the top of the used range down
  int pos = total;
  for (int c = 0; c < m; ++c) {
    pos -= w[c];
    // a column must not straddle the 64-bit word boundary
    if (pos < 64 && pos + w[c] > 64) fail(QSX_ERR_UNSUPPORTED, "label column straddles key words");
    s.in_hi[c] = pos >= 64;
    s.shift[c] = pos >= 64 ? pos - 64 : pos;
  }
  *nbits = total;
  return s;
}

// realign so no column straddles bit 64: pad the low word
PackSpec make_spec_safe(const std::vector<int32_t>& R, int n, int m, int* nbits) {
  std::vector<int> w(m, 0);
  std::vector<int32_t> lo(m, 0);
  for (int c = 0; c < m; ++c) {
    int32_t l = 0, h = 0;
    for (int i = 0; i < n; ++i) {
      int32_t v = R[(size_t)i * m + c];
      if (i == 0 || v < l) l = v;
      if (i == 0 || v > h) h = v;
    }
    lo[c] = l;
    uint64_t span = (uint64_t)((int64_t)h - (int64_t)l);
    int b = 0;
    while (b < 63 && (span >> b)) ++b;
    w[c] = b;
  }
  if (m > 32) fail(QSX_ERR_UNSUPPORTED, "label record of %d ints exceeds 32", m);
  // fill from least significant (last column) upwards, skipping to bit 64
  // when a column would straddle the word boundary
  std::vector<int> pos(m);
  int p = 0;
  for (int c = m - 1; c >= 0; --c) {
    if (p < 64 && p + w[c] > 64) p = 64;
    pos[c] = p;
    p += w[c];
  }
  if (p > 128) fail(QSX_ERR_UNSUPPORTED, "label records need %d bits (> 128)", p);
  PackSpec s{};
  s.ncols = m;
  for (int c = 0; c < m; ++c) {
    s.lo[c] = lo[c];
    s.in_hi[c] = pos[c] >= 64;
    s.shift[c] = pos[c] >= 64 ? pos[c] - 64 : pos[c];
  }
  *nbits = p;
  return s;
}

struct SortedKeys {
  DevBuf h, l, v, th, tl, tv;
  int n;
  explicit SortedKeys(int n_)
      : h(sizeof(uint64_t) * std::max(1, n_)), l(sizeof(uint64_t) * std::max(1, n_)), v(sizeof(uint32_t) * std::max(1, n_)),
        th(sizeof(uint64_t) * std::max(1, n_)), tl(sizeof(uint64_t) * std::max(1, n_)), tv(sizeof(uint32_t) * std::max(1, n_)),
        n(n_) {}
};

// pack + stable sort records that are already on the device
void pack_and_sort(const int32_t* drec, int n, int m, const PackSpec& spec, int nbits, SortedKeys& K, hipStream_t s) {
  if (n == 0) return;
  int tb = 256;
  hipLaunchKernelGGL(k_pack_keys, dim3((n + tb - 1) / tb), dim3(tb), 0, s, drec, n, m, spec, K.h.as<uint64_t>(),
                     K.l.as<uint64_t>(), K.v.as<uint32_t>());
  int npass = (nbits + 7) / 8;
  if (npass > 0)
    hipLaunchKernelGGL(k_radix_sort_1wg, dim3(1), dim3(SORT_THREADS), 0, s, K.h.as<uint64_t>(), K.l.as<uint64_t>(),
                       K.v.as<uint32_t>(), K.th.as<uint64_t>(), K.tl.as<uint64_t>(), K.tv.as<uint32_t>(), n, npass);
  QSX_HIP(hipGetLastError());
}

}  // namespace

void device_group_records(const std::vector<int32_t>& R, int n, int m, std::vector<int32_t>& order,
                          std::vector<int64_t>& seg, hipStream_t s) {
  order.resize(n);
  seg.clear();
  if (n == 0) {
    seg.push_back(0);
    return;
  }
  if (m == 0) {  // records of no ints are all equal
    for (int i = 0; i < n; ++i) order[i] = i;
    seg = {0, (int64_t)n};
    return;
  }
  int nbits = 0;
  PackSpec spec = make_spec_safe(R, n, m, &nbits);
  DevBuf dR(sizeof(int32_t) * R.size());
  QSX_HIP(hipMemcpyAsync(dR.p, R.data(), sizeof(int32_t) * R.size(), hipMemcpyHostToDevice, s));
  SortedKeys K(n);
  pack_and_sort(dR.as<int32_t>(), n, m, spec, nbits, K, s);
  DevBuf head(sizeof(uint32_t) * n);
  const int tb = 256;
  hipLaunchKernelGGL(k_key_heads, dim3((n + tb - 1) / tb), dim3(tb), 0, s, K.h.as<uint64_t>(), K.l.as<uint64_t>(), n,
                     head.as<uint32_t>());
  QSX_HIP(hipGetLastError());
  std::vector<uint32_t> hv(n), pv(n);
  QSX_HIP(hipMemcpyAsync(hv.data(), head.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
  QSX_HIP(hipMemcpyAsync(pv.data(), K.v.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
  QSX_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < n; ++i) {
    order[i] = (int32_t)pv[i];
    if (hv[i]) seg.push_back(i);
  }
  seg.push_back(n);
}

LabelResult device_label_engine(const std::vector<int32_t>& QAc, const std::vector<int32_t>& QAk, int na,
                                const std::vector<int32_t>& QBc, const std::vector<int32_t>& QBk, int nb, int kc,
                                int ka, int kb, hipStream_t s) {
  LabelResult R;
  R.mc = ka + kb;
  R.seg.assign(1, 0);
  if (na == 0 || nb == 0) return R;
  // a common pack spec for the contracted keys of A and B (same columns)
  std::vector<int32_t> both(QAc);
  both.insert(both.end(), QBc.begin(), QBc.end());
  int nbits_c = 0;
  PackSpec spec_c = make_spec_safe(both, na + nb, kc, &nbits_c);

  DevBuf dQAc(sizeof(int32_t) * std::max<size_t>(1, QAc.size())), dQBc(sizeof(int32_t) * std::max<size_t>(1, QBc.size()));
  DevBuf dQAk(sizeof(int32_t) * std::max<size_t>(1, QAk.size())), dQBk(sizeof(int32_t) * std::max<size_t>(1, QBk.size()));
  QSX_HIP(hipMemcpyAsync(dQAc.p, QAc.data(), sizeof(int32_t) * QAc.size(), hipMemcpyHostToDevice, s));
  QSX_HIP(hipMemcpyAsync(dQBc.p, QBc.data(), sizeof(int32_t) * QBc.size(), hipMemcpyHostToDevice, s));
  if (!QAk.empty()) QSX_HIP(hipMemcpyAsync(dQAk.p, QAk.data(), sizeof(int32_t) * QAk.size(), hipMemcpyHostToDevice, s));
  if (!QBk.empty()) QSX_HIP(hipMemcpyAsync(dQBk.p, QBk.data(), sizeof(int32_t) * QBk.size(), hipMemcpyHostToDevice, s));

  SortedKeys KA(na), KB(nb);
  pack_and_sort(dQAc.as<int32_t>(), na, kc, spec_c, nbits_c, KA, s);
  pack_and_sort(dQBc.as<int32_t>(), nb, kc, spec_c, nbits_c, KB, s);

  DevBuf lo(sizeof(uint32_t) * na), cnt(sizeof(uint32_t) * na), off(sizeof(uint32_t) * na), tot(sizeof(uint32_t));
  int tb = 256;
  hipLaunchKernelGGL(k_match_count, dim3((na + tb - 1) / tb), dim3(tb), 0, s, KA.h.as<uint64_t>(), KA.l.as<uint64_t>(),
                     na, KB.h.as<uint64_t>(), KB.l.as<uint64_t>(), nb, lo.as<uint32_t>(), cnt.as<uint32_t>());
  hipLaunchKernelGGL(k_scan_1wg, dim3(1), dim3(1024), 0, s, cnt.as<uint32_t>(), off.as<uint32_t>(), na,
                     tot.as<uint32_t>());
  QSX_HIP(hipGetLastError());
  uint32_t np = 0;
  QSX_HIP(hipMemcpyAsync(&np, tot.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  QSX_HIP(hipStreamSynchronize(s));
  if (np == 0) return R;
  DevBuf pa(sizeof(int32_t) * np), pb(sizeof(int32_t) * np);
  hipLaunchKernelGGL(k_match_emit, dim3((na + tb - 1) / tb), dim3(tb), 0, s, KA.v.as<uint32_t>(), na,
                     KB.v.as<uint32_t>(), lo.as<uint32_t>(), cnt.as<uint32_t>(), off.as<uint32_t>(), pa.as<int32_t>(),
                     pb.as<int32_t>());
  QSX_HIP(hipGetLastError());
  int mc = ka + kb;
  if (mc == 0) {  // contraction to a scalar: one output block with all pairs
    R.Ia.resize(np);
    R.Ib.resize(np);
    QSX_HIP(hipMemcpyAsync(R.Ia.data(), pa.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost, s));
    QSX_HIP(hipMemcpyAsync(R.Ib.data(), pb.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost, s));
    QSX_HIP(hipStreamSynchronize(s));
    R.seg.push_back(np);
    return R;
  }
  // QC records per pair, packed with one spec over the output columns
  DevBuf qc(sizeof(int32_t) * (size_t)np * mc);
  hipLaunchKernelGGL(k_gather_qc, dim3((np + tb - 1) / tb), dim3(tb), 0, s, pa.as<int32_t>(), pb.as<int32_t>(), (int)np,
                     dQAk.as<int32_t>(), ka, dQBk.as<int32_t>(), kb, qc.as<int32_t>());
  QSX_HIP(hipGetLastError());
  // column ranges of QC follow from those of QAk and QBk
  std::vector<int32_t> rng_rec;  // two synthetic records: column minima and maxima
  {
    std::vector<int32_t> mn(mc), mx(mc);
    for (int c = 0; c < ka; ++c) {
      int32_t l = QAk[c], h = QAk[c];
      for (int i = 0; i < na; ++i) {
        l = std::min(l, QAk[(size_t)i * ka + c]);
        h = std::max(h, QAk[(size_t)i * ka + c]);
      }
      mn[c] = l;
      mx[c] = h;
    }
    for (int c = 0; c < kb; ++c) {
      int32_t l = QBk[c], h = QBk[c];
      for (int i = 0; i < nb; ++i) {
        l = std::min(l, QBk[(size_t)i * kb + c]);
        h = std::max(h, QBk[(size_t)i * kb + c]);
      }
      mn[ka + c] = l;
      mx[ka + c] = h;
    }
    rng_rec = mn;
    rng_rec.insert(rng_rec.end(), mx.begin(), mx.end());
  }
  int nbits_q = 0;
  PackSpec spec_q = make_spec_safe(rng_rec, 2, mc, &nbits_q);
  SortedKeys KQ((int)np);
  pack_and_sort(qc.as<int32_t>(), (int)np, mc, spec_q, nbits_q, KQ, s);
  DevBuf head(sizeof(uint32_t) * np), hoff(sizeof(uint32_t) * np), ga(sizeof(int32_t) * np), gb(sizeof(int32_t) * np);
  hipLaunchKernelGGL(k_group_heads, dim3((np + tb - 1) / tb), dim3(tb), 0, s, KQ.h.as<uint64_t>(), KQ.l.as<uint64_t>(),
                     KQ.v.as<uint32_t>(), (int)np, pa.as<int32_t>(), pb.as<int32_t>(), head.as<uint32_t>(),
                     ga.as<int32_t>(), gb.as<int32_t>());
  hipLaunchKernelGGL(k_scan_1wg, dim3(1), dim3(1024), 0, s, head.as<uint32_t>(), hoff.as<uint32_t>(), (int)np,
                     tot.as<uint32_t>());
  QSX_HIP(hipGetLastError());
  uint32_t nc = 0;
  QSX_HIP(hipMemcpyAsync(&nc, tot.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  QSX_HIP(hipStreamSynchronize(s));
  DevBuf seg(sizeof(int64_t) * (nc + 1)), qout(sizeof(int32_t) * (size_t)nc * mc);
  hipLaunchKernelGGL(k_group_emit, dim3((np + tb - 1) / tb), dim3(tb), 0, s, head.as<uint32_t>(), hoff.as<uint32_t>(),
                     KQ.v.as<uint32_t>(), (int)np, qc.as<int32_t>(), mc, seg.as<int64_t>(), qout.as<int32_t>());
  QSX_HIP(hipGetLastError());
  R.Ia.resize(np);
  R.Ib.resize(np);
  R.seg.resize(nc + 1);
  R.QC.resize((size_t)nc * mc);
  QSX_HIP(hipMemcpyAsync(R.Ia.data(), ga.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost, s));
  QSX_HIP(hipMemcpyAsync(R.Ib.data(), gb.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost, s));
  QSX_HIP(hipMemcpyAsync(R.seg.data(), seg.p, sizeof(int64_t) * nc, hipMemcpyDeviceToHost, s));
  QSX_HIP(hipMemcpyAsync(R.QC.data(), qout.p, sizeof(int32_t) * (size_t)nc * mc, hipMemcpyDeviceToHost, s));
  QSX_HIP(hipStreamSynchronize(s));
  R.seg[nc] = np;
  return R;
}

}  // namespace qsx
