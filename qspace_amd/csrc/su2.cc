// su2.cc -- host SU(2) coefficient engine: exact Clebsch-Gordan
// coefficients, block CGT conventions and X-symbols (OM = 1).
//
// Reference conventions restated (moondg/qspace):
//  * irrep basis index i = j - m, i.e. index 0 <-> m = +j (CG::get_SU2mat,
//    Source/clebsch_aux.cc:1048-1082);
//  * a block's CGT is the canonical CGT (legs sorted by QSet::Sort,
//    Source/clebsch.cc:1435-1505: majority direction first, then ascending
//    labels, ties stable) normalised to |C|_F = 1 with the first non-zero
//    entry in column-major order positive (cdata::NormSignC,
//    Source/clebsch.cc:1928-1985; column-major = wbsparray::Compress with
//    lex=-1, Source/wbsparray.cc:3485-3520), permuted back to the block's
//    leg order (cgp);
//  * x3 = overlap of the contracted CGTs of A and B with the CGT of C
//    (X3Map::contractCGR / contract_x3, Source/clebsch_aux.cc:4073-4510);
//  * external norm sqrt(d) (CRef::normExt, Source/clebsch.cc:4738-4760).
//
// Outer multiplicity (rank >= 4, one SU(2) symmetry): a block's CGT space
// has dimension OM >= 1 and its data carry a trailing OM dimension (the
// extra data dimension of the reference's QS_FULL_OM blocks, ExpandOM /
// FuseOM / tensorRefs in Source/QSpace_aux.cc:242-345, matrix-valued cgw in
// Source/clebsch_aux.cc:4464-4471).  The basis is this library's own
// convention (the reference takes its OM bases from RC_STORE fusion trees),
// stated in su2_om_basis below; the contraction coefficients are the
// overlaps X[a][b][c] = <C_c | contract(A_a, B_b)> of unit basis tensors.
// For rank <= 3 the basis is the single canonical CGT above.
//
// Method (ours): CG coefficients by the Racah sum evaluated in exact integer
// arithmetic (prime-exponent bookkeeping + a small big-integer type), so they
// carry only the final rounding at any spin.  Since C_AB = x3 * C_C for
// OM = 1, x3 is read off at ONE non-zero entry of C_C: T_AB there is a short
// sum over the contracted magnetic quantum numbers (m-conservation prunes
// it to O(min contracted dim) terms) -- no dense CGT is ever formed.

#include <algorithm>
#include <cmath>
#include <functional>
#include <memory>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <shared_mutex>
#include <unordered_map>
#include <vector>
#include <string>

#include "qsx_internal.hpp"

namespace qsx {
namespace {

// ---- minimal unsigned big integer (base 2^32, little endian) --------------
struct Big {
  std::vector<uint32_t> w;  // empty = 0
  void set(uint64_t v) {
    w.clear();
    while (v) {
      w.push_back((uint32_t)v);
      v >>= 32;
    }
  }
  bool zero() const { return w.empty(); }
  void mul(uint32_t f) {
    uint64_t carry = 0;
    for (auto& x : w) {
      uint64_t t = (uint64_t)x * f + carry;
      x = (uint32_t)t;
      carry = t >> 32;
    }
    if (carry) w.push_back((uint32_t)carry);
  }
  void mulbig(const Big& o) {
    if (zero() || o.zero()) {
      w.clear();
      return;
    }
    std::vector<uint32_t> r(w.size() + o.w.size(), 0);
    for (size_t i = 0; i < w.size(); ++i) {
      uint64_t carry = 0;
      for (size_t j = 0; j < o.w.size(); ++j) {
        uint64_t t = (uint64_t)w[i] * o.w[j] + r[i + j] + carry;
        r[i + j] = (uint32_t)t;
        carry = t >> 32;
      }
      size_t k = i + o.w.size();
      while (carry) {
        uint64_t t = (uint64_t)r[k] + carry;
        r[k] = (uint32_t)t;
        carry = t >> 32;
        ++k;
      }
    }
    while (!r.empty() && !r.back()) r.pop_back();
    w.swap(r);
  }
  static int cmp(const Big& a, const Big& b) {
    if (a.w.size() != b.w.size()) return a.w.size() < b.w.size() ? -1 : 1;
    for (size_t i = a.w.size(); i-- > 0;)
      if (a.w[i] != b.w[i]) return a.w[i] < b.w[i] ? -1 : 1;
    return 0;
  }
  void add(const Big& o) {
    if (o.w.size() > w.size()) w.resize(o.w.size(), 0);
    uint64_t carry = 0;
    for (size_t i = 0; i < w.size(); ++i) {
      uint64_t t = (uint64_t)w[i] + (i < o.w.size() ? o.w[i] : 0) + carry;
      w[i] = (uint32_t)t;
      carry = t >> 32;
    }
    if (carry) w.push_back((uint32_t)carry);
  }
  void sub(const Big& o) {  // requires *this >= o
    int64_t borrow = 0;
    for (size_t i = 0; i < w.size(); ++i) {
      int64_t t = (int64_t)w[i] - (i < o.w.size() ? o.w[i] : 0) - borrow;
      borrow = t < 0;
      if (t < 0) t += ((int64_t)1 << 32);
      w[i] = (uint32_t)t;
    }
    while (!w.empty() && !w.back()) w.pop_back();
  }
  // value = mant * 2^exp with mant in [0.5,1) as long double
  void frexp_ld(long double& mant, long& e) const {
    if (zero()) {
      mant = 0;
      e = 0;
      return;
    }
    size_t n = w.size();
    long double v = 0;
    int take = (int)std::min<size_t>(3, n);
    for (int i = 0; i < take; ++i) v = v * 4294967296.0L + (long double)w[n - 1 - i];
    int ex;
    mant = frexpl(v, &ex);
    e = ex + 32L * (long)(n - take);
  }
};

// signed big integer accumulator
struct SBig {
  Big mag;
  bool neg = false;
  void addsigned(const Big& v, bool vneg) {
    if (v.zero()) return;
    if (mag.zero()) {
      mag = v;
      neg = vneg;
      return;
    }
    if (neg == vneg) {
      mag.add(v);
      return;
    }
    int c = Big::cmp(mag, v);
    if (c >= 0) {
      mag.sub(v);
      if (mag.zero()) neg = false;
    } else {
      Big t = v;
      t.sub(mag);
      mag = t;
      neg = vneg;
    }
  }
};

std::vector<int> g_primes;
std::mutex g_prime_mu;

void ensure_primes(int n) {
  std::lock_guard<std::mutex> lk(g_prime_mu);
  int have = g_primes.empty() ? 1 : g_primes.back();
  if (have >= n) return;
  int lim = std::max(n, 2 * have) + 16;
  std::vector<char> sieve(lim + 1, 1);
  g_primes.clear();
  for (int i = 2; i <= lim; ++i) {
    if (!sieve[i]) continue;
    g_primes.push_back(i);
    for (long j = (long)i * i; j <= lim; j += i) sieve[j] = 0;
  }
}

// exponent of prime p in n! (Legendre)
inline int fact_exp(int n, int p) {
  int e = 0;
  while (n) {
    n /= p;
    e += n;
  }
  return e;
}

void add_fact(std::vector<int>& ex, int n, int sign, int np) {
  for (int i = 0; i < np && g_primes[i] <= n; ++i) ex[i] += sign * fact_exp(n, g_primes[i]);
}

Big prime_product(const std::vector<int>& ex, int np, bool positive_part) {
  Big b;
  b.set(1);
  uint64_t acc = 1;
  for (int i = 0; i < np; ++i) {
    int e = positive_part ? ex[i] : -ex[i];
    for (int k = 0; k < e; ++k) {
      uint64_t p = (uint64_t)g_primes[i];
      if (acc * p >= (1ull << 32)) {
        b.mul((uint32_t)acc);
        acc = 1;
      }
      acc *= p;
    }
  }
  if (acc > 1) b.mul((uint32_t)acc);
  return b;
}

}  // namespace

// <j1 m1; j2 m2 | j3 m3>, two-j/two-m integer arguments, Condon-Shortley phase
double su2_cg_exact(int tj1, int tm1, int tj2, int tm2, int tj3, int tm3) {
  if (tj1 < 0 || tj2 < 0 || tj3 < 0) return 0;
  if (std::abs(tm1) > tj1 || std::abs(tm2) > tj2 || std::abs(tm3) > tj3) return 0;
  if (((tj1 + tm1) & 1) || ((tj2 + tm2) & 1) || ((tj3 + tm3) & 1)) return 0;
  if (tm1 + tm2 != tm3) return 0;
  if (tj3 > tj1 + tj2 || tj3 < std::abs(tj1 - tj2) || ((tj1 + tj2 + tj3) & 1)) return 0;
  const int a1 = (tj3 + tj1 - tj2) / 2, a2 = (tj3 - tj1 + tj2) / 2, a3 = (tj1 + tj2 - tj3) / 2;
  const int a4 = (tj1 + tj2 + tj3) / 2 + 1;
  const int b1 = (tj3 + tm3) / 2, b2 = (tj3 - tm3) / 2, b3 = (tj1 - tm1) / 2, b4 = (tj1 + tm1) / 2;
  const int b5 = (tj2 - tm2) / 2, b6 = (tj2 + tm2) / 2;
  const int c1 = (tj3 - tj2 + tm1) / 2, c2 = (tj3 - tj1 - tm2) / 2;
  const int kmin = std::max({0, -c1, -c2}), kmax = std::min({a3, b3, b6});
  if (kmin > kmax) return 0;
  ensure_primes(a4 + tj3 + 2);
  int np = 0;
  while (np < (int)g_primes.size() && g_primes[np] <= std::max(a4, tj3 + 1)) ++np;
  // denominators D_k = k! (a3-k)! (b3-k)! (b6-k)! (c1+k)! (c2+k)!
  const int nk = kmax - kmin + 1;
  std::vector<std::vector<int>> ek(nk, std::vector<int>(np, 0));
  std::vector<int> E(np, 0);
  for (int t = 0; t < nk; ++t) {
    int k = kmin + t;
    add_fact(ek[t], k, 1, np);
    add_fact(ek[t], a3 - k, 1, np);
    add_fact(ek[t], b3 - k, 1, np);
    add_fact(ek[t], b6 - k, 1, np);
    add_fact(ek[t], c1 + k, 1, np);
    add_fact(ek[t], c2 + k, 1, np);
    for (int i = 0; i < np; ++i) E[i] = std::max(E[i], ek[t][i]);
  }
  SBig S;
  std::vector<int> diff(np);
  for (int t = 0; t < nk; ++t) {
    for (int i = 0; i < np; ++i) diff[i] = E[i] - ek[t][i];
    Big term = prime_product(diff, np, true);
    S.addsigned(term, ((kmin + t) & 1) != 0);
  }
  if (S.mag.zero()) return 0;
  // CG^2 = (tj3+1) a1! a2! a3! / a4! * b1!..b6! * S^2 / prod p^(2E)
  std::vector<int> g(np, 0);
  {
    int v = tj3 + 1;
    for (int i = 0; i < np && v > 1; ++i)
      while (v % g_primes[i] == 0) {
        v /= g_primes[i];
        g[i] += 1;
      }
  }
  add_fact(g, a1, 1, np);
  add_fact(g, a2, 1, np);
  add_fact(g, a3, 1, np);
  add_fact(g, a4, -1, np);
  add_fact(g, b1, 1, np);
  add_fact(g, b2, 1, np);
  add_fact(g, b3, 1, np);
  add_fact(g, b4, 1, np);
  add_fact(g, b5, 1, np);
  add_fact(g, b6, 1, np);
  for (int i = 0; i < np; ++i) g[i] -= 2 * E[i];
  Big num = prime_product(g, np, true);  // positive exponents only
  std::vector<int> gneg(np);
  for (int i = 0; i < np; ++i) gneg[i] = g[i] < 0 ? g[i] : 0;
  for (int i = 0; i < np; ++i) g[i] = g[i] > 0 ? g[i] : 0;
  num = prime_product(g, np, true);
  Big den = prime_product(gneg, np, false);
  Big s2 = S.mag;
  s2.mulbig(S.mag);
  num.mulbig(s2);
  long double mn, md;
  long en, ed;
  num.frexp_ld(mn, en);
  den.frexp_ld(md, ed);
  long double ratio = ldexpl(mn / md, (int)(en - ed));
  double v = (double)sqrtl(ratio);
  return S.neg ? -v : v;
}

namespace {

// canonical description of one block CGT
enum Kind { K_SCALAR = 0, K_ID2 = 1, K_SING2 = 2, K_CG3 = 3, K_SING3 = 4 };

struct CanonCGT {
  int r = 0;
  int ord[3] = {0, 0, 0};  // canonical position k -> block leg ord[k]
  int tj[3] = {0, 0, 0};   // canonical labels
  int kind = K_SCALAR;
  double scale = 1.0;      // sigma / norm
};

// raw (unnormalised) canonical entry at canonical two-m values
double raw_entry(const CanonCGT& c, const int* tm) {
  switch (c.kind) {
    case K_SCALAR: return 1.0;
    case K_ID2: return tm[0] == tm[1] ? 1.0 : 0.0;
    case K_SING2:
      if (tm[0] != -tm[1]) return 0.0;
      return (((c.tj[0] - tm[0]) / 2) & 1) ? -1.0 : 1.0;
    case K_CG3: return su2_cg_exact(c.tj[0], tm[0], c.tj[1], tm[1], c.tj[2], tm[2]);
    case K_SING3: {
      double v = su2_cg_exact(c.tj[0], tm[0], c.tj[1], tm[1], c.tj[2], -tm[2]);
      return (((c.tj[2] - tm[2]) / 2) & 1) ? -v : v;
    }
  }
  return 0.0;
}

// m-selection rule of the canonical CGT (cheap reject before raw_entry)
inline bool allowed(const CanonCGT& c, const int* tm) {
  switch (c.kind) {
    case K_SCALAR: return true;
    case K_ID2: return tm[0] == tm[1];
    case K_SING2: return tm[0] == -tm[1];
    case K_CG3: return tm[0] + tm[1] == tm[2];
    case K_SING3: return tm[0] + tm[1] + tm[2] == 0;
  }
  return false;
}

// canonical leg order (QSet::Sort): legs of the majority direction first
// (flag 0; incoming on a tie), each group by ascending label with ties kept
// in leg order; equal groups are swapped when the first differing label of
// the first group is the larger.  ord[k] = block leg at canonical position
// k, fl[k] its flag.
void canon_order(const std::vector<int>& tj, const std::vector<int>& dir, std::vector<int>& ord,
                 std::vector<int>& fl) {
  const int r = (int)tj.size();
  int nin = 0;
  for (int d : dir) nin += d > 0;
  const bool flip = nin < (r + 1) / 2;
  auto flag = [&](int k) { return (dir[k] > 0 ? 0 : 1) ^ (flip ? 1 : 0); };
  ord.resize(r);
  for (int k = 0; k < r; ++k) ord[k] = k;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
    if (flag(a) != flag(b)) return flag(a) < flag(b);
    return tj[a] < tj[b];
  });
  if (2 * nin == r) {
    int l = r / 2;
    for (int i = 0; i < l; ++i) {
      int a = tj[ord[i]], b = tj[ord[i + l]];
      if (a != b) {
        if (a > b) std::rotate(ord.begin(), ord.begin() + l, ord.end());
        break;
      }
    }
  }
  fl.resize(r);
  for (int k = 0; k < r; ++k) fl[k] = flag(ord[k]);
}

CanonCGT make_canon(const std::vector<int>& tj, const std::vector<int>& dir) {
  CanonCGT c;
  int r = (int)tj.size();
  c.r = r;
  if (r == 0) return c;
  if (r == 1) {
    if (tj[0] != 0) fail(QSX_ERR_CG, "rank-1 SU(2) block with 2S=%d", tj[0]);
    c.kind = K_SCALAR;
    return c;
  }
  if (r > 3) fail(QSX_ERR_UNSUPPORTED, "SU(2) CGT of rank %d has an outer multiplicity basis (su2_om_basis)", r);
  std::vector<int> ord, flv;
  canon_order(tj, dir, ord, flv);
  int fl[3];
  for (int k = 0; k < r; ++k) {
    c.ord[k] = ord[k];
    c.tj[k] = tj[ord[k]];
    fl[k] = flv[k];
  }
  double nrm;
  if (r == 2) {
    if (c.tj[0] != c.tj[1]) fail(QSX_ERR_CG, "rank-2 SU(2) block with 2S=(%d,%d)", c.tj[0], c.tj[1]);
    c.kind = fl[0] != fl[1] ? K_ID2 : K_SING2;
    nrm = std::sqrt((double)(c.tj[0] + 1));
  } else {
    if (fl[0] == fl[1] && fl[1] != fl[2]) c.kind = K_CG3;
    else if (fl[0] == fl[1] && fl[1] == fl[2]) c.kind = K_SING3;
    else fail(QSX_ERR_CG, "unexpected canonical direction pattern");
    int a = c.tj[0], b = c.tj[1], d = c.tj[2];
    if (d > a + b || d < std::abs(a - b) || ((a + b + d) & 1))
      fail(QSX_ERR_CG, "SU(2) block labels (%d,%d,%d) violate the triangle rule", a, b, d);
    nrm = std::sqrt((double)(d + 1));
  }
  // sign of the first entry (column-major over canonical indices) whose
  // magnitude exceeds 1e-14 of the norm
  double sgn = 0;
  const double thr = 1e-14 * nrm;
  if (c.kind == K_ID2) sgn = 1;
  else if (c.kind == K_SING2) sgn = (c.tj[0] & 1) ? -1 : 1;  // entry (m=-j, m'=+j)
  else {
    const int d0 = c.tj[0] + 1, d1 = c.tj[1] + 1, d2 = c.tj[2] + 1;
    for (int i2 = 0; i2 < d2 && !sgn; ++i2)
      for (int i1 = 0; i1 < d1 && !sgn; ++i1) {
        int tm[3];
        tm[1] = c.tj[1] - 2 * i1;
        tm[2] = c.tj[2] - 2 * i2;
        tm[0] = c.kind == K_CG3 ? tm[2] - tm[1] : -tm[2] - tm[1];
        if (std::abs(tm[0]) > c.tj[0]) continue;
        (void)d0;
        double v = raw_entry(c, tm);
        if (std::fabs(v) > thr) sgn = v < 0 ? -1 : 1;
      }
  }
  if (!sgn) fail(QSX_ERR_CG, "vanishing SU(2) CGT");
  c.scale = sgn / nrm;
  return c;
}

// block entry at block-order two-m values
inline double block_entry(const CanonCGT& c, const int* tm_block) {
  int tm[3];
  for (int k = 0; k < c.r; ++k) tm[k] = tm_block[c.ord[k]];
  if (!allowed(c, tm)) return 0.0;
  return c.scale * raw_entry(c, tm);
}

struct VecHash {
  size_t operator()(const std::vector<int>& v) const {
    uint64_t h = 1469598103934665603ull;
    for (int x : v) {
      h ^= (uint64_t)(uint32_t)x;
      h *= 1099511628211ull;
    }
    return (size_t)h;
  }
};

std::unordered_map<std::vector<int>, double, VecHash> g_x3;
std::shared_mutex g_x3_mu;  // lookups shared, inserts exclusive

double x3_compute(const std::vector<int>& tja, const std::vector<int>& dira, const std::vector<int>& tjb,
                  const std::vector<int>& dirb, const std::vector<int>& ia, const std::vector<int>& ib) {
  const int ra = (int)tja.size(), rb = (int)tjb.size(), k = (int)ia.size();
  std::vector<char> ca(ra, 0), cb(rb, 0);
  for (int i = 0; i < k; ++i) {
    ca[ia[i]] = 1;
    cb[ib[i]] = 1;
    if (tja[ia[i]] != tjb[ib[i]]) fail(QSX_ERR_CG, "contracted SU(2) labels differ");
    if (dira[ia[i]] == dirb[ib[i]]) fail(QSX_ERR_ITAG, "contracted SU(2) legs must have opposite directions");
  }
  std::vector<int> ka, kb, tjc, dirc;
  for (int i = 0; i < ra; ++i)
    if (!ca[i]) { ka.push_back(i); tjc.push_back(tja[i]); dirc.push_back(dira[i]); }
  for (int i = 0; i < rb; ++i)
    if (!cb[i]) { kb.push_back(i); tjc.push_back(tjb[i]); dirc.push_back(dirb[i]); }
  const int rc = (int)tjc.size();
  // triangle / rank-2 checks of the output labels: a forbidden result
  // CGT means x3 = 0 (CR_CTR_ZERO)
  if (rc == 2 && tjc[0] != tjc[1]) return 0.0;
  if (rc == 1 && tjc[0] != 0) return 0.0;
  if (rc == 3) {
    int a = tjc[0], b = tjc[1], d = tjc[2];
    if (d > a + b || d < std::abs(a - b) || ((a + b + d) & 1)) return 0.0;
  }
  CanonCGT CA = make_canon(tja, dira), CB = make_canon(tjb, dirb), CC = make_canon(tjc, dirc);
  // candidate output entries in column-major order with |T_C| clearly non-zero
  std::vector<int> dc(rc);
  size_t Nc = 1;
  for (int i = 0; i < rc; ++i) {
    dc[i] = tjc[i] + 1;
    Nc *= dc[i];
  }
  std::vector<int> tmA(ra), tmB(rb), tmC(std::max(rc, 1));
  std::vector<int> dk(k);
  size_t Nk = 1;
  for (int i = 0; i < k; ++i) {
    dk[i] = tja[ia[i]] + 1;
    Nk *= dk[i];
  }
  auto tab_at = [&](void) {
    // T_AB at the output two-m values in tmC: sum over contracted m's
    for (size_t q = 0; q < ka.size(); ++q) tmA[ka[q]] = tmC[q];
    for (size_t q = 0; q < kb.size(); ++q) tmB[kb[q]] = tmC[ka.size() + q];
    double acc = 0;
    for (size_t lin = 0; lin < Nk; ++lin) {
      size_t rem = lin;
      for (int i = 0; i < k; ++i) {
        int mi = (int)(rem % dk[i]);
        rem /= dk[i];
        int tm = tja[ia[i]] - 2 * mi;
        tmA[ia[i]] = tm;
        tmB[ib[i]] = tm;
      }
      double a = block_entry(CA, tmA.data());
      if (a == 0.0) continue;
      double b = block_entry(CB, tmB.data());
      if (b == 0.0) continue;
      acc += a * b;
    }
    return acc;
  };
  double x = 0, xs = 0;
  int found = 0;
  for (size_t lin = 0; lin < Nc && found < 2; ++lin) {
    size_t rem = lin;
    for (int i = 0; i < rc; ++i) {
      tmC[i] = tjc[i] - 2 * (int)(rem % dc[i]);
      rem /= dc[i];
    }
    double tc = block_entry(CC, tmC.data());
    if (std::fabs(tc) < 1e-3 / std::sqrt((double)Nc)) continue;
    double tab = tab_at();
    double v = tab / tc;
    if (!found) x = v;
    else xs = v;
    ++found;
  }
  if (!found) fail(QSX_ERR_CG, "no usable entry of the result CGT");
  if (found == 2 && std::fabs(x - xs) > 1e-9 * std::max(1.0, std::fabs(x)))
    fail(QSX_ERR_UNSUPPORTED, "contracted CGT is not proportional to a single result CGT (OM>1)");
  if (std::fabs(x) < 1e-14) x = 0;
  return x;
}

}  // namespace

// su2_cg_exact asks for primes up to (tj1+tj2+tj3)/2 + tj3 + 3 <= 2.5 max + 3;
// the table then ends at the largest prime <= its limit, which is above half
// the limit (Bertrand): a limit of 6 max + 64 covers every such request
void su2_reserve_labels(int max_tj) { ensure_primes(6 * std::max(0, max_tj) + 64); }

double su2_x3(const std::vector<int>& tja, const std::vector<int>& dira, const std::vector<int>& tjb,
              const std::vector<int>& dirb, const std::vector<int>& ia, const std::vector<int>& ib) {
  std::vector<int> key;
  key.reserve(4 + 2 * (tja.size() + tjb.size() + ia.size()));
  key.push_back((int)tja.size());
  key.insert(key.end(), tja.begin(), tja.end());
  key.insert(key.end(), dira.begin(), dira.end());
  key.push_back((int)tjb.size());
  key.insert(key.end(), tjb.begin(), tjb.end());
  key.insert(key.end(), dirb.begin(), dirb.end());
  key.push_back((int)ia.size());
  key.insert(key.end(), ia.begin(), ia.end());
  key.insert(key.end(), ib.begin(), ib.end());
  {
    std::shared_lock<std::shared_mutex> lk(g_x3_mu);
    auto it = g_x3.find(key);
    if (it != g_x3.end()) return it->second;
  }
  double x = x3_compute(tja, dira, tjb, dirb, ia, ib);
  std::unique_lock<std::shared_mutex> lk(g_x3_mu);
  g_x3.emplace(std::move(key), x);
  return x;
}

// A block's CGT is the canonical CGT of its leg set in the block's leg
// order; the canonical order breaks label ties by leg position, so after a
// leg permutation the CGT implied by the new leg order is the transposed
// CGT up to a sign (-1 when two legs of equal label and direction swap
// order and the coupling is odd).  Returns that sign: the ratio of the new
// implied CGT to the transposed old one at one non-zero entry.
// the sign depends only on (labels, directions, permutation): memoised,
// since a tensor's blocks repeat few label sets (4645 blocks of the D* =
// 32000 environment have a few hundred)
static std::mutex g_psign_mu;
static std::unordered_map<std::string, double> g_psign;

static double perm_sign_eval(const std::vector<int>& tj, const std::vector<int>& dir, const std::vector<int>& perm);

double su2_perm_sign(const std::vector<int>& tj, const std::vector<int>& dir, const std::vector<int>& perm) {
  const int r = (int)tj.size();
  if (r < 2) return 1.0;
  std::string key;
  key.reserve(3 * r * sizeof(int));
  key.append((const char*)tj.data(), sizeof(int) * r);
  key.append((const char*)dir.data(), sizeof(int) * r);
  key.append((const char*)perm.data(), sizeof(int) * r);
  {
    std::lock_guard<std::mutex> lk(g_psign_mu);
    auto it = g_psign.find(key);
    if (it != g_psign.end()) return it->second;
  }
  const double v = perm_sign_eval(tj, dir, perm);
  std::lock_guard<std::mutex> lk(g_psign_mu);
  g_psign.emplace(std::move(key), v);
  return v;
}

static double perm_sign_eval(const std::vector<int>& tj, const std::vector<int>& dir, const std::vector<int>& perm) {
  const int r = (int)tj.size();
  std::vector<int> tj2(r), dir2(r);
  for (int k = 0; k < r; ++k) {
    tj2[k] = tj[perm[k]];
    dir2[k] = dir[perm[k]];
  }
  const CanonCGT a = make_canon(tj, dir), b = make_canon(tj2, dir2);
  int tm[3] = {0, 0, 0}, tm2[3] = {0, 0, 0};
  const int n0 = tj[0] + 1, n1 = tj[1] + 1, n2 = r > 2 ? tj[2] + 1 : 1;
  for (int i2 = 0; i2 < n2; ++i2)
    for (int i1 = 0; i1 < n1; ++i1)
      for (int i0 = 0; i0 < n0; ++i0) {
        tm[0] = tj[0] - 2 * i0;
        tm[1] = tj[1] - 2 * i1;
        if (r > 2) tm[2] = tj[2] - 2 * i2;
        const double v = block_entry(a, tm);
        if (std::fabs(v) < 1e-12) continue;
        for (int k = 0; k < r; ++k) tm2[k] = tm[perm[k]];
        const double ratio = block_entry(b, tm2) / v;
        if (std::fabs(std::fabs(ratio) - 1.0) > 1e-9)
          fail(QSX_ERR_CG, "permuted SU(2) CGT is not the transposed one up to sign (ratio %g)", ratio);
        return ratio < 0 ? -1.0 : 1.0;
      }
  fail(QSX_ERR_CG, "vanishing SU(2) CGT");
}

double su2_norm_ext(const std::vector<int>& tj, const std::vector<int>& dir) {
  int r = (int)tj.size();
  if (r == 0 || r > 3) return 1.0;
  bool scalar = true;
  for (int t : tj) scalar &= (t == 0);
  if (scalar) return 1.0;
  if (r == 3) {
    int nin = 0;
    for (int d : dir) nin += d > 0;
    if (nin == 0 || nin == 3) return 1.0;
    int odd = -1;
    for (int k = 0; k < 3; ++k)
      if ((nin == 2 && dir[k] < 0) || (nin == 1 && dir[k] > 0)) odd = k;
    return std::sqrt((double)(tj[odd] + 1));
  }
  int d = 0;
  for (int t : tj) d = std::max(d, t + 1);
  return std::sqrt((double)d);
}


// ---------------------------------------------------------------- outer multiplicity
// Basis of the SU(2) CGT space of one block (labels tj, directions dir in
// block leg order), this library's convention:
//  * canonical leg order and flags as for rank <= 3 (canon_order); group 0 =
//    the leading canonical legs of the first flag (n0 legs), group 1 = the
//    rest (n1 legs);
//  * n1 > 0: each group is coupled left to right by CG coefficients,
//    J_1 = j_1, (J_{i-1} M_{i-1}) x (j_i m_i) -> (J_i M_i), and the two
//    groups' totals are identified (J, M) = (J', M'):
//      C[m] = prod_i CG_0(...) * prod_i CG_1(...);
//  * n1 = 0 (one direction): legs 1..r-1 of the group are coupled as above
//    to (J, M) = (j_r, -m_r) with the factor (-1)^(j_r - m_r);
//  * one basis tensor per intermediate path (J_2..J_{n0-1}, J'_2..J'_{n1-1})
//    with matching totals, ordered lexicographically (group 0's path first,
//    each intermediate ascending); each normalised to |C|_F = 1 with its
//    first entry above 1e-14 of the norm, in column-major order over the
//    canonical legs, positive.
// For rank <= 3 this is the single canonical CGT (block_entry) and
// intermediate paths do not occur.
namespace {

struct OmKey {
  std::vector<int> v;
  bool operator==(const OmKey& o) const { return v == o.v; }
};
struct OmKeyHash {
  size_t operator()(const OmKey& k) const { return VecHash()(k.v); }
};

std::unordered_map<OmKey, std::shared_ptr<const SU2Basis>, OmKeyHash> g_om;
std::shared_mutex g_om_mu;
std::unordered_map<OmKey, std::shared_ptr<const SU2XTensor>, OmKeyHash> g_xom;
std::shared_mutex g_xom_mu;

// couplings of one group: every m-tuple of its legs with its total M and
// coefficient, for one intermediate path (pathJ[i] = J after leg i)
struct ChainEntry {
  std::vector<int> tm;  // two-m per group leg
  int TM;
  double v;
};

void chain_paths(const std::vector<int>& tjs, std::vector<std::vector<int>>& out) {
  out.clear();
  const int n = (int)tjs.size();
  if (n == 0) return;
  std::vector<int> path(n);
  path[0] = tjs[0];
  std::function<void(int)> rec = [&](int i) {
    if (i == n) {
      out.push_back(path);
      return;
    }
    const int a = path[i - 1], b = tjs[i];
    for (int J = std::abs(a - b); J <= a + b; J += 2) {
      path[i] = J;
      rec(i + 1);
    }
  };
  rec(1);
}

void chain_entries(const std::vector<int>& tjs, const std::vector<int>& path, std::vector<ChainEntry>& out) {
  out.clear();
  const int n = (int)tjs.size();
  std::vector<int> tm(n);
  std::function<void(int, int, double)> rec = [&](int i, int M, double v) {
    if (i == n) {
      out.push_back(ChainEntry{tm, M, v});
      return;
    }
    for (int m = -tjs[i]; m <= tjs[i]; m += 2) {
      tm[i] = m;
      if (i == 0) {
        rec(1, m, 1.0);
        continue;
      }
      const int M2 = M + m;
      if (std::abs(M2) > path[i]) continue;
      const double c = su2_cg_exact(path[i - 1], M, tjs[i], m, path[i], M2);
      if (c == 0.0) continue;
      rec(i + 1, M2, v * c);
    }
  };
  rec(0, 0, 1.0);
}

std::shared_ptr<const SU2Basis> om_basis_compute(const std::vector<int>& tj, const std::vector<int>& dir) {
  auto B = std::make_shared<SU2Basis>();
  const int r = (int)tj.size();
  B->r = r;
  B->tj = tj;
  B->dims.resize(r);
  int64_t tot = 1;
  for (int k = 0; k < r; ++k) {
    if (tj[k] < 0) fail(QSX_ERR_CG, "negative SU(2) label");
    B->dims[k] = tj[k] + 1;
    tot *= tj[k] + 1;
  }
  B->numel = tot;
  if (r == 0) {
    B->om = 1;
    B->ent.push_back({{0, 1.0}});
    return B;
  }
  std::vector<int> ord, fl;
  canon_order(tj, dir, ord, fl);
  int n0 = 1;
  while (n0 < r && fl[n0] == fl[0]) ++n0;
  const int n1 = r - n0;
  // legs of the chains (canonical positions) and the singlet leg
  std::vector<int> g0, g1;
  int last = -1;
  if (n1 > 0) {
    for (int k = 0; k < n0; ++k) g0.push_back(k);
    for (int k = n0; k < r; ++k) g1.push_back(k);
  } else {
    for (int k = 0; k + 1 < r; ++k) g0.push_back(k);
    last = r - 1;
  }
  auto labels = [&](const std::vector<int>& g) {
    std::vector<int> v;
    for (int k : g) v.push_back(tj[ord[k]]);
    return v;
  };
  const std::vector<int> t0 = labels(g0), t1 = labels(g1);
  std::vector<std::vector<int>> P0, P1;
  chain_paths(t0, P0);
  if (n1 > 0) chain_paths(t1, P1);
  else if (r == 1) P0.assign(1, std::vector<int>{tj[0]});
  // canonical column-major strides, block strides
  std::vector<int64_t> cst(r), bst(r);
  {
    int64_t a = 1, b = 1;
    for (int k = 0; k < r; ++k) {
      cst[k] = a;
      a *= tj[ord[k]] + 1;
      bst[k] = b;
      b *= tj[k] + 1;
    }
  }
  std::vector<ChainEntry> E0, E1;
  auto emit = [&](const std::vector<std::pair<std::vector<int>, double>>& raw) {
    // raw: two-m per canonical position -> value
    double nrm2 = 0;
    for (auto& e : raw) nrm2 += e.second * e.second;
    const double nrm = std::sqrt(nrm2);
    if (!(nrm > 0)) return;
    int64_t first = -1;
    double sgn = 1;
    for (auto& e : raw) {
      if (std::fabs(e.second) <= 1e-14 * nrm) continue;
      int64_t cl = 0;
      for (int k = 0; k < r; ++k) cl += cst[k] * ((tj[ord[k]] - e.first[k]) / 2);
      if (first < 0 || cl < first) {
        first = cl;
        sgn = e.second < 0 ? -1 : 1;
      }
    }
    std::vector<std::pair<int64_t, double>> ent;
    ent.reserve(raw.size());
    for (auto& e : raw) {
      int64_t bl = 0;
      for (int k = 0; k < r; ++k) bl += bst[ord[k]] * ((tj[ord[k]] - e.first[k]) / 2);
      ent.push_back({bl, sgn * e.second / nrm});
    }
    std::sort(ent.begin(), ent.end());
    B->ent.push_back(std::move(ent));
  };
  if (r == 1) {
    if (tj[0] == 0) B->ent.push_back({{0, 1.0}});
    B->om = (int)B->ent.size();
    return B;
  }
  std::vector<std::pair<std::vector<int>, double>> raw;
  for (const auto& p0 : P0) {
    if (n1 == 0) {
      if (p0.back() != tj[ord[last]]) continue;
      chain_entries(t0, p0, E0);
      raw.clear();
      for (const auto& e : E0) {
        std::vector<int> tm(r);
        for (int i = 0; i < (int)g0.size(); ++i) tm[g0[i]] = e.tm[i];
        tm[last] = -e.TM;
        const double ph = (((tj[ord[last]] - tm[last]) / 2) & 1) ? -1.0 : 1.0;
        raw.push_back({tm, ph * e.v});
      }
      emit(raw);
      continue;
    }
    for (const auto& p1 : P1) {
      if (p1.back() != p0.back()) continue;
      chain_entries(t0, p0, E0);
      chain_entries(t1, p1, E1);
      std::unordered_map<int, std::vector<int>> byM;
      for (int i = 0; i < (int)E1.size(); ++i) byM[E1[i].TM].push_back(i);
      raw.clear();
      for (const auto& e : E0) {
        auto it = byM.find(e.TM);
        if (it == byM.end()) continue;
        for (int i1 : it->second) {
          std::vector<int> tm(r);
          for (int i = 0; i < n0; ++i) tm[g0[i]] = e.tm[i];
          for (int i = 0; i < n1; ++i) tm[g1[i]] = E1[i1].tm[i];
          raw.push_back({tm, e.v * E1[i1].v});
        }
      }
      emit(raw);
    }
  }
  B->om = (int)B->ent.size();
  return B;
}

}  // namespace

std::shared_ptr<const SU2Basis> su2_om_basis(const std::vector<int>& tj, const std::vector<int>& dir) {
  OmKey key;
  key.v = tj;
  key.v.insert(key.v.end(), dir.begin(), dir.end());
  {
    std::shared_lock<std::shared_mutex> lk(g_om_mu);
    auto it = g_om.find(key);
    if (it != g_om.end()) return it->second;
  }
  int mx = 0;
  for (int t : tj) mx += t;
  su2_reserve_labels(mx);
  auto B = om_basis_compute(tj, dir);
  std::unique_lock<std::shared_mutex> lk(g_om_mu);
  return g_om.emplace(std::move(key), B).first->second;
}

int su2_om(const std::vector<int>& tj, const std::vector<int>& dir) { return su2_om_basis(tj, dir)->om; }

std::vector<int> su2_canon_order(const std::vector<int>& tj, const std::vector<int>& dir) {
  std::vector<int> ord, fl;
  canon_order(tj, dir, ord, fl);
  return ord;
}

// X[a][b][c] = <C_c | P(contract(A_a, B_b))>: A's legs ia contracted with
// B's legs ib, the kept legs [A | B] permuted by perm (new leg l = old leg
// perm[l]), C the basis of the permuted output
std::shared_ptr<const SU2XTensor> su2_xom(const std::vector<int>& tja, const std::vector<int>& dira,
                                          const std::vector<int>& tjb, const std::vector<int>& dirb,
                                          const std::vector<int>& ia, const std::vector<int>& ib,
                                          const std::vector<int>& perm) {
  OmKey key;
  auto put = [&](const std::vector<int>& v) {
    key.v.push_back((int)v.size());
    key.v.insert(key.v.end(), v.begin(), v.end());
  };
  put(tja);
  put(dira);
  put(tjb);
  put(dirb);
  put(ia);
  put(ib);
  put(perm);
  {
    std::shared_lock<std::shared_mutex> lk(g_xom_mu);
    auto it = g_xom.find(key);
    if (it != g_xom.end()) return it->second;
  }
  const int ra = (int)tja.size(), rb = (int)tjb.size(), k = (int)ia.size();
  std::vector<char> ca(ra, 0), cb(rb, 0);
  for (int i = 0; i < k; ++i) {
    ca[ia[i]] = 1;
    cb[ib[i]] = 1;
    if (tja[ia[i]] != tjb[ib[i]]) fail(QSX_ERR_CG, "contracted SU(2) labels differ");
    if (dira[ia[i]] == dirb[ib[i]]) fail(QSX_ERR_ITAG, "contracted SU(2) legs must have opposite directions");
  }
  std::vector<int> ka, kb, tj0, dir0;
  for (int i = 0; i < ra; ++i)
    if (!ca[i]) { ka.push_back(i); tj0.push_back(tja[i]); dir0.push_back(dira[i]); }
  for (int i = 0; i < rb; ++i)
    if (!cb[i]) { kb.push_back(i); tj0.push_back(tjb[i]); dir0.push_back(dirb[i]); }
  const int rc = (int)tj0.size();
  if ((int)perm.size() != rc) fail(QSX_ERR_ARG, "su2_xom: permutation length");
  std::vector<int> tjc(rc), dirc(rc);
  for (int l = 0; l < rc; ++l) {
    tjc[l] = tj0[perm[l]];
    dirc[l] = dir0[perm[l]];
  }
  auto A = su2_om_basis(tja, dira), B = su2_om_basis(tjb, dirb), Cb = su2_om_basis(tjc, dirc);
  auto X = std::make_shared<SU2XTensor>();
  X->oa = A->om;
  X->ob = B->om;
  X->oc = Cb->om;
  X->x.assign((size_t)X->oa * X->ob * X->oc, 0.0);
  if (X->oc > 0 && X->oa > 0 && X->ob > 0) {
    if (Cb->numel > ((int64_t)1 << 26)) fail(QSX_ERR_UNSUPPORTED, "SU(2) output CGT of %lld entries", (long long)Cb->numel);
    // per-leg digits of a block-order linear index
    auto digits = [](int64_t lin, const std::vector<int>& d, int* out) {
      for (size_t q = 0; q < d.size(); ++q) {
        out[q] = (int)(lin % d[q]);
        lin /= d[q];
      }
    };
    // output strides in the permuted order: position of old kept leg o
    std::vector<int64_t> cst(rc);
    {
      std::vector<int64_t> s(rc);
      int64_t a = 1;
      for (int l = 0; l < rc; ++l) {
        s[l] = a;
        a *= Cb->dims[l];
      }
      for (int l = 0; l < rc; ++l) cst[perm[l]] = s[l];
    }
    // B entries bucketed by the contracted digits
    std::vector<std::unordered_map<int64_t, std::vector<std::pair<int64_t, double>>>> Bk(X->ob);
    std::vector<int> dg(std::max(ra, rb) + 1);
    for (int b = 0; b < X->ob; ++b)
      for (const auto& e : B->ent[b]) {
        digits(e.first, B->dims, dg.data());
        int64_t key2 = 0, kept = 0;
        for (int i = 0; i < k; ++i) key2 = key2 * (tjb[ib[i]] + 1) + dg[ib[i]];
        for (size_t q = 0; q < kb.size(); ++q) kept += cst[ka.size() + q] * dg[kb[q]];
        Bk[b][key2].push_back({kept, e.second});
      }
    std::vector<double> T((size_t)Cb->numel);
    for (int a = 0; a < X->oa; ++a)
      for (int b = 0; b < X->ob; ++b) {
        std::fill(T.begin(), T.end(), 0.0);
        bool any = false;
        for (const auto& e : A->ent[a]) {
          digits(e.first, A->dims, dg.data());
          int64_t key2 = 0, kept = 0;
          for (int i = 0; i < k; ++i) key2 = key2 * (tja[ia[i]] + 1) + dg[ia[i]];
          auto it = Bk[b].find(key2);
          if (it == Bk[b].end()) continue;
          for (size_t q = 0; q < ka.size(); ++q) kept += cst[q] * dg[ka[q]];
          for (const auto& f : it->second) {
            T[(size_t)(kept + f.first)] += e.second * f.second;
            any = true;
          }
        }
        if (!any) continue;
        for (int c = 0; c < X->oc; ++c) {
          double acc = 0;
          for (const auto& e : Cb->ent[c]) acc += e.second * T[(size_t)e.first];
          if (std::fabs(acc) < 1e-14) acc = 0;
          X->x[((size_t)a * X->ob + b) * X->oc + c] = acc;
        }
      }
  }
  std::unique_lock<std::shared_mutex> lk(g_xom_mu);
  return g_xom.emplace(std::move(key), X).first->second;
}

}  // namespace qsx
