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
// Method (ours): CG coefficients by the Racah sum evaluated in exact integer
// arithmetic (prime-exponent bookkeeping + a small big-integer type), so they
// carry only the final rounding at any spin.  Since C_AB = x3 * C_C for
// OM = 1, x3 is read off at ONE non-zero entry of C_C: T_AB there is a short
// sum over the contracted magnetic quantum numbers (m-conservation prunes
// it to O(min contracted dim) terms) -- no dense CGT is ever formed.

#include <algorithm>
#include <cmath>
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
  if (r > 3) fail(QSX_ERR_UNSUPPORTED, "SU(2) CGT of rank %d (outer multiplicity) not supported", r);
  int nin = 0;
  for (int d : dir) nin += d > 0;
  const bool flip = nin < (r + 1) / 2;
  auto flag = [&](int k) { return (dir[k] > 0 ? 0 : 1) ^ (flip ? 1 : 0); };
  std::vector<int> ord(r);
  for (int k = 0; k < r; ++k) ord[k] = k;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
    if (flag(a) != flag(b)) return flag(a) < flag(b);
    return tj[a] < tj[b];
  });
  if (2 * nin == r) {  // equal groups (rank 2 in/out): rotate if first > second
    int l = r / 2;
    for (int i = 0; i < l; ++i) {
      int a = tj[ord[i]], b = tj[ord[i + l]];
      if (a != b) {
        if (a > b) std::rotate(ord.begin(), ord.begin() + l, ord.end());
        break;
      }
    }
  }
  int fl[3];
  for (int k = 0; k < r; ++k) {
    c.ord[k] = ord[k];
    c.tj[k] = tj[ord[k]];
    fl[k] = flag(ord[k]);
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

}  // namespace qsx
