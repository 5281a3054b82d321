// qsops.hpp -- helpers shared by the host sides of the whole-QSpace
// operations (qsops.cc: permuteQS / plusQS; ortho.cc: orthoQS).
#pragma once

#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "contract.hpp"

namespace qsx {

inline uint64_t tag_name(uint64_t t) { return t & ~(uint64_t)0x80; }
inline bool tag_out(uint64_t t) { return (t & 0x80) != 0; }

inline std::vector<int> leg_dirs(const std::vector<uint64_t>& tags, int r) {
  std::vector<int> dir(r, 1);
  for (int i = 0; i < r && !tags.empty(); ++i) dir[i] = tag_out(tags[i]) ? -1 : 1;
  return dir;
}

inline std::vector<double> packed_host(const TensorDesc& T) {
  std::vector<double> buf(std::max<int64_t>(1, T.total()));
  if (T.total() && !T.hdata) fail(QSX_ERR_ARG, "tensor without data");
  for (int i = 0; i < T.n; ++i) {
    if (!T.numel(i)) continue;
    if (!T.hdata[i]) fail(QSX_ERR_ARG, "null data block %d", i + 1);
    memcpy(buf.data() + T.offs[i], T.hdata[i], sizeof(double) * T.numel(i));
  }
  return buf;
}

inline qsx_tensor_out* alloc_out(int n, int r, int q, int nsym, const std::string& qtype, int64_t total) {
  auto* o = (qsx_tensor_out*)calloc(1, sizeof(qsx_tensor_out));
  o->nblk = n;
  o->rank = r;
  o->qdim = q;
  o->nsym = nsym;
  snprintf(o->qtype, sizeof o->qtype, "%s", qtype.c_str());
  o->Q = (int32_t*)calloc(std::max<size_t>(1, (size_t)n * r * q), sizeof(int32_t));
  o->dims = (int64_t*)calloc(std::max<size_t>(1, (size_t)n * r), sizeof(int64_t));
  o->offs = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  o->cgw = (double*)calloc(std::max<size_t>(1, (size_t)n * nsym), sizeof(double));
  o->itags = (uint64_t*)calloc(std::max(1, r), sizeof(uint64_t));
  o->data = (double*)calloc(std::max<int64_t>(1, total), sizeof(double));
  return o;
}

// events around the device work of one operation
struct Timer {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  explicit Timer(hipStream_t s) {
    QSX_HIP(hipEventCreate(&e0));
    QSX_HIP(hipEventCreate(&e1));
    QSX_HIP(hipEventRecord(e0, s));
  }
  float stop(hipStream_t s) {
    QSX_HIP(hipEventRecord(e1, s));
    QSX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    QSX_HIP(hipEventElapsedTime(&ms, e0, e1));
    return ms;
  }
  ~Timer() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

struct DevMem {
  double* p = nullptr;
  explicit DevMem(int64_t n) { p = (double*)dev_alloc(sizeof(double) * std::max<int64_t>(1, n)); }
  ~DevMem() { dev_free(p); }
};

inline void set_err(char* err, size_t errlen, const std::string& m) {
  if (err && errlen) snprintf(err, errlen, "%s", m.c_str());
}

template <class F>
int guarded(char* err, size_t errlen, F&& f) {
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

}  // namespace qsx
