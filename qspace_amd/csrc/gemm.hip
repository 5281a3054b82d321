// gemm.hip -- data half of the contraction on gfx950.
//
// What it replaces: the per-output-block loop of contractDATA_group
// (Source/QSpace_aux.cc:27-355) -- for each pair (ia,ib) of output block ic
// a dgemm into a temporary Ci (wbarray::contract, Source/wbarray.cc:3386-3500,
// DZGEMM Source/wbarray_blas.cc:121-160) and Ck += afac*Ci (wbarray::Plus,
// Source/wbarray.cc:252-285).
//
// Here one launch covers every output tile of a contraction ("grouped
// GEMM"): a workgroup owns one TMxTN tile of one output block and walks that
// block's pairs in the reference order, accumulating afac*A*B in registers
// with v_mfma_f64_16x16x4f64; the tile is stored once at the end.  Each
// output element is produced by exactly one workgroup in a fixed order: no
// atomics, deterministic results.  Operand tiles are staged through LDS,
// with the afac scale folded into the A-tile store.  Operands may be stored
// MxK or KxM (A) and KxN or NxK (B): the "reference" matricisations that
// toMatrixRef (Source/wbarray.cc:4085-4140) takes without a permute.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "qsx_internal.hpp"

namespace qsx {

typedef double d4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int TK = 16;

template <int TM, int TN>
struct Smem {
  double As[TK][TM + 1];
  double Bs[TK][TN + 1];
};

// Accumulate the contributions of pairs [p0,p1) into the wave's fragments.
template <int TM, int TN>
__device__ __forceinline__ void tile_accumulate(Smem<TM, TN>& sm, d4 (&acc)[TM / 32][TN / 32], const PairDesc* __restrict__ pairs,
                                                int p0, int p1, int M, int N, int m0, int n0,
                                                const double* __restrict__ A, const double* __restrict__ B) {
  constexpr int FM = TM / 32, FN = TN / 32;  // 16x16 fragments per wave (2x2 waves)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int fr = lane & 15, fk = lane >> 4;
  for (int p = p0; p < p1; ++p) {
    const PairDesc pd = pairs[p];
    const double* __restrict__ a = A + pd.a_off;
    const double* __restrict__ b = B + pd.b_off;
    const bool ta = pd.flags & 1, tb = pd.flags & 2;
    const double af = pd.afac;
    for (int k0 = 0; k0 < pd.K; k0 += TK) {
#pragma unroll
      for (int j = 0; j < (TM * TK) / 256; ++j) {
        int idx = tid + 256 * j;
        int i, k;
        if (!ta) { i = idx % TM; k = idx / TM; }
        else { k = idx % TK; i = idx / TK; }
        int gi = m0 + i, gk = k0 + k;
        double v = 0.0;
        if (gi < M && gk < pd.K) v = ta ? a[gk + (int64_t)pd.lda * gi] : a[gi + (int64_t)pd.lda * gk];
        sm.As[k][i] = v * af;
      }
#pragma unroll
      for (int j = 0; j < (TN * TK) / 256; ++j) {
        int idx = tid + 256 * j;
        int c, k;
        if (!tb) { k = idx % TK; c = idx / TK; }
        else { c = idx % TN; k = idx / TN; }
        int gc = n0 + c, gk = k0 + k;
        double v = 0.0;
        if (gc < N && gk < pd.K) v = tb ? b[gc + (int64_t)pd.ldb * gk] : b[gk + (int64_t)pd.ldb * gc];
        sm.Bs[k][c] = v;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < TK; kk += 4) {
        double af_[FM], bf_[FN];
#pragma unroll
        for (int f = 0; f < FM; ++f) af_[f] = sm.As[kk + fk][wm * (TM / 2) + f * 16 + fr];
#pragma unroll
        for (int f = 0; f < FN; ++f) bf_[f] = sm.Bs[kk + fk][wn * (TN / 2) + f * 16 + fr];
#pragma unroll
        for (int fm = 0; fm < FM; ++fm)
#pragma unroll
          for (int fn = 0; fn < FN; ++fn)
            acc[fm][fn] = __builtin_amdgcn_mfma_f64_16x16x4f64(af_[fm], bf_[fn], acc[fm][fn], 0, 0, 0);
      }
      __syncthreads();
    }
  }
}

template <int TM, int TN>
__device__ __forceinline__ void tile_store(const d4 (&acc)[TM / 32][TN / 32], double* __restrict__ C, int M, int N,
                                           int m0, int n0, int beta) {
  constexpr int FM = TM / 32, FN = TN / 32;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
#pragma unroll
  for (int fm = 0; fm < FM; ++fm)
#pragma unroll
    for (int fn = 0; fn < FN; ++fn)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int row = m0 + wm * (TM / 2) + fm * 16 + (lane >> 4) + 4 * v;  // D layout of v_mfma_f64_16x16x4f64
        int col = n0 + wn * (TN / 2) + fn * 16 + (lane & 15);
        if (row < M && col < N) {
          double* c = C + row + (int64_t)M * col;
          *c = beta ? *c + acc[fm][fn][v] : acc[fm][fn][v];
        }
      }
}

// plain path: one launch per pair; the grid covers the tiles of its output block
template <int TM, int TN>
__global__ __launch_bounds__(256) void k_pair_gemm_f64(PairDesc pd, int64_t c_off, int M, int N, int beta,
                                                       const double* __restrict__ A, const double* __restrict__ B,
                                                       double* __restrict__ C) {
  __shared__ Smem<TM, TN> sm;
  __shared__ PairDesc one;
  if (threadIdx.x == 0) one = pd;
  __syncthreads();
  const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
  d4 acc[TM / 32][TN / 32];
#pragma unroll
  for (int i = 0; i < TM / 32; ++i)
#pragma unroll
    for (int j = 0; j < TN / 32; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  tile_accumulate<TM, TN>(sm, acc, &one, 0, 1, M, N, m0, n0, A, B);
  tile_store<TM, TN>(acc, C + c_off, M, N, m0, n0, beta);
}

// flags[b] = 1 if block b is all exact zeros.  One workgroup per block;
// each thread reads 8 elements per round (independent loads in flight), and
// the workgroup stops at the first round that finds a non-zero (the usual
// case: a non-zero block costs one round)
__global__ __launch_bounds__(256) void k_block_zero_flags(const int64_t* __restrict__ offs,
                                                          const double* __restrict__ data,
                                                          int32_t* __restrict__ flags) {
  constexpr int U = 8;
  const int64_t o0 = offs[blockIdx.x], o1 = offs[blockIdx.x + 1];
  int nz = 0;
  for (int64_t base = o0; base < o1; base += (int64_t)U * 256) {
    double x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = base + (int64_t)u * 256 + threadIdx.x;
      x[u] = i < o1 ? data[i] : 0.0;
    }
    int any = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) any |= x[u] != 0.0;
    if (__syncthreads_or(any)) {
      nz = 1;
      break;
    }
  }
  if (threadIdx.x == 0) flags[blockIdx.x] = nz ? 0 : 1;
}

__global__ void k_copy(double* __restrict__ d, const double* __restrict__ s, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) d[i] = s[i];
}

__global__ void k_fill_zero(double* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

}  // namespace

void launch_pair_gemm_f64(const PairDesc& p, int64_t c_off, int M, int N, int beta, const double* A, const double* B,
                          double* C, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  dim3 g((M + 63) / 64, (N + 63) / 64);
  hipLaunchKernelGGL((k_pair_gemm_f64<64, 64>), g, dim3(256), 0, s, p, c_off, M, N, beta, A, B, C);
  QSX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- no-allocation scratch
// Device memory of the entry points that allocate nothing
// (qsx_block_zero_flags_noalloc): a file-scope array, one per device, held
// until the process ends.  Host code never uses the array's name as an
// address (on the host it is a stand-in, not device memory): the device
// address comes from hipGetSymbolAddress, taken once per device with that
// device current, and is what kernels and copies receive.  Every call on a
// device shares that device's array; calls that overlap on one device race.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

void noalloc_scratch_setup() {
  int dev = -1;
  QSX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) fail(QSX_ERR_DEVICE, "device %d beyond the library's table", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    fail(QSX_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
  g_scratch_ptr[dev] = p;
}

void* noalloc_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[dev];
}

void launch_block_zero_flags(const int64_t* doffs, int nblk, const double* data, int32_t* flags, hipStream_t s) {
  if (nblk <= 0) return;
  hipLaunchKernelGGL(k_block_zero_flags, dim3(nblk), dim3(256), 0, s, doffs, data, flags);
  QSX_HIP(hipGetLastError());
}

void launch_copy(double* d, const double* src, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  int g = (int)std::min<int64_t>(8192, (n + 255) / 256);
  hipLaunchKernelGGL(k_copy, dim3(g), dim3(256), 0, s, d, src, n);
  QSX_HIP(hipGetLastError());
}

void launch_fill_zero(double* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  int g = (int)std::min<int64_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(k_fill_zero, dim3(g), dim3(256), 0, s, p, n);
  QSX_HIP(hipGetLastError());
}

}  // namespace qsx
