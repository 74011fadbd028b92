// Pauli sum -> CSR on the device.
//
// Input: the terms grouped by X mask.  Group g (X masks ascending) holds the
// terms t in [toff[g], toff[g+1]) in their input order; term t has the mask
// sm[t] = Y|Z bits and an FP64 complex coefficient coef[t] with the Y phase
// (-i)^popcount(Y) already folded in.  For row i, group g contributes
//     v(i, g) = sum_t coef[t] * (-1)^popcount(i & sm[t])
// summed in FP64 in term order (the product by the sign is written out as
// the complex product (c.re + i c.im)(s + 0i) so signed zeros come out as
// they do in an FP64 complex multiply), at column i ^ xm[g], kept when
// v != 0 and rounded to complex64.  Columns of a row follow the group order.
//
// Three kernels, all one row per thread:
//   k_csr_count  per-block nonzero totals;
//   k_csr_scan   one workgroup, exclusive scan of the block totals -> block
//                offsets, plus the grand total;
//   k_csr_fill   recomputes its row, block-local exclusive scan of the row
//                counts, writes row_ptr, col_idx and values.
#include "qsim_internal.hpp"

namespace qsim {
namespace {

constexpr int kCsrThreads = 256;

__device__ __forceinline__ int row_count(uint64_t row, int n_groups, const int* __restrict__ toff,
                                         const uint64_t* __restrict__ sm,
                                         const double2* __restrict__ coef) {
  int c = 0;
  for (int g = 0; g < n_groups; ++g) {
    double re, im;
    c += pauli_group_value<uint64_t>(row, toff[g], toff[g + 1], sm, coef, re, im) ? 1 : 0;
  }
  return c;
}

// Exclusive scan of v over the 256 threads of a block; returns the prefix
// and the block total in *total.
__device__ __forceinline__ int64_t block_exscan(int64_t v, int64_t* total) {
  __shared__ int64_t wsum[kCsrThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t incl = v;
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int64_t base = 0, all = 0;
  for (int k = 0; k < kCsrThreads / 64; ++k) {
    if (k < w) base += wsum[k];
    all += wsum[k];
  }
  *total = all;
  return base + incl - v;
}

__global__ void __launch_bounds__(kCsrThreads)
k_csr_count(uint64_t n_rows, int n_groups, const int* __restrict__ toff,
            const uint64_t* __restrict__ sm, const double2* __restrict__ coef,
            int64_t* __restrict__ block_sum) {
  const uint64_t row = (uint64_t)blockIdx.x * kCsrThreads + threadIdx.x;
  const int c = row < n_rows ? row_count(row, n_groups, toff, sm, coef) : 0;
  int64_t total;
  block_exscan(c, &total);
  if (threadIdx.x == 0) block_sum[blockIdx.x] = total;
}

// In place: block_sum[b] -> sum of block_sum[0..b); total -> *nnz.
__global__ void __launch_bounds__(kCsrThreads)
k_csr_scan(uint64_t n_blocks, int64_t* __restrict__ block_sum, int64_t* __restrict__ nnz) {
  int64_t carry = 0;
  for (uint64_t base = 0; base < n_blocks; base += kCsrThreads) {
    const uint64_t b = base + threadIdx.x;
    const int64_t v = b < n_blocks ? block_sum[b] : 0;
    int64_t total;
    const int64_t pre = block_exscan(v, &total);
    __syncthreads();   // every thread has read wsum before the next chunk rewrites it
    if (b < n_blocks) block_sum[b] = carry + pre;
    carry += total;
  }
  if (threadIdx.x == 0) *nnz = carry;
}

__global__ void __launch_bounds__(kCsrThreads)
k_csr_fill(uint64_t n_rows, int n_groups, const uint64_t* __restrict__ xm,
           const int* __restrict__ toff, const uint64_t* __restrict__ sm,
           const double2* __restrict__ coef, const int64_t* __restrict__ block_off,
           int32_t* __restrict__ row_ptr, int32_t* __restrict__ col_idx,
           float2* __restrict__ values) {
  const uint64_t row = (uint64_t)blockIdx.x * kCsrThreads + threadIdx.x;
  const int c = row < n_rows ? row_count(row, n_groups, toff, sm, coef) : 0;
  int64_t total;
  int64_t off = block_off[blockIdx.x] + block_exscan(c, &total);
  if (row >= n_rows) return;
  row_ptr[row] = (int32_t)off;
  for (int g = 0; g < n_groups; ++g) {
    double re, im;
    if (pauli_group_value<uint64_t>(row, toff[g], toff[g + 1], sm, coef, re, im)) {
      col_idx[off] = (int32_t)(row ^ xm[g]);
      values[off] = make_float2((float)re, (float)im);
      ++off;
    }
  }
  if (row == n_rows - 1) row_ptr[n_rows] = (int32_t)off;
}

int check_groups(int n_qubits, int n_groups, const void* xm, const void* toff, const void* sm,
                 const void* coef, const void* ws) {
  if (n_qubits < 1 || n_qubits > 31) return fail(QSIM_ERANGE, "n_qubits=%d outside 1..31", n_qubits);
  if (n_groups < 0) return fail(QSIM_EINVAL, "n_groups < 0");
  if (!ws || (n_groups > 0 && (!xm || !toff || !sm || !coef)))
    return fail(QSIM_EINVAL, "NULL pointer argument");
  return QSIM_OK;
}

}  // namespace
}  // namespace qsim

using namespace qsim;

extern "C" {

int64_t qsim_pauli_csr_workspace_bytes(int n_qubits) {
  if (n_qubits < 1 || n_qubits > 31) return -1;
  const uint64_t n_blocks = ceil_div(1ull << n_qubits, kCsrThreads);
  return (int64_t)((n_blocks + 1) * sizeof(int64_t));
}

int qsim_pauli_csr_count(void* stream, int n_qubits, int n_groups, const void* xm, const void* toff,
                         const void* sm, const void* coef, void* ws, int64_t* nnz_out) {
  clear_error();
  if (int rc = check_groups(n_qubits, n_groups, xm, toff, sm, coef, ws)) return rc;
  if (!nnz_out) return fail(QSIM_EINVAL, "NULL nnz_out");
  const uint64_t n_rows = 1ull << n_qubits;
  const uint64_t n_blocks = ceil_div(n_rows, kCsrThreads);
  int64_t* block_sum = (int64_t*)ws;
  int64_t* total = block_sum + n_blocks;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_csr_count, dim3((unsigned)n_blocks), dim3(kCsrThreads), 0, s, n_rows, n_groups,
                     (const int*)toff, (const uint64_t*)sm, (const double2*)coef, block_sum);
  if (int rc = launch_status("k_csr_count")) return rc;
  hipLaunchKernelGGL(k_csr_scan, dim3(1), dim3(kCsrThreads), 0, s, n_blocks, block_sum, total);
  if (int rc = launch_status("k_csr_scan")) return rc;
  int64_t h = 0;
  if (hipMemcpyAsync(&h, total, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(QSIM_EHIP, "reading the nonzero count failed");
  if (h > INT32_MAX) return fail(QSIM_ERANGE, "%lld nonzeros exceed int32 CSR indices", (long long)h);
  *nnz_out = h;
  return QSIM_OK;
}

int qsim_pauli_csr_fill(void* stream, int n_qubits, int n_groups, const void* xm, const void* toff,
                        const void* sm, const void* coef, const void* ws, int64_t nnz,
                        void* row_ptr, void* col_idx, void* values) {
  clear_error();
  if (int rc = check_groups(n_qubits, n_groups, xm, toff, sm, coef, ws)) return rc;
  if (!row_ptr || (nnz > 0 && (!col_idx || !values))) return fail(QSIM_EINVAL, "NULL output pointer");
  if (nnz < 0 || nnz > INT32_MAX) return fail(QSIM_ERANGE, "nnz=%lld", (long long)nnz);
  const uint64_t n_rows = 1ull << n_qubits;
  const uint64_t n_blocks = ceil_div(n_rows, kCsrThreads);
  hipLaunchKernelGGL(k_csr_fill, dim3((unsigned)n_blocks), dim3(kCsrThreads), 0, as_stream(stream),
                     n_rows, n_groups, (const uint64_t*)xm, (const int*)toff, (const uint64_t*)sm,
                     (const double2*)coef, (const int64_t*)ws, (int32_t*)row_ptr, (int32_t*)col_idx,
                     (float2*)values);
  return launch_status("k_csr_fill");
}

}  // extern "C"
