// Chebyshev evolution kernels for gfx950: CSR SpMV/SpMM and the three-term
// recurrence with compensated accumulation.
//
// Row sums use TwoProduct on every real product and Neumaier's improved Kahan
// summation (Neumaier 1974) for the running sum, so the order of magnitudes
// along a row does not matter.  The batched form keeps n_rhs right-hand sides
// RHS-contiguous (element (row, j) at row*n_rhs + j): with n_rhs = 32 one
// wave64 covers two rows, every column gather reads 256 contiguous bytes, and
// the recurrence update and the Kahan accumulation are fused into the SpMM
// epilogue, so one launch per Chebyshev term touches A, T_k, T_{k-1}, the
// accumulator and its compensation exactly once.
#include "qsim_internal.hpp"

namespace qsim {

// Nonzeros whose column indices, values and gathers are issued together
// before any of their products is accumulated: the gathers of a row are
// independent, so batching them overlaps their latencies (the summation
// order, and so every result bit, is unchanged).  Measured per Chebyshev
// term on the 22-qubit chain at 32 RHS: 1 -> 9.6 ms, 2 -> 6.9, 4 -> 6.2,
// 8 -> 6.9, 16 -> 8.0 (more in flight costs occupancy).
constexpr int kGather = 4;

// Neumaier step: s + c holds the running sum with the low part in c.
__device__ __forceinline__ void neu_add(float& s, float& c, float x) {
  const float t = s + x;
  if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
  else c += (x - t) + s;
  s = t;
}

// Adds the complex product v * xj to (sr + cr, si + ci) with TwoProduct
// errors folded into the compensation terms.
__device__ __forceinline__ void cmac_acc(float2 v, float2 xj, float& sr, float& cr, float& si,
                                         float& ci) {
  float p, e;
  two_prod(v.x, xj.x, p, e); neu_add(sr, cr, p); cr += e;
  two_prod(-v.y, xj.y, p, e); neu_add(sr, cr, p); cr += e;
  two_prod(v.x, xj.y, p, e); neu_add(si, ci, p); ci += e;
  two_prod(v.y, xj.x, p, e); neu_add(si, ci, p); ci += e;
}

__global__ __launch_bounds__(kThreads) void k_spmv(const int32_t* __restrict__ rp,
                                                   const int32_t* __restrict__ ci,
                                                   const float2* __restrict__ val, int64_t n_rows,
                                                   const float2* __restrict__ x,
                                                   float2* __restrict__ y, float2 alpha,
                                                   float2 beta) {
  const int64_t row = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (row >= n_rows) return;
  float sr = 0.f, cr = 0.f, si = 0.f, cc = 0.f;
  const int32_t b = rp[row], e = rp[row + 1];
  int32_t k = b;
  for (; k + kGather <= e; k += kGather) {
    int32_t c[kGather];
    float2 v[kGather], xs[kGather];
#pragma unroll
    for (int u = 0; u < kGather; ++u) { c[u] = ci[k + u]; v[u] = val[k + u]; }
#pragma unroll
    for (int u = 0; u < kGather; ++u) xs[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < kGather; ++u) cmac_acc(v[u], xs[u], sr, cr, si, cc);
  }
  for (; k < e; ++k) cmac_acc(val[k], x[ci[k]], sr, cr, si, cc);
  const float2 dot = make_float2(sr + cr, si + cc);
  const float2 r = cmulc(alpha, dot);
  float2 out = r;
  if (beta.x != 0.f || beta.y != 0.f) {
    const float2 yb = cmulc(beta, y[row]);
    out = make_float2(r.x + yb.x, r.y + yb.y);
  }
  y[row] = out;
}

__global__ __launch_bounds__(kThreads) void k_step(const float2* __restrict__ s,
                                                   const float2* __restrict__ tp,
                                                   float2* __restrict__ tn, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float2 a = s[i], b = tp[i];
  tn[i] = make_float2(2.f * a.x - b.x, 2.f * a.y - b.y);
}

__global__ __launch_bounds__(kThreads) void k_accum(float2* __restrict__ acc,
                                                    const float2* __restrict__ t, int64_t n,
                                                    float2 c) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float2 v = cmulc(c, t[i]);
  const float2 a = acc[i];
  acc[i] = make_float2(a.x + v.x, a.y + v.y);
}

__global__ __launch_bounds__(kThreads) void k_step_accum(const float2* __restrict__ s,
                                                         const float2* __restrict__ tp,
                                                         float2* __restrict__ tn,
                                                         float2* __restrict__ acc, int64_t n,
                                                         float2 c) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float2 a = s[i], b = tp[i];
  const float2 t = make_float2(2.f * a.x - b.x, 2.f * a.y - b.y);
  tn[i] = t;
  const float2 v = cmulc(c, t);
  const float2 z = acc[i];
  acc[i] = make_float2(z.x + v.x, z.y + v.y);
}

// Fused batched Chebyshev term.  Thread = (row, j), j fastest.
__global__ __launch_bounds__(kThreads) void k_cheb_spmm(const int32_t* __restrict__ rp,
                                                        const int32_t* __restrict__ ci,
                                                        const float2* __restrict__ val,
                                                        int64_t n_rows, int n_rhs, float scale,
                                                        const float2* __restrict__ tc,
                                                        float2* __restrict__ tpn,
                                                        float2* __restrict__ acc,
                                                        float2* __restrict__ comp, float2 c,
                                                        int first) {
  const int64_t e_idx = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t row = e_idx / n_rhs;
  const int j = (int)(e_idx - row * n_rhs);
  if (row >= n_rows) return;
  float sr = 0.f, cr = 0.f, si = 0.f, cc = 0.f;
  const int32_t b = rp[row], e = rp[row + 1];
  int32_t k = b;
  for (; k + kGather <= e; k += kGather) {
    int32_t c[kGather];
    float2 v[kGather], xs[kGather];
#pragma unroll
    for (int u = 0; u < kGather; ++u) { c[u] = ci[k + u]; v[u] = val[k + u]; }
#pragma unroll
    for (int u = 0; u < kGather; ++u) xs[u] = tc[(int64_t)c[u] * n_rhs + j];
#pragma unroll
    for (int u = 0; u < kGather; ++u) cmac_acc(v[u], xs[u], sr, cr, si, cc);
  }
  for (; k < e; ++k) cmac_acc(val[k], tc[(int64_t)ci[k] * n_rhs + j], sr, cr, si, cc);
  // w = scale * (A t_cur): the exact product errors of the scaling are kept
  float p, q;
  float wr, wi;
  two_prod(sr, scale, p, q); wr = p + (q + cr * scale);
  two_prod(si, scale, p, q); wi = p + (q + cc * scale);
  float2 tn;
  if (first) {
    tn = make_float2(wr, wi);
  } else {
    const float2 tp = tpn[e_idx];
    tn = make_float2(2.f * wr - tp.x, 2.f * wi - tp.y);
  }
  tpn[e_idx] = tn;
  const float2 v = cmulc(c, tn);
  float2 a = acc[e_idx], k2 = comp[e_idx];
  neu_add(a.x, k2.x, v.x);
  neu_add(a.y, k2.y, v.y);
  acc[e_idx] = a;
  comp[e_idx] = k2;
}

__global__ __launch_bounds__(kThreads) void k_cheb_init(float2* __restrict__ acc,
                                                        float2* __restrict__ comp,
                                                        const float2* __restrict__ t, int64_t n,
                                                        float2 c) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  acc[i] = cmulc(c, t[i]);
  comp[i] = make_float2(0.f, 0.f);
}

__global__ __launch_bounds__(kThreads) void k_cheb_finish(float2* __restrict__ out,
                                                          const float2* __restrict__ acc,
                                                          const float2* __restrict__ comp,
                                                          int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float2 a = acc[i], k = comp[i];
  out[i] = make_float2(a.x + k.x, a.y + k.y);
}

static inline float2 f2(double re, double im) { return make_float2((float)re, (float)im); }

}  // namespace qsim

using namespace qsim;

extern "C" {

int qsim_spmv_csr(void* stream, const void* row_ptr, const void* col_idx, const void* values,
                  int64_t n_rows, const void* x, void* y, double alpha_re, double alpha_im,
                  double beta_re, double beta_im) {
  if (n_rows < 0) return fail(QSIM_EINVAL, "n_rows < 0");
  if (n_rows == 0) return QSIM_OK;
  if (!row_ptr || !col_idx || !values || !x || !y) return fail(QSIM_EINVAL, "NULL pointer argument");
  hipLaunchKernelGGL(k_spmv, dim3(ceil_div(n_rows, kThreads)), dim3(kThreads), 0, as_stream(stream),
                     (const int32_t*)row_ptr, (const int32_t*)col_idx, (const float2*)values, n_rows,
                     (const float2*)x, (float2*)y, f2(alpha_re, alpha_im), f2(beta_re, beta_im));
  return launch_status("spmv_csr");
}

int qsim_chebyshev_step(void* stream, const void* s, const void* t_prev, void* t_next, int64_t n) {
  if (n < 0) return fail(QSIM_EINVAL, "n < 0");
  if (n == 0) return QSIM_OK;
  if (!s || !t_prev || !t_next) return fail(QSIM_EINVAL, "NULL pointer argument");
  hipLaunchKernelGGL(k_step, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, as_stream(stream),
                     (const float2*)s, (const float2*)t_prev, (float2*)t_next, n);
  return launch_status("chebyshev_step");
}

int qsim_chebyshev_accum(void* stream, void* accum, const void* t, int64_t n, double c_re,
                         double c_im) {
  if (n < 0) return fail(QSIM_EINVAL, "n < 0");
  if (n == 0) return QSIM_OK;
  if (!accum || !t) return fail(QSIM_EINVAL, "NULL pointer argument");
  hipLaunchKernelGGL(k_accum, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, as_stream(stream),
                     (float2*)accum, (const float2*)t, n, f2(c_re, c_im));
  return launch_status("chebyshev_accum");
}

int qsim_chebyshev_step_accum(void* stream, const void* s, const void* t_prev, void* t_next,
                              void* accum, int64_t n, double c_re, double c_im) {
  if (n < 0) return fail(QSIM_EINVAL, "n < 0");
  if (n == 0) return QSIM_OK;
  if (!s || !t_prev || !t_next || !accum) return fail(QSIM_EINVAL, "NULL pointer argument");
  hipLaunchKernelGGL(k_step_accum, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0,
                     as_stream(stream), (const float2*)s, (const float2*)t_prev, (float2*)t_next,
                     (float2*)accum, n, f2(c_re, c_im));
  return launch_status("chebyshev_step_accum");
}

int qsim_cheb_spmm_step(void* stream, const void* row_ptr, const void* col_idx, const void* values,
                        int64_t n_rows, int n_rhs, double scale, const void* t_cur,
                        void* t_prev_next, void* acc, void* comp, double c_re, double c_im,
                        int first) {
  if (n_rows < 0 || n_rhs < 1) return fail(QSIM_EINVAL, "bad sizes (n_rows=%lld, n_rhs=%d)",
                                           (long long)n_rows, n_rhs);
  if (n_rows == 0) return QSIM_OK;
  if (!row_ptr || !col_idx || !values || !t_cur || !t_prev_next || !acc || !comp)
    return fail(QSIM_EINVAL, "NULL pointer argument");
  const int64_t total = n_rows * (int64_t)n_rhs;
  hipLaunchKernelGGL(k_cheb_spmm, dim3(ceil_div(total, kThreads)), dim3(kThreads), 0,
                     as_stream(stream), (const int32_t*)row_ptr, (const int32_t*)col_idx,
                     (const float2*)values, n_rows, n_rhs, (float)scale, (const float2*)t_cur,
                     (float2*)t_prev_next, (float2*)acc, (float2*)comp, f2(c_re, c_im), first);
  return launch_status("cheb_spmm_step");
}

int qsim_cheb_init_acc(void* stream, void* acc, void* comp, const void* t, int64_t n, double c_re,
                       double c_im) {
  if (n < 0) return fail(QSIM_EINVAL, "n < 0");
  if (n == 0) return QSIM_OK;
  if (!acc || !comp || !t) return fail(QSIM_EINVAL, "NULL pointer argument");
  hipLaunchKernelGGL(k_cheb_init, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, as_stream(stream),
                     (float2*)acc, (float2*)comp, (const float2*)t, n, f2(c_re, c_im));
  return launch_status("cheb_init_acc");
}

int qsim_cheb_finish(void* stream, void* out, const void* acc, const void* comp, int64_t n) {
  if (n < 0) return fail(QSIM_EINVAL, "n < 0");
  if (n == 0) return QSIM_OK;
  if (!out || !acc || !comp) return fail(QSIM_EINVAL, "NULL pointer argument");
  hipLaunchKernelGGL(k_cheb_finish, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0,
                     as_stream(stream), (float2*)out, (const float2*)acc, (const float2*)comp, n);
  return launch_status("cheb_finish");
}

}  // extern "C"
