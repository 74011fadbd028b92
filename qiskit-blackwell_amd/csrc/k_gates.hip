// Per-gate streaming kernels for gfx950: state init, dense 1Q/2Q, diagonal,
// multi-controlled, permutation and batched gates, plus the library's error
// plumbing.
//
// Streaming layout.  Every kernel moves amplitudes as float4 (two adjacent
// complex64, 16 B per lane), so one wave64 load instruction covers 1 KiB of
// contiguous HBM.  A "work item" is a fixed set of float4 slots; each thread
// owns U items spaced by the block size, all loads are issued before any
// arithmetic so that 2*U (1Q) or 4*U (2Q) 16-byte loads are in flight per
// lane.  Index math uses 32-bit integers for states up to 2^32 amplitudes and
// 64-bit above (template parameter I).
//
// Pair addressing for a 2x2 gate on target t (amplitude pairs differing in
// bit t): in float4 units the two slots of item w are
//     f0 = ins0(w, s),  f1 = f0 + 2^s,  s = max(t - 1, 0).
// For t >= 1 the pairs are (f0.lo, f1.lo) and (f0.hi, f1.hi); for t = 0 they
// are the two halves of each slot.  The index math follows the reference's
// insert-zero rule (bwk/csrc/apply_gate_sm120.cu:100-104) in float4 units.
#include "qsim_internal.hpp"
#include <stdio.h>
#include <string.h>
#include <algorithm>

namespace qsim {

// ============================================================ error plumbing
static thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}
void clear_error() { g_err[0] = 0; }

int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(QSIM_EHIP, "%s: launch failed: %s", what, hipGetErrorString(e));
  return QSIM_OK;
}

int check_n(int n) {
  if (n < 1 || n > QSIM_MAX_QUBITS)
    return fail(QSIM_ERANGE, "n_qubits=%d outside 1..%d", n, QSIM_MAX_QUBITS);
  return QSIM_OK;
}
int check_q(int q, int n, const char* what) {
  if (q < 0 || q >= n) return fail(QSIM_EINVAL, "%s=%d outside 0..%d", what, q, n - 1);
  return QSIM_OK;
}

int build_mc(McSpec& s, int n, const int64_t* controls, int n_controls, const int* targets,
             int n_targets) {
  if (n_controls < 0 || n_controls > QSIM_MAX_CONTROLS)
    return fail(QSIM_EINVAL, "controls must have 0-%d elements (got %d)", QSIM_MAX_CONTROLS,
                n_controls);
  if (n_controls > 0 && !controls) return fail(QSIM_EINVAL, "controls is NULL");
  s.n_pos = 0;
  s.cmask = 0;
  uint64_t seen = 0;
  for (int j = 0; j < n_targets; ++j) {
    int t = targets[j];
    if (t < 0 || t >= n) return fail(QSIM_EINVAL, "target %d outside 0..%d", t, n - 1);
    if (seen >> t & 1) return fail(QSIM_EINVAL, "target qubits must differ");
    seen |= 1ull << t;
    s.pos[s.n_pos++] = t;
  }
  for (int j = 0; j < n_controls; ++j) {
    int64_t c = controls[j];
    if (c < 0 || c >= n) return fail(QSIM_EINVAL, "control %lld outside 0..%d", (long long)c, n - 1);
    if (seen >> c & 1) {
      for (int k = 0; k < n_targets; ++k)
        if (targets[k] == c) return fail(QSIM_EINVAL, "target_qubit must not appear in controls");
      return fail(QSIM_EINVAL, "duplicate control %lld", (long long)c);
    }
    seen |= 1ull << c;
    s.cmask |= 1ull << c;
    s.pos[s.n_pos++] = (int)c;
  }
  if (s.n_pos > n) return fail(QSIM_EINVAL, "more affected qubits than the state has");
  std::sort(s.pos, s.pos + s.n_pos);
  return QSIM_OK;
}

// ============================================================ state init
template <typename I>
__global__ __launch_bounds__(kThreads) void k_init(float4* __restrict__ st, I n4, I per_state4) {
  const I f = (I)blockIdx.x * kThreads + threadIdx.x;
  if (f >= n4) return;
  const bool first = (f & (per_state4 - 1)) == 0;
  st[f] = make_float4(first ? 1.f : 0.f, 0.f, 0.f, 0.f);
}

static int init_impl(void* stream, void* states, uint64_t n_states, int n) {
  if (int rc = check_n(n)) return rc;
  if (!states) return fail(QSIM_EINVAL, "state is NULL");
  if (n_states == 0) return QSIM_OK;
  const uint64_t per4 = (1ull << n) / 2;
  const uint64_t n4 = per4 * n_states;
  const uint64_t blocks = ceil_div(n4, kThreads);
  if (n4 <= 0xffffffffull)
    hipLaunchKernelGGL(k_init<uint32_t>, dim3(blocks), dim3(kThreads), 0, as_stream(stream),
                       (float4*)states, (uint32_t)n4, (uint32_t)per4);
  else
    hipLaunchKernelGGL(k_init<uint64_t>, dim3(blocks), dim3(kThreads), 0, as_stream(stream),
                       (float4*)states, n4, per4);
  return launch_status("state_init");
}

// ============================================================ 1Q dense
// Coefficients come either from device memory (gdev != nullptr, the
// reference's calling convention) or by value.
__device__ __forceinline__ void load_m2(const float2* gdev, const M2& gv, float2 g[4]) {
  if (gdev) {
#pragma unroll
    for (int k = 0; k < 4; ++k) g[k] = gdev[k];
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) g[k] = gv.g[k];
  }
}

// Batched states need no extra index math: with s <= n-2 the insert-zero of
// the flat item index w = state*2^(n-2) + local shifts the state number up by
// one bit, which is exactly the state stride 2^(n-1) in float4 units.
template <typename I, int U>
__global__ __launch_bounds__(kThreads) void k_gate1q(float4* __restrict__ st, I n_items, int t,
                                                     const float2* __restrict__ gdev, M2 gv) {
  float2 g[4];
  load_m2(gdev, gv, g);
  const int s = t > 0 ? t - 1 : 0;
  const I step = I(1) << s;
  const I base = (I)blockIdx.x * (kThreads * U) + threadIdx.x;
  I f0[U];
  float4 v0[U], v1[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    I w = base + (I)u * kThreads;
    if (w >= n_items) w = n_items - 1;  // duplicate the last item; store is masked below
    f0[u] = ins0(w, s);
    v0[u] = st[f0[u]];
    v1[u] = st[f0[u] + step];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const I w = base + (I)u * kThreads;
    float4 r0, r1;
    if (t == 0) {
      r0 = cat4(crow2(g[0], lo2(v0[u]), g[1], hi2(v0[u])), crow2(g[2], lo2(v0[u]), g[3], hi2(v0[u])));
      r1 = cat4(crow2(g[0], lo2(v1[u]), g[1], hi2(v1[u])), crow2(g[2], lo2(v1[u]), g[3], hi2(v1[u])));
    } else {
      r0 = cat4(crow2(g[0], lo2(v0[u]), g[1], lo2(v1[u])), crow2(g[0], hi2(v0[u]), g[1], hi2(v1[u])));
      r1 = cat4(crow2(g[2], lo2(v0[u]), g[3], lo2(v1[u])), crow2(g[2], hi2(v0[u]), g[3], hi2(v1[u])));
    }
    if (w < n_items) {
      st[f0[u]] = r0;
      st[f0[u] + step] = r1;
    }
  }
}

// Tiny states (n = 1): one thread per pair, float2 access.
__global__ void k_gate1q_tiny(float2* __restrict__ st, uint64_t n_states, int n, int t,
                              const float2* __restrict__ gdev, M2 gv) {
  float2 g[4];
  load_m2(gdev, gv, g);
  const uint64_t pairs = (1ull << n) / 2;
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pairs * n_states) return;
  const uint64_t sidx = p / pairs, pl = p % pairs;
  float2* s = st + (sidx << n);
  const uint64_t i0 = ins0(pl, t), i1 = i0 | (1ull << t);
  const float2 a0 = s[i0], a1 = s[i1];
  s[i0] = crow2(g[0], a0, g[1], a1);
  s[i1] = crow2(g[2], a0, g[3], a1);
}

constexpr int kU1 = 2;

int gate1q_impl(void* stream, void* states, uint64_t n_states, int n, const float2* gdev,
                const M2& gv, int t) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_q(t, n, "target_qubit")) return rc;
  if (!states) return fail(QSIM_EINVAL, "state is NULL");
  if (n_states == 0) return QSIM_OK;
  hipStream_t s = as_stream(stream);
  if (n < 2) {
    const uint64_t pairs = n_states * ((1ull << n) / 2);
    hipLaunchKernelGGL(k_gate1q_tiny, dim3(ceil_div(pairs, 64)), dim3(64), 0, s, (float2*)states,
                       n_states, n, t, gdev, gv);
    return launch_status("apply_gate");
  }
  const uint64_t items = (1ull << n) / 4 * n_states;
  const uint64_t blocks = ceil_div(items, (uint64_t)kThreads * kU1);
  if (items * 2 <= 0xffffffffull)   // largest float4 index fits 32 bits
    hipLaunchKernelGGL((k_gate1q<uint32_t, kU1>), dim3(blocks), dim3(kThreads), 0, s,
                       (float4*)states, (uint32_t)items, t, gdev, gv);
  else
    hipLaunchKernelGGL((k_gate1q<uint64_t, kU1>), dim3(blocks), dim3(kThreads), 0, s,
                       (float4*)states, items, t, gdev, gv);
  return launch_status("apply_gate");
}

// ============================================================ 2Q dense
// Gate row/column index r = 2*bit(hi) + bit(lo), lo < hi
// (bwk/csrc/apply_gate2q_sm120.cu:11-20, 245-246).  Item w covers two groups
// through four float4 slots F[0..3]:
//   lo >= 1: F = f, f+2^(lo-1), f+2^(hi-1), f+both, f = ins0(ins0(w,lo-1),hi-1);
//            group A = the .lo halves, group B = the .hi halves.
//   lo == 0: F[0]=ins0(2w,hi-1), F[1]=F[0]+2^(hi-1), F[2]=ins0(2w+1,hi-1),
//            F[3]=F[2]+2^(hi-1); group A = (F0.lo,F0.hi,F1.lo,F1.hi), B likewise.
__device__ __forceinline__ void load_m4(const float2* gdev, const M4& gv, float2 g[16]) {
  if (gdev) {
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k] = gdev[k];
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k] = gv.g[k];
  }
}

__device__ __forceinline__ void mat4(const float2 g[16], float2& a0, float2& a1, float2& a2,
                                     float2& a3) {
  const float2 b0 = crow4(g + 0, a0, a1, a2, a3);
  const float2 b1 = crow4(g + 4, a0, a1, a2, a3);
  const float2 b2 = crow4(g + 8, a0, a1, a2, a3);
  const float2 b3 = crow4(g + 12, a0, a1, a2, a3);
  a0 = b0; a1 = b1; a2 = b2; a3 = b3;
}

template <typename I>
__global__ __launch_bounds__(kThreads) void k_gate2q(float4* __restrict__ st, I n_items, int lo,
                                                     int hi, const float2* __restrict__ gdev,
                                                     M4 gv) {
  float2 g[16];
  load_m4(gdev, gv, g);
  I w = (I)blockIdx.x * kThreads + threadIdx.x;
  if (w >= n_items) return;
  I F[4];
  const I mh = I(1) << (hi - 1);
  if (lo > 0) {
    const I ml = I(1) << (lo - 1);
    F[0] = ins0(ins0(w, lo - 1), hi - 1);
    F[1] = F[0] + ml; F[2] = F[0] + mh; F[3] = F[0] + ml + mh;
  } else {
    F[0] = ins0(2 * w, hi - 1);     F[1] = F[0] + mh;
    F[2] = ins0(2 * w + 1, hi - 1); F[3] = F[2] + mh;
  }
  float4 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = st[F[k]];
  float2 a[8];
  if (lo > 0) {
    a[0] = lo2(v[0]); a[1] = lo2(v[1]); a[2] = lo2(v[2]); a[3] = lo2(v[3]);
    a[4] = hi2(v[0]); a[5] = hi2(v[1]); a[6] = hi2(v[2]); a[7] = hi2(v[3]);
  } else {
    a[0] = lo2(v[0]); a[1] = hi2(v[0]); a[2] = lo2(v[1]); a[3] = hi2(v[1]);
    a[4] = lo2(v[2]); a[5] = hi2(v[2]); a[6] = lo2(v[3]); a[7] = hi2(v[3]);
  }
  mat4(g, a[0], a[1], a[2], a[3]);
  mat4(g, a[4], a[5], a[6], a[7]);
  if (lo > 0) {
    v[0] = cat4(a[0], a[4]); v[1] = cat4(a[1], a[5]); v[2] = cat4(a[2], a[6]); v[3] = cat4(a[3], a[7]);
  } else {
    v[0] = cat4(a[0], a[1]); v[1] = cat4(a[2], a[3]); v[2] = cat4(a[4], a[5]); v[3] = cat4(a[6], a[7]);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) st[F[k]] = v[k];
}

__global__ void k_gate2q_tiny(float2* __restrict__ st, int lo, int hi,
                              const float2* __restrict__ gdev, M4 gv) {
  float2 g[16];
  load_m4(gdev, gv, g);
  float2 a0 = st[0], a1 = st[1u << lo], a2 = st[1u << hi], a3 = st[(1u << lo) | (1u << hi)];
  mat4(g, a0, a1, a2, a3);
  st[0] = a0; st[1u << lo] = a1; st[1u << hi] = a2; st[(1u << lo) | (1u << hi)] = a3;
}

int gate2q_impl(void* stream, void* state, int n, const float2* gdev, const M4& gv, int q0,
                int q1) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_q(q0, n, "qubit0")) return rc;
  if (int rc = check_q(q1, n, "qubit1")) return rc;
  if (q0 == q1) return fail(QSIM_EINVAL, "qubit0 and qubit1 must differ");
  if (!state) return fail(QSIM_EINVAL, "state is NULL");
  const int lo = std::min(q0, q1), hi = std::max(q0, q1);
  hipStream_t s = as_stream(stream);
  if (n == 2) {
    hipLaunchKernelGGL(k_gate2q_tiny, dim3(1), dim3(1), 0, s, (float2*)state, lo, hi, gdev, gv);
    return launch_status("apply_gate_2q");
  }
  const uint64_t items = (1ull << n) / 8;
  const uint64_t blocks = ceil_div(items, kThreads);
  if (n <= 32)
    hipLaunchKernelGGL(k_gate2q<uint32_t>, dim3(blocks), dim3(kThreads), 0, s, (float4*)state,
                       (uint32_t)items, lo, hi, gdev, gv);
  else
    hipLaunchKernelGGL(k_gate2q<uint64_t>, dim3(blocks), dim3(kThreads), 0, s, (float4*)state,
                       items, lo, hi, gdev, gv);
  return launch_status("apply_gate_2q");
}

// ============================================================ diagonals
// kind 0: diag(d0, d1) on bit t;  kind 1: diag(d00, d01, d10, d11) on (lo, hi).
template <typename I, int U, int KIND>
__global__ __launch_bounds__(kThreads) void k_diag(float4* __restrict__ st, I n4, int t, int hi,
                                                   const float2* __restrict__ ddev, M2 dv) {
  constexpr int kind = KIND;
  float2 d0, d1, d2, d3;
  if (ddev) {
    d0 = ddev[0]; d1 = ddev[1];
    if (KIND == 1) { d2 = ddev[2]; d3 = ddev[3]; } else { d2 = d0; d3 = d1; }
  } else {
    d0 = dv.g[0]; d1 = dv.g[1]; d2 = dv.g[2]; d3 = dv.g[3];
  }
  const I base = (I)blockIdx.x * (kThreads * U) + threadIdx.x;
  float4 v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    I f = base + (I)u * kThreads;
    v[u] = st[f < n4 ? f : n4 - 1];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const I f = base + (I)u * kThreads;
    const I i0 = 2 * f, i1 = i0 + 1;
    // kind 0 uses (d0, d1) only; for kind 1 bit(hi) picks the (d2, d3) half.
    const bool h0 = kind == 1 && ((i0 >> hi) & 1), h1 = kind == 1 && ((i1 >> hi) & 1);
    const bool b0 = (i0 >> t) & 1, b1 = (i1 >> t) & 1;
    const float2 e0 = h0 ? (b0 ? d3 : d2) : (b0 ? d1 : d0);
    const float2 e1 = h1 ? (b1 ? d3 : d2) : (b1 ? d1 : d0);
    const float4 r = cat4(cmulc(lo2(v[u]), e0), cmulc(hi2(v[u]), e1));
    if (f < n4) st[f] = r;
  }
}

// Phase on bit t = 1 only: t >= 1 touches slot ins0(w, t-1) | 2^(t-1); t = 0
// touches the .hi half of every slot (the cache line is read either way).
template <typename I, int U>
__global__ __launch_bounds__(kThreads) void k_phase(float4* __restrict__ st, I n_items, int t,
                                                    float2 ph) {
  const I base = (I)blockIdx.x * (kThreads * U) + threadIdx.x;
  const int s = t - 1;
  I f[U];
  float4 v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    I w = base + (I)u * kThreads;
    if (w >= n_items) w = n_items - 1;
    f[u] = t > 0 ? (ins0(w, s) | (I(1) << s)) : w;
    v[u] = st[f[u]];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const I w = base + (I)u * kThreads;
    float4 r;
    if (t > 0) r = cat4(cmulc(lo2(v[u]), ph), cmulc(hi2(v[u]), ph));
    else r = cat4(lo2(v[u]), cmulc(hi2(v[u]), ph));
    if (w < n_items) st[f[u]] = r;
  }
}

constexpr int kUd = 4;

static int diag_impl(void* stream, void* state, int n, int kind, int q0, int q1,
                     const float2* ddev, const M2& dv) {
  if (int rc = check_n(n)) return rc;
  if (!state) return fail(QSIM_EINVAL, "state is NULL");
  int t = q0, hi = 0;
  if (kind == 0) {
    if (int rc = check_q(q0, n, "target_qubit")) return rc;
  } else {
    if (int rc = check_q(q0, n, "qubit0")) return rc;
    if (int rc = check_q(q1, n, "qubit1")) return rc;
    if (q0 == q1) return fail(QSIM_EINVAL, "qubit0 and qubit1 must differ");
    t = std::min(q0, q1);
    hi = std::max(q0, q1);
  }
  const uint64_t n4 = (1ull << n) / 2;
  const uint64_t blocks = ceil_div(n4, (uint64_t)kThreads * kUd);
  hipStream_t s = as_stream(stream);
  if (n <= 32) {
    if (kind == 0)
      hipLaunchKernelGGL((k_diag<uint32_t, kUd, 0>), dim3(blocks), dim3(kThreads), 0, s,
                         (float4*)state, (uint32_t)n4, t, hi, ddev, dv);
    else
      hipLaunchKernelGGL((k_diag<uint32_t, kUd, 1>), dim3(blocks), dim3(kThreads), 0, s,
                         (float4*)state, (uint32_t)n4, t, hi, ddev, dv);
  } else {
    if (kind == 0)
      hipLaunchKernelGGL((k_diag<uint64_t, kUd, 0>), dim3(blocks), dim3(kThreads), 0, s,
                         (float4*)state, n4, t, hi, ddev, dv);
    else
      hipLaunchKernelGGL((k_diag<uint64_t, kUd, 1>), dim3(blocks), dim3(kThreads), 0, s,
                         (float4*)state, n4, t, hi, ddev, dv);
  }
  return launch_status("apply_diagonal");
}

static int phase_impl(void* stream, void* state, int n, float2 ph, int t) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_q(t, n, "target_qubit")) return rc;
  if (!state) return fail(QSIM_EINVAL, "state is NULL");
  const uint64_t items = t > 0 ? (1ull << n) / 4 : (1ull << n) / 2;
  const uint64_t blocks = ceil_div(items, (uint64_t)kThreads * kUd);
  if (n <= 32)
    hipLaunchKernelGGL((k_phase<uint32_t, kUd>), dim3(blocks), dim3(kThreads), 0,
                       as_stream(stream), (float4*)state, (uint32_t)items, t, ph);
  else
    hipLaunchKernelGGL((k_phase<uint64_t, kUd>), dim3(blocks), dim3(kThreads), 0,
                       as_stream(stream), (float4*)state, items, t, ph);
  return launch_status("apply_diagonal_phase");
}

// ============================================================ multi-controlled
// The affected qubits (controls + targets) are sorted ascending; item p's base
// index inserts a zero at each of them in turn and then sets the control bits
// (bwk/csrc/apply_mcgate_sm120.cu:64-91 describes the same insert-zeros rule).
struct McArgs {
  int pos[10];
  int n_pos;
  int t0, t1;        // targets (t1 unused for the 1-target ops)
  int op;            // 0: dense 2x2, 1: X, 2: dense 4x4, 3: SWAP
  uint64_t cmask;
};

template <typename I>
__device__ __forceinline__ I mc_base(I p, const McArgs& a) {
  I r = p;
#pragma unroll
  for (int j = 0; j < 10; ++j)
    if (j < a.n_pos) r = ins0(r, a.pos[j]);
  return r | (I)a.cmask;
}

// Scalar float2 path (works for any qubit set).  Each thread handles U items.
template <typename I, int U>
__global__ __launch_bounds__(kThreads) void k_mc(float2* __restrict__ st, I n_items, McArgs a,
                                                 const float2* __restrict__ gdev) {
  float2 g[16];
  if (a.op == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) g[k] = gdev[k];
  } else if (a.op == 2) {
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k] = gdev[k];
  }
  const I base = (I)blockIdx.x * (kThreads * U) + threadIdx.x;
  const I m0 = I(1) << a.t0;
  const I m1 = (a.op >= 2) ? (I(1) << a.t1) : I(0);
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const I p = base + (I)u * kThreads;
    if (p >= n_items) return;
    const I i0 = mc_base(p, a);
    if (a.op < 2) {
      const float2 x0 = st[i0], x1 = st[i0 | m0];
      if (a.op == 1) {
        st[i0] = x1; st[i0 | m0] = x0;
      } else {
        st[i0] = crow2(g[0], x0, g[1], x1);
        st[i0 | m0] = crow2(g[2], x0, g[3], x1);
      }
    } else {
      // r = 2*bit(t1) + bit(t0), t0 < t1
      float2 x0 = st[i0], x1 = st[i0 | m0], x2 = st[i0 | m1], x3 = st[i0 | m0 | m1];
      if (a.op == 3) {
        st[i0 | m0] = x2; st[i0 | m1] = x1;
      } else {
        mat4(g, x0, x1, x2, x3);
        st[i0] = x0; st[i0 | m0] = x1; st[i0 | m1] = x2; st[i0 | m0 | m1] = x3;
      }
    }
  }
}

// float4 path: valid when the lowest affected qubit is >= 1, so items 2w and
// 2w+1 are adjacent amplitudes and share one 16-byte slot.  The host passes
// positions, targets and control mask already shifted to float4 units.
template <typename I, int U>
__global__ __launch_bounds__(kThreads) void k_mc4(float4* __restrict__ st, I n_items, McArgs a,
                                                  const float2* __restrict__ gdev) {
  float2 g[16];
  if (a.op == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) g[k] = gdev[k];
  } else if (a.op == 2) {
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k] = gdev[k];
  }
  const I base = (I)blockIdx.x * (kThreads * U) + threadIdx.x;
  const I m0 = I(1) << a.t0;
  const I m1 = (a.op >= 2) ? (I(1) << a.t1) : I(0);
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const I w = base + (I)u * kThreads;
    if (w >= n_items) return;
    const I f0 = mc_base(w, a);
    if (a.op < 2) {
      const float4 x0 = st[f0], x1 = st[f0 | m0];
      if (a.op == 1) {
        st[f0] = x1; st[f0 | m0] = x0;
      } else {
        st[f0] = cat4(crow2(g[0], lo2(x0), g[1], lo2(x1)), crow2(g[0], hi2(x0), g[1], hi2(x1)));
        st[f0 | m0] = cat4(crow2(g[2], lo2(x0), g[3], lo2(x1)), crow2(g[2], hi2(x0), g[3], hi2(x1)));
      }
    } else {
      float4 x0 = st[f0], x1 = st[f0 | m0], x2 = st[f0 | m1], x3 = st[f0 | m0 | m1];
      if (a.op == 3) {
        st[f0 | m0] = x2; st[f0 | m1] = x1;
      } else {
        float2 l0 = lo2(x0), l1 = lo2(x1), l2 = lo2(x2), l3 = lo2(x3);
        float2 h0 = hi2(x0), h1 = hi2(x1), h2 = hi2(x2), h3 = hi2(x3);
        mat4(g, l0, l1, l2, l3);
        mat4(g, h0, h1, h2, h3);
        st[f0] = cat4(l0, h0); st[f0 | m0] = cat4(l1, h1);
        st[f0 | m1] = cat4(l2, h2); st[f0 | m0 | m1] = cat4(l3, h3);
      }
    }
  }
}

constexpr int kUm = 2;

static int mc_impl(void* stream, void* state, int n, const float2* gdev, const int64_t* controls,
                   int n_controls, int t0, int t1, int op) {
  if (int rc = check_n(n)) return rc;
  if (!state) return fail(QSIM_EINVAL, "state is NULL");
  if ((op == 0 || op == 2) && !gdev) return fail(QSIM_EINVAL, "gate is NULL");
  const int nt = op >= 2 ? 2 : 1;
  int tg[2] = {t0, t1};
  if (op >= 2) {
    if (t0 == t1) return fail(QSIM_EINVAL, "target0 and target1 must differ");
    tg[0] = std::min(t0, t1);
    tg[1] = std::max(t0, t1);
  }
  McSpec spec;
  if (int rc = build_mc(spec, n, controls, n_controls, tg, nt)) return rc;
  McArgs a;
  memset(&a, 0, sizeof a);
  for (int j = 0; j < spec.n_pos; ++j) a.pos[j] = spec.pos[j];
  a.n_pos = spec.n_pos;
  a.cmask = spec.cmask;
  a.t0 = tg[0];
  a.t1 = tg[1];
  a.op = op;
  hipStream_t s = as_stream(stream);
  const bool wide = n > 32;
  if (spec.pos[0] >= 1 && n - spec.n_pos >= 1) {
    const uint64_t items = 1ull << (n - spec.n_pos - 1);
    for (int j = 0; j < a.n_pos; ++j) a.pos[j] -= 1;
    a.cmask >>= 1;
    a.t0 -= 1;
    a.t1 = op >= 2 ? a.t1 - 1 : 0;
    const uint64_t blocks = ceil_div(items, (uint64_t)kThreads * kUm);
    if (!wide)
      hipLaunchKernelGGL((k_mc4<uint32_t, kUm>), dim3(blocks), dim3(kThreads), 0, s,
                         (float4*)state, (uint32_t)items, a, gdev);
    else
      hipLaunchKernelGGL((k_mc4<uint64_t, kUm>), dim3(blocks), dim3(kThreads), 0, s,
                         (float4*)state, items, a, gdev);
  } else {
    const uint64_t items = 1ull << (n - spec.n_pos);
    const uint64_t blocks = ceil_div(items, (uint64_t)kThreads * kUm);
    if (!wide)
      hipLaunchKernelGGL((k_mc<uint32_t, kUm>), dim3(blocks), dim3(kThreads), 0, s,
                         (float2*)state, (uint32_t)items, a, gdev);
    else
      hipLaunchKernelGGL((k_mc<uint64_t, kUm>), dim3(blocks), dim3(kThreads), 0, s,
                         (float2*)state, items, a, gdev);
  }
  return launch_status("apply_mc");
}

// ============================================================ host helpers
static M2 m2_from_doubles(const double* v, int count) {
  M2 m;
  for (int k = 0; k < 4; ++k) m.g[k] = k < count ? make_float2((float)v[2 * k], (float)v[2 * k + 1])
                                                 : make_float2(0.f, 0.f);
  return m;
}

}  // namespace qsim

using namespace qsim;

// ============================================================ C ABI
extern "C" {

const char* qsim_version(void) { return "qsim-gfx950 0.1.0"; }
int qsim_abi_version(void) { return 1; }
const char* qsim_last_error(void) { return g_err; }

int qsim_device_count(int* count) {
  if (!count) return fail(QSIM_EINVAL, "count is NULL");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *count = (e == hipSuccess) ? c : 0;
  return QSIM_OK;
}

int qsim_state_init(void* stream, void* state, int n_qubits) {
  return init_impl(stream, state, 1, n_qubits);
}
int qsim_state_init_batched(void* stream, void* states, int64_t n_states, int n_qubits) {
  if (n_states < 0) return fail(QSIM_EINVAL, "n_states < 0");
  return init_impl(stream, states, (uint64_t)n_states, n_qubits);
}

int qsim_apply_gate(void* stream, void* state, int n_qubits, const void* gate_dev, int target) {
  if (!gate_dev) return fail(QSIM_EINVAL, "gate is NULL");
  M2 dummy = {};
  return gate1q_impl(stream, state, 1, n_qubits, (const float2*)gate_dev, dummy, target);
}
int qsim_apply_gate_h(void* stream, void* state, int n_qubits, const double* gate8, int target) {
  if (!gate8) return fail(QSIM_EINVAL, "gate is NULL");
  return gate1q_impl(stream, state, 1, n_qubits, nullptr, m2_from_doubles(gate8, 4), target);
}
int qsim_apply_gate_batched(void* stream, void* states, int64_t n_states, int n_qubits,
                            const void* gate_dev, int target) {
  if (!gate_dev) return fail(QSIM_EINVAL, "gate is NULL");
  if (n_states < 0) return fail(QSIM_EINVAL, "n_states < 0");
  M2 dummy = {};
  return gate1q_impl(stream, states, (uint64_t)n_states, n_qubits, (const float2*)gate_dev, dummy,
                     target);
}

int qsim_apply_gate_2q(void* stream, void* state, int n_qubits, const void* gate_dev, int q0, int q1) {
  if (!gate_dev) return fail(QSIM_EINVAL, "gate is NULL");
  M4 dummy = {};
  return gate2q_impl(stream, state, n_qubits, (const float2*)gate_dev, dummy, q0, q1);
}
int qsim_apply_gate_2q_h(void* stream, void* state, int n_qubits, const double* gate32, int q0,
                         int q1) {
  if (!gate32) return fail(QSIM_EINVAL, "gate is NULL");
  M4 m;
  for (int k = 0; k < 16; ++k) m.g[k] = make_float2((float)gate32[2 * k], (float)gate32[2 * k + 1]);
  return gate2q_impl(stream, state, n_qubits, nullptr, m, q0, q1);
}

int qsim_apply_diagonal(void* stream, void* state, int n_qubits, const void* diag_dev, int target) {
  if (!diag_dev) return fail(QSIM_EINVAL, "diag is NULL");
  M2 dummy = {};
  return diag_impl(stream, state, n_qubits, 0, target, 0, (const float2*)diag_dev, dummy);
}
int qsim_apply_diagonal_h(void* stream, void* state, int n_qubits, const double* diag4, int target) {
  if (!diag4) return fail(QSIM_EINVAL, "diag is NULL");
  return diag_impl(stream, state, n_qubits, 0, target, 0, nullptr, m2_from_doubles(diag4, 2));
}
int qsim_apply_diagonal_phase(void* stream, void* state, int n_qubits, double phase_re,
                              double phase_im, int target) {
  return phase_impl(stream, state, n_qubits, make_float2((float)phase_re, (float)phase_im), target);
}
int qsim_apply_diagonal_2q(void* stream, void* state, int n_qubits, const void* diag_dev, int q0,
                           int q1) {
  if (!diag_dev) return fail(QSIM_EINVAL, "diag is NULL");
  M2 dummy = {};
  return diag_impl(stream, state, n_qubits, 1, q0, q1, (const float2*)diag_dev, dummy);
}

int qsim_apply_mcgate(void* stream, void* state, int n_qubits, const void* gate_dev,
                      const int64_t* controls, int n_controls, int target) {
  return mc_impl(stream, state, n_qubits, (const float2*)gate_dev, controls, n_controls, target, 0, 0);
}
int qsim_apply_mc2qgate(void* stream, void* state, int n_qubits, const void* gate_dev,
                        const int64_t* controls, int n_controls, int t0, int t1) {
  return mc_impl(stream, state, n_qubits, (const float2*)gate_dev, controls, n_controls, t0, t1, 2);
}
int qsim_apply_mcx(void* stream, void* state, int n_qubits, const int64_t* controls,
                   int n_controls, int target) {
  return mc_impl(stream, state, n_qubits, nullptr, controls, n_controls, target, 0, 1);
}
int qsim_apply_mcswap(void* stream, void* state, int n_qubits, const int64_t* controls,
                      int n_controls, int q0, int q1) {
  return mc_impl(stream, state, n_qubits, nullptr, controls, n_controls, q0, q1, 3);
}

}  // extern "C"
