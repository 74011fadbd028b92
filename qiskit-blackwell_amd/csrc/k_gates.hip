// Per-gate streaming kernels for gfx950: state init, dense 1Q/2Q, diagonal,
// multi-controlled, permutation and batched gates, plus the library's error
// plumbing.
//
// Streaming layout.  Amplitudes move as float4 slots (two adjacent complex64,
// 16 B per lane) and every load/store instruction of a wave covers 64
// consecutive slots (1 KiB), with nontemporal hints: the state is touched once
// per gate and must not evict the L2/MALL working set.  How a gate finds the
// partner of an amplitude depends on the target bit t:
//   t = 0      the pair sits inside one float4;
//   t = 1..6   the pair sits in lane ^ 2^(t-1) of the same wave (cross-lane
//              shuffle), so the wave still reads one contiguous KiB;
//   t >= 7     two lane-contiguous streams 2^(t-1) slots apart:
//              f0 = ins0(w, t-1), f1 = f0 + 2^(t-1).
// The insert-zero rule is the reference's pair enumeration
// (bwk/csrc/apply_gate_sm120.cu:100-104) expressed in float4 slots.  Grids
// of the stream modes are renumbered XCD-major (xcd_block).  Index math is
// 32-bit for states up to 2^32 amplitudes and 64-bit above (template I).
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
template <typename I, bool NT>
__global__ __launch_bounds__(kThreads) void k_init(float4* __restrict__ st, I n4, I per_state4, PassOrder po) {
  bool tail = false;
  const I f = pass_blk<I, NT, false>(po, tail) * kThreads + threadIdx.x;
  if (f >= n4) return;
  const bool first = (f & (per_state4 - 1)) == 0;
  st4p<NT>(st + f, make_float4(first ? 1.f : 0.f, 0.f, 0.f, 0.f), tail);
}

static int init_impl(void* stream, void* states, uint64_t n_states, int n) {
  if (int rc = check_n(n)) return rc;
  if (!states) return fail(QSIM_EINVAL, "state is NULL");
  if (n_states == 0) return QSIM_OK;
  const uint64_t per4 = (1ull << n) / 2;
  const uint64_t n4 = per4 * n_states;
  const uint64_t blocks = ceil_div(n4, kThreads);
  if (n4 > 0xffffffffull)
    hipLaunchKernelGGL((k_init<uint64_t, true>), grid1d(blocks), dim3(kThreads), 0, as_stream(stream),
                       (float4*)states, n4, per4, next_pass(n4 * 16));
  else if (nt_for_bytes(n4 * 16))
    hipLaunchKernelGGL((k_init<uint32_t, true>), grid1d(blocks), dim3(kThreads), 0, as_stream(stream),
                       (float4*)states, (uint32_t)n4, (uint32_t)per4, next_pass(n4 * 16));
  else
    hipLaunchKernelGGL((k_init<uint32_t, false>), grid1d(blocks), dim3(kThreads), 0, as_stream(stream),
                       (float4*)states, (uint32_t)n4, (uint32_t)per4, PassOrder{0, 0});
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

// MODE 0: t = 0, MODE 1: t = 1..6 (cross-lane), MODE 2: t >= 7 (streams).
// n_work = float4 slots (modes 0, 1) or slot pairs (mode 2).  Batched states
// need no extra index math: the flat slot index has the state number above
// bit n-2 and t < n keeps every pair inside one state.
template <typename I, int MODE, bool NT>
__global__ __launch_bounds__(kThreads) void k_gate1q(float4* __restrict__ st, I n_work, int t,
                                                     const float2* __restrict__ gdev, M2 gv, PassOrder po) {
  bool tail = false;
  float2 g[4];
  load_m2(gdev, gv, g);
  const I w0 = pass_blk<I, NT, (MODE >= 2)>(po, tail) * kThreads + threadIdx.x;
  const bool live = w0 < n_work;
  const I w = live ? w0 : n_work - 1;
  if (MODE == 0) {
    const float4 v = ld4<NT>(st + w);
    const float4 r = cat4(crow2(g[0], lo2(v), g[1], hi2(v)), crow2(g[2], lo2(v), g[3], hi2(v)));
    if (live) st4p<NT>(st + w, r, tail);
  } else if (MODE == 1) {
    // lane l holds slot w with w & 63 == l; bit (t-1) of the slot index is
    // bit t of the amplitude index.  Partner lanes of live lanes are live
    // (n_work is a power of two larger than 2^(t-1)).
    const float4 v = ld4<NT>(st + w);
    const int m = 1 << (t - 1);
    float4 p;
    with_lane_bit(m, [&](auto M) { p = xor4c<decltype(M)::value>(v); });
    const bool low = (threadIdx.x & m) == 0;
    const float2 ga = low ? g[0] : g[3], gb = low ? g[1] : g[2];
    const float4 r = cat4(crow2(ga, lo2(v), gb, lo2(p)), crow2(ga, hi2(v), gb, hi2(p)));
    if (live) st4p<NT>(st + w, r, tail);
  } else {
    const int s = t - 1;
    const I f0 = ins0(w, s), f1 = f0 + (I(1) << s);
    const float4 v0 = ld4<NT>(st + f0), v1 = ld4<NT>(st + f1);
    const float4 r0 = cat4(crow2(g[0], lo2(v0), g[1], lo2(v1)), crow2(g[0], hi2(v0), g[1], hi2(v1)));
    const float4 r1 = cat4(crow2(g[2], lo2(v0), g[3], lo2(v1)), crow2(g[2], hi2(v0), g[3], hi2(v1)));
    if (live) { st4p<NT>(st + f0, r0, tail); st4p<NT>(st + f1, r1, tail); }
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
  const uint64_t slots = (1ull << n) / 2 * n_states;
  // a second pair 32 KiB away per thread for far targets measured 2-6
  // points slower at 30 qubits (t = 19: 74 % against 81 % of the HBM peak),
  // so every stream target runs one pair per thread
  const int mode = t == 0 ? 0 : (t < 7 ? 1 : 2);
  const uint64_t work = mode == 2 ? slots / 2 : slots;
  const PassOrder po = nt_for_bytes(slots * 16) ? next_pass(slots * 16) : PassOrder{0, 0};
  const dim3 grid = grid1d(ceil_div(work, kThreads)), block(kThreads);
  if (slots > 0xffffffffull) {
    if (mode == 0) hipLaunchKernelGGL((k_gate1q<uint64_t, 0, true>), grid, block, 0, s, (float4*)states, work, t, gdev, gv, po);
    if (mode == 1) hipLaunchKernelGGL((k_gate1q<uint64_t, 1, true>), grid, block, 0, s, (float4*)states, work, t, gdev, gv, po);
    if (mode == 2) hipLaunchKernelGGL((k_gate1q<uint64_t, 2, true>), grid, block, 0, s, (float4*)states, work, t, gdev, gv, po);
  } else if (nt_for_bytes(slots * 16)) {
    const uint32_t w = (uint32_t)work;
    if (mode == 0) hipLaunchKernelGGL((k_gate1q<uint32_t, 0, true>), grid, block, 0, s, (float4*)states, w, t, gdev, gv, po);
    if (mode == 1) hipLaunchKernelGGL((k_gate1q<uint32_t, 1, true>), grid, block, 0, s, (float4*)states, w, t, gdev, gv, po);
    if (mode == 2) hipLaunchKernelGGL((k_gate1q<uint32_t, 2, true>), grid, block, 0, s, (float4*)states, w, t, gdev, gv, po);
  } else {
    const uint32_t w = (uint32_t)work;
    if (mode == 0) hipLaunchKernelGGL((k_gate1q<uint32_t, 0, false>), grid, block, 0, s, (float4*)states, w, t, gdev, gv, po);
    if (mode == 1) hipLaunchKernelGGL((k_gate1q<uint32_t, 1, false>), grid, block, 0, s, (float4*)states, w, t, gdev, gv, po);
    if (mode == 2) hipLaunchKernelGGL((k_gate1q<uint32_t, 2, false>), grid, block, 0, s, (float4*)states, w, t, gdev, gv, po);
  }
  return launch_status("apply_gate");
}

// ============================================================ 2Q dense
// Gate row/column index r = 2*bit(hi) + bit(lo), lo < hi
// (bwk/csrc/apply_gate2q_sm120.cu:11-20, 245-246).  S = number of targets
// >= 7, each of which becomes a lane-contiguous stream; a target in 1..6 is
// reached by a cross-lane shuffle and target 0 inside the float4.  Each lane
// owns 2^S slots and computes the rows of the amplitudes it holds.
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

// Lane-permuted rows, selected without dynamic register indexing.  A lane
// whose amplitudes sit at matrix row r keeps its inputs in lane-relative
// order x_k (column k ^ r), so it needs h[k] = g[r][k ^ r].  The compensated
// dot sums the columns as ((c0 + c1) + (c2 + c3)), a tree that an XOR of the
// column index maps onto itself, so the permuted form rounds identically.
// perm_row: r = 2*b1 + b0.
__device__ __forceinline__ void perm_row(const float2 g[16], bool b1, bool b0, float2 h[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float2 r0 = b0 ? g[4 + (k ^ 1)] : g[k];                 // rows 1 / 0
    const float2 r1 = b0 ? g[12 + (k ^ 3)] : g[8 + (k ^ 2)];      // rows 3 / 2
    h[k] = b1 ? r1 : r0;
  }
}
// perm_rows2: rows b0 and 2 + b0 with columns k ^ b0.
__device__ __forceinline__ void perm_rows2(const float2 g[16], bool b0, float2 ha[4], float2 hb[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    ha[k] = b0 ? g[4 + (k ^ 1)] : g[k];
    hb[k] = b0 ? g[12 + (k ^ 1)] : g[8 + k];
  }
}
// perm_rows0: rows 2*b1 and 2*b1 + 1 with columns k ^ 2*b1.
__device__ __forceinline__ void perm_rows0(const float2 g[16], bool b1, float2 ha[4], float2 hb[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    ha[k] = b1 ? g[8 + (k ^ 2)] : g[k];
    hb[k] = b1 ? g[12 + (k ^ 2)] : g[4 + k];
  }
}

template <typename I, int S, bool NT>
__global__ __launch_bounds__(kThreads) void k_gate2q(float4* __restrict__ st, I n_work, int lo,
                                                     int hi, const float2* __restrict__ gdev,
                                                     M4 gv, PassOrder po) {
  float2 g[16];
  load_m4(gdev, gv, g);
  bool tail = false;
  // streamed states take the XCD-major pass order for every S, so that a
  // both-low-target gate starts where the stream gates before it ended (the
  // 30-qubit SU(4) step measured 2.8 ms shorter than with natural order for S = 0)
  const I blk = NT ? pass_block<I, true>(po, tail) : (S > 0 ? xcd_block<I>() : gblock<I>());
  const I w0 = blk * kThreads + threadIdx.x;
  const bool live = w0 < n_work;
  const I w = live ? w0 : n_work - 1;
  const int lane = threadIdx.x & 63;
  if (S == 2) {
    const I ml = I(1) << (lo - 1), mh = I(1) << (hi - 1);
    const I f0 = ins0(ins0(w, lo - 1), hi - 1), f1 = f0 + ml, f2 = f0 + mh, f3 = f1 + mh;
    const float4 v0 = ld4<NT>(st + f0), v1 = ld4<NT>(st + f1), v2 = ld4<NT>(st + f2), v3 = ld4<NT>(st + f3);
    float2 a0 = lo2(v0), a1 = lo2(v1), a2 = lo2(v2), a3 = lo2(v3);
    float2 b0 = hi2(v0), b1 = hi2(v1), b2 = hi2(v2), b3 = hi2(v3);
    mat4(g, a0, a1, a2, a3);
    mat4(g, b0, b1, b2, b3);
    if (live) {
      st4p<NT>(st + f0, cat4(a0, b0), tail); st4p<NT>(st + f1, cat4(a1, b1), tail);
      st4p<NT>(st + f2, cat4(a2, b2), tail); st4p<NT>(st + f3, cat4(a3, b3), tail);
    }
  } else if (S == 1) {
    // hi is a stream: slot f0 has bit(hi) = 0, slot f1 bit(hi) = 1
    const I f0 = ins0(w, hi - 1), f1 = f0 + (I(1) << (hi - 1));
    const float4 v0 = ld4<NT>(st + f0), v1 = ld4<NT>(st + f1);
    float4 o0, o1;
    if (lo == 0) {
      float2 a0 = lo2(v0), a1 = hi2(v0), a2 = lo2(v1), a3 = hi2(v1);
      mat4(g, a0, a1, a2, a3);
      o0 = cat4(a0, a1);
      o1 = cat4(a2, a3);
    } else {
      // Inputs in lane-relative order: x_k has matrix column k ^ bl (self,
      // lo partner, then the same two with bit hi set); the rows are
      // permuted to match once per lane, so no per-amplitude select remains.
      const int m = 1 << (lo - 1);
      const bool bl = (lane & m) != 0;
      float4 p0, p1;
      with_lane_bit(m, [&](auto M) {
        p0 = xor4c<decltype(M)::value>(v0);
        p1 = xor4c<decltype(M)::value>(v1);
      });
      float2 ha[4], hb[4];
      perm_rows2(g, bl, ha, hb);
      float2 y[2][2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float2 x0 = h ? hi2(v0) : lo2(v0), x1 = h ? hi2(p0) : lo2(p0);
        const float2 x2 = h ? hi2(v1) : lo2(v1), x3 = h ? hi2(p1) : lo2(p1);
        y[h][0] = crow4(ha, x0, x1, x2, x3);
        y[h][1] = crow4(hb, x0, x1, x2, x3);
      }
      o0 = cat4(y[0][0], y[1][0]);
      o1 = cat4(y[0][1], y[1][1]);
    }
    if (live) { st4p<NT>(st + f0, o0, tail); st4p<NT>(st + f1, o1, tail); }
  } else {
    const float4 v = ld4<NT>(st + w);
    const int mh = 1 << (hi - 1);
    const bool bh = (lane & mh) != 0;
    float4 o;
    if (lo == 0) {
      // x = (lo2(v), hi2(v), lo2(p), hi2(p)) have columns k ^ 2*bh
      const float4 p = shfl_xor4_lds(v, mh);
      float2 ha[4], hb[4];
      perm_rows0(g, bh, ha, hb);
      o = cat4(crow4(ha, lo2(v), hi2(v), lo2(p), hi2(p)),
               crow4(hb, lo2(v), hi2(v), lo2(p), hi2(p)));
    } else {
      // x = (self, lo partner, hi partner, both) have columns k ^ r,
      // r = 2*bh + bl the row of the amplitudes this lane holds
      const int ml = 1 << (lo - 1);
      const bool bl = (lane & ml) != 0;
      const float4 pl = shfl_xor4_lds(v, ml), ph = shfl_xor4_lds(v, mh), pb = shfl_xor4_lds(v, ml | mh);
      float2 h[4];
      perm_row(g, bh, bl, h);
      o = cat4(crow4(h, lo2(v), lo2(pl), lo2(ph), lo2(pb)),
               crow4(h, hi2(v), hi2(pl), hi2(ph), hi2(pb)));
    }
    if (live) st4p<NT>(st + w, o, tail);
  }
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
  const int S = (lo >= 7) + (hi >= 7);
  const uint64_t work = ((1ull << n) / 2) >> S;
  const dim3 grid = grid1d(ceil_div(work, kThreads)), block(kThreads);
  const PassOrder po = nt_for_qubits(n) ? next_pass(8ull << n) : PassOrder{0, 0};
  if (n > 32) {
    if (S == 0) hipLaunchKernelGGL((k_gate2q<uint64_t, 0, true>), grid, block, 0, s, (float4*)state, work, lo, hi, gdev, gv, po);
    if (S == 1) hipLaunchKernelGGL((k_gate2q<uint64_t, 1, true>), grid, block, 0, s, (float4*)state, work, lo, hi, gdev, gv, po);
    if (S == 2) hipLaunchKernelGGL((k_gate2q<uint64_t, 2, true>), grid, block, 0, s, (float4*)state, work, lo, hi, gdev, gv, po);
  } else if (nt_for_qubits(n)) {
    const uint32_t w = (uint32_t)work;
    if (S == 0) hipLaunchKernelGGL((k_gate2q<uint32_t, 0, true>), grid, block, 0, s, (float4*)state, w, lo, hi, gdev, gv, po);
    if (S == 1) hipLaunchKernelGGL((k_gate2q<uint32_t, 1, true>), grid, block, 0, s, (float4*)state, w, lo, hi, gdev, gv, po);
    if (S == 2) hipLaunchKernelGGL((k_gate2q<uint32_t, 2, true>), grid, block, 0, s, (float4*)state, w, lo, hi, gdev, gv, po);
  } else {
    const uint32_t w = (uint32_t)work;
    if (S == 0) hipLaunchKernelGGL((k_gate2q<uint32_t, 0, false>), grid, block, 0, s, (float4*)state, w, lo, hi, gdev, gv, po);
    if (S == 1) hipLaunchKernelGGL((k_gate2q<uint32_t, 1, false>), grid, block, 0, s, (float4*)state, w, lo, hi, gdev, gv, po);
    if (S == 2) hipLaunchKernelGGL((k_gate2q<uint32_t, 2, false>), grid, block, 0, s, (float4*)state, w, lo, hi, gdev, gv, po);
  }
  return launch_status("apply_gate_2q");
}

// ============================================================ diagonals
// kind 0: diag(d0, d1) on bit t;  kind 1: diag(d00, d01, d10, d11) on (lo, hi).
template <typename I, int U, int KIND, bool NT>
__global__ __launch_bounds__(kThreads) void k_diag(float4* __restrict__ st, I n4, int t, int hi,
                                                   const float2* __restrict__ ddev, M2 dv, PassOrder po) {
  bool tail = false;
  constexpr int kind = KIND;
  float2 d0, d1, d2, d3;
  if (ddev) {
    d0 = ddev[0]; d1 = ddev[1];
    if (KIND == 1) { d2 = ddev[2]; d3 = ddev[3]; } else { d2 = d0; d3 = d1; }
  } else {
    d0 = dv.g[0]; d1 = dv.g[1]; d2 = dv.g[2]; d3 = dv.g[3];
  }
  const I base = pass_blk<I, NT, false>(po, tail) * (kThreads * U) + threadIdx.x;
  float4 v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    I f = base + (I)u * kThreads;
    v[u] = ld4<NT>(st + (f < n4 ? f : n4 - 1));
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
    if (f < n4) st4p<NT>(st + f, r, tail);
  }
}

// Phase on bit t = 1 only: t >= 1 touches slot ins0(w, t-1) | 2^(t-1); t = 0
// touches the .hi half of every slot (the cache line is read either way).
// Phase on the bit-t = 1 half.  MODE 0: t = 0, every slot (its odd
// amplitude changes).  MODE 1: t = 1..3, every slot: the slots with bit t set
// come in runs of 16-64 B, so whole lines are rewritten (the untouched slots
// store what they loaded; measured faster than touching the runs alone).
// MODE 2: t >= 4, only the slots with bit t set, in whole 128 B lines.  All
// accesses are nontemporal.
template <typename I, int U, int MODE, bool NT>
__global__ __launch_bounds__(kThreads) void k_phase(float4* __restrict__ st, I n_items, int t,
                                                    float2 ph, PassOrder po) {
  bool tail = false;
  const I base = pass_blk<I, NT, false>(po, tail) * (kThreads * U) + threadIdx.x;
  const int s = t - 1;
  I f[U];
  float4 v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    I w = base + (I)u * kThreads;
    if (w >= n_items) w = n_items - 1;
    f[u] = MODE == 2 ? (ins0(w, s) | (I(1) << s)) : w;
    v[u] = ld4<NT>(st + f[u]);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const I w = base + (I)u * kThreads;
    float4 r;
    if (MODE == 0) r = cat4(lo2(v[u]), cmulc(hi2(v[u]), ph));
    else if (MODE == 1 && !((f[u] >> s) & 1)) r = v[u];
    else r = cat4(cmulc(lo2(v[u]), ph), cmulc(hi2(v[u]), ph));
    if (w < n_items) st4p<NT>(st + f[u], r, tail);
  }
}

constexpr int kUd = 1;

template <typename I, int KIND, bool NT>
static void launch_diag(hipStream_t s, uint64_t blocks, void* state, uint64_t n4, int t, int hi,
                        const float2* ddev, const M2& dv) {
  hipLaunchKernelGGL((k_diag<I, kUd, KIND, NT>), grid1d(blocks), dim3(kThreads), 0, s,
                     (float4*)state, (I)n4, t, hi, ddev, dv,
                     NT ? next_pass(16 * n4) : PassOrder{0, 0});
}

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
  if (n > 32) {
    if (kind == 0) launch_diag<uint64_t, 0, true>(s, blocks, state, n4, t, hi, ddev, dv);
    else launch_diag<uint64_t, 1, true>(s, blocks, state, n4, t, hi, ddev, dv);
  } else if (nt_for_qubits(n)) {
    if (kind == 0) launch_diag<uint32_t, 0, true>(s, blocks, state, n4, t, hi, ddev, dv);
    else launch_diag<uint32_t, 1, true>(s, blocks, state, n4, t, hi, ddev, dv);
  } else {
    if (kind == 0) launch_diag<uint32_t, 0, false>(s, blocks, state, n4, t, hi, ddev, dv);
    else launch_diag<uint32_t, 1, false>(s, blocks, state, n4, t, hi, ddev, dv);
  }
  return launch_status("apply_diagonal");
}

template <typename I, int MODE>
static void launch_phase(hipStream_t s, void* state, uint64_t items, int t, float2 ph, bool nt) {
  const dim3 grid = grid1d(ceil_div(items, (uint64_t)kThreads * kUd));
  if (nt)
    hipLaunchKernelGGL((k_phase<I, kUd, MODE, true>), grid, dim3(kThreads), 0, s, (float4*)state,
                       (I)items, t, ph, next_pass(16 * items));
  else
    hipLaunchKernelGGL((k_phase<I, kUd, MODE, false>), grid, dim3(kThreads), 0, s, (float4*)state,
                       (I)items, t, ph, PassOrder{0, 0});
}

static int phase_impl(void* stream, void* state, int n, float2 ph, int t) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_q(t, n, "target_qubit")) return rc;
  if (!state) return fail(QSIM_EINVAL, "state is NULL");
  const int mode = t == 0 ? 0 : (t < 4 ? 1 : 2);
  const uint64_t items = mode == 2 ? (1ull << n) / 4 : (1ull << n) / 2;
  hipStream_t s = as_stream(stream);
  const bool nt = nt_for_qubits(n);
  if (n <= 32) {
    if (mode == 0) launch_phase<uint32_t, 0>(s, state, items, t, ph, nt);
    if (mode == 1) launch_phase<uint32_t, 1>(s, state, items, t, ph, nt);
    if (mode == 2) launch_phase<uint32_t, 2>(s, state, items, t, ph, nt);
  } else {
    if (mode == 0) launch_phase<uint64_t, 0>(s, state, items, t, ph, nt);
    if (mode == 1) launch_phase<uint64_t, 1>(s, state, items, t, ph, nt);
    if (mode == 2) launch_phase<uint64_t, 2>(s, state, items, t, ph, nt);
  }
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
  const I base = gblock<I>() * (kThreads * U) + threadIdx.x;
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
    } else if (a.op == 3) {
      const float2 x1 = st[i0 | m0], x2 = st[i0 | m1];
      st[i0 | m0] = x2; st[i0 | m1] = x1;
    } else {
      // r = 2*bit(t1) + bit(t0), t0 < t1
      float2 x0 = st[i0], x1 = st[i0 | m0], x2 = st[i0 | m1], x3 = st[i0 | m0 | m1];
      {
        mat4(g, x0, x1, x2, x3);
        st[i0] = x0; st[i0 | m0] = x1; st[i0 | m1] = x2; st[i0 | m0 | m1] = x3;
      }
    }
  }
}

// float4 path: valid when the lowest affected qubit is >= 1, so items 2w and
// 2w+1 are adjacent amplitudes and share one 16-byte slot.  The host passes
// positions, targets and control mask already shifted to float4 units.
// NT: nontemporal accesses, used when every access run is a whole 128 B line
// (lowest affected qubit >= 4) and the state is larger than the MALL;
// shorter runs keep cached accesses, which merge the partial lines
// (nontemporal 32 B runs measured 30-40 % slower).
template <typename I, int U, bool NT>
__global__ __launch_bounds__(kThreads) void k_mc4(float4* __restrict__ st, I n_items, McArgs a,
                                                  const float2* __restrict__ gdev, PassOrder po) {
  bool tail = false;
  float2 g[16];
  if (a.op == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) g[k] = gdev[k];
  } else if (a.op == 2) {
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k] = gdev[k];
  }
  const I base = pass_blk<I, NT, false>(po, tail) * (kThreads * U) + threadIdx.x;
  const I m0 = I(1) << a.t0;
  const I m1 = (a.op >= 2) ? (I(1) << a.t1) : I(0);
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const I w = base + (I)u * kThreads;
    if (w >= n_items) return;
    const I f0 = mc_base(w, a);
    if (a.op < 2) {
      const float4 x0 = ld4<NT>(st + (f0)), x1 = ld4<NT>(st + (f0 | m0));
      if (a.op == 1) {
        st4p<NT>(st + (f0), x1, tail); st4p<NT>(st + (f0 | m0), x0, tail);
      } else {
        st4p<NT>(st + (f0), cat4(crow2(g[0], lo2(x0), g[1], lo2(x1)), crow2(g[0], hi2(x0), g[1], hi2(x1))), tail);
        st4p<NT>(st + (f0 | m0), cat4(crow2(g[2], lo2(x0), g[3], lo2(x1)), crow2(g[2], hi2(x0), g[3], hi2(x1))), tail);
      }
    } else if (a.op == 3) {
      // SWAP moves only the (1, 0) and (0, 1) amplitudes
      const float4 x1 = ld4<NT>(st + (f0 | m0)), x2 = ld4<NT>(st + (f0 | m1));
      st4p<NT>(st + (f0 | m0), x2, tail); st4p<NT>(st + (f0 | m1), x1, tail);
    } else {
      float4 x0 = ld4<NT>(st + (f0)), x1 = ld4<NT>(st + (f0 | m0)), x2 = ld4<NT>(st + (f0 | m1)), x3 = ld4<NT>(st + (f0 | m0 | m1));
      {
        float2 l0 = lo2(x0), l1 = lo2(x1), l2 = lo2(x2), l3 = lo2(x3);
        float2 h0 = hi2(x0), h1 = hi2(x1), h2 = hi2(x2), h3 = hi2(x3);
        mat4(g, l0, l1, l2, l3);
        mat4(g, h0, h1, h2, h3);
        st4p<NT>(st + (f0), cat4(l0, h0), tail); st4p<NT>(st + (f0 | m0), cat4(l1, h1), tail);
        st4p<NT>(st + (f0 | m1), cat4(l2, h2), tail); st4p<NT>(st + (f0 | m0 | m1), cat4(l3, h3), tail);
      }
    }
  }
}

// Controlled 1-target gates (X and dense 2x2) in the lane-contiguous layout
// of k_gate1q: one wave owns a window of 64 float4 slots (128 amplitudes,
// index bits 0..6), or a pair of windows 2^(t-7) apart when t >= 7.
//   * controls >= 7 choose the windows: the window index gets zeros inserted
//     at those bits (and at the target in MODE 2) and then has them set, so
//     only windows with every high control set are visited;
//   * controls 1..6 are slot bits inside the window.  On slot bits 3..5 a
//     control switches whole 128 B lines on or off, and on slot bit 2 whole
//     64 B halves, so only active lanes load and store.  On slot bits 0..1
//     (FULL) it splits every line into 16-32 B runs, and the wave rewrites
//     whole lines instead (inactive lanes store what they loaded): such
//     partial-line traffic costs more than the untouched half;
//   * a control on qubit 0 restricts the update to the odd amplitude (.zw)
//     of each slot.
struct Mc1Args {
  int t;
  int nh;                              // window-level zero insertions
  int hpos[QSIM_MAX_CONTROLS + 1];     // ascending, in window-index bits (qubit - 7)
  uint64_t hset;                       // window-index bits of the high controls
  uint32_t lmask;                      // slot bits of the controls 1..6
  int c0;                              // control on qubit 0
};

template <typename I>
__device__ __forceinline__ I mc1_window(I item, const Mc1Args& a) {
  I w = item;
#pragma unroll
  for (int j = 0; j < QSIM_MAX_CONTROLS + 1; ++j)
    if (j < a.nh) w = ins0(w, a.hpos[j]);
  return w | (I)a.hset;
}

__device__ __forceinline__ float4 pick4(bool c, float4 a, float4 b) { return c ? a : b; }

// OP 0: dense 2x2 (compensated), OP 1: X.
template <typename I, int MODE, int OP, int FULL, bool NT>
__global__ __launch_bounds__(kThreads) void k_mc1(float4* __restrict__ st, I n_items, Mc1Args a,
                                                  const float2* __restrict__ gdev, M2 gv, PassOrder po) {
  bool tail = false;
  float2 g[4];
  if (OP == 0) load_m2(gdev, gv, g);
  const I item = pass_blk<I, NT, true>(po, tail) * (kThreads / 64) + (threadIdx.x >> 6);
  if (item >= n_items) return;      // uniform per wave
  const int lane = threadIdx.x & 63;
  const I w = mc1_window(item, a);
  const I s0 = (w << 6) | (I)lane;
  const bool act = ((uint32_t)lane & a.lmask) == a.lmask;
  if (!FULL && !act && MODE != 1) return;
  if (MODE == 0) {
    const float4 v = ld4<NT>(st + s0);
    float4 r;
    if (OP == 1) r = cat4(hi2(v), lo2(v));
    else r = cat4(crow2(g[0], lo2(v), g[1], hi2(v)), crow2(g[2], lo2(v), g[3], hi2(v)));
    st4p<NT>(st + s0, FULL ? pick4(act, r, v) : r, tail);
  } else if (MODE == 1) {
    // the partner lane has the same control bits, so it is active with this one
    const bool ld = FULL || act;
    const float4 v = ld ? ld4<NT>(st + s0) : make_float4(0.f, 0.f, 0.f, 0.f);
    const int m = 1 << (a.t - 1);
    float4 p;
    with_lane_bit(m, [&](auto M) { p = xor4c<decltype(M)::value>(v); });
    if (!ld) return;
    float4 r;
    if (OP == 1) {
      r = p;
    } else {
      const bool low = (lane & m) == 0;
      const float2 ga = low ? g[0] : g[3], gb = low ? g[1] : g[2];
      r = cat4(crow2(ga, lo2(v), gb, lo2(p)), crow2(ga, hi2(v), gb, hi2(p)));
    }
    if (a.c0) r = cat4(lo2(v), hi2(r));
    st4p<NT>(st + s0, FULL ? pick4(act, r, v) : r, tail);
  } else {
    const I s1 = s0 + (I(1) << (a.t - 1));
    const float4 v0 = ld4<NT>(st + s0), v1 = ld4<NT>(st + s1);
    float4 r0, r1;
    if (OP == 1) {
      r0 = v1; r1 = v0;
    } else {
      r0 = cat4(crow2(g[0], lo2(v0), g[1], lo2(v1)), crow2(g[0], hi2(v0), g[1], hi2(v1)));
      r1 = cat4(crow2(g[2], lo2(v0), g[3], lo2(v1)), crow2(g[2], hi2(v0), g[3], hi2(v1)));
    }
    if (a.c0) { r0 = cat4(lo2(v0), hi2(r0)); r1 = cat4(lo2(v1), hi2(r1)); }
    st4p<NT>(st + s0, FULL ? pick4(act, r0, v0) : r0, tail);
    st4p<NT>(st + s1, FULL ? pick4(act, r1, v1) : r1, tail);
  }
}

template <typename I, int MODE, int OP, bool NT>
static void launch_mc1(hipStream_t s, void* state, uint64_t items, const Mc1Args& a, bool full,
                       const float2* gdev, const M2& gv) {
  const dim3 grid = grid1d(ceil_div(items, kThreads / 64)), block(kThreads);
  // bytes a pass touches: a 1 KiB window per item, two in mode 2
  const PassOrder po = NT ? next_pass(items * (MODE == 2 ? 2048 : 1024)) : PassOrder{0, 0};
  if (full)
    hipLaunchKernelGGL((k_mc1<I, MODE, OP, 1, NT>), grid, block, 0, s, (float4*)state, (I)items, a, gdev, gv, po);
  else
    hipLaunchKernelGGL((k_mc1<I, MODE, OP, 0, NT>), grid, block, 0, s, (float4*)state, (I)items, a, gdev, gv, po);
}

template <typename I, int OP, bool NT>
static void launch_mc1_mode(hipStream_t s, void* state, uint64_t items, const Mc1Args& a, int mode,
                            bool full, const float2* gdev, const M2& gv) {
  if (mode == 0) launch_mc1<I, 0, OP, NT>(s, state, items, a, full, gdev, gv);
  if (mode == 1) launch_mc1<I, 1, OP, NT>(s, state, items, a, full, gdev, gv);
  if (mode == 2) launch_mc1<I, 2, OP, NT>(s, state, items, a, full, gdev, gv);
}

// Returns false when the state is too small for 64-slot windows.
static bool mc1_launch(hipStream_t s, void* state, int n, const McSpec& spec, int t, int op,
                       const float2* gdev) {
  Mc1Args a;
  memset(&a, 0, sizeof a);
  a.t = t;
  const int mode = t == 0 ? 0 : (t < 7 ? 1 : 2);
  for (int j = 0; j < spec.n_pos; ++j) {
    const int q = spec.pos[j];
    if (q == t) {
      if (mode == 2) a.hpos[a.nh++] = q - 7;
      continue;
    }
    if (q == 0) a.c0 = 1;
    else if (q < 7) a.lmask |= 1u << (q - 1);
    else { a.hpos[a.nh++] = q - 7; a.hset |= 1ull << (q - 7); }
  }
  if (n < 8 || n - 7 - a.nh < 0) return false;
  const uint64_t items = 1ull << (n - 7 - a.nh);
  // controls on qubits 1..2 leave active runs of 16-32 B: the wave rewrites
  // whole lines; a control on qubit 3 leaves 64 B runs, which the active
  // lanes read and write alone (measured 11 % faster than whole lines)
  const bool full = (a.lmask & 3u) != 0;
  const bool wide = (1ull << (n - 1)) > 0xffffffffull;
  M2 gv = {};
  const bool nt = nt_for_qubits(n);
  if (op == 0) {
    if (wide) launch_mc1_mode<uint64_t, 0, true>(s, state, items, a, mode, full, gdev, gv);
    else if (nt) launch_mc1_mode<uint32_t, 0, true>(s, state, items, a, mode, full, gdev, gv);
    else launch_mc1_mode<uint32_t, 0, false>(s, state, items, a, mode, full, gdev, gv);
  } else {
    if (wide) launch_mc1_mode<uint64_t, 1, true>(s, state, items, a, mode, full, gdev, gv);
    else if (nt) launch_mc1_mode<uint32_t, 1, true>(s, state, items, a, mode, full, gdev, gv);
    else launch_mc1_mode<uint32_t, 1, false>(s, state, items, a, mode, full, gdev, gv);
  }
  return true;
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
  if (op < 2 && mc1_launch(s, state, n, spec, tg[0], op, gdev)) return launch_status("apply_mc");
  const bool wide = n > 32;
  if (spec.pos[0] >= 1 && n - spec.n_pos >= 1) {
    const uint64_t items = 1ull << (n - spec.n_pos - 1);
    for (int j = 0; j < a.n_pos; ++j) a.pos[j] -= 1;
    a.cmask >>= 1;
    a.t0 -= 1;
    a.t1 = op >= 2 ? a.t1 - 1 : 0;
    const uint64_t blocks = ceil_div(items, (uint64_t)kThreads * kUm);
    const bool nt = a.pos[0] >= 3 && nt_for_qubits(n);
    const PassOrder po = nt ? next_pass(items * (op >= 2 ? 64 : 32)) : PassOrder{0, 0};
    if (!wide && nt)
      hipLaunchKernelGGL((k_mc4<uint32_t, kUm, true>), grid1d(blocks), dim3(kThreads), 0, s,
                         (float4*)state, (uint32_t)items, a, gdev, po);
    else if (!wide)
      hipLaunchKernelGGL((k_mc4<uint32_t, kUm, false>), grid1d(blocks), dim3(kThreads), 0, s,
                         (float4*)state, (uint32_t)items, a, gdev, po);
    else if (nt)
      hipLaunchKernelGGL((k_mc4<uint64_t, kUm, true>), grid1d(blocks), dim3(kThreads), 0, s,
                         (float4*)state, items, a, gdev, po);
    else
      hipLaunchKernelGGL((k_mc4<uint64_t, kUm, false>), grid1d(blocks), dim3(kThreads), 0, s,
                         (float4*)state, items, a, gdev, po);
  } else {
    const uint64_t items = 1ull << (n - spec.n_pos);
    const uint64_t blocks = ceil_div(items, (uint64_t)kThreads * kUm);
    if (!wide)
      hipLaunchKernelGGL((k_mc<uint32_t, kUm>), grid1d(blocks), dim3(kThreads), 0, s,
                         (float2*)state, (uint32_t)items, a, gdev);
    else
      hipLaunchKernelGGL((k_mc<uint64_t, kUm>), grid1d(blocks), dim3(kThreads), 0, s,
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
int qsim_abi_version(void) { return 2; }
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
