// Pauli-string expectation values, Pauli-string application and the
// stochastic noise channels (one Kraus branch per call) for gfx950.
//
// A Pauli string is given by two masks: x (qubits carrying X or Y) and z
// (qubits carrying Z or Y).  With y = popcount(x & z) (the number of Y's),
//   P = i^y X^x Z^z,   (P psi)_i = i^y (-1)^popcount((i ^ x) & z) psi_{i ^ x}.
// Its expectation pairs every index i0 whose highest x bit is clear with its
// partner j = i0 ^ x.  With zc = conj(psi_i0) psi_j and s = (-1)^popcount(i0 & z):
//   <psi|P|psi> = sum_{i0} s * ( (-1)^(y/2)     * 2 Re zc   if y is even,
//                                (-1)^((y-1)/2) * 2 Im zc   if y is odd ),
// and for x = 0 it is sum_i s(i) |psi_i|^2.  Products of FP32 amplitudes are
// exact in FP64; every sum is FP64 in a fixed order (per lane, then a wave
// butterfly, then the waves of a workgroup, then the workgroups in index
// order), so results are deterministic.  The reference reduces in FP32 with
// atomicAdd and handles a Z string or one X/Y qubit only
// (bwk/csrc/expectation_sm120.cu:45-163).
//
// Memory layout follows the gate kernels: a wave owns a window of 64
// float4 slots (128 amplitudes, 1 KiB).  When x < 128 the partner of every
// amplitude is in the same window (a cross-lane swap by x >> 1, plus a swap
// inside the float4 when x & 1); when x >= 128 the wave owns the window pair
// (w, w ^ (x >> 7)) with bit h - 7 of w clear, h the highest bit of x.  Each
// amplitude is read once, by whole-line 1 KiB loads.
#include <algorithm>
#include <vector>

#include "qsim_internal.hpp"

namespace qsim {

constexpr int kExpMaxTerms = 8;   // terms sharing one x mask per state pass

struct ExpTerms {
  uint64_t z[kExpMaxTerms];
  double fac[kExpMaxTerms];       // +-1 (x = 0) or +-2 (x != 0)
  int out[kExpMaxTerms];          // output slot, -1 = discard (padding)
  int im;                         // every term of the pass takes Im(zc) (else Re(zc))
};

__device__ __forceinline__ double wsum64(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += lane_xor(v, m);
  return v;
}

// Pass 1.  Units are windows (KIND 0: x == 0; KIND 1: 0 < x < 128) or window
// pairs (KIND 2: x >= 128); workgroup b reduces units [b*chunk, (b+1)*chunk),
// wave k of it takes units k, k+4, ...  part[b*NT + j] = the workgroup's sum
// for term j.
//
// Amplitude i = 128 w + 2 lane + c (window w, lane, c = 0/1 in the float4),
// so the sign (-1)^popcount(i & z) factors into a lane constant (bits 0..6,
// folded with the term's factor and, for KIND 1, with the "highest x bit
// clear" selection into k_c) and a window sign that is the same for the
// whole wave (computed on the scalar unit).  Per slot and term that leaves
// t = v_0 k_0 + v_1 k_1 and acc += ws * t, with v_c = |a_c|^2 (KIND 0) or
// Re / Im of conj(a_c) * partner_c.
template <int NT, int KIND, bool NTL>
__global__ __launch_bounds__(kThreads) void k_expect(const float4* __restrict__ st, uint64_t n4,
                                                     uint64_t n_units, uint64_t chunk, uint64_t x,
                                                     int hw, ExpTerms a, double* __restrict__ part) {
  __shared__ double sh[NT][4];
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t beg = (uint64_t)blockIdx.x * chunk;
  const uint64_t end = beg + chunk < n_units ? beg + chunk : n_units;
  const int xl = (int)((x >> 1) & 63), csw = (int)(x & 1);
  const int h = x ? 63 - __clzll(x) : 0;
  const uint64_t xw = x >> 7;
  double acc[NT], k0[NT], k1[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    acc[j] = 0.0;
    const uint32_t zl = (uint32_t)(a.z[j] & 127);
    const uint32_t i0 = 2 * lane, i1 = 2 * lane + 1;
    const bool t0 = KIND != 1 || !((i0 >> h) & 1), t1 = KIND != 1 || !((i1 >> h) & 1);
    k0[j] = t0 ? ((__popc(i0 & zl) & 1) ? -a.fac[j] : a.fac[j]) : 0.0;
    k1[j] = t1 ? ((__popc(i1 & zl) & 1) ? -a.fac[j] : a.fac[j]) : 0.0;
  }
  constexpr int U = KIND == 2 ? 2 : 4;   // units in flight per wave
  const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint64_t u0 = beg + wid; u0 < end; u0 += 4 * U) {
    float4 v[U], p[U];
    uint64_t wk[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint64_t u = u0 + 4 * k;
      wk[k] = KIND == 2 ? ins0(u, hw) : u;
      const uint64_t sl = wk[k] * 64 + lane;
      const bool ok = u < end && sl < n4;
      v[k] = ok ? ld4<NTL>(st + sl) : zero;
      if (KIND == 2) p[k] = ok ? ld4<NTL>(st + ((wk[k] ^ xw) * 64 + lane)) : zero;
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      if (KIND == 1) p[k] = shfl_xor4(v[k], xl);
      if (KIND == 2) p[k] = shfl_xor4(p[k], xl);
      if (u0 + 4 * k >= end) continue;
      double v0, v1;
      if (KIND == 0) {
        v0 = (double)v[k].x * v[k].x + (double)v[k].y * v[k].y;
        v1 = (double)v[k].z * v[k].z + (double)v[k].w * v[k].w;
      } else {
        const float2 a0 = lo2(v[k]), a1 = hi2(v[k]);
        const float2 b0 = csw ? hi2(p[k]) : lo2(p[k]), b1 = csw ? lo2(p[k]) : hi2(p[k]);
        if (a.im) {
          v0 = (double)a0.x * b0.y - (double)a0.y * b0.x;
          v1 = (double)a1.x * b1.y - (double)a1.y * b1.x;
        } else {
          v0 = (double)a0.x * b0.x + (double)a0.y * b0.y;
          v1 = (double)a1.x * b1.x + (double)a1.y * b1.y;
        }
      }
      const uint64_t wb = wk[k] * 128;
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const double ws = (__popcll(wb & a.z[j]) & 1) ? -1.0 : 1.0;
        // k and ws are 0, +-1 or +-2, so the products are exact and the
        // fused forms round exactly as the separate ones would
        const double t = __builtin_fma(v1, k1[j], v0 * k0[j]);
        acc[j] = __builtin_fma(ws, t, acc[j]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const double r = wsum64(acc[j]);
    if (lane == 0) sh[j][wid] = r;
  }
  __syncthreads();
  if (threadIdx.x < NT) {
    const int j = threadIdx.x;
    part[(uint64_t)blockIdx.x * NT + j] = (sh[j][0] + sh[j][1]) + (sh[j][2] + sh[j][3]);
  }
}

// Pass 2: one workgroup sums each of the NT columns of part[nb][NT] in index
// order and writes term j to out[a.out[j]] (padding terms are dropped).
template <int NT>
__global__ __launch_bounds__(kThreads) void k_expect_final(const double* __restrict__ part, int nb,
                                                           ExpTerms a, double* __restrict__ out) {
  __shared__ double sh[NT][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double acc[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) acc[j] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kThreads) {
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] += part[(int64_t)b * NT + j];
  }
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const double r = wsum64(acc[j]);
    if (lane == 0) sh[j][wid] = r;
  }
  __syncthreads();
  if (threadIdx.x < NT) {
    const int j = threadIdx.x;
    if (a.out[j] >= 0) out[a.out[j]] = (sh[j][0] + sh[j][1]) + (sh[j][2] + sh[j][3]);
  }
}

template <int NT, int KIND>
static void launch_expect(hipStream_t s, const void* state, int n, uint64_t x, const ExpTerms& a,
                          double* part, double* out) {
  const uint64_t n4 = (1ull << n) / 2;
  const uint64_t windows = ceil_div(n4, 64);
  const uint64_t n_units = KIND == 2 ? windows / 2 : windows;
  const int hw = KIND == 2 ? (63 - __builtin_clzll(x)) - 7 : 0;
  const uint64_t want = ceil_div(n_units, 16);   // 4 waves x 4 units per workgroup
  const int blocks = (int)(want < (uint64_t)kRedBlocks ? (want ? want : 1) : kRedBlocks);
  const uint64_t chunk = ceil_div(n_units, blocks);
  // NTL: nontemporal loads above the MALL size
  if (nt_for_qubits(n))
    hipLaunchKernelGGL((k_expect<NT, KIND, true>), dim3(blocks), dim3(kThreads), 0, s,
                       (const float4*)state, n4, n_units, chunk, x, hw, a, part);
  else
    hipLaunchKernelGGL((k_expect<NT, KIND, false>), dim3(blocks), dim3(kThreads), 0, s,
                       (const float4*)state, n4, n_units, chunk, x, hw, a, part);
  hipLaunchKernelGGL((k_expect_final<NT>), dim3(1), dim3(kThreads), 0, s, part, blocks, a, out);
}

template <int NT>
static void launch_expect_kind(hipStream_t s, const void* state, int n, uint64_t x,
                               const ExpTerms& a, double* part, double* out) {
  if (x == 0) launch_expect<NT, 0>(s, state, n, x, a, part, out);
  else if (x < 128) launch_expect<NT, 1>(s, state, n, x, a, part, out);
  else launch_expect<NT, 2>(s, state, n, x, a, part, out);
}

// Workspace (doubles): pass-1 partials [kRedBlocks][kExpMaxTerms].
constexpr int64_t kExpWsDoubles = (int64_t)kRedBlocks * kExpMaxTerms;

static int expectation_impl(hipStream_t s, const void* state, int n, int64_t n_terms,
                            const uint64_t* xm, const uint64_t* zm, double* ws, double* out) {
  if (int rc = check_n(n)) return rc;
  if (n_terms < 0) return fail(QSIM_EINVAL, "n_terms < 0");
  if (n_terms == 0) return QSIM_OK;
  if (!state || !xm || !zm || !ws || !out) return fail(QSIM_EINVAL, "NULL pointer argument");
  const uint64_t lim = 1ull << n;
  for (int64_t k = 0; k < n_terms; ++k)
    if (xm[k] >= lim || zm[k] >= lim)
      return fail(QSIM_EINVAL, "Pauli term %lld: mask has a qubit >= n_qubits (%d)", (long long)k, n);
  // terms sharing an x mask and the parity of their Y count (Re or Im of
  // the pair products) share a state pass: visit them grouped, in input
  // order within a group
  std::vector<int64_t> order(n_terms);
  for (int64_t k = 0; k < n_terms; ++k) order[k] = k;
  auto key = [xm, zm](int64_t k) {
    return std::make_pair(xm[k], __builtin_popcountll(xm[k] & zm[k]) & 1);
  };
  std::stable_sort(order.begin(), order.end(),
                   [&key](int64_t a, int64_t b) { return key(a) < key(b); });
  int rc = QSIM_OK;
  for (int64_t g = 0; g < n_terms && rc == QSIM_OK;) {
    const auto k0 = key(order[g]);
    const uint64_t x = k0.first;
    int cnt = 0;
    ExpTerms a = {};
    a.im = k0.second;
    while (g + cnt < n_terms && cnt < kExpMaxTerms && key(order[g + cnt]) == k0) {
      const int64_t k = order[g + cnt];
      const int y = __builtin_popcountll(x & zm[k]);
      a.z[cnt] = zm[k];
      a.out[cnt] = (int)k;
      a.fac[cnt] = x == 0 ? 1.0 : (((y >> 1) & 1) ? -2.0 : 2.0);
      ++cnt;
    }
    for (int j = cnt; j < kExpMaxTerms; ++j) { a.out[j] = -1; a.fac[j] = 0.0; }
    if (cnt == 1) launch_expect_kind<1>(s, state, n, x, a, ws, out);
    else if (cnt == 2) launch_expect_kind<2>(s, state, n, x, a, ws, out);
    else if (cnt <= 4) launch_expect_kind<4>(s, state, n, x, a, ws, out);
    else launch_expect_kind<8>(s, state, n, x, a, ws, out);
    rc = launch_status("expectation");
    g += cnt;
  }
  return rc;
}

// ------------------------------------------------------------ apply P
// psi <- P psi, exact: every output amplitude is an input amplitude with its
// components swapped and/or negated.  KIND as for k_expect; one wave per
// window (KIND 0, 1) or window pair (KIND 2).  KIND 3 is x = 0 with no z bit
// below 7: the sign is one per window, (-1)^popcount(w & (z >> 7)), so only
// the half of the windows whose sign is -1 is visited (unit u inserts the
// highest bit hw of z >> 7 into u, set so that the parity is odd).
__device__ __forceinline__ float2 quarter_turn(float2 a, int k) {
  switch (k & 3) {
    case 0: return a;
    case 1: return make_float2(-a.y, a.x);     // * i
    case 2: return make_float2(-a.x, -a.y);    // * -1
    default: return make_float2(a.y, -a.x);    // * -i
  }
}

// New slot s from partner float4 p: amplitude i = 2s + c takes
// i^y (-1)^popcount((i ^ x) & z) psi_{i ^ x}.
__device__ __forceinline__ float4 pauli_slot(float4 p, uint64_t s, uint64_t x, uint64_t z, int y,
                                             int csw) {
  float2 r[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const uint64_t i = 2 * s + c;
    const float2 b = (c ^ csw) ? hi2(p) : lo2(p);
    r[c] = quarter_turn(b, y + 2 * (int)(__popcll((i ^ x) & z) & 1));
  }
  return cat4(r[0], r[1]);
}

template <int KIND, bool NT>
__global__ __launch_bounds__(kThreads) void k_pauli_apply(float4* __restrict__ st, uint64_t n4,
                                                          uint64_t n_units, uint64_t x, uint64_t z,
                                                          int y, int hw) {
  const uint64_t u = xcd_block<uint64_t>() * 4 + (threadIdx.x >> 6);
  if (u >= n_units) return;
  const int lane = threadIdx.x & 63;
  if (KIND == 3) {
    const uint64_t w0 = ins0(u, hw);
    const uint64_t w = (__popcll(w0 & (z >> 7)) & 1) ? w0 : (w0 | (1ull << hw));
    const uint64_t s = w * 64 + lane;
    const float4 v = ld4<NT>(st + s);
    st4<NT>(st + s, make_float4(-v.x, -v.y, -v.z, -v.w));
    return;
  }
  const int xl = (int)((x >> 1) & 63), csw = (int)(x & 1);
  if (KIND == 2) {
    const uint64_t w0 = ins0(u, hw), w1 = w0 ^ (x >> 7);
    const uint64_t s0 = w0 * 64 + lane, s1 = w1 * 64 + lane;
    const float4 v0 = ld4<NT>(st + s0), v1 = ld4<NT>(st + s1);
    const float4 p0 = shfl_xor4(v1, xl), p1 = shfl_xor4(v0, xl);
    st4<NT>(st + s0, pauli_slot(p0, s0, x, z, y, csw));
    st4<NT>(st + s1, pauli_slot(p1, s1, x, z, y, csw));
    return;
  }
  const uint64_t s = u * 64 + lane;
  const bool ok = s < n4;
  const float4 v = ok ? ld4<NT>(st + s) : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 p = KIND == 1 ? shfl_xor4(v, xl) : v;
  if (ok) st4<NT>(st + s, pauli_slot(p, s, x, z, y, csw));
}

template <bool NT>
static void launch_pauli(hipStream_t s, void* state, uint64_t n4, uint64_t windows, uint64_t x,
                         uint64_t z, int y) {
  if (x == 0 && (z & 127) == 0) {
    const int hw = (63 - __builtin_clzll(z >> 7));
    hipLaunchKernelGGL((k_pauli_apply<3, NT>), grid1d(ceil_div(windows / 2, 4)), dim3(kThreads), 0, s,
                       (float4*)state, n4, windows / 2, x, z, y, hw);
  } else if (x == 0) {
    hipLaunchKernelGGL((k_pauli_apply<0, NT>), grid1d(ceil_div(windows, 4)), dim3(kThreads), 0, s,
                       (float4*)state, n4, windows, x, z, y, 0);
  } else if (x < 128) {
    hipLaunchKernelGGL((k_pauli_apply<1, NT>), grid1d(ceil_div(windows, 4)), dim3(kThreads), 0, s,
                       (float4*)state, n4, windows, x, z, y, 0);
  } else {
    const int hw = (63 - __builtin_clzll(x)) - 7;
    hipLaunchKernelGGL((k_pauli_apply<2, NT>), grid1d(ceil_div(windows / 2, 4)), dim3(kThreads), 0, s,
                       (float4*)state, n4, windows / 2, x, z, y, hw);
  }
}

static int pauli_apply_impl(hipStream_t s, void* state, int n, uint64_t x, uint64_t z) {
  if (int rc = check_n(n)) return rc;
  if (!state) return fail(QSIM_EINVAL, "state is NULL");
  const uint64_t lim = 1ull << n;
  if (x >= lim || z >= lim) return fail(QSIM_EINVAL, "Pauli mask has a qubit >= n_qubits (%d)", n);
  if (x == 0 && z == 0) return QSIM_OK;
  const uint64_t n4 = (1ull << n) / 2;
  const uint64_t windows = ceil_div(n4, 64);
  const int y = __builtin_popcountll(x & z) & 3;
  if (nt_for_qubits(n)) launch_pauli<true>(s, state, n4, windows, x, z, y);
  else launch_pauli<false>(s, state, n4, windows, x, z, y);
  return launch_status("apply_pauli_string");
}

// ------------------------------------------------------------ damping channels
// One Kraus branch of amplitude damping (kind 0: E0 = diag(1, sqrt(1-g)),
// E1 = [[0, sqrt(g)], [0, 0]]) or dephasing (kind 1: E0 as before,
// E1 = diag(0, sqrt(g))) on qubit t.  From res = (total, S0) of the FP64
// reduction, S1 = total - S0 is the weight with bit t = 1; branch 1 has
// probability g * S1 / total.  The chosen branch is written as a 2x2 matrix
// that also renormalises the state to unit norm (E0: 1/sqrt(S0 + (1-g) S1);
// E1: 1/sqrt(g S1) times sqrt(g)), all in FP64, and applied by the dense
// 2x2 gate kernel.
__global__ void k_channel_branch(const double* __restrict__ res, int kind, double g, double u,
                                 float2* __restrict__ mat, int* __restrict__ branch) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double tot = res[0], s0 = res[1];
  const double s1 = tot - s0 > 0.0 ? tot - s0 : 0.0;
  const double p1 = tot > 0.0 ? g * s1 / tot : 0.0;
  const int b = u < p1 ? 1 : 0;
  float2 m[4];
  for (int k = 0; k < 4; ++k) m[k] = make_float2(0.f, 0.f);
  if (b == 0) {
    const double nrm = s0 + (1.0 - g) * s1;
    const double r = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
    m[0] = make_float2((float)r, 0.f);
    m[3] = make_float2((float)(r * sqrt(1.0 - g)), 0.f);
  } else {
    const double r = s1 > 0.0 ? 1.0 / sqrt(s1) : 0.0;
    m[kind == 0 ? 1 : 3] = make_float2((float)r, 0.f);
  }
  for (int k = 0; k < 4; ++k) mat[k] = m[k];
  *branch = b;
}

static int damping_impl(void* stream, void* state, int n, int t, double g, double u, void* workspace,
                        int* branch, int kind, const char* what) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_q(t, n, "target_qubit")) return rc;
  if (!state || !workspace || !branch) return fail(QSIM_EINVAL, "NULL pointer argument");
  if (!(g >= 0.0 && g <= 1.0)) return fail(QSIM_EINVAL, "%s: probability %g outside [0, 1]", what, g);
  hipStream_t s = as_stream(stream);
  double* ws = (double*)workspace;
  if (int rc = reduce_impl(s, state, n, 1, t, ws)) return rc;
  // the pass-1 partials are consumed by pass 2, so their space holds the
  // branch matrix; the branch index goes next to the results
  float2* mat = (float2*)ws;
  int* dev_branch = (int*)(ws + 2 * kRedBlocks + 4);
  hipLaunchKernelGGL(k_channel_branch, dim3(1), dim3(64), 0, s, ws + 2 * kRedBlocks, kind, g, u,
                     mat, dev_branch);
  if (int rc = launch_status(what)) return rc;
  M2 unused = {};
  if (int rc = gate1q_impl(stream, state, 1, n, mat, unused, t)) return rc;
  int host = -1;
  if (hipMemcpyAsync(&host, dev_branch, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
    return fail(QSIM_EHIP, "%s: copy failed", what);
  if (hipStreamSynchronize(s) != hipSuccess) return fail(QSIM_EHIP, "%s: sync failed", what);
  *branch = host;
  return QSIM_OK;
}

}  // namespace qsim

using namespace qsim;

extern "C" {

int64_t qsim_expectation_workspace_bytes(int n_qubits) {
  (void)n_qubits;
  return kExpWsDoubles * (int64_t)sizeof(double);
}

int qsim_expectation_paulis(void* stream, const void* state, int n_qubits, int64_t n_terms,
                            const uint64_t* x_masks, const uint64_t* z_masks, void* workspace,
                            void* out_dev) {
  clear_error();
  return expectation_impl(as_stream(stream), state, n_qubits, n_terms, x_masks, z_masks,
                          (double*)workspace, (double*)out_dev);
}

int qsim_expectation_z(void* stream, const void* state, int n_qubits, uint64_t z_mask,
                       void* workspace, void* out_dev) {
  clear_error();
  const uint64_t x = 0;
  return expectation_impl(as_stream(stream), state, n_qubits, 1, &x, &z_mask, (double*)workspace,
                          (double*)out_dev);
}

int qsim_expectation_pauli(void* stream, const void* state, int n_qubits, uint64_t z_mask,
                           int xy_qubit, int op, void* workspace, void* out_dev) {
  clear_error();
  if (op < 0 || op > 2) return fail(QSIM_EINVAL, "op must be 0=I/Z, 1=X, 2=Y (got %d)", op);
  if (op == 0) return qsim_expectation_z(stream, state, n_qubits, z_mask, workspace, out_dev);
  if (int rc = check_n(n_qubits)) return rc;
  if (int rc = check_q(xy_qubit, n_qubits, "xy_qubit")) return rc;
  // the Z string acts on the other qubits: a z bit on xy_qubit is ignored
  const uint64_t bit = 1ull << xy_qubit;
  const uint64_t x = bit, z = (z_mask & ~bit) | (op == 2 ? bit : 0);
  return expectation_impl(as_stream(stream), state, n_qubits, 1, &x, &z, (double*)workspace,
                          (double*)out_dev);
}

int qsim_apply_pauli_string(void* stream, void* state, int n_qubits, uint64_t x_mask,
                            uint64_t z_mask) {
  clear_error();
  return pauli_apply_impl(as_stream(stream), state, n_qubits, x_mask, z_mask);
}

int qsim_apply_depolarizing(void* stream, void* state, int n_qubits, int target, double p,
                            double random_val, int* branch) {
  clear_error();
  if (int rc = check_n(n_qubits)) return rc;
  if (int rc = check_q(target, n_qubits, "target_qubit")) return rc;
  if (!state || !branch) return fail(QSIM_EINVAL, "NULL pointer argument");
  if (!(p >= 0.0 && p <= 1.0)) return fail(QSIM_EINVAL, "depolarizing: p = %g outside [0, 1]", p);
  // identity on [0, 1-p); then X, Y, Z on three intervals of width p/3
  int b = 0;
  if (!(random_val < 1.0 - p)) {
    const double r = random_val - (1.0 - p), third = p / 3.0;
    b = r < third ? 1 : (r < 2.0 * third ? 2 : 3);
  }
  *branch = b;
  if (b == 0) return QSIM_OK;
  const uint64_t bit = 1ull << target;
  const uint64_t x = b == 3 ? 0 : bit, z = b == 1 ? 0 : bit;
  return pauli_apply_impl(as_stream(stream), state, n_qubits, x, z);
}

int qsim_apply_amplitude_damping(void* stream, void* state, int n_qubits, int target, double gamma,
                                 double random_val, void* workspace, int* branch) {
  clear_error();
  return damping_impl(stream, state, n_qubits, target, gamma, random_val, workspace, branch, 0,
                      "amplitude_damping");
}

int qsim_apply_dephasing(void* stream, void* state, int n_qubits, int target, double lam,
                         double random_val, void* workspace, int* branch) {
  clear_error();
  return damping_impl(stream, state, n_qubits, target, lam, random_val, workspace, branch, 1,
                      "dephasing");
}

}  // extern "C"
