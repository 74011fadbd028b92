// Internal helpers shared by the gfx950 kernels of libqsim_gfx950.so.
//
// Numerics: every kernel is compiled with -ffp-contract=off, so a*b+c is
// never silently fused; the only FMAs are the explicit __builtin_fmaf calls
// below.  That keeps the error-free transforms exact:
//   TwoProduct (Dekker 1971, FMA form): p = fl(a*b), e = fma(a, b, -p)
//   gives a*b = p + e exactly.
// A compensated dot product adds the rounded products and, separately, their
// exact product errors, then combines the two (the "Dot with TwoProduct"
// scheme of Ogita, Rump and Oishi, "Accurate sum and dot product", SISC 2005,
// without the TwoSum stage, which costs 6 more VALU ops per term; MI355X has
// ~8x less VALU throughput per HBM byte than the FP32 product count of a
// full Dot2 would need at 8 TB/s).
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <type_traits>
#include <stdint.h>
#include <stdarg.h>
#include "qsim_gfx950.h"

namespace qsim {

// ---------------------------------------------------------------- host side
int  fail(int code, const char* fmt, ...);     // records message, returns code
void clear_error();
int  launch_status(const char* what);          // hipGetLastError after a launch

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Validates the qubit count of a state (1..QSIM_MAX_QUBITS).
int check_n(int n_qubits);
// Validates one qubit index against n_qubits.
int check_q(int q, int n_qubits, const char* what);

constexpr int kThreads = 256;

// FP64 two-pass reduction of |a|^2 (k_reduce.hip).  Workspace (doubles):
// [0, 2*kRedBlocks) pass-1 partials, then res[8]: res[0] = total and, for
// mode 1, res[1] = the part with bit t = 0.  Launches only (asynchronous).
constexpr int kRedBlocks = 1024;
constexpr int64_t kRedWsDoubles = 2 * kRedBlocks + 8;
int reduce_impl(hipStream_t s, const void* state, int n, int mode, int t, double* ws);

// Dense 2x2 gate on qubit t of n_states contiguous states (k_gates.hip); the
// matrix comes from gdev (device complex64[4]) when non-NULL, else from gv.
struct M2;
int gate1q_impl(void* stream, void* states, uint64_t n_states, int n, const float2* gdev,
                const M2& gv, int t);

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Grid of `blocks` workgroups: one row while it fits 2^23 blocks (2^31
// work-items of 256), else rows of 2^23 blocks (kernels index blocks with
// gblock() and guard the tail of the last row).
inline dim3 grid1d(uint64_t blocks) {
  constexpr uint64_t kRow = 1ull << 23;
  if (blocks <= kRow) return dim3((unsigned)(blocks ? blocks : 1));
  return dim3((unsigned)kRow, (unsigned)ceil_div(blocks, kRow));
}

// Sorted affected-qubit list plus control mask for the multi-controlled kernels.
struct McSpec {
  int      pos[10];   // ascending positions of all affected qubits
  int      n_pos;
  uint64_t cmask;     // OR of 1 << control
};
// Builds an McSpec from controls + targets; checks ranges and distinctness.
int build_mc(McSpec& s, int n_qubits, const int64_t* controls, int n_controls,
             const int* targets, int n_targets);

// ---------------------------------------------------------------- device side
// Insert a zero bit at position q of x (bits >= q move up by one).
template <typename I>
__device__ __forceinline__ I ins0(I x, int q) {
  const I lo = x & ((I(1) << q) - 1);
  return ((x ^ lo) << 1) | lo;
}

__device__ __forceinline__ void two_prod(float a, float b, float& p, float& e) {
  p = a * b;
  e = __builtin_fmaf(a, b, -p);
}

// Compensated x0*y0 + x1*y1 + x2*y2 + x3*y3.
__device__ __forceinline__ float cdot4(float x0, float y0, float x1, float y1,
                                       float x2, float y2, float x3, float y3) {
  float p0, e0, p1, e1, p2, e2, p3, e3;
  two_prod(x0, y0, p0, e0);
  two_prod(x1, y1, p1, e1);
  two_prod(x2, y2, p2, e2);
  two_prod(x3, y3, p3, e3);
  const float s = (p0 + p1) + (p2 + p3);
  const float e = (e0 + e1) + (e2 + e3);
  return s + e;
}

// Compensated complex g0*a0 + g1*a1 (one row of a 2x2 gate).
__device__ __forceinline__ float2 crow2(float2 g0, float2 a0, float2 g1, float2 a1) {
  float2 r;
  r.x = cdot4(g0.x, a0.x, -g0.y, a0.y, g1.x, a1.x, -g1.y, a1.y);
  r.y = cdot4(g0.x, a0.y, g0.y, a0.x, g1.x, a1.y, g1.y, a1.x);
  return r;
}

// Compensated complex sum of four products (one row of a 4x4 gate):
// eight real products per component, split into two 4-term halves whose
// rounded sums and exact product errors are combined at the end.
__device__ __forceinline__ float2 crow4(const float2* g, float2 a0, float2 a1,
                                        float2 a2, float2 a3) {
  float p[8], e[8];
  two_prod(g[0].x, a0.x, p[0], e[0]);  two_prod(-g[0].y, a0.y, p[1], e[1]);
  two_prod(g[1].x, a1.x, p[2], e[2]);  two_prod(-g[1].y, a1.y, p[3], e[3]);
  two_prod(g[2].x, a2.x, p[4], e[4]);  two_prod(-g[2].y, a2.y, p[5], e[5]);
  two_prod(g[3].x, a3.x, p[6], e[6]);  two_prod(-g[3].y, a3.y, p[7], e[7]);
  float2 r;
  r.x = (((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]))) +
        (((e[0] + e[1]) + (e[2] + e[3])) + ((e[4] + e[5]) + (e[6] + e[7])));
  two_prod(g[0].x, a0.y, p[0], e[0]);  two_prod(g[0].y, a0.x, p[1], e[1]);
  two_prod(g[1].x, a1.y, p[2], e[2]);  two_prod(g[1].y, a1.x, p[3], e[3]);
  two_prod(g[2].x, a2.y, p[4], e[4]);  two_prod(g[2].y, a2.x, p[5], e[5]);
  two_prod(g[3].x, a3.y, p[6], e[6]);  two_prod(g[3].y, a3.x, p[7], e[7]);
  r.y = (((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]))) +
        (((e[0] + e[1]) + (e[2] + e[3])) + ((e[4] + e[5]) + (e[6] + e[7])));
  return r;
}

// Compensated complex product a*d.
// Entry of one X-mask group of a Pauli sum in row `row`: sum over the
// group's terms t in [t0, t1) of coef[t] (-1)^popcount(row & sm[t]), in FP64
// and term order (each term enters with its sign flipped or not, which is
// exact); false when the sum is exactly zero.  The device CSR builder and the
// matrix-free Chebyshev term both evaluate entries with this function, so
// they see the same values.  R is the row index type (uint32_t when every
// row index fits in 32 bits).
template <typename R>
__device__ __forceinline__ bool pauli_group_value(R row, int t0, int t1, const uint64_t* __restrict__ sm,
                                                  const double2* __restrict__ coef, double& re,
                                                  double& im) {
  re = 0.0;
  im = 0.0;
  for (int t = t0; t < t1; ++t) {
    const uint32_t pc = sizeof(R) == 4 ? (uint32_t)__popc((uint32_t)row & (uint32_t)sm[t])
                                       : (uint32_t)__popcll((uint64_t)row & sm[t]);
    // +-1.0 with the parity of pc in the sign bit; c * (+-1) is exact, so
    // each FMA rounds exactly as re +- c would
    const double sgn = __hiloint2double((int)((pc << 31) | 0x3ff00000u), 0);
    const double2 c = coef[t];
    re = __builtin_fma(c.x, sgn, re);
    im = __builtin_fma(c.y, sgn, im);
  }
  return re != 0.0 || im != 0.0;
}

__device__ __forceinline__ float2 cmulc(float2 a, float2 d) {
  float p0, e0, p1, e1;
  float2 r;
  two_prod(a.x, d.x, p0, e0);
  two_prod(-a.y, d.y, p1, e1);
  r.x = (p0 + p1) + (e0 + e1);
  two_prod(a.x, d.y, p0, e0);
  two_prod(a.y, d.x, p1, e1);
  r.y = (p0 + p1) + (e0 + e1);
  return r;
}

__device__ __forceinline__ float2 lo2(float4 v) { return make_float2(v.x, v.y); }
__device__ __forceinline__ float2 hi2(float4 v) { return make_float2(v.z, v.w); }
__device__ __forceinline__ float4 cat4(float2 a, float2 b) { return make_float4(a.x, a.y, b.x, b.y); }

// Nontemporal 16-byte global accesses (global_load/store_dwordx4 ... nt):
// streaming data that is touched once per kernel does not displace the L2/MALL
// working set.  Every caller accesses whole cache lines from one instruction
// (lane-contiguous float4 slots), so no line is fetched twice.
typedef float f32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_nt(const float4* p) {
  const f32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const f32x4_t*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_nt(float4* p, float4 v) {
  const f32x4_t w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, reinterpret_cast<f32x4_t*>(p));
}

// Cache policy by working-set size.  A state of at most 256 MiB (2^25
// amplitudes) fits the 256 MB MALL (Infinity Cache) of MI355X and stays
// resident from one launch to the next when it is read and written with
// ordinary cached accesses; nontemporal accesses would send every pass to
// HBM.  Larger states stream through with nontemporal accesses, so that one
// pass does not displace what the next one needs.  Measured on a 4-stream 2Q
// update: cached 8.1-8.4 TB/s at 24-25 qubits against 5.6-6.3 TB/s
// nontemporal; from 26 qubits on nontemporal is 5-10 % faster.
constexpr uint64_t kMallBytes = 1ull << 28;
inline bool nt_for_bytes(uint64_t bytes) { return bytes > kMallBytes; }
inline bool nt_for_qubits(int n_qubits) { return n_qubits > 25; }
template <bool NT>
__device__ __forceinline__ float4 ld4(const float4* p) {
  if constexpr (NT) return ld_nt(p);
  else return *p;
}
template <bool NT>
__device__ __forceinline__ void st4(float4* p, float4 v) {
  if constexpr (NT) st_nt(p, v);
  else *p = v;
}

// Value of lane (lane ^ m) for a wave-uniform mask m < 64, with every lane
// of the wave active.  Each set bit of m is one exchange step, and the
// steps compose because lane -> lane ^ b permutations commute:
//   xor 1, 2  DPP quad_perm [1,0,3,2] / [2,3,0,1]   (v_mov_b32_dpp)
//   xor 8     DPP row_ror:8 (rotation by 8 within a row of 16 is xor 8)
//   xor 4, 16 ds_swizzle bitmask mode (and 0x1f, xor mask), no LDS memory
//   xor 32    v_permlane32_swap between the two 32-lane halves
// None of them needs an address register or LDS traffic, unlike the
// ds_bpermute that __shfl_xor lowers to.  A mask known at compile time
// folds to its steps; a kernel argument takes uniform scalar branches.
__device__ __forceinline__ int lane_xor_i(int x, int m) {
  if (m & 1) x = __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);
  if (m & 2) x = __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);
  if (m & 4) x = __builtin_amdgcn_ds_swizzle(x, 0x101F);
  if (m & 8) x = __builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);
  if (m & 16) x = __builtin_amdgcn_ds_swizzle(x, 0x401F);
  if (m & 32) {
    // r[0]: this register with lanes 32..63 taken from lanes 0..31 of the
    // other operand; r[1]: the other operand with lanes 0..31 taken from
    // lanes 32..63 of this one.  Both operands are x.
    const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    x = __lane_id() < 32 ? (int)r[1] : (int)r[0];
  }
  return x;
}
__device__ __forceinline__ float lane_xor(float v, int m) {
  return __int_as_float(lane_xor_i(__float_as_int(v), m));
}
__device__ __forceinline__ double lane_xor(double v, int m) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const int lo = lane_xor_i((int)(uint32_t)u, m), hi = lane_xor_i((int)(uint32_t)(u >> 32), m);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Cross-lane exchange of a float4 with lane ^ M for a compile-time mask.
template <int M>
__device__ __forceinline__ float4 xor4c(float4 v) {
  float4 r;
  r.x = lane_xor(v.x, M);
  r.y = lane_xor(v.y, M);
  r.z = lane_xor(v.z, M);
  r.w = lane_xor(v.w, M);
  return r;
}
// Calls f(std::integral_constant<int, M>) for the wave-uniform single-bit
// lane mask m (1, 2, ..., 32) known only at run time: one uniform branch
// selects a body in which the exchanges have compile-time masks, so every
// exchange inside it is straight-line DPP / swizzle / permlane code.
template <class F>
__device__ __forceinline__ void with_lane_bit(int m, F&& f) {
  switch (m) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    default: f(std::integral_constant<int, 32>{}); break;
  }
}
// Cross-lane exchange of a float4 through ds_bpermute (the LDS crossbar,
// no LDS memory): one address for the four components, and no VALU work.
// Used where the VALU is the scarcer pipe: the 4x4 gate with both targets
// below 7 exchanges three times per slot, and there DPP moves measured
// 2.5-5 % slower at 30 qubits.
__device__ __forceinline__ float4 shfl_xor4_lds(float4 v, int m) {
  float4 r;
  r.x = __shfl_xor(v.x, m, 64);
  r.y = __shfl_xor(v.y, m, 64);
  r.z = __shfl_xor(v.z, m, 64);
  r.w = __shfl_xor(v.w, m, 64);
  return r;
}
// Cross-lane exchange of a float4 with lane ^ m, m < 64 wave-uniform and
// known only at run time: a single-bit mask takes one uniform branch; a
// mask of several bits is exchanged one bit at a time.
__device__ __forceinline__ float4 shfl_xor4(float4 v, int m) {
#pragma unroll
  for (int b = 1; b < 64; b <<= 1)
    if (m & b) with_lane_bit(b, [&](auto M) { v = xor4c<decltype(M)::value>(v); });
  return v;
}

// Flat block index of a 1D grid laid out as grid1d() does (gridDim.y rows of
// gridDim.x blocks): an AMD dispatch holds at most 2^32 - 1 work-items per
// dimension, so launches over states of more than 2^31 slots use rows.
template <typename I>
__device__ __forceinline__ I gblock() {
  return (I)blockIdx.y * (I)gridDim.x + (I)blockIdx.x;
}

// XCD-aware block order.  Workgroups are dispatched round-robin over the 8
// XCDs (block b -> XCD b % 8); renumbering so that XCD x walks the x-th
// eighth of the grid keeps each XCD's L2 and DRAM pages on one contiguous
// region instead of 4 KiB stripes shared by all eight.
template <typename I>
__device__ __forceinline__ I xcd_block() {
  const I nb = (I)gridDim.x * (I)gridDim.y, b = gblock<I>();
  if (nb & 7) return b;
  return (b & 7) * (nb >> 3) + (b >> 3);
}

// Pass order across consecutive launches of streaming kernels.  A pass over
// a state larger than the MALL ends with its last-processed part written
// through the cache (cached stores), and the next pass runs in the opposite
// direction, so it starts on what is still in the MALL.  tail1024 is the
// part of each pass (in 1/1024) stored with cached writes: the MALL size
// over the state size.  The order changes nothing in the results; it is
// flipped by a process-wide counter at every such launch.
struct PassOrder {
  int rev;
  int tail1024;
};
inline PassOrder next_pass(uint64_t state_bytes) {
  static std::atomic<unsigned> counter{0};
  PassOrder p;
  p.rev = (int)(counter.fetch_add(1, std::memory_order_relaxed) & 1u);
  const uint64_t t = state_bytes ? (kMallBytes * 1024) / state_bytes : 1024;
  p.tail1024 = (int)(t < 1024 ? t : 1024);
  return p;
}
// Address-order block of this workgroup under the pass order, and whether it
// lies in the pass's tail.  XCD: the XCD-major numbering of xcd_block (each
// XCD walks its own contiguous eighth, which it then walks backwards in a
// reversed pass).
template <typename I, bool XCD>
__device__ __forceinline__ I pass_block(const PassOrder& po, bool& tail) {
  const I nb = (I)gridDim.x * (I)gridDim.y, b = (I)blockIdx.y * (I)gridDim.x + (I)blockIdx.x;
  const bool x = XCD && (nb & 7) == 0;
  const I len = x ? nb >> 3 : nb;
  const I pos = x ? b >> 3 : b;
  const I bx = x ? (b & 7) * len + pos : b;
  tail = (uint64_t)pos * 1024 >= (uint64_t)len * (uint64_t)(1024 - po.tail1024);
  return po.rev ? nb - 1 - bx : bx;
}
// Block index of a streaming kernel: under the pass order when NT (states
// larger than the MALL), else the plain numbering (XCD-major when XCD).
template <typename I, bool NT, bool XCD>
__device__ __forceinline__ I pass_blk(const PassOrder& po, bool& tail) {
  if constexpr (NT) return pass_block<I, XCD>(po, tail);
  else return XCD ? xcd_block<I>() : gblock<I>();
}
// Store of a streaming kernel: nontemporal, except in the pass's tail.
template <bool NT>
__device__ __forceinline__ void st4p(float4* p, float4 v, bool tail) {
  if constexpr (NT) {
    if (tail) *p = v;
    else st_nt(p, v);
  } else {
    *p = v;
  }
}


// Uniform k of a seeded SplitMix64 stream (Steele, Lea and Flood, "Fast
// splittable pseudorandom number generators", OOPSLA 2014): the 64-bit state
// seed + (k+1) * golden gamma through the published mix, top 53 bits scaled
// to [0, 1).
__device__ __forceinline__ double splitmix_u(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}

// Gate coefficients passed by value (kernel arguments live in SGPRs).
struct M2 { float2 g[4]; };
struct M4 { float2 g[16]; };

}  // namespace qsim
