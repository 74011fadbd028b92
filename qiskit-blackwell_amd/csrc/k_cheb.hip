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

// ---------------------------------------------------------------- 32 RHS
// The batched term at 32 right-hand sides, the width of one RHS-contiguous
// row (32 x 8 B = 256 B).  One wave owns four consecutive CSR rows: lanes
// 16 q .. 16 q + 15 row r + q, each lane two RHS (one 16-byte load per
// gathered row).  Per nonzero a quarter-wave needs the column index and the
// value (the same for its 16 lanes) and one 256-byte gathered row of T_k, so
// one gather instruction fetches four rows (1 KiB).
//  * Why four rows: the gathers are bound by the number of vector memory
//    instructions in flight, not by their bytes; measured on the 22-qubit
//    chain at 32 RHS (FP64 sums), two rows per wave with 8-byte lanes take
//    4.2 ms per term and four rows with 16-byte lanes 3.7 ms, bit for bit the
//    same result.  2, 4 or 8 gathers in flight per lane measured the same
//    (1 is 5 % slower).
//  * Row data: the wave's four rows are consecutive in the CSR arrays, so it
//    stages them into its own LDS slice with coalesced full-wave loads
//    (nontemporal: the CSR is read once per term) and reads them back as
//    quarter-wave broadcasts.
//  * Row sum, ACC = kAccF64: the product of two FP32 numbers is exact in FP64
//    (24 + 24 bits < 53), so every product is converted and accumulated with
//    FP64 FMAs.  CDNA4 issues FP64 FMA at the non-packed FP32 rate, so this
//    costs 8 VALU operations per complex multiply-add against ~44 for the
//    TwoProduct + Neumaier double-float form (ACC = kAccF32c); the FP64 sum is
//    also the more accurate of the two (53-bit accumulator, exact terms).
//  * Epilogue (both forms): T_{k+1} = 2 scale (A T_k) - T_{k-1}, rounded to
//    complex64 once, and acc += c_k T_{k+1} as a Neumaier pair (acc, comp) in
//    FP32, as in the generic kernel.  T_{k-1}, acc and comp are loaded before
//    the row's gathers and streamed (nontemporal when the 32-RHS block
//    exceeds the MALL), so the cache keeps the gathered rows of T_k.
//  * Block order: order_shift s > 0 hands each XCD contiguous chunks of 2^s
//    rows, chunk x of every group of 8 chunks going to XCD x, so that a
//    gather through a low column bit stays in the XCD's own L2 while all 8
//    XCDs still walk one front through the MALL.  s = 0 is the dispatch
//    order (blocks round-robin over the XCDs).  Shifts 0, 6, 9, 12 and 19
//    measured within 4 % of each other.
constexpr int kAccF64 = 0;
constexpr int kAccF32c = 1;
constexpr int kWaves = kThreads / 64;
constexpr int kRowsPerWave = 4;
constexpr int kLanesPerRow = 64 / kRowsPerWave;            // 16
constexpr int kRhsPerLane = 32 / kLanesPerRow;             // 2
constexpr int kBlockRowsLog2 = 4;                          // 4 waves x 4 rows per workgroup
static_assert(kWaves * kRowsPerWave == 1 << kBlockRowsLog2, "rows per workgroup");
constexpr int kStageNnz = 256;      // nonzeros of a wave's four rows staged per chunk
constexpr int kSpmmOrderShift = 9;
constexpr int kGather32 = 2;        // gathers in flight per lane

typedef float f32x2_t __attribute__((ext_vector_type(2)));
template <bool NT>
__device__ __forceinline__ float2 ld2(const float2* p) {
  if constexpr (NT) {
    const f32x2_t v = __builtin_nontemporal_load(reinterpret_cast<const f32x2_t*>(p));
    return make_float2(v.x, v.y);
  } else {
    return *p;
  }
}
template <bool NT>
__device__ __forceinline__ void st2(float2* p, float2 v) {
  if constexpr (NT) {
    const f32x2_t w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<f32x2_t*>(p));
  } else {
    *p = v;
  }
}

// Wave-level ordering of LDS writes and reads by other lanes of the same wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Logical block of this workgroup under the chunked XCD order (a bijection
// on [0, gridDim.x) when gridDim.x is a multiple of 8 * chunk blocks; the
// host passes s = 0 otherwise).
__device__ __forceinline__ int64_t spmm_block(int s) {
  const int64_t b = blockIdx.x;
  if (s <= 0) return b;
  const int g = s > kBlockRowsLog2 ? s - kBlockRowsLog2 : 0;   // log2(blocks per chunk)
  const int64_t xcd = b & 7, i = b >> 3;
  const int64_t grp = i >> g, w = i & ((1ll << g) - 1);
  return ((grp * 8 + xcd) << g) + w;
}

template <int ACC>
struct RowSum;
template <>
struct RowSum<kAccF64> {
  double r = 0.0, i = 0.0;
  __device__ __forceinline__ void add(float2 v, float2 x) { add_d(v.x, v.y, x); }
  // vr, vi: an entry already rounded to complex64, held in FP64
  __device__ __forceinline__ void add_d(double vr, double vi, float2 x) {
    const double xr = x.x, xi = x.y;
    r = __builtin_fma(vr, xr, r);
    r = __builtin_fma(-vi, xi, r);
    i = __builtin_fma(vr, xi, i);
    i = __builtin_fma(vi, xr, i);
  }
  // scale * sum, rounded once to FP32 after the recurrence update
  __device__ __forceinline__ float2 next(double scale, float2 tp, bool first) const {
    const double wr = r * scale, wi = i * scale;
    if (first) return make_float2((float)wr, (float)wi);
    return make_float2((float)(2.0 * wr - (double)tp.x), (float)(2.0 * wi - (double)tp.y));
  }
};
template <>
struct RowSum<kAccF32c> {
  float sr = 0.f, cr = 0.f, si = 0.f, ci = 0.f;
  __device__ __forceinline__ void add(float2 v, float2 x) { cmac_acc(v, x, sr, cr, si, ci); }
  __device__ __forceinline__ float2 next(double scale_d, float2 tp, bool first) const {
    const float scale = (float)scale_d;
    float p, q, wr, wi;
    two_prod(sr, scale, p, q); wr = p + (q + cr * scale);
    two_prod(si, scale, p, q); wi = p + (q + ci * scale);
    if (first) return make_float2(wr, wi);
    return make_float2(2.f * wr - tp.x, 2.f * wi - tp.y);
  }
};

// Split term (PASS 1 and 2).  For a matrix of N = 2^n rows whose entries
// mostly connect rows that differ either only below bit s or only at and
// above it (a Pauli Hamiltonian's X masks), the SpMM is done in two
// passes, each of whose gathers stay within 2^s (2^(n-s)) neighbouring rows
// of the order it walks, so the gathered 256-byte rows of T_k are reused
// from L2 instead of being fetched again for every row that needs them:
//   PASS 1  rows in natural order, the entries with (row ^ col) < 2^s:
//           w1 = (A_low T_k) rounded to complex64;
//   PASS 2  positions p in order, row r = rot(p) (the low n-s bits of p are
//           the high bits of r), the other entries, stored in p order:
//           w = w1 + A_high T_k (FP64), then the same recurrence epilogue.
// Entries whose mask has bits on both sides of s (at most a few per row)
// go to pass 2 and are gathered from farther away.  On a 22-qubit chain at
// 32 RHS this cuts the bytes fetched per term from 24 GB to 10 GB but gains
// only a few per cent over PASS 0 (the gathers' issue and latency, not their
// bytes, bound both forms), so callers opt in to it.
__device__ __forceinline__ int64_t split_row(int64_t p, int n, int s) {
  const int lo_bits = n - s;
  return ((p & ((int64_t(1) << lo_bits) - 1)) << s) | (p >> lo_bits);
}

// kRhsPerLane complex RHS of one lane: float2 or float4 loads and stores.
template <int EPL>
struct CVec {
  float2 c[EPL];
};
template <bool NT, int EPL>
__device__ __forceinline__ CVec<EPL> ldv(const float2* p) {
  CVec<EPL> r;
  if constexpr (EPL == 1) {
    r.c[0] = ld2<NT>(p);
  } else {
    typedef float f32x4_t __attribute__((ext_vector_type(4)));
    f32x4_t v;
    if constexpr (NT) v = __builtin_nontemporal_load(reinterpret_cast<const f32x4_t*>(p));
    else v = *reinterpret_cast<const f32x4_t*>(p);
    r.c[0] = make_float2(v.x, v.y);
    r.c[1] = make_float2(v.z, v.w);
  }
  return r;
}
template <bool NT, int EPL>
__device__ __forceinline__ void stv(float2* p, const CVec<EPL>& r) {
  if constexpr (EPL == 1) {
    st2<NT>(p, r.c[0]);
  } else {
    typedef float f32x4_t __attribute__((ext_vector_type(4)));
    const f32x4_t v = {r.c[0].x, r.c[0].y, r.c[1].x, r.c[1].y};
    if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<f32x4_t*>(p));
    else *reinterpret_cast<f32x4_t*>(p) = v;
  }
}

template <int ACC, bool NT, int PASS>
__global__ __launch_bounds__(kThreads) void k_cheb_spmm32(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const float2* __restrict__ val,
    int64_t n_rows, double scale, const float2* __restrict__ tc, float2* __restrict__ tpn,
    float2* __restrict__ acc, float2* __restrict__ comp, float2* __restrict__ w1, float2 c,
    int first, int order_shift, int n_bits, int split) {
  constexpr int RW = kRowsPerWave, LPR = kLanesPerRow, EPL = kRhsPerLane, G = kGather32;
  __shared__ int32_t s_ci[kWaves][kStageNnz];
  __shared__ float2 s_v[kWaves][kStageNnz];
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63, qr = lane / LPR, jj = lane % LPR;
  const int64_t pos0 = (spmm_block(order_shift) * kWaves + wave) * RW;   // first row (position)
  if (pos0 >= n_rows) return;
  const int n_here = (int)(n_rows - pos0 < RW ? n_rows - pos0 : RW);
  const int32_t my_rp = lane <= n_here ? rp[pos0 + lane] : 0;
  const char* tcb = reinterpret_cast<const char*>(tc);
  const uint32_t lane_off = (uint32_t)jj * (8u * EPL);
  const int32_t gb = __builtin_amdgcn_readlane(my_rp, 0);        // the wave's entries [gb, ge)
  const int32_t ge = __builtin_amdgcn_readlane(my_rp, n_here);
  const bool live = qr < n_here;
  const int my_r = live ? qr : 0;   // lanes past the last row shadow the first and store nothing
  const int32_t my_b = __shfl(my_rp, my_r), my_e = live ? __shfl(my_rp, my_r + 1) : my_b;
  const int64_t pos = pos0 + my_r;
  const int64_t row = PASS == 2 ? split_row(pos, n_bits, split) : pos;
  const int64_t e_idx = row * 32 + jj * EPL;
  CVec<EPL> tp{}, a{}, k2{}, wv{};
  if constexpr (PASS != 1) {
    tp = ldv<NT, EPL>(tpn + e_idx);
    a = ldv<NT, EPL>(acc + e_idx);
    k2 = ldv<NT, EPL>(comp + e_idx);
    if constexpr (PASS == 2) wv = ldv<NT, EPL>(w1 + e_idx);
  }
  RowSum<ACC> sum[EPL];
  for (int32_t c0 = gb; c0 < ge; c0 += kStageNnz) {
    const int32_t cend = c0 + kStageNnz < ge ? c0 + kStageNnz : ge;
    wave_sync();
    for (int32_t k = c0 + lane; k < cend; k += 64) {
      s_ci[wave][k - c0] = __builtin_nontemporal_load(ci + k);
      s_v[wave][k - c0] = ld2<true>(val + k);
    }
    wave_sync();
    const int32_t lo = my_b > c0 ? my_b : c0, hi = my_e < cend ? my_e : cend;
    const int32_t* sc = s_ci[wave] - c0;
    const float2* sv = s_v[wave] - c0;
    for (int32_t k = lo; k < hi; k += G) {
      CVec<EPL> x[G];
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const int32_t idx = k + u < hi ? k + u : hi - 1;
        // positions past hi gather the row's last entry again (no branch
        // between the loads) and add it with weight 0, an exact no-op
        x[u] = ldv<false, EPL>(
            reinterpret_cast<const float2*>(tcb + ((uint32_t)sc[idx] * 256u + lane_off)));
      }
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const float2 v = k + u < hi ? sv[k + u] : make_float2(0.f, 0.f);
#pragma unroll
        for (int e = 0; e < EPL; ++e) sum[e].add(v, x[u].c[e]);
      }
    }
  }
  if (!live) return;
  if constexpr (PASS == 1) {
    CVec<EPL> o;
#pragma unroll
    for (int e = 0; e < EPL; ++e) o.c[e] = make_float2((float)sum[e].r, (float)sum[e].i);
    stv<NT, EPL>(w1 + e_idx, o);
    return;
  }
  CVec<EPL> tn;
#pragma unroll
  for (int e = 0; e < EPL; ++e) {
    if constexpr (PASS == 2) {
      sum[e].r += (double)wv.c[e].x;
      sum[e].i += (double)wv.c[e].y;
    }
    tn.c[e] = sum[e].next(scale, first ? make_float2(0.f, 0.f) : tp.c[e], first != 0);
    const float2 v = cmulc(c, tn.c[e]);
    neu_add(a.c[e].x, k2.c[e].x, v.x);
    neu_add(a.c[e].y, k2.c[e].y, v.y);
  }
  stv<NT, EPL>(tpn + e_idx, tn);
  stv<NT, EPL>(acc + e_idx, a);
  stv<NT, EPL>(comp + e_idx, k2);
}

// ---------------------------------------------------------------- CSR split
// Splits a CSR of N = 2^n rows at bit s: entries with (row ^ col) < 2^s go
// to the low CSR (rows in natural order), the others to the high CSR, whose
// row p is row rot(p) of the input (split_row).  Both keep each row's entry
// order.  One row (position) per thread: k_split_count sums the per-block
// counts of both outputs, k_split_scan (one workgroup) turns them into block
// offsets, k_split_fill writes row pointers, columns and values.
constexpr int kSplitThreads = 256;

__device__ __forceinline__ int64_t blk_exscan(int64_t v, int64_t* total) {
  __shared__ int64_t wsum[kSplitThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t incl = v;
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int64_t base = 0, all = 0;
  for (int k = 0; k < kSplitThreads / 64; ++k) {
    if (k < w) base += wsum[k];
    all += wsum[k];
  }
  __syncthreads();
  *total = all;
  return base + incl - v;
}

__device__ __forceinline__ void split_counts(const int32_t* __restrict__ rp,
                                             const int32_t* __restrict__ ci, int64_t i, int n,
                                             int s, int& nlo, int& nhi) {
  nlo = 0;
  for (int32_t k = rp[i]; k < rp[i + 1]; ++k) nlo += ((uint32_t)(ci[k] ^ (int32_t)i) >> s) == 0;
  const int64_t r = split_row(i, n, s);
  nhi = 0;
  for (int32_t k = rp[r]; k < rp[r + 1]; ++k) nhi += ((uint32_t)(ci[k] ^ (int32_t)r) >> s) != 0;
}

__global__ __launch_bounds__(kSplitThreads) void k_split_count(const int32_t* __restrict__ rp,
                                                               const int32_t* __restrict__ ci,
                                                               int64_t n_rows, int n, int s,
                                                               int64_t* __restrict__ bsum) {
  const int64_t i = (int64_t)blockIdx.x * kSplitThreads + threadIdx.x;
  int nlo = 0, nhi = 0;
  if (i < n_rows) split_counts(rp, ci, i, n, s, nlo, nhi);
  int64_t tl, th;
  blk_exscan(nlo, &tl);
  blk_exscan(nhi, &th);
  if (threadIdx.x == 0) {
    bsum[2 * blockIdx.x] = tl;
    bsum[2 * blockIdx.x + 1] = th;
  }
}

__global__ __launch_bounds__(kSplitThreads) void k_split_scan(uint64_t n_blocks,
                                                              int64_t* __restrict__ bsum,
                                                              int64_t* __restrict__ nnz2) {
  int64_t carry[2] = {0, 0};
  for (uint64_t base = 0; base < n_blocks; base += kSplitThreads) {
    const uint64_t b = base + threadIdx.x;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int64_t v = b < n_blocks ? bsum[2 * b + side] : 0;
      int64_t total;
      const int64_t pre = blk_exscan(v, &total);
      if (b < n_blocks) bsum[2 * b + side] = carry[side] + pre;
      carry[side] += total;
    }
  }
  if (threadIdx.x == 0) {
    nnz2[0] = carry[0];
    nnz2[1] = carry[1];
  }
}

__global__ __launch_bounds__(kSplitThreads) void k_split_fill(
    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const float2* __restrict__ val,
    int64_t n_rows, int n, int s, const int64_t* __restrict__ boff, int32_t* __restrict__ rp_lo,
    int32_t* __restrict__ ci_lo, float2* __restrict__ val_lo, int32_t* __restrict__ rp_hi,
    int32_t* __restrict__ ci_hi, float2* __restrict__ val_hi) {
  const int64_t i = (int64_t)blockIdx.x * kSplitThreads + threadIdx.x;
  int nlo = 0, nhi = 0;
  if (i < n_rows) split_counts(rp, ci, i, n, s, nlo, nhi);
  int64_t tl, th;
  const int64_t olo = boff[2 * blockIdx.x] + blk_exscan(nlo, &tl);
  const int64_t ohi = boff[2 * blockIdx.x + 1] + blk_exscan(nhi, &th);
  if (i >= n_rows) return;
  rp_lo[i] = (int32_t)olo;
  rp_hi[i] = (int32_t)ohi;
  if (i == n_rows - 1) {
    rp_lo[n_rows] = (int32_t)(olo + nlo);
    rp_hi[n_rows] = (int32_t)(ohi + nhi);
  }
  int64_t o = olo;
  for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
    if (((uint32_t)(ci[k] ^ (int32_t)i) >> s) == 0) {
      ci_lo[o] = ci[k];
      val_lo[o] = val[k];
      ++o;
    }
  const int64_t r = split_row(i, n, s);
  o = ohi;
  for (int32_t k = rp[r]; k < rp[r + 1]; ++k)
    if (((uint32_t)(ci[k] ^ (int32_t)r) >> s) != 0) {
      ci_hi[o] = ci[k];
      val_hi[o] = val[k];
      ++o;
    }
}

// Fused batched Chebyshev term for any number of right-hand sides (and for
// 32-RHS blocks beyond 4 GiB).  Thread = (row, j), j fastest; the row data
// come per lane from global memory.
template <int ACC>
__global__ __launch_bounds__(kThreads) void k_cheb_spmm(const int32_t* __restrict__ rp,
                                                        const int32_t* __restrict__ ci,
                                                        const float2* __restrict__ val,
                                                        int64_t n_rows, int n_rhs, double scale,
                                                        const float2* __restrict__ tc,
                                                        float2* __restrict__ tpn,
                                                        float2* __restrict__ acc,
                                                        float2* __restrict__ comp, float2 c,
                                                        int first) {
  const int64_t e_idx = gblock<int64_t>() * kThreads + threadIdx.x;
  const int64_t row = e_idx / n_rhs;
  const int j = (int)(e_idx - row * n_rhs);
  if (row >= n_rows) return;
  RowSum<ACC> sum;
  const int32_t b = rp[row], e = rp[row + 1];
  int32_t k = b;
  for (; k + kGather <= e; k += kGather) {
    int32_t cc[kGather];
    float2 v[kGather], xs[kGather];
#pragma unroll
    for (int u = 0; u < kGather; ++u) { cc[u] = ci[k + u]; v[u] = val[k + u]; }
#pragma unroll
    for (int u = 0; u < kGather; ++u) xs[u] = tc[(int64_t)cc[u] * n_rhs + j];
#pragma unroll
    for (int u = 0; u < kGather; ++u) sum.add(v[u], xs[u]);
  }
  for (; k < e; ++k) sum.add(val[k], tc[(int64_t)ci[k] * n_rhs + j]);
  const float2 tp = first ? make_float2(0.f, 0.f) : tpn[e_idx];
  const float2 tn = sum.next(scale, tp, first != 0);
  tpn[e_idx] = tn;
  const float2 v = cmulc(c, tn);
  float2 a = acc[e_idx], k2 = comp[e_idx];
  neu_add(a.x, k2.x, v.x);
  neu_add(a.y, k2.y, v.y);
  acc[e_idx] = a;
  comp[e_idx] = k2;
}

// Fused Chebyshev term for one right-hand side (the single-state
// chebyshev_evolve and run_hamiltonian path).  kSpmvLanes lanes share a row:
// lane l of the group takes entries b + l, b + l + kSpmvLanes, ..., so each
// load instruction reads contiguous runs of the CSR arrays instead of one
// scattered entry per lane, and the partial row sums are merged across the
// group by lane exchanges (FP64 sums add; compensated pairs merge with one
// Neumaier step).  The group's first lane runs the recurrence epilogue.
// Measured on the 22-qubit chain (33.5 entries per row): one lane per row
// (the generic kernel) 1.56 ms per term; 4, 8, 16, 32 and 64 lanes per row
// 1.26, 0.87, 0.70, 1.14 and 1.56 ms.
#ifndef QSIM_SPMV_LANES
#define QSIM_SPMV_LANES 16
#endif
constexpr int kSpmvLanes = QSIM_SPMV_LANES;

template <int ACC>
__device__ __forceinline__ void merge_lanes(RowSum<ACC>& sum) {
#pragma unroll
  for (int m = 1; m < kSpmvLanes; m <<= 1) {
    if constexpr (ACC == kAccF64) {
      sum.r += lane_xor(sum.r, m);
      sum.i += lane_xor(sum.i, m);
    } else {
      const float sr = lane_xor(sum.sr, m), cr = lane_xor(sum.cr, m);
      const float si = lane_xor(sum.si, m), ci = lane_xor(sum.ci, m);
      neu_add(sum.sr, sum.cr, sr);
      sum.cr += cr;
      neu_add(sum.si, sum.ci, si);
      sum.ci += ci;
    }
  }
}

template <int ACC, bool NT>
__global__ __launch_bounds__(kThreads) void k_cheb_spmv1(const int32_t* __restrict__ rp,
                                                         const int32_t* __restrict__ ci,
                                                         const float2* __restrict__ val,
                                                         int64_t n_rows, double scale,
                                                         const float2* __restrict__ tc,
                                                         float2* __restrict__ tpn,
                                                         float2* __restrict__ acc,
                                                         float2* __restrict__ comp, float2 c,
                                                         int first) {
  const int64_t row = (gblock<int64_t>() * kThreads + threadIdx.x) / kSpmvLanes;
  const int sub = threadIdx.x % kSpmvLanes;
  const bool live = row < n_rows;
  RowSum<ACC> sum;
  if (live) {
    const int32_t b = rp[row], e = rp[row + 1];
    int32_t k = b + sub;
    for (; k + kSpmvLanes < e; k += 2 * kSpmvLanes) {
      const int32_t c0 = __builtin_nontemporal_load(ci + k), c1 = __builtin_nontemporal_load(ci + k + kSpmvLanes);
      const float2 v0 = ld2<true>(val + k), v1 = ld2<true>(val + k + kSpmvLanes);
      const float2 x0 = tc[c0], x1 = tc[c1];
      sum.add(v0, x0);
      sum.add(v1, x1);
    }
    if (k < e) sum.add(ld2<true>(val + k), tc[__builtin_nontemporal_load(ci + k)]);
  }
  merge_lanes<ACC>(sum);   // every lane of the group, live or not, takes part
  if (!live || sub != 0) return;
  const float2 tp = first ? make_float2(0.f, 0.f) : ld2<NT>(tpn + row);
  const float2 tn = sum.next(scale, tp, first != 0);
  st2<NT>(tpn + row, tn);
  const float2 v = cmulc(c, tn);
  float2 a = ld2<NT>(acc + row), k2 = ld2<NT>(comp + row);
  neu_add(a.x, k2.x, v.x);
  neu_add(a.y, k2.y, v.y);
  st2<NT>(acc + row, a);
  st2<NT>(comp + row, k2);
}

// Matrix-free single-state term over a Pauli sum grouped by X mask (the
// tables of qsim_pauli_csr_count): row r's entries are the nonzero group
// values v_g(r) (pauli_group_value, rounded to complex64 exactly as the
// device CSR stores them) at columns r ^ xm[g], summed in group order as
// FP64 sums of exact products.  One thread per row; the group and term loops
// are uniform across the wave, so the tables come in through scalar loads.
// gconst (optional) gives, per group, the entry already rounded to complex64
// when it is the same in every row (no term has a Y|Z mask), or NaN: such a
// group costs its gather and four FMAs, nothing else.
// No index or value stream is read: per term the only HBM traffic is the
// five state vectors, and the gathers of T_k hit the caches (T_k is 32 MiB
// at 22 qubits).  Each term of an entry costs an AND, a popcount, one
// shift-or that puts the parity into the sign of 1.0, and two FP64 FMAs
// (per-term selects of +-c took 18 instructions); a 22-qubit chain term
// takes 0.15 ms (with gconst) against 0.70 ms for the CSR term.  Walking
// XCD-contiguous row ranges measured the same.
constexpr int kPauliRun = 8;   // 4: same time, 16: 3 % slower, 1: 10 % slower (22-qubit chain)
// rows per thread, kThreads apart: the scalar table loads and the loop
// control serve two rows (22-qubit chain: 1 row 0.206 ms, 2 rows 0.164 ms,
// 4 rows 0.182 ms per term, same bits)
constexpr int kPauliRows = 2;
// Entries of one group g for RR rows (uniform group, per-row values).
template <typename R, int RR>
__device__ __forceinline__ void pauli_group_rows(const R (&r)[RR], int t0, int t1,
                                                 const uint64_t* __restrict__ sm,
                                                 const double2* __restrict__ coef, double (&re)[RR],
                                                 double (&im)[RR]) {
#pragma unroll
  for (int q = 0; q < RR; ++q) re[q] = im[q] = 0.0;
  for (int t = t0; t < t1; ++t) {
    const uint64_t m = sm[t];
    const double2 c = coef[t];
#pragma unroll
    for (int q = 0; q < RR; ++q) {
      const uint32_t pc = sizeof(R) == 4 ? (uint32_t)__popc((uint32_t)r[q] & (uint32_t)m)
                                         : (uint32_t)__popcll((uint64_t)r[q] & m);
      const double sgn = __hiloint2double((int)((pc << 31) | 0x3ff00000u), 0);
      re[q] = __builtin_fma(c.x, sgn, re[q]);
      im[q] = __builtin_fma(c.y, sgn, im[q]);
    }
  }
}

template <typename R, bool NT>
__global__ __launch_bounds__(kThreads) void k_cheb_pauli1(int n_groups, const uint64_t* __restrict__ xm,
                                                          const int* __restrict__ toff,
                                                          const uint64_t* __restrict__ sm,
                                                          const double2* __restrict__ coef,
                                                          const double2* __restrict__ gconst,
                                                          int64_t n_rows, double scale,
                                                          const float2* __restrict__ tc,
                                                          float2* __restrict__ tpn,
                                                          float2* __restrict__ acc,
                                                          float2* __restrict__ comp, float2 c,
                                                          int first) {
  constexpr int RR = kPauliRows;
  const int64_t row0 = gblock<int64_t>() * (kThreads * RR) + threadIdx.x;
  if (row0 >= n_rows) return;
  const R rmask = (R)(n_rows - 1);        // n_rows = 2^n: every column stays in range
  R r[RR];
#pragma unroll
  for (int q = 0; q < RR; ++q) r[q] = (R)(row0 + q * kThreads) & rmask;   // rows past the end are not stored
  RowSum<kAccF64> sum[RR];
  // groups in runs of kPauliRun: the run's gathers are issued together (an
  // entry that turns out to be zero is gathered and not used), then the
  // values are formed and added in group order
  int g = 0;
  for (; g + kPauliRun <= n_groups; g += kPauliRun) {
    float2 x[kPauliRun][RR];
#pragma unroll
    for (int j = 0; j < kPauliRun; ++j)
#pragma unroll
      for (int q = 0; q < RR; ++q) x[j][q] = tc[(r[q] ^ (R)xm[g + j]) & rmask];
#pragma unroll
    for (int j = 0; j < kPauliRun; ++j) {
      if (gconst) {
        const double2 v = gconst[g + j];
        if (v.x == v.x) {                   // the same entry in every row (uniform branch)
          if (v.x != 0.0 || v.y != 0.0) {
#pragma unroll
            for (int q = 0; q < RR; ++q) sum[q].add_d(v.x, v.y, x[j][q]);
          }
          continue;
        }
      }
      double re[RR], im[RR];
      pauli_group_rows<R, RR>(r, toff[g + j], toff[g + j + 1], sm, coef, re, im);
#pragma unroll
      for (int q = 0; q < RR; ++q)
        if (re[q] != 0.0 || im[q] != 0.0) sum[q].add(make_float2((float)re[q], (float)im[q]), x[j][q]);
    }
  }
  for (; g < n_groups; ++g) {
    if (gconst) {
      const double2 v = gconst[g];
      if (v.x == v.x) {
        if (v.x != 0.0 || v.y != 0.0) {
#pragma unroll
          for (int q = 0; q < RR; ++q) sum[q].add_d(v.x, v.y, tc[(r[q] ^ (R)xm[g]) & rmask]);
        }
        continue;
      }
    }
    double re[RR], im[RR];
    pauli_group_rows<R, RR>(r, toff[g], toff[g + 1], sm, coef, re, im);
#pragma unroll
    for (int q = 0; q < RR; ++q)
      if (re[q] != 0.0 || im[q] != 0.0)
        sum[q].add(make_float2((float)re[q], (float)im[q]), tc[(r[q] ^ (R)xm[g]) & rmask]);
  }
#pragma unroll
  for (int q = 0; q < RR; ++q) {
    const int64_t row = row0 + q * kThreads;
    if (row >= n_rows) break;
    const float2 tp = first ? make_float2(0.f, 0.f) : ld2<NT>(tpn + row);
    const float2 tn = sum[q].next(scale, tp, first != 0);
    st2<NT>(tpn + row, tn);
    const float2 v = cmulc(c, tn);
    float2 a = ld2<NT>(acc + row), k2 = ld2<NT>(comp + row);
    neu_add(a.x, k2.x, v.x);
    neu_add(a.y, k2.y, v.y);
    st2<NT>(acc + row, a);
    st2<NT>(comp + row, k2);
  }
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

// Order shift for a grid of `blocks` 32-RHS workgroups (2^kBlockRowsLog2
// rows each): the requested shift when the chunks tile the grid exactly,
// else 0.
static int spmm_order(int64_t blocks, int shift) {
  const int64_t chunk = shift > kBlockRowsLog2 ? int64_t(1) << (shift - kBlockRowsLog2) : 1;
  return blocks % (8 * chunk) == 0 ? shift : 0;
}

static int log2_exact(int64_t v) {
  if (v <= 0 || (v & (v - 1))) return -1;
  int n = 0;
  while ((int64_t(1) << n) < v) ++n;
  return n;
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
                        int first, int accumulate) {
  if (n_rows < 0 || n_rhs < 1) return fail(QSIM_EINVAL, "bad sizes (n_rows=%lld, n_rhs=%d)",
                                           (long long)n_rows, n_rhs);
  if (accumulate != QSIM_CHEB_ACC_FP64 && accumulate != QSIM_CHEB_ACC_FP32C)
    return fail(QSIM_EINVAL, "unknown accumulate mode %d", accumulate);
  if (n_rows == 0) return QSIM_OK;
  if (!row_ptr || !col_idx || !values || !t_cur || !t_prev_next || !acc || !comp)
    return fail(QSIM_EINVAL, "NULL pointer argument");
  hipStream_t s = as_stream(stream);
  const float2 cf = f2(c_re, c_im);
  const bool f64 = accumulate == QSIM_CHEB_ACC_FP64;
  // the 32-RHS kernel addresses T_k with 32-bit byte offsets (<= 2^24 rows)
  if (n_rhs == 32 && n_rows <= (int64_t(1) << 24)) {
    const int64_t blocks = (int64_t)ceil_div((uint64_t)n_rows, kWaves * kRowsPerWave);
    const int order = spmm_order(blocks, kSpmmOrderShift);
    const bool nt = nt_for_bytes((uint64_t)n_rows * 256u);
    auto* kern = f64 ? (nt ? k_cheb_spmm32<kAccF64, true, 0> : k_cheb_spmm32<kAccF64, false, 0>)
                     : (nt ? k_cheb_spmm32<kAccF32c, true, 0> : k_cheb_spmm32<kAccF32c, false, 0>);
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kThreads), 0, s, (const int32_t*)row_ptr,
                       (const int32_t*)col_idx, (const float2*)values, n_rows, scale,
                       (const float2*)t_cur, (float2*)t_prev_next, (float2*)acc, (float2*)comp,
                       (float2*)nullptr, cf, first, order, 0, 0);
    return launch_status("cheb_spmm_step");
  }
  if (n_rhs == 1) {
    const int64_t threads = n_rows * kSpmvLanes;
    const bool nt = nt_for_bytes((uint64_t)n_rows * 8u);
    auto* kern = f64 ? (nt ? k_cheb_spmv1<kAccF64, true> : k_cheb_spmv1<kAccF64, false>)
                     : (nt ? k_cheb_spmv1<kAccF32c, true> : k_cheb_spmv1<kAccF32c, false>);
    hipLaunchKernelGGL(kern, grid1d(ceil_div(threads, kThreads)), dim3(kThreads), 0, s,
                       (const int32_t*)row_ptr, (const int32_t*)col_idx, (const float2*)values, n_rows,
                       scale, (const float2*)t_cur, (float2*)t_prev_next, (float2*)acc, (float2*)comp,
                       cf, first);
    return launch_status("cheb_spmm_step");
  }
  const int64_t total = n_rows * (int64_t)n_rhs;
  auto* kern = f64 ? k_cheb_spmm<kAccF64> : k_cheb_spmm<kAccF32c>;
  hipLaunchKernelGGL(kern, grid1d(ceil_div(total, kThreads)), dim3(kThreads), 0, s,
                     (const int32_t*)row_ptr, (const int32_t*)col_idx, (const float2*)values,
                     n_rows, n_rhs, scale, (const float2*)t_cur, (float2*)t_prev_next, (float2*)acc,
                     (float2*)comp, cf, first);
  return launch_status("cheb_spmm_step");
}

int qsim_cheb_pauli_step(void* stream, int n_qubits, int n_groups, const void* xm, const void* toff,
                         const void* sm, const void* coef, const void* gconst, double scale, const void* t_cur,
                         void* t_prev_next, void* acc, void* comp, double c_re, double c_im,
                         int first) {
  if (n_qubits < 1 || n_qubits > QSIM_MAX_QUBITS)
    return fail(QSIM_ERANGE, "n_qubits=%d outside 1..%d", n_qubits, QSIM_MAX_QUBITS);
  if (n_groups < 0) return fail(QSIM_EINVAL, "n_groups < 0");
  if ((n_groups > 0 && (!xm || !toff || !sm || !coef)) || !t_cur || !t_prev_next || !acc || !comp)
    return fail(QSIM_EINVAL, "NULL pointer argument");
  const int64_t n_rows = int64_t(1) << n_qubits;
  const bool nt = nt_for_bytes((uint64_t)n_rows * 8u);
  auto* kern = n_qubits > 32 ? (nt ? k_cheb_pauli1<uint64_t, true> : k_cheb_pauli1<uint64_t, false>)
                             : (nt ? k_cheb_pauli1<uint32_t, true> : k_cheb_pauli1<uint32_t, false>);
  hipLaunchKernelGGL(kern, grid1d(ceil_div((uint64_t)n_rows, kThreads * kPauliRows)), dim3(kThreads), 0,
                     as_stream(stream), n_groups, (const uint64_t*)xm, (const int*)toff,
                     (const uint64_t*)sm, (const double2*)coef, (const double2*)gconst, n_rows, scale,
                     (const float2*)t_cur,
                     (float2*)t_prev_next, (float2*)acc, (float2*)comp, f2(c_re, c_im), first);
  return launch_status("cheb_pauli_step");
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

int64_t qsim_csr_split_workspace_bytes(int64_t n_rows) {
  if (n_rows < 1) return -1;
  const int64_t blocks = (int64_t)ceil_div((uint64_t)n_rows, kSplitThreads);
  return (2 * blocks + 2) * (int64_t)sizeof(int64_t);
}

static int check_split(int64_t n_rows, int split_bit, int& n) {
  n = log2_exact(n_rows);
  if (n < 2 || n_rows > (int64_t(1) << 30))
    return fail(QSIM_EINVAL, "the CSR split needs 2^n rows, 4 <= rows <= 2^30 (got %lld)", (long long)n_rows);
  if (split_bit < 1 || split_bit >= n)
    return fail(QSIM_EINVAL, "split_bit %d outside 1..%d", split_bit, n - 1);
  return QSIM_OK;
}

int qsim_csr_split_count(void* stream, const void* row_ptr, const void* col_idx, int64_t n_rows,
                         int split_bit, void* workspace, int64_t* nnz_lo, int64_t* nnz_hi) {
  int n = 0;
  if (int rc = check_split(n_rows, split_bit, n)) return rc;
  if (!row_ptr || !col_idx || !workspace || !nnz_lo || !nnz_hi) return fail(QSIM_EINVAL, "NULL pointer argument");
  hipStream_t s = as_stream(stream);
  const uint64_t blocks = ceil_div((uint64_t)n_rows, kSplitThreads);
  int64_t* bsum = (int64_t*)workspace;
  int64_t* tot = bsum + 2 * blocks;
  hipLaunchKernelGGL(k_split_count, dim3((unsigned)blocks), dim3(kSplitThreads), 0, s,
                     (const int32_t*)row_ptr, (const int32_t*)col_idx, n_rows, n, split_bit, bsum);
  if (int rc = launch_status("csr_split: count")) return rc;
  hipLaunchKernelGGL(k_split_scan, dim3(1), dim3(kSplitThreads), 0, s, blocks, bsum, tot);
  if (int rc = launch_status("csr_split: scan")) return rc;
  int64_t h[2] = {0, 0};
  hipError_t e = hipMemcpyAsync(h, tot, sizeof h, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(QSIM_EHIP, "csr_split: %s", hipGetErrorString(e));
  if (h[0] > INT32_MAX || h[1] > INT32_MAX) return fail(QSIM_ERANGE, "split CSR beyond int32 nonzeros");
  *nnz_lo = h[0];
  *nnz_hi = h[1];
  return QSIM_OK;
}

int qsim_csr_split_fill(void* stream, const void* row_ptr, const void* col_idx, const void* values,
                        int64_t n_rows, int split_bit, const void* workspace, void* rp_lo,
                        void* ci_lo, void* val_lo, void* rp_hi, void* ci_hi, void* val_hi) {
  int n = 0;
  if (int rc = check_split(n_rows, split_bit, n)) return rc;
  if (!row_ptr || !col_idx || !values || !workspace || !rp_lo || !ci_lo || !val_lo || !rp_hi ||
      !ci_hi || !val_hi)
    return fail(QSIM_EINVAL, "NULL pointer argument");
  const uint64_t blocks = ceil_div((uint64_t)n_rows, kSplitThreads);
  hipLaunchKernelGGL(k_split_fill, dim3((unsigned)blocks), dim3(kSplitThreads), 0, as_stream(stream),
                     (const int32_t*)row_ptr, (const int32_t*)col_idx, (const float2*)values, n_rows, n,
                     split_bit, (const int64_t*)workspace, (int32_t*)rp_lo, (int32_t*)ci_lo,
                     (float2*)val_lo, (int32_t*)rp_hi, (int32_t*)ci_hi, (float2*)val_hi);
  return launch_status("csr_split: fill");
}

int qsim_cheb_spmm_step_split(void* stream, const void* rp_lo, const void* ci_lo, const void* val_lo,
                              const void* rp_hi, const void* ci_hi, const void* val_hi, int64_t n_rows,
                              int split_bit, int n_rhs, double scale, const void* t_cur,
                              void* t_prev_next, void* acc, void* comp, void* w1, double c_re,
                              double c_im, int first) {
  int n = 0;
  if (int rc = check_split(n_rows, split_bit, n)) return rc;
  if (n_rhs != 32) return fail(QSIM_EINVAL, "the split term runs 32 right-hand sides (got %d)", n_rhs);
  if (n_rows > (int64_t(1) << 24)) return fail(QSIM_ERANGE, "the split term handles at most 2^24 rows");
  if (!rp_lo || !ci_lo || !val_lo || !rp_hi || !ci_hi || !val_hi || !t_cur || !t_prev_next || !acc ||
      !comp || !w1)
    return fail(QSIM_EINVAL, "NULL pointer argument");
  hipStream_t s = as_stream(stream);
  const float2 cf = f2(c_re, c_im);
  const int64_t blocks = (int64_t)ceil_div((uint64_t)n_rows, kWaves * kRowsPerWave);
  // each XCD walks one contiguous eighth of the rows (positions) in both passes
  const int order = spmm_order(blocks, n >= 3 ? n - 3 : 0);
  const bool nt = nt_for_bytes((uint64_t)n_rows * 256u);
  auto* k1 = nt ? k_cheb_spmm32<kAccF64, true, 1> : k_cheb_spmm32<kAccF64, false, 1>;
  auto* k2 = nt ? k_cheb_spmm32<kAccF64, true, 2> : k_cheb_spmm32<kAccF64, false, 2>;
  hipLaunchKernelGGL(k1, dim3((unsigned)blocks), dim3(kThreads), 0, s, (const int32_t*)rp_lo,
                     (const int32_t*)ci_lo, (const float2*)val_lo, n_rows, scale, (const float2*)t_cur,
                     (float2*)t_prev_next, (float2*)acc, (float2*)comp, (float2*)w1, cf, first, order,
                     n, split_bit);
  if (int rc = launch_status("cheb_spmm_step_split: low pass")) return rc;
  hipLaunchKernelGGL(k2, dim3((unsigned)blocks), dim3(kThreads), 0, s, (const int32_t*)rp_hi,
                     (const int32_t*)ci_hi, (const float2*)val_hi, n_rows, scale, (const float2*)t_cur,
                     (float2*)t_prev_next, (float2*)acc, (float2*)comp, (float2*)w1, cf, first, order,
                     n, split_bit);
  return launch_status("cheb_spmm_step_split: high pass");
}

}  // extern "C"
