// Reductions, measurement and shot sampling for gfx950.
//
// Every sum is FP64 and deterministic: pass 1 gives each workgroup a fixed
// contiguous range of float4 slots and reduces it in a fixed order (per-lane
// serial sum, then a wave64 butterfly, then the four waves in order); pass 2
// is a single workgroup that adds the partials in index order.  No float
// atomics are used, so the same state always gives the same bits (the
// reference accumulates FP32 with atomicAdd: bwk/csrc/measure_sm120.cu:46-68,
// bwk/csrc/renorm_sm120.cu:33-49).
//
// Sampling never materialises a CDF of the whole state.  The state is cut
// into leaves of 256 amplitudes.  One wave computes a leaf's |a|^2 in FP64
// and takes an inclusive scan across the wave in a fixed order; the last
// value is the leaf total.  An exclusive FP64 scan of the leaf totals gives
// the leaf prefix.  For each shot x = u * total: a binary search on the leaf
// prefix finds the leaf, and a wave scan of that leaf finds the smallest
// index with nonzero probability whose cumulative probability exceeds
// x - prefix.  Where rounding puts x past the leaf's scanned total, the leaf's
// last nonzero amplitude is taken, and the nonzero test keeps a rounding step
// in the cumulative sums from ever selecting a zero-probability amplitude.
// (Rule: smallest i with cdf(i) > u * total; the reference returns the
// smallest i with cdf[i] >= u on an FP32 CDF, bwk/csrc/sample_sm120.cu:48-69.)
#include "qsim_internal.hpp"
#include <mutex>

namespace qsim {

constexpr int kRedUnroll = 4;      // slots in flight per lane in pass 1
constexpr int kLeaf = 256;         // amplitudes per sampling leaf
constexpr int kScanSeg = 4096;     // leaf totals per scan segment

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += lane_xor(v, m);
  return v;
}

// Block of 256 threads: fixed-order sum; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = ((sh[0] + sh[1]) + (sh[2] + sh[3]));
  __syncthreads();
  return r;
}

__device__ __forceinline__ double p2(float2 a) {
  const double x = a.x, y = a.y;
  return x * x + y * y;
}

// Pass 1.  mode 0: sum |a|^2.  mode 1: also the part with bit t == 0
// (out[2*b] = total, out[2*b+1] = part with bit t = 0).
template <int MODE, bool NT>
__global__ __launch_bounds__(kThreads) void k_red1(const float4* __restrict__ st, uint64_t n4,
                                                   uint64_t chunk, int t, double* __restrict__ out) {
  __shared__ double sh[4];
  const uint64_t beg = (uint64_t)blockIdx.x * chunk;
  const uint64_t end = beg + chunk < n4 ? beg + chunk : n4;
  double s = 0.0, s0 = 0.0;
  // kRedUnroll slots per lane are loaded before any is summed (more bytes in
  // flight per wave); the summation order stays fixed
  uint64_t f = beg + threadIdx.x;
  for (; f + (kRedUnroll - 1) * kThreads < end; f += kRedUnroll * kThreads) {
    float4 v[kRedUnroll];
#pragma unroll
    for (int u = 0; u < kRedUnroll; ++u) v[u] = ld4<NT>(st + f + u * kThreads);
#pragma unroll
    for (int u = 0; u < kRedUnroll; ++u) {
      const double a = p2(lo2(v[u])), b = p2(hi2(v[u]));
      s += a + b;
      if (MODE == 1) {
        const uint64_t i0 = 2 * (f + u * kThreads);
        if (t == 0) s0 += a;
        else if (!((i0 >> t) & 1)) s0 += a + b;
      }
    }
  }
  for (; f < end; f += kThreads) {
    const float4 v = ld4<NT>(st + f);
    const double a = p2(lo2(v)), b = p2(hi2(v));
    s += a + b;
    if (MODE == 1) {
      const uint64_t i0 = 2 * f;
      if (t == 0) s0 += a;
      else if (!((i0 >> t) & 1)) s0 += a + b;
    }
  }
  const double r = block_sum(s, sh);
  if (MODE == 1) {
    const double r0 = block_sum(s0, sh);
    if (threadIdx.x == 0) { out[2 * blockIdx.x] = r; out[2 * blockIdx.x + 1] = r0; }
  } else {
    if (threadIdx.x == 0) out[blockIdx.x] = r;
  }
}

// Pass 2: one workgroup adds nb partials of width w (1 or 2) in order.
// res[0] = sum of column 0, res[1] = sum of column 1.
__global__ __launch_bounds__(kThreads) void k_red2(const double* __restrict__ part, int nb, int w,
                                                   double* __restrict__ res) {
  __shared__ double sh[4];
  double a = 0.0, b = 0.0;
  for (int j = threadIdx.x; j < nb; j += kThreads) {
    a += part[j * w];
    if (w == 2) b += part[j * w + 1];
  }
  const double ra = block_sum(a, sh);
  const double rb = block_sum(b, sh);
  if (threadIdx.x == 0) { res[0] = ra; res[1] = rb; }
}

struct RedPlan {
  uint64_t n4, chunk;
  int blocks;
};
static RedPlan red_plan(int n) {
  RedPlan p;
  p.n4 = (1ull << n) / 2;
  if (p.n4 == 0) p.n4 = 1;  // n = 0 is rejected earlier; keep arithmetic sane
  const uint64_t want = ceil_div(p.n4, 4 * kThreads);
  p.blocks = (int)(want < (uint64_t)kRedBlocks ? (want ? want : 1) : kRedBlocks);
  p.chunk = ceil_div(p.n4, p.blocks);
  return p;
}

int reduce_impl(hipStream_t s, const void* state, int n, int mode, int t, double* ws) {
  const RedPlan p = red_plan(n);
  double* res = ws + 2 * kRedBlocks;
  const bool nt = nt_for_qubits(n);
  if (mode == 0 && nt)
    hipLaunchKernelGGL((k_red1<0, true>), dim3(p.blocks), dim3(kThreads), 0, s, (const float4*)state,
                       p.n4, p.chunk, t, ws);
  else if (mode == 0)
    hipLaunchKernelGGL((k_red1<0, false>), dim3(p.blocks), dim3(kThreads), 0, s, (const float4*)state,
                       p.n4, p.chunk, t, ws);
  else if (nt)
    hipLaunchKernelGGL((k_red1<1, true>), dim3(p.blocks), dim3(kThreads), 0, s, (const float4*)state,
                       p.n4, p.chunk, t, ws);
  else
    hipLaunchKernelGGL((k_red1<1, false>), dim3(p.blocks), dim3(kThreads), 0, s, (const float4*)state,
                       p.n4, p.chunk, t, ws);
  if (int rc = launch_status("reduce pass 1")) return rc;
  hipLaunchKernelGGL(k_red2, dim3(1), dim3(kThreads), 0, s, ws, p.blocks, mode == 0 ? 1 : 2, res);
  return launch_status("reduce pass 2");
}

// ---------------------------------------------------------------- measure
// res[0] = total, res[1] = p0 (bit t = 0).  The outcome is decided on the
// device: 0 if random_val * total < p0, else 1; the kept half is scaled by
// 1/sqrt(p_outcome) in FP64, the other half is zeroed.
template <bool NT>
__global__ __launch_bounds__(kThreads) void k_collapse(float4* __restrict__ st, uint64_t n4, int t,
                                                       double u, const double* __restrict__ res,
                                                       int* __restrict__ outcome) {
  const double tot = res[0], p0 = res[1];
  const int o = (u * tot < p0) ? 0 : 1;
  const double po = o == 0 ? p0 : tot - p0;
  const double sc = po > 0.0 ? 1.0 / sqrt(po) : 0.0;
  const uint64_t f = gblock<uint64_t>() * kThreads + threadIdx.x;
  if (f == 0) *outcome = o;
  if (f >= n4) return;
  float4 v = ld4<NT>(st + f);
  const uint64_t i0 = 2 * f;
  const int b0 = (int)((i0 >> t) & 1), b1 = (int)(((i0 + 1) >> t) & 1);
  const double k0 = b0 == o ? sc : 0.0, k1 = b1 == o ? sc : 0.0;
  v.x = (float)(v.x * k0); v.y = (float)(v.y * k0);
  v.z = (float)(v.z * k1); v.w = (float)(v.w * k1);
  st4<NT>(st + f, v);
}

template <bool NT>
__global__ __launch_bounds__(kThreads) void k_scale(float4* __restrict__ st, uint64_t n4,
                                                    const double* __restrict__ res) {
  const double tot = res[0];
  const double sc = tot > 0.0 ? 1.0 / sqrt(tot) : 0.0;
  const uint64_t f = gblock<uint64_t>() * kThreads + threadIdx.x;
  if (f >= n4) return;
  float4 v = ld4<NT>(st + f);
  v.x = (float)(v.x * sc); v.y = (float)(v.y * sc);
  v.z = (float)(v.z * sc); v.w = (float)(v.w * sc);
  st4<NT>(st + f, v);
}

template <bool NT>
__global__ __launch_bounds__(kThreads) void k_probs(const float4* __restrict__ st, uint64_t n4,
                                                    float2* __restrict__ probs) {
  const uint64_t f = gblock<uint64_t>() * kThreads + threadIdx.x;
  if (f >= n4) return;
  const float4 v = ld4<NT>(st + f);
  probs[f] = make_float2((float)p2(lo2(v)), (float)p2(hi2(v)));
}

// ---------------------------------------------------------------- sampling
// Wave-level view of one leaf of 256 amplitudes: lane l holds slot l
// (amplitudes 2l, 2l+1) and slot 64 + l (amplitudes 128 + 2l, 129 + 2l), so
// each of the two loads reads 1 KiB contiguous.  q[0..3] are their |a|^2 in
// FP64, c[0..3] the inclusive cumulative sums within the leaf, in amplitude
// order: the first half by a fixed-order Hillis-Steele scan of the lane pairs,
// the second half the same way on top of the first half's last cumulative
// value (taken from lane 63, so the sequence is exactly non-decreasing across
// the halves).  nl = amplitudes in the leaf (256, or 2^n for tiny states);
// missing amplitudes count as zero.
struct LeafScan {
  double q[4], c[4];
};

__device__ __forceinline__ double wave_incl_scan(double v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double y = __shfl_up(v, d, 64);
    if (lane >= d) v += y;
  }
  return v;
}

__device__ __forceinline__ void leaf_scan(const float4* __restrict__ leaf, int lane, int nl,
                                          LeafScan& L) {
  L.q[0] = L.q[1] = L.q[2] = L.q[3] = 0.0;
  if (2 * lane < nl) {
    const float4 v0 = leaf[lane];
    L.q[0] = p2(lo2(v0)); L.q[1] = p2(hi2(v0));
  }
  if (128 + 2 * lane < nl) {
    const float4 v1 = leaf[64 + lane];
    L.q[2] = p2(lo2(v1)); L.q[3] = p2(hi2(v1));
  }
  const double pa = L.q[0] + L.q[1];
  const double ia = wave_incl_scan(pa, lane);
  double prev = __shfl_up(ia, 1, 64);
  double base = lane == 0 ? 0.0 : prev;
  L.c[0] = base + L.q[0];
  L.c[1] = L.c[0] + L.q[1];
  const double half = __shfl(L.c[1], 63, 64);
  const double pb = L.q[2] + L.q[3];
  const double ib = wave_incl_scan(pb, lane);
  prev = __shfl_up(ib, 1, 64);
  base = half + (lane == 0 ? 0.0 : prev);
  L.c[2] = base + L.q[2];
  L.c[3] = L.c[2] + L.q[3];
}

// Leaf totals: one wave per leaf, the four |a|^2 of each lane (the
// leaf_scan layout, 1 KiB per load) summed across the wave by the fixed
// butterfly of wave_sum.  The in-leaf search recomputes the cumulative sums
// with leaf_scan; where rounding makes the two disagree, an x past the scan's
// total takes the leaf's last nonzero amplitude.
template <bool NT>
__global__ __launch_bounds__(kThreads) void k_leaf_totals(const float4* __restrict__ st,
                                                          uint64_t n_amps, uint64_t n_leaves,
                                                          double* __restrict__ tot) {
  const int lane = threadIdx.x & 63;
  const uint64_t leaf = gblock<uint64_t>() * 4 + (threadIdx.x >> 6);
  if (leaf >= n_leaves) return;
  const int nl = (int)(n_amps < (uint64_t)kLeaf ? n_amps : kLeaf);
  const float4* lf = st + leaf * (kLeaf / 2);
  double v = 0.0;
  if (2 * lane < nl) {
    const float4 a = ld4<NT>(lf + lane);
    v = p2(lo2(a)) + p2(hi2(a));
  }
  if (128 + 2 * lane < nl) {
    const float4 b = ld4<NT>(lf + 64 + lane);
    v += p2(lo2(b)) + p2(hi2(b));
  }
  v = wave_sum(v);
  if (lane == 0) tot[leaf] = v;
}

// Exclusive scan of the leaf totals in segments of kScanSeg: pass A writes
// each segment's sum, pass B scans the segment sums (one workgroup, serial
// over chunks of 256 with a fixed-order block scan), pass C writes the
// exclusive prefix of every leaf.
__device__ __forceinline__ double block_incl_scan(double v, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double y = __shfl_up(v, d, 64);
    if (lane >= d) v += y;
  }
  if (lane == 63) sh[wid] = v;
  __syncthreads();
  double off = 0.0;
  for (int k = 0; k < wid; ++k) off += sh[k];
  __syncthreads();
  return v + off;
}

__global__ __launch_bounds__(kThreads) void k_seg_sum(const double* __restrict__ tot, uint64_t nl,
                                                      double* __restrict__ seg) {
  __shared__ double sh[4];
  const uint64_t beg = (uint64_t)blockIdx.x * kScanSeg;
  double s = 0.0;
  for (int k = threadIdx.x; k < kScanSeg; k += kThreads) {
    const uint64_t j = beg + k;
    if (j < nl) s += tot[j];
  }
  const double r = block_sum(s, sh);
  if (threadIdx.x == 0) seg[blockIdx.x] = r;
}

__global__ __launch_bounds__(kThreads) void k_seg_scan(double* __restrict__ seg, uint64_t ns,
                                                       double* __restrict__ total) {
  __shared__ double sh[4];
  double carry = 0.0;
  for (uint64_t b = 0; b < ns; b += kThreads) {
    const uint64_t j = b + threadIdx.x;
    const double v = j < ns ? seg[j] : 0.0;
    const double incl = block_incl_scan(v, sh);
    if (j < ns) seg[j] = carry + (incl - v);  // exclusive
    __shared__ double last;
    if (threadIdx.x == kThreads - 1) last = incl;
    __syncthreads();
    carry += last;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kThreads) void k_leaf_prefix(const double* __restrict__ tot,
                                                          uint64_t nl,
                                                          const double* __restrict__ seg,
                                                          double* __restrict__ pre) {
  __shared__ double sh[4];
  const uint64_t beg = (uint64_t)blockIdx.x * kScanSeg;
  double carry = seg[blockIdx.x];
  for (int b = 0; b < kScanSeg && beg + b < nl; b += kThreads) {
    const uint64_t j = beg + b + threadIdx.x;
    const double v = j < nl ? tot[j] : 0.0;
    const double incl = block_incl_scan(v, sh);
    if (j < nl) pre[j] = carry + (incl - v);
    __shared__ double last;
    if (threadIdx.x == kThreads - 1) last = incl;
    __syncthreads();
    carry += last;
    __syncthreads();
  }
}

// States of at most kScanSeg leaves (n <= 20): the three scan passes in one
// workgroup, with the same operations in the same order (the total is the
// segment sum of k_seg_sum; the single segment's exclusive offset is 0).
__device__ __forceinline__ void scan_small(const double* __restrict__ tot, uint64_t nl,
                                           double* __restrict__ pre, double* __restrict__ total,
                                           double* sh) {
  double s = 0.0;
  for (int k = threadIdx.x; k < kScanSeg; k += kThreads)
    if ((uint64_t)k < nl) s += tot[k];
  const double r = block_sum(s, sh);
  double carry = 0.0;
  for (int b = 0; b < kScanSeg && (uint64_t)b < nl; b += kThreads) {
    const uint64_t j = b + threadIdx.x;
    const double v = j < nl ? tot[j] : 0.0;
    const double incl = block_incl_scan(v, sh);
    if (j < nl) pre[j] = carry + (incl - v);
    __shared__ double last;
    if (threadIdx.x == kThreads - 1) last = incl;
    __syncthreads();
    carry += last;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = r;
}

__global__ __launch_bounds__(kThreads) void k_scan_small(const double* __restrict__ tot, uint64_t nl,
                                                         double* __restrict__ pre,
                                                         double* __restrict__ total) {
  __shared__ double sh[4];
  scan_small(tot, nl, pre, total, sh);
}

// Last amplitude of the leaf with nonzero probability (-1 if none).
__device__ __forceinline__ int64_t leaf_last_nonzero(const LeafScan& L, int lane) {
  const uint64_t hi = __ballot(L.q[2] > 0.0 || L.q[3] > 0.0);
  if (hi) {
    const int fl = 63 - __clzll((unsigned long long)hi);
    return 128 + 2 * fl + (__shfl(L.q[3] > 0.0 ? 1 : 0, fl, 64));
  }
  const uint64_t lo = __ballot(L.q[0] > 0.0 || L.q[1] > 0.0);
  if (!lo) return -1;
  const int fl = 63 - __clzll((unsigned long long)lo);
  return 2 * fl + (__shfl(L.q[1] > 0.0 ? 1 : 0, fl, 64));
}

// Index sampled for x = u * total by one wave.  Leaf j = largest j with
// pre[j] <= x; a leaf with zero total reached by rounding steps back to the
// previous nonzero leaf, and an x past the leaf's own total (rounding again)
// takes the leaf's last nonzero amplitude, so a zero-probability index is
// never returned.
__device__ __forceinline__ int64_t sample_one(const float4* __restrict__ st, uint64_t n_amps,
                                              uint64_t nl, const double* __restrict__ pre,
                                              const double* __restrict__ tot, double x, int lane) {
  uint64_t lo = 0, hi = nl - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= x) lo = mid; else hi = mid - 1;
  }
  uint64_t leaf = lo;
  while (leaf > 0 && !(tot[leaf] > 0.0)) --leaf;
  const double y = x - pre[leaf];
  const int nla = (int)(n_amps < (uint64_t)kLeaf ? n_amps : kLeaf);
  LeafScan L;
  leaf_scan(st + leaf * (kLeaf / 2), lane, nla, L);
  // smallest index with nonzero probability whose cumulative exceeds y:
  // first half (amplitudes 2l, 2l+1), then second half (128 + 2l, 129 + 2l)
  const bool h0 = L.q[0] > 0.0 && L.c[0] > y, h1 = L.q[1] > 0.0 && L.c[1] > y;
  const bool h2 = L.q[2] > 0.0 && L.c[2] > y, h3 = L.q[3] > 0.0 && L.c[3] > y;
  const uint64_t hit_lo = __ballot(h0 || h1), hit_hi = __ballot(h2 || h3);
  int64_t local;
  if (hit_lo) {
    const int fl = __ffsll((unsigned long long)hit_lo) - 1;
    local = 2 * fl + __shfl(h0 ? 0 : 1, fl, 64);
  } else if (hit_hi) {
    const int fl = __ffsll((unsigned long long)hit_hi) - 1;
    local = 128 + 2 * fl + __shfl(h2 ? 0 : 1, fl, 64);
  } else {
    local = leaf_last_nonzero(L, lane);
    if (local < 0) local = 0;  // all-zero state: index 0 (as the reference)
  }
  return (int64_t)(leaf * kLeaf) + local;
}

// One wave per shot.
__global__ __launch_bounds__(kThreads) void k_sample(const float4* __restrict__ st, uint64_t n_amps,
                                                     uint64_t nl, const double* __restrict__ pre,
                                                     const double* __restrict__ tot,
                                                     const double* __restrict__ total,
                                                     const double* __restrict__ u, uint64_t shots,
                                                     int64_t* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const uint64_t shot = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (shot >= shots) return;
  const int64_t idx = sample_one(st, n_amps, nl, pre, tot, u[shot] * total[0], lane);
  if (lane == 0) out[shot] = idx;
}

// ---------------------------------------------------------------- batched sampling
// Many states at once (the serving HBM lane): qsim_sample_rec r names a
// state of 8..20 qubits, its output and seed, and the prefix offsets of its
// leaves and shots.  Each state's leaf totals, scan and searches are exactly
// those of qsim_sample on it alone, with u_k = splitmix_u(seed, k) drawn in
// the kernel as qsim_uniforms_splitmix draws them.
__device__ __forceinline__ int find_rec(const qsim_sample_rec* __restrict__ recs, int n,
                                        uint64_t v, bool by_leaf) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const uint64_t key = by_leaf ? recs[mid].leaf_off : recs[mid].shot_off;
    if (key <= v) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kThreads) void k_leaf_totals_batch(const qsim_sample_rec* __restrict__ recs,
                                                                int n_recs, uint64_t n_leaves,
                                                                double* __restrict__ tot) {
  const int lane = threadIdx.x & 63;
  const uint64_t leaf = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (leaf >= n_leaves) return;
  const qsim_sample_rec& r = recs[find_rec(recs, n_recs, leaf, true)];
  const float4* lf = reinterpret_cast<const float4*>(r.state) + (leaf - r.leaf_off) * (kLeaf / 2);
  const float4 a = lf[lane], b = lf[64 + lane];   // small states, cached (MALL-resident)
  double v = p2(lo2(a)) + p2(hi2(a));
  v += p2(lo2(b)) + p2(hi2(b));
  v = wave_sum(v);
  if (lane == 0) tot[leaf] = v;
}

__global__ __launch_bounds__(kThreads) void k_scan_small_batch(const qsim_sample_rec* __restrict__ recs,
                                                               const double* __restrict__ tot,
                                                               double* __restrict__ pre,
                                                               double* __restrict__ totals) {
  __shared__ double sh[4];
  const qsim_sample_rec& r = recs[blockIdx.x];
  const uint64_t nl = 1ull << (r.n_qubits - 8);
  scan_small(tot + r.leaf_off, nl, pre + r.leaf_off, totals + blockIdx.x, sh);
}

__global__ __launch_bounds__(kThreads) void k_sample_batch(const qsim_sample_rec* __restrict__ recs,
                                                           int n_recs, uint64_t n_shots,
                                                           const double* __restrict__ tot,
                                                           const double* __restrict__ pre,
                                                           const double* __restrict__ totals) {
  const int lane = threadIdx.x & 63;
  const uint64_t shot = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (shot >= n_shots) return;
  const int ri = find_rec(recs, n_recs, shot, false);
  const qsim_sample_rec& r = recs[ri];
  const uint64_t k = shot - r.shot_off;
  const double x = splitmix_u(r.seed, k) * totals[ri];
  const uint64_t n_amps = 1ull << r.n_qubits;
  const int64_t idx = sample_one(reinterpret_cast<const float4*>(r.state), n_amps, n_amps / kLeaf,
                                 pre + r.leaf_off, tot + r.leaf_off, x, lane);
  if (lane == 0) reinterpret_cast<int64_t*>(r.out)[k] = idx;
}

struct SamplePlan {
  uint64_t n_amps, nl, ns;
};
static SamplePlan sample_plan(int n) {
  SamplePlan p;
  p.n_amps = 1ull << n;
  p.nl = ceil_div(p.n_amps, kLeaf);
  p.ns = ceil_div(p.nl, kScanSeg);
  return p;
}

__global__ __launch_bounds__(kThreads) void k_splitmix(uint64_t seed, int64_t count,
                                                      double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (k < count) out[k] = splitmix_u(seed, (uint64_t)k);
}

// ------------------------------------------------------------ no-alloc sampling
// Scratch of qsim_sample_noalloc: the sampling workspace (leaf totals, leaf
// prefixes, segment sums, total) of a state of up to kNoAllocMaxQubits
// qubits, as a file-scope device array.  Every GPU the library runs on holds
// its own copy (4 MiB) until the process ends.  The host never passes the
// array's name to a kernel: noalloc_scratch() looks up this device's
// address once with hipGetSymbolAddress and hands that out.
constexpr int kNoAllocMaxQubits = 26;
constexpr uint64_t kNoAllocLeaves = (1ull << kNoAllocMaxQubits) / kLeaf;
constexpr uint64_t kNoAllocDoubles = 2 * kNoAllocLeaves + kNoAllocLeaves / kScanSeg + 2;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on the device that owns `stream` (the current
// device for the null stream), looked up once per device with that device
// current.
static int noalloc_scratch(hipStream_t stream, void** out) {
  int dev = 0;
  hipError_t e = stream ? hipStreamGetDevice(stream, &dev) : hipGetDevice(&dev);
  if (e != hipSuccess) return fail(QSIM_EHIP, "no device for the stream: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(QSIM_ERANGE, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int prev = 0;
    if ((e = hipGetDevice(&prev)) != hipSuccess)
      return fail(QSIM_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (prev != dev && (e = hipSetDevice(dev)) != hipSuccess)
      return fail(QSIM_EHIP, "hipSetDevice(%d): %s", dev, hipGetErrorString(e));
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != dev) hipSetDevice(prev);
    if (e != hipSuccess || !p)
      return fail(QSIM_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return QSIM_OK;
}

}  // namespace qsim

using namespace qsim;

extern "C" {

int64_t qsim_reduce_workspace_bytes(int n_qubits) {
  (void)n_qubits;
  return kRedWsDoubles * (int64_t)sizeof(double);
}

int qsim_norm2(void* stream, const void* state, int n_qubits, void* workspace, void* out_dev) {
  if (int rc = check_n(n_qubits)) return rc;
  if (!state || !workspace || !out_dev) return fail(QSIM_EINVAL, "NULL pointer argument");
  hipStream_t s = as_stream(stream);
  double* ws = (double*)workspace;
  if (int rc = reduce_impl(s, state, n_qubits, 0, 0, ws)) return rc;
  if (hipMemcpyAsync(out_dev, ws + 2 * kRedBlocks, sizeof(double), hipMemcpyDeviceToDevice, s) !=
      hipSuccess)
    return fail(QSIM_EHIP, "norm2: copy failed");
  return QSIM_OK;
}

int qsim_measure_probs(void* stream, const void* state, int n_qubits, void* probs_dev) {
  if (int rc = check_n(n_qubits)) return rc;
  if (!state || !probs_dev) return fail(QSIM_EINVAL, "NULL pointer argument");
  const uint64_t n4 = (1ull << n_qubits) / 2;
  if (nt_for_qubits(n_qubits))
    hipLaunchKernelGGL(k_probs<true>, grid1d(ceil_div(n4, kThreads)), dim3(kThreads), 0,
                       as_stream(stream), (const float4*)state, n4, (float2*)probs_dev);
  else
    hipLaunchKernelGGL(k_probs<false>, grid1d(ceil_div(n4, kThreads)), dim3(kThreads), 0,
                       as_stream(stream), (const float4*)state, n4, (float2*)probs_dev);
  return launch_status("measure_probs");
}

int qsim_measure_qubit(void* stream, void* state, int n_qubits, int target, double random_val,
                       void* workspace, int* outcome) {
  if (int rc = check_n(n_qubits)) return rc;
  if (int rc = check_q(target, n_qubits, "target_qubit")) return rc;
  if (!state || !workspace || !outcome) return fail(QSIM_EINVAL, "NULL pointer argument");
  hipStream_t s = as_stream(stream);
  double* ws = (double*)workspace;
  if (int rc = reduce_impl(s, state, n_qubits, 1, target, ws)) return rc;
  const uint64_t n4 = (1ull << n_qubits) / 2;
  int* dev_out = (int*)(ws + 2 * kRedBlocks + 4);
  if (nt_for_qubits(n_qubits))
    hipLaunchKernelGGL(k_collapse<true>, grid1d(ceil_div(n4, kThreads)), dim3(kThreads), 0, s,
                       (float4*)state, n4, target, random_val, ws + 2 * kRedBlocks, dev_out);
  else
    hipLaunchKernelGGL(k_collapse<false>, grid1d(ceil_div(n4, kThreads)), dim3(kThreads), 0, s,
                       (float4*)state, n4, target, random_val, ws + 2 * kRedBlocks, dev_out);
  if (int rc = launch_status("measure_qubit collapse")) return rc;
  int host = -1;
  if (hipMemcpyAsync(&host, dev_out, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
    return fail(QSIM_EHIP, "measure_qubit: copy failed");
  if (hipStreamSynchronize(s) != hipSuccess) return fail(QSIM_EHIP, "measure_qubit: sync failed");
  *outcome = host;
  return QSIM_OK;
}

int qsim_renormalize(void* stream, void* state, int n_qubits, void* workspace) {
  if (int rc = check_n(n_qubits)) return rc;
  if (!state || !workspace) return fail(QSIM_EINVAL, "NULL pointer argument");
  hipStream_t s = as_stream(stream);
  double* ws = (double*)workspace;
  if (int rc = reduce_impl(s, state, n_qubits, 0, 0, ws)) return rc;
  const uint64_t n4 = (1ull << n_qubits) / 2;
  if (nt_for_qubits(n_qubits))
    hipLaunchKernelGGL(k_scale<true>, grid1d(ceil_div(n4, kThreads)), dim3(kThreads), 0, s,
                       (float4*)state, n4, ws + 2 * kRedBlocks);
  else
    hipLaunchKernelGGL(k_scale<false>, grid1d(ceil_div(n4, kThreads)), dim3(kThreads), 0, s,
                       (float4*)state, n4, ws + 2 * kRedBlocks);
  return launch_status("renormalize");
}

// Layout (doubles): tot[nl] | pre[nl] | seg[ns] | total[1]
int64_t qsim_sample_workspace_bytes(int n_qubits) {
  if (n_qubits < 1 || n_qubits > QSIM_MAX_QUBITS) return -1;
  const SamplePlan p = sample_plan(n_qubits);
  return (int64_t)(2 * p.nl + p.ns + 2) * (int64_t)sizeof(double);
}

int qsim_sample(void* stream, const void* state, int n_qubits, const void* uniforms_dev,
                int64_t n_shots, void* workspace, void* out_dev) {
  if (int rc = check_n(n_qubits)) return rc;
  if (n_shots < 0) return fail(QSIM_EINVAL, "n_shots < 0");
  if (n_shots == 0) return QSIM_OK;
  if (!state || !uniforms_dev || !workspace || !out_dev)
    return fail(QSIM_EINVAL, "NULL pointer argument");
  hipStream_t s = as_stream(stream);
  const SamplePlan p = sample_plan(n_qubits);
  double* tot = (double*)workspace;
  double* pre = tot + p.nl;
  double* seg = pre + p.nl;
  double* total = seg + p.ns;
  if (nt_for_qubits(n_qubits))
    hipLaunchKernelGGL(k_leaf_totals<true>, grid1d(ceil_div(p.nl, 4)), dim3(kThreads), 0, s,
                       (const float4*)state, p.n_amps, p.nl, tot);
  else
    hipLaunchKernelGGL(k_leaf_totals<false>, grid1d(ceil_div(p.nl, 4)), dim3(kThreads), 0, s,
                       (const float4*)state, p.n_amps, p.nl, tot);
  if (int rc = launch_status("sample: leaf totals")) return rc;
  if (p.ns == 1) {
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kThreads), 0, s, tot, p.nl, pre, total);
    if (int rc = launch_status("sample: leaf prefix")) return rc;
  } else {
    hipLaunchKernelGGL(k_seg_sum, dim3(p.ns), dim3(kThreads), 0, s, tot, p.nl, seg);
    if (int rc = launch_status("sample: segment sums")) return rc;
    hipLaunchKernelGGL(k_seg_scan, dim3(1), dim3(kThreads), 0, s, seg, p.ns, total);
    if (int rc = launch_status("sample: segment scan")) return rc;
    hipLaunchKernelGGL(k_leaf_prefix, dim3(p.ns), dim3(kThreads), 0, s, tot, p.nl, seg, pre);
    if (int rc = launch_status("sample: leaf prefix")) return rc;
  }
  hipLaunchKernelGGL(k_sample, dim3(ceil_div((uint64_t)n_shots, 4)), dim3(kThreads), 0, s,
                     (const float4*)state, p.n_amps, p.nl, pre, tot, total,
                     (const double*)uniforms_dev,
                     (uint64_t)n_shots, (int64_t*)out_dev);
  return launch_status("sample: search");
}

int qsim_sample_noalloc(void* stream, const void* state, int n_qubits, const void* uniforms_dev,
                        int64_t n_shots, void* out_dev) {
  if (int rc = check_n(n_qubits)) return rc;
  if (n_qubits > kNoAllocMaxQubits)
    return fail(QSIM_ERANGE, "qsim_sample_noalloc holds scratch for at most %d qubits (got %d); "
                "use qsim_sample with a workspace", kNoAllocMaxQubits, n_qubits);
  if (n_shots < 0) return fail(QSIM_EINVAL, "n_shots < 0");
  if (n_shots == 0) return QSIM_OK;
  if (!state || !uniforms_dev || !out_dev) return fail(QSIM_EINVAL, "NULL pointer argument");
  if (qsim_sample_workspace_bytes(n_qubits) > (int64_t)sizeof(scratch))
    return fail(QSIM_ERANGE, "sampling workspace of %d qubits exceeds the static scratch", n_qubits);
  void* ws = nullptr;
  if (int rc = noalloc_scratch(as_stream(stream), &ws)) return rc;
  return qsim_sample(stream, state, n_qubits, uniforms_dev, n_shots, ws, out_dev);
}

// Layout (doubles): tot[n_leaves] | pre[n_leaves] | totals[n_states]
int64_t qsim_sample_batch_workspace_bytes(int64_t n_leaves, int64_t n_states) {
  if (n_leaves < 0 || n_states < 0) return -1;
  return (2 * n_leaves + n_states + 1) * (int64_t)sizeof(double);
}

int qsim_sample_batch(void* stream, const qsim_sample_rec* recs_host, const void* recs_dev,
                      int64_t n_states, void* workspace) {
  clear_error();
  if (n_states < 0 || n_states > (1 << 24)) return fail(QSIM_ERANGE, "n_states = %lld out of range",
                                                      (long long)n_states);
  if (n_states == 0) return QSIM_OK;
  if (!recs_host || !recs_dev || !workspace) return fail(QSIM_EINVAL, "NULL pointer argument");
  uint64_t leaves = 0, shots = 0;
  for (int64_t i = 0; i < n_states; ++i) {
    const qsim_sample_rec& r = recs_host[i];
    if (r.n_qubits < 8 || r.n_qubits > 20)
      return fail(QSIM_ERANGE, "record %lld: batched sampling needs 8 <= n_qubits <= 20 (got %d)",
                  (long long)i, r.n_qubits);
    if (!r.state || (r.shots && !r.out)) return fail(QSIM_EINVAL, "record %lld: NULL state or output",
                                                     (long long)i);
    if (r.leaf_off != leaves || r.shot_off != shots)
      return fail(QSIM_EINVAL, "record %lld: leaf_off/shot_off must be the prefix sums of "
                  "2^(n_qubits-8) and shots", (long long)i);
    leaves += 1ull << (r.n_qubits - 8);
    shots += r.shots;
  }
  hipStream_t s = as_stream(stream);
  const qsim_sample_rec* rd = (const qsim_sample_rec*)recs_dev;
  double* tot = (double*)workspace;
  double* pre = tot + leaves;
  double* totals = pre + leaves;
  hipLaunchKernelGGL(k_leaf_totals_batch, dim3((unsigned)ceil_div(leaves, 4)), dim3(kThreads), 0, s, rd,
                     (int)n_states, leaves, tot);
  if (int rc = launch_status("sample_batch: leaf totals")) return rc;
  hipLaunchKernelGGL(k_scan_small_batch, dim3((unsigned)n_states), dim3(kThreads), 0, s, rd, tot, pre,
                     totals);
  if (int rc = launch_status("sample_batch: scans")) return rc;
  if (shots == 0) return QSIM_OK;
  hipLaunchKernelGGL(k_sample_batch, dim3((unsigned)ceil_div(shots, 4)), dim3(kThreads), 0, s, rd,
                     (int)n_states, shots, tot, pre, totals);
  return launch_status("sample_batch: search");
}

int qsim_uniforms_splitmix(void* stream, uint64_t seed, int64_t count, void* uniforms_dev) {
  clear_error();
  if (count < 0) return fail(QSIM_EINVAL, "count < 0");
  if (count == 0) return QSIM_OK;
  if (!uniforms_dev) return fail(QSIM_EINVAL, "NULL pointer argument");
  hipLaunchKernelGGL(k_splitmix, dim3((unsigned)ceil_div((uint64_t)count, kThreads)), dim3(kThreads), 0,
                     as_stream(stream), seed, count, (double*)uniforms_dev);
  return launch_status("uniforms_splitmix");
}

}  // extern "C"
