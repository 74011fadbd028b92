// LDS-tile executors for gfx950: many gates per HBM pass.
//
// A tile is 2^m amplitudes whose indices vary over a chosen set of m state
// qubits (the tile mask) with the other bits fixed.  A workgroup copies its
// tile from HBM into LDS (160 KB per CU on gfx950, so m <= 14), applies a
// program of ops to it, and writes it back: the state crosses HBM once per
// program instead of once per gate.  The lowest L qubits of the mask are
// always qubits 0..L-1 (L >= 1), so a tile is read as runs of 2^L contiguous
// amplitudes with 16-byte float4 accesses.
//
// The same op executor serves three callers:
//   * qsim_tile_program: one program over every tile of one HBM state;
//   * qsim_apply_gates_fused: the reference's fused 1Q list (its tiled
//     kernel, bwk/csrc/apply_gate_sm120.cu:121-161, uses one block per tile
//     of 2^(max_t+1) and degenerates to a single block for high targets);
//   * qsim_run_instances: whole circuits of uneven size (n <= 14), one
//     workgroup per circuit, state resident in LDS for the whole program,
//     optionally sampled in-kernel with a counter-based RNG.
// Register-resident fusion (qsim_apply_gates_reg_tiled and its batched form)
// keeps 128 amplitudes per wave64 in VGPRs and exchanges partners with
// cross-lane shuffles, covering targets 0..7.
#include "qsim_internal.hpp"
#include <string.h>
#include <algorithm>

namespace qsim {

constexpr int kTileMaxBits = 14;
constexpr int kTileThreads = 256;

// ------------------------------------------------------------ op executor
__device__ __forceinline__ bool ctl_ok(uint32_t i, uint32_t cmask) { return (i & cmask) == cmask; }

// Compensated complex dot of a row of length K with vector a (K <= 64).
template <int K>
__device__ __forceinline__ float2 crowk(const float2* __restrict__ row, const float2* a) {
  float sr = 0.f, er = 0.f, si = 0.f, ei = 0.f;
#pragma unroll
  for (int c = 0; c < K; ++c) {
    const float2 g = row[c];
    float p, e;
    two_prod(g.x, a[c].x, p, e); sr += p; er += e;
    two_prod(-g.y, a[c].y, p, e); sr += p; er += e;
    two_prod(g.x, a[c].y, p, e); si += p; ei += e;
    two_prod(g.y, a[c].x, p, e); si += p; ei += e;
  }
  return make_float2(sr + er, si + ei);
}

// Dense 2^K block, one group per thread iteration; in-place safe because a
// thread reads its whole group into registers before writing any of it.
template <int K>
__device__ void op_mk(float2* sh, int m, const qsim_op& op, const float2* __restrict__ cf) {
  constexpr int D = 1 << K;
  const uint32_t groups = 1u << (m - K);
  int sq[K];
#pragma unroll
  for (int j = 0; j < K; ++j) sq[j] = op.q[j];
  // sorted positions for the insert-zero expansion
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int b = a + 1; b < K; ++b)
      if (sq[b] < sq[a]) { int t = sq[a]; sq[a] = sq[b]; sq[b] = t; }
  uint32_t bitpos[K];
#pragma unroll
  for (int j = 0; j < K; ++j) bitpos[j] = op.q[j];
  for (uint32_t g = threadIdx.x; g < groups; g += blockDim.x) {
    uint32_t base = g;
#pragma unroll
    for (int j = 0; j < K; ++j) base = ins0(base, sq[j]);
    if (!ctl_ok(base, op.cmask)) continue;
    float2 a[D];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      uint32_t o = base;
#pragma unroll
      for (int j = 0; j < K; ++j) o |= ((uint32_t)(r >> j) & 1u) << bitpos[j];
      a[r] = sh[o];
    }
#pragma unroll 1
    for (int r = 0; r < D; ++r) {
      uint32_t o = base;
#pragma unroll
      for (int j = 0; j < K; ++j) o |= ((uint32_t)(r >> j) & 1u) << bitpos[j];
      sh[o] = crowk<D>(cf + op.coef + r * D, a);
    }
  }
}

__device__ void apply_op(float2* sh, int m, const qsim_op& op, const float2* __restrict__ cf) {
  const uint32_t N = 1u << m;
  const int nt = blockDim.x;
  switch (op.kind) {
    case QSIM_OP_M1: {
      const int q = op.q[0];
      const float2 g0 = cf[op.coef], g1 = cf[op.coef + 1], g2 = cf[op.coef + 2], g3 = cf[op.coef + 3];
      for (uint32_t p = threadIdx.x; p < N / 2; p += nt) {
        const uint32_t i0 = ins0(p, q), i1 = i0 | (1u << q);
        if (!ctl_ok(i0, op.cmask)) continue;
        const float2 a0 = sh[i0], a1 = sh[i1];
        sh[i0] = crow2(g0, a0, g1, a1);
        sh[i1] = crow2(g2, a0, g3, a1);
      }
      break;
    }
    case QSIM_OP_M2: {
      const int qa = op.q[0], qb = op.q[1];
      const int lo = qa < qb ? qa : qb, hi = qa < qb ? qb : qa;
      float2 g[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) g[k] = cf[op.coef + k];
      for (uint32_t p = threadIdx.x; p < N / 4; p += nt) {
        const uint32_t b = ins0(ins0(p, lo), hi);
        if (!ctl_ok(b, op.cmask)) continue;
        const uint32_t i1 = b | (1u << qa), i2 = b | (1u << qb), i3 = i1 | (1u << qb);
        const float2 a0 = sh[b], a1 = sh[i1], a2 = sh[i2], a3 = sh[i3];
        sh[b] = crow4(g + 0, a0, a1, a2, a3);
        sh[i1] = crow4(g + 4, a0, a1, a2, a3);
        sh[i2] = crow4(g + 8, a0, a1, a2, a3);
        sh[i3] = crow4(g + 12, a0, a1, a2, a3);
      }
      break;
    }
    case QSIM_OP_D1:
    case QSIM_OP_D2:
    case QSIM_OP_DK: {
      const int nq = op.kind == QSIM_OP_D1 ? 1 : (op.kind == QSIM_OP_D2 ? 2 : op.nq);
      for (uint32_t i = threadIdx.x; i < N; i += nt) {
        if (!ctl_ok(i, op.cmask)) continue;
        uint32_t sel = 0;
#pragma unroll
        for (int j = 0; j < 6; ++j)
          if (j < nq) sel |= ((i >> op.q[j]) & 1u) << j;
        sh[i] = cmulc(sh[i], cf[op.coef + sel]);
      }
      break;
    }
    case QSIM_OP_PH: {
      const int q = op.q[0];
      const float2 ph = cf[op.coef];
      for (uint32_t p = threadIdx.x; p < N / 2; p += nt) {
        const uint32_t i1 = ins0(p, q) | (1u << q);
        if (!ctl_ok(i1, op.cmask)) continue;
        sh[i1] = cmulc(sh[i1], ph);
      }
      break;
    }
    case QSIM_OP_X: {
      const int q = op.q[0];
      for (uint32_t p = threadIdx.x; p < N / 2; p += nt) {
        const uint32_t i0 = ins0(p, q), i1 = i0 | (1u << q);
        if (!ctl_ok(i0, op.cmask)) continue;
        const float2 a0 = sh[i0], a1 = sh[i1];
        sh[i0] = a1;
        sh[i1] = a0;
      }
      break;
    }
    case QSIM_OP_SWAP: {
      const int qa = op.q[0], qb = op.q[1];
      const int lo = qa < qb ? qa : qb, hi = qa < qb ? qb : qa;
      for (uint32_t p = threadIdx.x; p < N / 4; p += nt) {
        const uint32_t b = ins0(ins0(p, lo), hi);
        if (!ctl_ok(b, op.cmask)) continue;
        const uint32_t i1 = b | (1u << qa), i2 = b | (1u << qb);
        const float2 a1 = sh[i1], a2 = sh[i2];
        sh[i1] = a2;
        sh[i2] = a1;
      }
      break;
    }
    case QSIM_OP_MK: {
      switch (op.nq) {
        case 1: op_mk<1>(sh, m, op, cf); break;
        case 2: op_mk<2>(sh, m, op, cf); break;
        case 3: op_mk<3>(sh, m, op, cf); break;
        case 4: op_mk<4>(sh, m, op, cf); break;
        case 5: op_mk<5>(sh, m, op, cf); break;
        default: break;  // rejected on the host
      }
      break;
    }
    default: break;
  }
}

// Coefficients an op reads from the table: up to kStaged of them are staged
// in LDS one op ahead; -1 means more than that (read from the table in place).
constexpr int kStaged = 16;
__device__ __forceinline__ int op_ncoef(const qsim_op& op) {
  switch (op.kind) {
    case QSIM_OP_M1: return 4;
    case QSIM_OP_M2: return 16;
    case QSIM_OP_D1: return 2;
    case QSIM_OP_D2: return 4;
    case QSIM_OP_DK: return op.nq <= 4 ? (1 << op.nq) : -1;
    case QSIM_OP_PH: return 1;
    case QSIM_OP_MK: return op.nq <= 2 ? (1 << (2 * op.nq)) : -1;
    default: return 0;
  }
}

// Runs ops get(0..n_ops-1) on the LDS state with a two-stage software
// pipeline: the descriptor of op k+2 and the coefficients of op k+1 are
// fetched from global memory while op k runs, so an op costs its LDS work
// instead of two dependent global-memory latencies.  Op k+1's coefficients
// land in the LDS slot that op k-1 used, before the barrier that ends op k.
template <class GetOp>
__device__ __forceinline__ void run_ops(float2* sh, int m, int n_ops, GetOp get,
                                        const float2* __restrict__ cf) {
  __shared__ float2 cslot[2][kStaged];
  if (n_ops <= 0) return;
  const int t = threadIdx.x;
  qsim_op cur = get(0);
  {
    const int nc = op_ncoef(cur);
    if (t < nc) cslot[0][t] = cf[cur.coef + t];
  }
  qsim_op nxt = n_ops > 1 ? get(1) : cur;
  __syncthreads();
  for (int k = 0; k < n_ops; ++k) {
    qsim_op nn = nxt;
    if (k + 2 < n_ops) nn = get(k + 2);
    const int ncn = k + 1 < n_ops ? op_ncoef(nxt) : 0;
    float2 pre = make_float2(0.f, 0.f);
    if (t < ncn) pre = cf[nxt.coef + t];
    if (op_ncoef(cur) >= 0) {
      qsim_op o = cur;
      o.coef = 0;
      apply_op(sh, m, o, cslot[k & 1]);
    } else {
      apply_op(sh, m, cur, cf);
    }
    if (t < ncn) cslot[(k + 1) & 1][t] = pre;
    __syncthreads();
    cur = nxt;
    nxt = nn;
  }
}

// ------------------------------------------------------------ tile geometry
struct TileGeom {
  int n, m, L;          // state qubits, tile bits, contiguous low bits
  int nh;               // m - L high tile bits
  uint8_t hp[kTileMaxBits];   // state positions of tile bits L..m-1
  int nc;               // n - m complement bits
  uint8_t cp[QSIM_MAX_QUBITS]; // state positions of the complement bits
};

// Loops run over the fixed array bounds with a predicate so that every
// access to the by-value geometry has a constant index (no private-memory
// copy of the kernel argument).
__device__ __forceinline__ uint64_t tile_base(uint64_t T, const TileGeom& g) {
  uint64_t b = 0;
#pragma unroll
  for (int k = 0; k < QSIM_MAX_QUBITS; ++k)
    if (k < g.nc) b |= ((T >> k) & 1ull) << g.cp[k];
  return b;
}
// global float4 slot of local float4 slot s (local amplitude 2s)
__device__ __forceinline__ uint64_t tile_slot(uint32_t s, uint64_t base, const TileGeom& g) {
  const uint32_t j = 2 * s;
  uint64_t off = j & ((1u << g.L) - 1u);
  const uint32_t jh = j >> g.L;
#pragma unroll
  for (int k = 0; k < kTileMaxBits; ++k)
    if (k < g.nh) off |= (uint64_t)((jh >> k) & 1u) << g.hp[k];
  return (base | off) >> 1;
}

__device__ void tile_load(float2* sh, const float4* __restrict__ st, uint64_t base,
                          const TileGeom& g) {
  const uint32_t ns = 1u << (g.m - 1);
  float4* sh4 = reinterpret_cast<float4*>(sh);
  for (uint32_t s = threadIdx.x; s < ns; s += blockDim.x) sh4[s] = st[tile_slot(s, base, g)];
}
__device__ void tile_store(const float2* sh, float4* __restrict__ st, uint64_t base,
                           const TileGeom& g) {
  const uint32_t ns = 1u << (g.m - 1);
  const float4* sh4 = reinterpret_cast<const float4*>(sh);
  for (uint32_t s = threadIdx.x; s < ns; s += blockDim.x) st[tile_slot(s, base, g)] = sh4[s];
}

__global__ __launch_bounds__(kTileThreads) void k_tile_program(float4* __restrict__ st, TileGeom g,
                                                              const qsim_op* __restrict__ ops,
                                                              int n_ops,
                                                              const float2* __restrict__ cf) {
  extern __shared__ float2 sh[];
  if (gblock<uint64_t>() >> g.nc) return;    // whole workgroup: past the last tile
  const uint64_t base = tile_base(gblock<uint64_t>(), g);
  tile_load(sh, st, base, g);
  __syncthreads();
  run_ops(sh, g.m, n_ops, [&](int k) { return ops[k]; }, cf);
  tile_store(sh, st, base, g);
}

// Fused 1Q list with the ops carried in the kernel arguments.
constexpr int kFusedMax = 64;
struct Fused1Q {
  uint8_t q[kFusedMax];   // tile-local target of gate k
  int first, count;       // gates [first, first+count) of the caller's list
};
__global__ __launch_bounds__(kTileThreads) void k_tile_fused1q(float4* __restrict__ st, TileGeom g,
                                                              Fused1Q f,
                                                              const float2* __restrict__ gates) {
  extern __shared__ float2 sh[];
  if (gblock<uint64_t>() >> g.nc) return;    // whole workgroup: past the last tile
  const uint64_t base = tile_base(gblock<uint64_t>(), g);
  tile_load(sh, st, base, g);
  __syncthreads();
  run_ops(sh, g.m, f.count, [&](int k) {
    qsim_op op;
    op.kind = QSIM_OP_M1;
    op.nq = 1;
    op.q[0] = f.q[k];
    op.cmask = 0;
    op.coef = (uint32_t)(4 * (f.first + k));
    return op;
  }, gates);
  tile_store(sh, st, base, g);
}

// Builds the geometry for a tile mask; the mask must contain qubit 0.
static int make_geom(TileGeom& g, int n, uint64_t mask) {
  memset(&g, 0, sizeof g);
  g.n = n;
  g.m = __builtin_popcountll(mask);
  if (mask >> n) return fail(QSIM_EINVAL, "tile_mask has bits above n_qubits");
  if (g.m < 1 || g.m > kTileMaxBits)
    return fail(QSIM_ERANGE, "tile of %d qubits outside 1..%d", g.m, kTileMaxBits);
  if (!(mask & 1)) return fail(QSIM_EINVAL, "tile_mask must contain qubit 0");
  int L = 0;
  while (L < n && (mask >> L & 1)) ++L;
  g.L = L;
  g.nh = 0;
  for (int q = L; q < n; ++q)
    if (mask >> q & 1) g.hp[g.nh++] = (uint8_t)q;
  g.nc = 0;
  for (int q = 0; q < n; ++q)
    if (!(mask >> q & 1)) g.cp[g.nc++] = (uint8_t)q;
  if (g.m < 2 && n >= 2) return fail(QSIM_ERANGE, "tile needs >= 2 qubits");
  return QSIM_OK;
}

// ------------------------------------------------------------ register tiles
// Wave w holds float4 slots [128w/2 ...): lane l owns slots b+l and b+64+l,
// i.e. amplitude bits: bit0 = half of the float4, bits1-6 = lane, bit7 =
// which slot.  Targets 1..6 exchange with lane ^ 2^(t-1).
template <int MASK>
__device__ __forceinline__ void xchg_apply(float2 v[4], const float2 g[4], int lane) {
  const bool low = (lane & MASK) == 0;
  const float2 ga = low ? g[0] : g[3];
  const float2 gb = low ? g[1] : g[2];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    float2 p;
    p.x = lane_xor(v[k].x, MASK);
    p.y = lane_xor(v[k].y, MASK);
    v[k] = crow2(ga, v[k], gb, p);
  }
}

__device__ __forceinline__ void reg_apply(float2 v[4], const float2 g[4], int t, int lane) {
  // v[0],v[1] = slot A (.lo,.hi); v[2],v[3] = slot B
  switch (t) {
    case 0: {
      const float2 a0 = crow2(g[0], v[0], g[1], v[1]), a1 = crow2(g[2], v[0], g[3], v[1]);
      const float2 b0 = crow2(g[0], v[2], g[1], v[3]), b1 = crow2(g[2], v[2], g[3], v[3]);
      v[0] = a0; v[1] = a1; v[2] = b0; v[3] = b1;
      break;
    }
    case 1: xchg_apply<1>(v, g, lane); break;
    case 2: xchg_apply<2>(v, g, lane); break;
    case 3: xchg_apply<4>(v, g, lane); break;
    case 4: xchg_apply<8>(v, g, lane); break;
    case 5: xchg_apply<16>(v, g, lane); break;
    case 6: xchg_apply<32>(v, g, lane); break;
    case 7: {
      const float2 a0 = crow2(g[0], v[0], g[1], v[2]), b0 = crow2(g[2], v[0], g[3], v[2]);
      const float2 a1 = crow2(g[0], v[1], g[1], v[3]), b1 = crow2(g[2], v[1], g[3], v[3]);
      v[0] = a0; v[2] = b0; v[1] = a1; v[3] = b1;
      break;
    }
    default: break;
  }
}

// n_slots = total float4 slots (all states); max_t = min(n, 8) - 1 bound on
// the targets: targets outside 0..max_t are skipped (the targets live in
// device memory, so the host cannot check them without a synchronisation).
__global__ __launch_bounds__(256) void k_reg_fused(float4* __restrict__ st, uint64_t n_slots,
                                                   const float2* __restrict__ gates,
                                                   const int32_t* __restrict__ targets,
                                                   int n_gates, int max_t) {
  const int lane = threadIdx.x & 63;
  const uint64_t wave = (gblock<uint64_t>() * 256 + threadIdx.x) >> 6;
  const uint64_t sa = wave * 128 + lane, sb = sa + 64;
  float4 A = sa < n_slots ? st[sa] : make_float4(0.f, 0.f, 0.f, 0.f);
  float4 B = sb < n_slots ? st[sb] : make_float4(0.f, 0.f, 0.f, 0.f);
  float2 v[4] = {lo2(A), hi2(A), lo2(B), hi2(B)};
  for (int k = 0; k < n_gates; ++k) {
    const int t = targets[k];
    if (t < 0 || t > max_t) continue;
    float2 g[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) g[j] = gates[4 * k + j];
    reg_apply(v, g, t, lane);
  }
  if (sa < n_slots) st[sa] = cat4(v[0], v[1]);
  if (sb < n_slots) st[sb] = cat4(v[2], v[3]);
}

static int reg_fused_impl(hipStream_t s, void* states, uint64_t n_states, int n, const void* gates,
                          const void* targets, int n_gates) {
  if (int rc = check_n(n)) return rc;
  if (!states || (n_gates > 0 && (!gates || !targets)))
    return fail(QSIM_EINVAL, "NULL pointer argument");
  if (n_gates < 0) return fail(QSIM_EINVAL, "n_gates < 0");
  if (n_gates == 0 || n_states == 0) return QSIM_OK;
  const uint64_t n_slots = n_states * ((1ull << n) / 2);
  const uint64_t waves = ceil_div(n_slots, 128);
  const int max_t = std::min(n, 8) - 1;
  hipLaunchKernelGGL(k_reg_fused, grid1d(ceil_div(waves, 4)), dim3(256), 0, s, (float4*)states, n_slots,
                     (const float2*)gates, (const int32_t*)targets, n_gates, max_t);
  return launch_status("apply_gates_reg_tiled");
}

// ------------------------------------------------------------ instances
// SplitMix64 (Steele, Lea and Flood, OOPSLA 2014) as a counter-based source:
// uniform u_k = (mix(seed + (k+1) * golden) >> 11) * 2^-53.
// Sampling of an LDS-resident state: thread c owns the contiguous chunk of
// 2^n / 256 amplitudes (n >= 8) or one amplitude (n < 8, threads >= 2^n own
// nothing).  Chunk totals are serial FP64 sums; their prefix is a serial scan
// by thread 0; a shot finds its chunk by binary search and its index by the
// same serial sum, so totals and search agree bit for bit.
__device__ void lds_sample(const float2* sh, int n, uint64_t seed, uint32_t shots,
                           int64_t* __restrict__ out, double* pre /* LDS, 257 doubles */) {
  const uint32_t N = 1u << n;
  const uint32_t chunk = N >= 256 ? N / 256 : 1;
  const uint32_t nch = N >= 256 ? 256 : N;
  const int c = threadIdx.x;
  double tot = 0.0;
  if (c < (int)nch)
    for (uint32_t i = 0; i < chunk; ++i) {
      const float2 a = sh[c * chunk + i];
      const double x = a.x, y = a.y;
      tot += x * x + y * y;
    }
  __syncthreads();  // everyone done reading before pre[] is written
  if (c < (int)nch) pre[c + 1] = tot;
  __syncthreads();
  if (c == 0) {
    pre[0] = 0.0;
    for (uint32_t k = 1; k <= nch; ++k) pre[k] += pre[k - 1];
  }
  __syncthreads();
  const double total = pre[nch];
  for (uint32_t k = threadIdx.x; k < shots; k += blockDim.x) {
    const double x = splitmix_u(seed, k) * total;
    uint32_t lo = 0, hi = nch - 1;   // largest chunk with pre[j] <= x
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (pre[mid] <= x) lo = mid; else hi = mid - 1;
    }
    while (lo > 0 && !(pre[lo + 1] > pre[lo])) --lo;
    const double y = x - pre[lo];
    double cum = 0.0;
    int64_t pick = -1, last = -1;
    for (uint32_t i = 0; i < chunk; ++i) {
      const float2 a = sh[lo * chunk + i];
      const double xa = a.x, ya = a.y;
      const double q = xa * xa + ya * ya;
      cum += q;
      if (q > 0.0) last = i;
      if (cum > y) { pick = i; break; }
    }
    if (pick < 0) pick = last >= 0 ? last : 0;
    out[k] = (int64_t)(lo * chunk) + pick;
  }
}

__global__ __launch_bounds__(kTileThreads) void k_instances(float2* __restrict__ states,
                                                           const qsim_instance* __restrict__ inst,
                                                           const qsim_op* __restrict__ ops,
                                                           const float2* __restrict__ cf,
                                                           int64_t* __restrict__ shots_out) {
  extern __shared__ float2 sh[];
  __shared__ double pre[257];
  const qsim_instance in = inst[blockIdx.x];
  const int n = in.n_qubits;
  const uint32_t N = 1u << n;
  float2* gst = states + in.state_off;
  if (in.init) {
    for (uint32_t i = threadIdx.x; i < N; i += blockDim.x)
      sh[i] = make_float2(i == 0 ? 1.f : 0.f, 0.f);
  } else {
    for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) sh[i] = gst[i];
  }
  __syncthreads();
  const qsim_op* iops = ops + in.op_off;
  run_ops(sh, n, (int)in.n_ops, [&](int k) { return iops[k]; }, cf);
  if (in.want & 1)
    for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) gst[i] = sh[i];
  if (in.shots) lds_sample(sh, n, in.seed, in.shots, shots_out + in.shot_off, pre);
}

// ------------------------------------------------------------ qv8 / qv4
// qv8: one wave per circuit, 256 amplitudes in LDS, lane = one 4-amplitude
// group per gate; row r = 2*bit(q1) + bit(q0), q0 = first listed qubit
// (bwk/csrc/qv8_sim_sm120.cu:88-90).
__global__ __launch_bounds__(64) void k_qv8(const float2* __restrict__ gm, const int32_t* __restrict__ gq,
                                            int64_t C, int G, float* __restrict__ probs) {
  __shared__ float2 sh[256];
  const int64_t c = blockIdx.x;
  const int lane = threadIdx.x;
  for (int i = lane; i < 256; i += 64) sh[i] = make_float2(i == 0 ? 1.f : 0.f, 0.f);
  __syncthreads();
  for (int k = 0; k < G; ++k) {
    const int q0 = gq[(c * G + k) * 2], q1 = gq[(c * G + k) * 2 + 1];
    const float2* g = gm + (c * G + k) * 16;
    float2 gg[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) gg[j] = g[j];
    const int lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;
    if (q0 != q1 && q0 >= 0 && q1 >= 0 && q0 < 8 && q1 < 8) {
      const uint32_t b = ins0(ins0((uint32_t)lane, lo), hi);
      const uint32_t i1 = b | (1u << q0), i2 = b | (1u << q1), i3 = i1 | (1u << q1);
      const float2 a0 = sh[b], a1 = sh[i1], a2 = sh[i2], a3 = sh[i3];
      sh[b] = crow4(gg + 0, a0, a1, a2, a3);
      sh[i1] = crow4(gg + 4, a0, a1, a2, a3);
      sh[i2] = crow4(gg + 8, a0, a1, a2, a3);
      sh[i3] = crow4(gg + 12, a0, a1, a2, a3);
    }
    __syncthreads();
  }
  for (int i = lane; i < 256; i += 64) {
    const float2 a = sh[i];
    const double x = a.x, y = a.y;
    probs[c * 256 + i] = (float)(x * x + y * y);
  }
}

// qv4: 16 circuits per wave (4 groups each).  gate_data row-major 4x4 as 16
// real parts then 16 imaginary parts; pair_ids index (0,1),(0,2),(0,3),
// (1,2),(1,3),(2,3) with r = 2*bit(qb) + bit(qa), qa < qb (qv4.py:14-21).
__constant__ uint8_t c_qv4_pairs[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
__global__ __launch_bounds__(64) void k_qv4(const float* __restrict__ gd, const int32_t* __restrict__ pid,
                                            int64_t C, int G, float* __restrict__ probs) {
  __shared__ float2 sh[16 * 16];
  const int lane = threadIdx.x;
  const int sub = lane >> 2;            // circuit within the wave
  const int grp = lane & 3;             // group within the circuit
  const int64_t c = (int64_t)blockIdx.x * 16 + sub;
  const bool live = c < C;
  float2* s = sh + sub * 16;
  for (int i = grp; i < 16; i += 4) s[i] = make_float2(i == 0 ? 1.f : 0.f, 0.f);
  __syncthreads();
  for (int k = 0; k < G; ++k) {
    if (live) {
      const int p = pid[c * G + k];
      if (p >= 0 && p < 6) {
        const float* g = gd + (c * G + k) * 32;
        float2 gg[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) gg[j] = make_float2(g[j], g[16 + j]);
        const int qa = c_qv4_pairs[p][0], qb = c_qv4_pairs[p][1];
        const uint32_t b = ins0(ins0((uint32_t)grp, qa), qb);
        const uint32_t i1 = b | (1u << qa), i2 = b | (1u << qb), i3 = i1 | (1u << qb);
        const float2 a0 = s[b], a1 = s[i1], a2 = s[i2], a3 = s[i3];
        s[b] = crow4(gg + 0, a0, a1, a2, a3);
        s[i1] = crow4(gg + 4, a0, a1, a2, a3);
        s[i2] = crow4(gg + 8, a0, a1, a2, a3);
        s[i3] = crow4(gg + 12, a0, a1, a2, a3);
      }
    }
    __syncthreads();
  }
  if (live)
    for (int i = grp; i < 16; i += 4) {
      const float2 a = s[i];
      const double x = a.x, y = a.y;
      probs[c * 16 + i] = (float)(x * x + y * y);
    }
}

}  // namespace qsim

using namespace qsim;

// Host check of a program that runs on a state (or tile) of m qubits: op
// kinds, target positions (in range, distinct), controls (in range, not on a
// target) and coefficient ranges.  `inst` names the instance in the message.
static int check_ops(const qsim_op* ops, uint32_t n_ops, int m, uint64_t n_coefs, int inst) {
  for (uint32_t k = 0; k < n_ops; ++k) {
    const qsim_op& o = ops[k];
    int nq = 0, ncoef = 0;
    switch (o.kind) {
      case QSIM_OP_M1: nq = 1; ncoef = 4; break;
      case QSIM_OP_M2: nq = 2; ncoef = 16; break;
      case QSIM_OP_D1: nq = 1; ncoef = 2; break;
      case QSIM_OP_D2: nq = 2; ncoef = 4; break;
      case QSIM_OP_PH: nq = 1; ncoef = 1; break;
      case QSIM_OP_X: nq = 1; ncoef = 0; break;
      case QSIM_OP_SWAP: nq = 2; ncoef = 0; break;
      case QSIM_OP_MK:
        nq = o.nq;
        if (nq < 1 || nq > 5) return fail(QSIM_EINVAL, "instance %d, op %u: dense block of %d qubits", inst, k, nq);
        ncoef = 1 << (2 * nq);
        break;
      case QSIM_OP_DK:
        nq = o.nq;
        if (nq < 1 || nq > 6) return fail(QSIM_EINVAL, "instance %d, op %u: diagonal of %d qubits", inst, k, nq);
        ncoef = 1 << nq;
        break;
      default: return fail(QSIM_EINVAL, "instance %d, op %u: unknown kind %d", inst, k, o.kind);
    }
    uint32_t tmask = 0;
    for (int j = 0; j < nq; ++j) {
      if (o.q[j] >= m) return fail(QSIM_EINVAL, "instance %d, op %u: qubit %d >= n_qubits %d", inst, k, o.q[j], m);
      if (tmask & (1u << o.q[j])) return fail(QSIM_EINVAL, "instance %d, op %u: repeated qubit %d", inst, k, o.q[j]);
      tmask |= 1u << o.q[j];
    }
    if ((m < 32 && (o.cmask >> m)) || (o.cmask & tmask))
      return fail(QSIM_EINVAL, "instance %d, op %u: control mask 0x%x invalid for %d qubits", inst, k, o.cmask, m);
    if ((uint64_t)o.coef + (uint64_t)ncoef > n_coefs)
      return fail(QSIM_EINVAL, "instance %d, op %u: coefficients [%u, +%d) beyond the %llu passed", inst, k,
                  o.coef, ncoef, (unsigned long long)n_coefs);
  }
  return QSIM_OK;
}

extern "C" {

int qsim_tile_program(void* stream, void* state, int n_qubits, uint64_t tile_mask,
                      const void* ops_dev, int n_ops, const void* coefs_dev) {
  if (int rc = check_n(n_qubits)) return rc;
  if (!state || !ops_dev || !coefs_dev) return fail(QSIM_EINVAL, "NULL pointer argument");
  if (n_ops <= 0) return n_ops == 0 ? QSIM_OK : fail(QSIM_EINVAL, "n_ops < 0");
  TileGeom g;
  if (int rc = make_geom(g, n_qubits, tile_mask)) return rc;
  const uint64_t tiles = 1ull << g.nc;
  const size_t lds = sizeof(float2) << g.m;
  hipLaunchKernelGGL(k_tile_program, grid1d(tiles), dim3(kTileThreads), lds, as_stream(stream),
                     (float4*)state, g, (const qsim_op*)ops_dev, n_ops, (const float2*)coefs_dev);
  return launch_status("tile_program");
}

// Groups consecutive gates whose distinct targets (plus the low qubits that
// fill the tile to 10 bits) fit one tile, one launch per group.
int qsim_apply_gates_fused(void* stream, void* state, int n_qubits, const void* gates_dev,
                           const int32_t* targets, int n_gates) {
  if (int rc = check_n(n_qubits)) return rc;
  if (n_gates < 0) return fail(QSIM_EINVAL, "n_gates < 0");
  if (n_gates == 0) return QSIM_OK;
  if (!state || !gates_dev || !targets) return fail(QSIM_EINVAL, "NULL pointer argument");
  for (int k = 0; k < n_gates; ++k)
    if (int rc = check_q(targets[k], n_qubits, "target")) return rc;
  const int mmax = std::min(n_qubits, 10);
  hipStream_t s = as_stream(stream);
  int k = 0;
  while (k < n_gates) {
    uint64_t tmask = 0;
    int cnt = 0;
    // extend while the union of targets plus qubit 0 fits mmax bits
    while (k + cnt < n_gates && cnt < kFusedMax) {
      const uint64_t nm = tmask | (1ull << targets[k + cnt]) | 1ull;
      if (__builtin_popcountll(nm) > mmax) break;
      tmask = nm;
      ++cnt;
    }
    // fill the tile with the lowest unused qubits up to mmax bits
    uint64_t mask = tmask | 1ull;
    for (int q = 0; q < n_qubits && __builtin_popcountll(mask) < mmax; ++q) mask |= 1ull << q;
    TileGeom g;
    if (int rc = make_geom(g, n_qubits, mask)) return rc;
    Fused1Q f;
    memset(&f, 0, sizeof f);
    f.first = k;
    f.count = cnt;
    for (int j = 0; j < cnt; ++j) {
      const int t = targets[k + j];
      f.q[j] = (uint8_t)__builtin_popcountll(mask & ((1ull << t) - 1));  // local position
    }
    const size_t lds = sizeof(float2) << g.m;
    hipLaunchKernelGGL(k_tile_fused1q, grid1d(1ull << g.nc), dim3(kTileThreads), lds, s,
                       (float4*)state, g, f, (const float2*)gates_dev);
    if (int rc = launch_status("apply_gates_fused")) return rc;
    k += cnt;
  }
  return QSIM_OK;
}

int qsim_apply_gates_reg_tiled(void* stream, void* state, int n_qubits, const void* gates_dev,
                               const void* targets_dev, int n_gates) {
  return reg_fused_impl(as_stream(stream), state, 1, n_qubits, gates_dev, targets_dev, n_gates);
}

int qsim_apply_gates_batched_fused(void* stream, void* states, int64_t n_states, int n_qubits,
                                   const void* gates_dev, const void* targets_dev, int n_gates) {
  if (n_states < 0) return fail(QSIM_EINVAL, "n_states < 0");
  return reg_fused_impl(as_stream(stream), states, (uint64_t)n_states, n_qubits, gates_dev,
                        targets_dev, n_gates);
}

int qsim_run_instances(void* stream, void* states_base, uint64_t n_state_slots,
                       const qsim_instance* inst_host, const void* inst_dev, int n_inst,
                       const qsim_op* ops_host, const void* ops_dev, uint64_t n_ops_total,
                       const void* coefs_dev, uint64_t n_coefs, void* shots_dev,
                       uint64_t n_shot_slots, int max_qubits) {
  clear_error();
  if (n_inst < 0) return fail(QSIM_EINVAL, "n_inst < 0");
  if (n_inst == 0) return QSIM_OK;
  if (!inst_host || !inst_dev || !ops_host || !ops_dev || !coefs_dev)
    return fail(QSIM_EINVAL, "NULL pointer argument");
  if (max_qubits < 1 || max_qubits > kTileMaxBits)
    return fail(QSIM_ERANGE, "instances hold at most %d qubits in LDS (got %d)", kTileMaxBits,
                max_qubits);
  for (int i = 0; i < n_inst; ++i) {
    const qsim_instance& in = inst_host[i];
    if (in.n_qubits < 1 || in.n_qubits > max_qubits)
      return fail(QSIM_EINVAL, "instance %d: n_qubits = %d outside 1..max_qubits (%d)", i,
                  in.n_qubits, max_qubits);
    if ((uint64_t)in.op_off + in.n_ops > n_ops_total)
      return fail(QSIM_EINVAL, "instance %d: ops [%u, %u + %u) beyond the %llu ops passed", i,
                  in.op_off, in.op_off, in.n_ops, (unsigned long long)n_ops_total);
    const uint64_t amps = 1ull << in.n_qubits;
    if ((!in.init || (in.want & 1)) &&
        (!states_base || in.state_off > n_state_slots || n_state_slots - in.state_off < amps))
      return fail(QSIM_EINVAL, "instance %d: state [%llu, +%llu) beyond the %llu state slots", i,
                  (unsigned long long)in.state_off, (unsigned long long)amps,
                  (unsigned long long)n_state_slots);
    if (in.shots && (!shots_dev || in.shot_off > n_shot_slots || n_shot_slots - in.shot_off < in.shots))
      return fail(QSIM_EINVAL, "instance %d: shots [%llu, +%u) beyond the %llu shot slots", i,
                  (unsigned long long)in.shot_off, in.shots, (unsigned long long)n_shot_slots);
    if (int rc = check_ops(ops_host + in.op_off, in.n_ops, in.n_qubits, n_coefs, i)) return rc;
  }
  const size_t lds = sizeof(float2) << max_qubits;
  hipLaunchKernelGGL(k_instances, dim3(n_inst), dim3(kTileThreads), lds, as_stream(stream),
                     (float2*)states_base, (const qsim_instance*)inst_dev, (const qsim_op*)ops_dev,
                     (const float2*)coefs_dev, (int64_t*)shots_dev);
  return launch_status("run_instances");
}

int qsim_qv8_simulate(void* stream, const void* gate_matrices, const void* gate_qubits,
                      int64_t n_circuits, int n_gates, void* probs_out) {
  if (n_circuits < 0 || n_gates < 0) return fail(QSIM_EINVAL, "negative size");
  if (n_circuits == 0) return QSIM_OK;
  if (!gate_matrices || !gate_qubits || !probs_out) return fail(QSIM_EINVAL, "NULL pointer argument");
  hipLaunchKernelGGL(k_qv8, dim3(n_circuits), dim3(64), 0, as_stream(stream),
                     (const float2*)gate_matrices, (const int32_t*)gate_qubits, n_circuits, n_gates,
                     (float*)probs_out);
  return launch_status("qv8_simulate");
}

int qsim_qv4_simulate(void* stream, const void* gate_data, const void* pair_ids, int64_t n_circuits,
                      int n_gates, void* probs_out) {
  if (n_circuits < 0 || n_gates < 0) return fail(QSIM_EINVAL, "negative size");
  if (n_circuits == 0) return QSIM_OK;
  if (!gate_data || !pair_ids || !probs_out) return fail(QSIM_EINVAL, "NULL pointer argument");
  hipLaunchKernelGGL(k_qv4, dim3(ceil_div(n_circuits, 16)), dim3(64), 0, as_stream(stream),
                     (const float*)gate_data, (const int32_t*)pair_ids, n_circuits, n_gates,
                     (float*)probs_out);
  return launch_status("qv4_simulate");
}

}  // extern "C"
