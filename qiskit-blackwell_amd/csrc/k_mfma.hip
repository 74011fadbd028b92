// Dense 5-qubit blocks on the gfx950 matrix cores (FP32 MFMA).
//
// A fused block U (32x32 complex) acts on every group of 32 amplitudes that
// differ only in the block's qubits q[0..4].  Over many groups this is the
// batched product Y = U X with X = [x_0 ... x_{G-1}] (32 x G complex), i.e.
// four real products: Re Y = Re U Re X - Im U Im X, Im Y = Im U Re X +
// Re U Im X.  One wave owns 32 groups (one 32x32 output tile for Re Y and one
// for Im Y) and runs 64 v_mfma_f32_32x32x2_f32 with FP32 accumulation.
//
// Operand layout of v_mfma_f32_32x32x2f32 (A 32x2, B 2x32, D 32x32): lane l
// supplies A[l%32][l/32] and B[l/32][l%32]; D[i][j] sits in lane
// j + 32*((i/4)%2), accumulator register (i%4) + 4*(i/8).  With the K slice
// {2s, 2s+1} lane (h = l/32, c = l%32) therefore needs U[c][2s+h] for A and
// x_c[2s+h] for B, and ends with rows i = 8*(r/4) + 4*h + r%4 of column c.
//
// The 32 groups of a wave are consecutive in the group index, whose bits are
// the non-block qubits in ascending order.  For a fixed row every B load of
// the wave reads the same amplitude of 32 consecutive groups: 256 contiguous
// bytes when qubits 0..4 are outside the block, a few cache lines otherwise.
// The A fragments (64 floats per lane) stay in VGPRs for the whole launch.
//
// Error budget: coefficients rounded to FP32 and 64-term FP32 accumulations
// give a relative error per amplitude of about sqrt(64) * 2^-24 per block,
// far inside the fidelity target, so no split-precision correction is run.
#include "qsim_internal.hpp"
#include <string.h>
#include <algorithm>

namespace qsim {

typedef float f32x16 __attribute__((ext_vector_type(16)));

struct BlockArgs {
  int n;
  int q[5];          // block qubits, row bit j <-> q[j]
  int nc;            // n - 5 non-block qubits
  uint8_t cp[QSIM_MAX_QUBITS];  // their positions, ascending
};

__device__ __forceinline__ uint64_t group_base(uint64_t g, const BlockArgs& a) {
  uint64_t b = 0;
#pragma unroll
  for (int k = 0; k < QSIM_MAX_QUBITS; ++k)
    if (k < a.nc) b |= ((g >> k) & 1ull) << a.cp[k];
  return b;
}

__device__ __forceinline__ uint64_t row_off(int r, const BlockArgs& a) {
  uint64_t o = 0;
#pragma unroll
  for (int j = 0; j < 5; ++j) o |= (uint64_t)((r >> j) & 1) << a.q[j];
  return o;
}

// Direct variant: operands straight from HBM into the MFMA lane layout,
// persistent waves with one set of 32 groups prefetched into registers while
// the MFMA chains of the current set run (load, MFMA and store phases of one
// wave overlap instead of alternating).  Row offsets are 32-bit (n <= 32).
template <int W>
__global__ __launch_bounds__(64 * W) void k_block5(float2* __restrict__ st, uint64_t n_sets,
                                                   BlockArgs a, const float2* __restrict__ U) {
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5, c = lane & 31;
  float ur[16], ui[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const float2 u = U[c * 32 + 2 * s + h];
    ur[s] = u.x;
    ui[s] = u.y;
  }
  uint32_t rof[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) rof[s] = (uint32_t)row_off(2 * s + h, a);
  const uint64_t stride = (uint64_t)gridDim.x * W;
  uint64_t set = (uint64_t)blockIdx.x * W + (threadIdx.x >> 6);
  if (set >= n_sets) return;
  uint64_t base = group_base(set * 32 + c, a);
  float2 x[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) x[s] = st[base | rof[s]];
  while (true) {
    const uint64_t nxt = set + stride;
    const bool more = nxt < n_sets;
    uint64_t nbase = 0;
    float2 y[16];
    if (more) {
      nbase = group_base(nxt * 32 + c, a);
#pragma unroll
      for (int s = 0; s < 16; ++s) y[s] = st[nbase | rof[s]];
    }
    f32x16 accr = {0.f}, acci = {0.f};
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      accr = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[s], x[s].x, accr, 0, 0, 0);
      accr = __builtin_amdgcn_mfma_f32_32x32x2f32(-ui[s], x[s].y, accr, 0, 0, 0);
      acci = __builtin_amdgcn_mfma_f32_32x32x2f32(ui[s], x[s].x, acci, 0, 0, 0);
      acci = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[s], x[s].y, acci, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = 8 * (r >> 2) + 4 * h + (r & 3);
      st[base | row_off(i, a)] = make_float2(accr[r], acci[r]);
    }
    if (!more) break;
    set = nxt;
    base = nbase;
#pragma unroll
    for (int s = 0; s < 16; ++s) x[s] = y[s];
  }
}

// Batched block step (the serving HBM lane): one launch applies one block
// to each of many instance states.  Record r (qsim_block_rec) names the
// state, the block matrix and the qubits, and covers the launch's sets
// [set_begin_r, set_begin_r + 2^(n_r - 10)) of 32 groups each.  Wave w takes
// sets w*spw .. w*spw + spw - 1; each set's record is found by a binary
// search over set_begin (the same for the whole wave), and its groups go
// through exactly the MFMA sequence of k_block5, so an instance's result
// does not depend on what else is in the batch.  The loads of a wave's next
// set are issued before the current set's MFMA chains when both sets belong
// to one record.  Row offsets are 32-bit (n <= 32, checked on the host).
__global__ __launch_bounds__(256) void k_block5_batch(const qsim_block_rec* __restrict__ recs,
                                                      int n_recs, uint64_t total_sets, int spw) {
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5, c = lane & 31;
  const uint64_t wave = (uint64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int cur = -1;
  float ur[16], ui[16];
  uint32_t rof[16];
  int qv[5], qs[5];
  float2* st = nullptr;
  uint64_t sb = 0, se = 0;          // sets [sb, se) belong to record cur
  float2 x[16], y[16];
  uint64_t base = 0;
  bool have = false;                // x already holds this set's amplitudes
  for (int k = 0; k < spw; ++k) {
    const uint64_t set = wave * spw + k;
    if (set >= total_sets) break;
    if (!have) {
      int lo = 0, hi = n_recs - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (recs[mid].set_begin <= set) lo = mid; else hi = mid - 1;
      }
      if (lo != cur) {
        cur = lo;
        const qsim_block_rec& r = recs[lo];
        st = reinterpret_cast<float2*>(r.state);
        sb = r.set_begin;
        se = lo + 1 < n_recs ? recs[lo + 1].set_begin : total_sets;
        const float2* U = reinterpret_cast<const float2*>(r.mat);
#pragma unroll
        for (int s = 0; s < 16; ++s) {
          const float2 u = U[c * 32 + 2 * s + h];
          ur[s] = u.x;
          ui[s] = u.y;
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) qv[j] = r.q[j];
#pragma unroll
        for (int s = 0; s < 16; ++s) {
          const int row = 2 * s + h;
          uint32_t o = 0;
#pragma unroll
          for (int j = 0; j < 5; ++j) o |= (uint32_t)((row >> j) & 1) << qv[j];
          rof[s] = o;
        }
        // ascending block positions, for inserting zeros into the group index
#pragma unroll
        for (int j = 0; j < 5; ++j) qs[j] = qv[j];
#pragma unroll
        for (int i = 0; i < 5; ++i)
#pragma unroll
          for (int j = 0; j < 4 - i; ++j)
            if (qs[j] > qs[j + 1]) { const int t = qs[j]; qs[j] = qs[j + 1]; qs[j + 1] = t; }
      }
      base = (set - sb) * 32 + c;
#pragma unroll
      for (int j = 0; j < 5; ++j) base = ins0(base, qs[j]);
#pragma unroll
      for (int s = 0; s < 16; ++s) x[s] = st[base | rof[s]];
    }
    // the wave's next set, when it belongs to the same record, is loaded
    // before this set's MFMA chains
    const uint64_t nset = set + 1;
    const bool pre = k + 1 < spw && nset < se;
    uint64_t nbase = 0;
    if (pre) {
      nbase = (nset - sb) * 32 + c;
#pragma unroll
      for (int j = 0; j < 5; ++j) nbase = ins0(nbase, qs[j]);
#pragma unroll
      for (int s = 0; s < 16; ++s) y[s] = st[nbase | rof[s]];
    }
    f32x16 accr = {0.f}, acci = {0.f};
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      accr = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[s], x[s].x, accr, 0, 0, 0);
      accr = __builtin_amdgcn_mfma_f32_32x32x2f32(-ui[s], x[s].y, accr, 0, 0, 0);
      acci = __builtin_amdgcn_mfma_f32_32x32x2f32(ui[s], x[s].x, acci, 0, 0, 0);
      acci = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[s], x[s].y, acci, 0, 0, 0);
    }
    // register r of lane (h, c) holds row i = 8 (r / 4) + 4 h + r % 4 of group c
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = 8 * (r >> 2) + 4 * h + (r & 3);
      uint32_t o = 0;
#pragma unroll
      for (int j = 0; j < 5; ++j) o |= (uint32_t)((i >> j) & 1) << qv[j];
      st[base | o] = make_float2(accr[r], acci[r]);
    }
    have = pre;
    if (pre) {
      base = nbase;
#pragma unroll
      for (int s = 0; s < 16; ++s) x[s] = y[s];
    }
  }
}

// Wide variant (n >= 11, n <= 32): 16-byte lanes over sets of 64 groups.
//   Q0 = 0 (qubit 0 outside the block): lane (h, c) holds groups 2c and 2c+1
//     of row 2s+h as one float4 (they are adjacent amplitudes); the even and
//     odd groups run as two MFMA column tiles.
//   Q0 = 1 (block row bit 0 = qubit 0): lane (h, c) holds rows 2s and 2s+1 of
//     group c + 32h as one float4; one cross-lane swap of the halves per K
//     slice gives the MFMA layout, and the results come out with rows i, i+1
//     of a group in the same lane, so stores need no exchange.
// Set order, IL: 0, each wave walks a contiguous range of sets (XCD-major
// workgroup order), so a wave's accesses to one row form runs of up to
// 2^(lowest block qubit) amplitudes; 1, all waves advance together through
// one window of (waves in the grid) sets, so at any moment the whole GPU
// reads one contiguous window of every row.  Measured at 30 qubits on
// MI355X, the window order is 4-19 % faster whenever the lowest block qubit
// is below 12, and 18-22 % slower when it is 12 (32 KiB runs per wave).
// PF: the next set's loads are issued before the current MFMA chains.
// NT: nontemporal loads and stores, used only when every access covers whole
// 128-byte lines (no block qubit among 0..3).
template <int PF, int NT, int Q0, int IL>
__global__ __launch_bounds__(256) void k_block5w(float2* __restrict__ st, uint64_t n_sets, BlockArgs a,
                                                 const float2* __restrict__ U) {
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5, c = lane & 31;
  float ur[16], ui[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const float2 u = U[c * 32 + 2 * s + h];
    ur[s] = u.x;
    ui[s] = u.y;
  }
  uint32_t rof[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) rof[s] = (uint32_t)row_off(Q0 ? 2 * s : 2 * s + h, a);
  uint32_t wof[16];
#pragma unroll
  for (int r = 0; r < 16; ++r)
    wof[r] = Q0 ? (uint32_t)row_off(8 * ((r >> 1) >> 1) + 4 * h + 2 * ((r >> 1) & 1), a)
                : (uint32_t)row_off(8 * (r >> 2) + 4 * h + (r & 3), a);
  const uint64_t tw = (uint64_t)gridDim.x * 4;
  uint64_t s0, s1, step;
  if (IL) {
    // all waves advance together through one window of tw sets
    s0 = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    s1 = n_sets;
    step = tw;
  } else {
    const uint64_t gw = (uint64_t)xcd_block<uint32_t>() * 4 + (threadIdx.x >> 6);
    const uint64_t per = (n_sets + tw - 1) / tw;
    s0 = gw * per;
    s1 = s0 + per < n_sets ? s0 + per : n_sets;
    step = 1;
  }
  if (s0 >= s1) return;
  const uint64_t gl = Q0 ? (uint64_t)(c + 32 * h) : (uint64_t)(2 * c);
  f32x4_t x[16], y[16];
  uint64_t base = group_base(s0 * 64 + gl, a);
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const f32x4_t* p = reinterpret_cast<const f32x4_t*>(st + (base | rof[s]));
    x[s] = NT ? __builtin_nontemporal_load(p) : *p;
  }
  for (uint64_t set = s0; set < s1; set += step) {
    const bool more = set + step < s1;
    uint64_t nbase = 0;
    if (PF && more) {
      nbase = group_base((set + step) * 64 + gl, a);
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const f32x4_t* p = reinterpret_cast<const f32x4_t*>(st + (nbase | rof[s]));
        y[s] = NT ? __builtin_nontemporal_load(p) : *p;
      }
    }
    f32x16 er = {0.f}, ei = {0.f}, orr = {0.f}, oi = {0.f};
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      float ar, ai, br, bi;   // operands of column tile 1 and column tile 2
      if (Q0) {
        const float sr = h ? x[s].x : x[s].z, si = h ? x[s].y : x[s].w;
        const float rr = lane_xor(sr, 32), ri = lane_xor(si, 32);
        ar = h ? rr : x[s].x;
        ai = h ? ri : x[s].y;
        br = h ? x[s].z : rr;
        bi = h ? x[s].w : ri;
      } else {
        ar = x[s].x; ai = x[s].y; br = x[s].z; bi = x[s].w;
      }
      er = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[s], ar, er, 0, 0, 0);
      er = __builtin_amdgcn_mfma_f32_32x32x2f32(-ui[s], ai, er, 0, 0, 0);
      ei = __builtin_amdgcn_mfma_f32_32x32x2f32(ui[s], ar, ei, 0, 0, 0);
      ei = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[s], ai, ei, 0, 0, 0);
      orr = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[s], br, orr, 0, 0, 0);
      orr = __builtin_amdgcn_mfma_f32_32x32x2f32(-ui[s], bi, orr, 0, 0, 0);
      oi = __builtin_amdgcn_mfma_f32_32x32x2f32(ui[s], br, oi, 0, 0, 0);
      oi = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[s], bi, oi, 0, 0, 0);
    }
    if (Q0) {
      // registers r, r+1 (r even) hold rows i, i+1 of group c (tile 1) and c + 32 (tile 2)
      const uint64_t b1 = group_base(set * 64 + c, a), b2 = group_base(set * 64 + 32 + c, a);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int r = 2 * k;
        const f32x4_t o1 = {er[r], ei[r], er[r + 1], ei[r + 1]};
        const f32x4_t o2 = {orr[r], oi[r], orr[r + 1], oi[r + 1]};
        f32x4_t* p1 = reinterpret_cast<f32x4_t*>(st + (b1 | wof[r]));
        f32x4_t* p2 = reinterpret_cast<f32x4_t*>(st + (b2 | wof[r]));
        if (NT) {
          __builtin_nontemporal_store(o1, p1);
          __builtin_nontemporal_store(o2, p2);
        } else {
          *p1 = o1;
          *p2 = o2;
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const f32x4_t o = {er[r], ei[r], orr[r], oi[r]};
        f32x4_t* p = reinterpret_cast<f32x4_t*>(st + (base | wof[r]));
        if (NT) __builtin_nontemporal_store(o, p);
        else *p = o;
      }
    }
    if (!more) break;
    if (PF) {
      base = nbase;
#pragma unroll
      for (int s = 0; s < 16; ++s) x[s] = y[s];
    } else {
      base = group_base((set + step) * 64 + gl, a);
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const f32x4_t* p = reinterpret_cast<const f32x4_t*>(st + (base | rof[s]));
        x[s] = NT ? __builtin_nontemporal_load(p) : *p;
      }
    }
  }
}

// LDS-staged variant (n >= 12).  A workgroup of 4 waves owns 128 groups:
// the tile is the 4096 amplitudes spanned by the 5 block qubits and the 7
// lowest non-block qubits.  Those 12 qubits always include qubits 0..6, so
// the tile is read and written as runs of >= 128 contiguous amplitudes with
// lane-contiguous 16-byte nontemporal accesses, whatever the block qubits.
// On the way in each amplitude is transposed to X[row][group] in LDS (row
// stride padded by one element so the scattered writes spread over banks);
// wave w then feeds its 32 groups to the MFMA chains from LDS, writes the
// results back in place (its own columns only), and after a barrier the tile
// is stored with the same coalesced mapping.
constexpr int kTileGroups = 128;
constexpr int kRowStride = kTileGroups + 1;

struct Block5Tile {
  int n;
  uint8_t tq[12];     // tile qubits, ascending (tq[0] = 0)
  uint8_t tcode[12];  // tile bit k: < 8 -> row bit tcode[k]; >= 8 -> group bit tcode[k] - 8
  int nh;             // non-block qubits above the tile
  uint8_t hp[QSIM_MAX_QUBITS];
};

__device__ __forceinline__ void tile_rc(uint32_t j, const Block5Tile& a, uint32_t& r, uint32_t& c,
                                        uint64_t& off) {
  r = 0; c = 0; off = 0;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const uint32_t bit = (j >> k) & 1u;
    off |= (uint64_t)bit << a.tq[k];
    const uint32_t code = a.tcode[k];
    if (code < 8) r |= bit << code;
    else c |= bit << (code - 8);
  }
}

template <bool NT>
__global__ __launch_bounds__(256) void k_block5_lds(float2* __restrict__ st, Block5Tile a,
                                                    const float2* __restrict__ U) {
  __shared__ float2 X[32 * kRowStride];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int h = lane >> 5, c = lane & 31;
  // tile base: workgroup index deposited on the non-block qubits above the tile
  uint64_t base = 0;
  const uint64_t b = blockIdx.x;
#pragma unroll
  for (int k = 0; k < QSIM_MAX_QUBITS; ++k)
    if (k < a.nh) base |= ((b >> k) & 1ull) << a.hp[k];
  float4* sp = reinterpret_cast<float4*>(st);
  // bit 0 of the tile index is qubit 0: its code tells where amplitude 2s+1 goes
  const uint32_t c0 = a.tcode[0];
  // all eight 16-byte loads of a thread are issued before the first LDS write
  float4 v[8];
  uint32_t lds0[8], lds1[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    uint32_t r, cc;
    uint64_t off;
    tile_rc(2u * (threadIdx.x + 256 * k), a, r, cc, off);
    v[k] = ld4<NT>(sp + ((base | off) >> 1));
    const uint32_t r1 = c0 < 8 ? (r | (1u << c0)) : r;
    const uint32_t cc1 = c0 < 8 ? cc : (cc | (1u << (c0 - 8)));
    lds0[k] = r * kRowStride + cc;
    lds1[k] = r1 * kRowStride + cc1;
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    X[lds0[k]] = lo2(v[k]);
    X[lds1[k]] = hi2(v[k]);
  }
  float ur[16], ui[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const float2 u = U[c * 32 + 2 * s + h];
    ur[s] = u.x;
    ui[s] = u.y;
  }
  __syncthreads();
  const int col = wv * 32 + c;
  float xr[16], xi[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const float2 x = X[(2 * s + h) * kRowStride + col];
    xr[s] = x.x;
    xi[s] = x.y;
  }
  f32x16 accr = {0.f}, acci = {0.f};
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    accr = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[s], xr[s], accr, 0, 0, 0);
    accr = __builtin_amdgcn_mfma_f32_32x32x2f32(-ui[s], xi[s], accr, 0, 0, 0);
    acci = __builtin_amdgcn_mfma_f32_32x32x2f32(ui[s], xr[s], acci, 0, 0, 0);
    acci = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[s], xi[s], acci, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = 8 * (r >> 2) + 4 * h + (r & 3);
    X[i * kRowStride + col] = make_float2(accr[r], acci[r]);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    uint32_t r, cc;
    uint64_t off;
    tile_rc(2u * (threadIdx.x + 256 * k), a, r, cc, off);
    const uint32_t r1 = c0 < 8 ? (r | (1u << c0)) : r;
    const uint32_t cc1 = c0 < 8 ? cc : (cc | (1u << (c0 - 8)));
    st4<NT>(sp + ((base | off) >> 1), cat4(X[r * kRowStride + cc], X[r1 * kRowStride + cc1]));
  }
}

// ------------------------------------------------------------ k = 4 and k = 6
// LDS-staged blocks of K = 4 or 6 qubits, the k_block5_lds scheme with the
// tile widened or narrowed: a workgroup of 4 waves owns 128 groups, the tile
// is the 2^(K+7) amplitudes spanned by the K block qubits and the 7 lowest
// non-block qubits (so it always holds qubits 0..6 and moves as runs of at
// least 128 contiguous amplitudes), transposed on the way in to
// X[row][group] in LDS (row stride padded by one element).  Wave w takes
// groups 32w .. 32w + 31.
//  * K = 6 (64 x 64 complex): v_mfma_f32_32x32x2_f32 over two 32-row output
//    tiles and 32 K slices; the A fragments (U[32 t + c][2 s + h], 128 floats
//    per lane) stay in VGPRs.  At 512 flop per amplitude (AI = 32 flop/B,
//    above the 19.7 ridge) it is bound by the MFMA rate.
//  * K = 4 (16 x 16 complex): v_mfma_f32_16x16x4_f32, whose operands are
//    A[i][k] from lane i % 16 + 16 k (k = 0..3), B[k][j] from lane
//    j % 16 + 16 k, and D[4 (l / 16) + r][l % 16] in register r of lane l;
//    two 16-group column tiles per wave, 4 K slices.  AI = 8 flop/B: bound
//    by HBM.
template <int K>
struct BlockKTile {
  int n;
  uint8_t tq[K + 7];     // tile qubits, ascending (tq[0] = 0)
  uint8_t tcode[K + 7];  // tile bit k: < 8 -> row bit tcode[k]; >= 8 -> group bit tcode[k] - 8
  int nh;                // non-block qubits above the tile
  uint8_t hp[QSIM_MAX_QUBITS];
};

template <int K>
__device__ __forceinline__ void tilek_rc(uint32_t j, const BlockKTile<K>& a, uint32_t& r, uint32_t& c,
                                         uint64_t& off) {
  r = 0; c = 0; off = 0;
#pragma unroll
  for (int k = 0; k < K + 7; ++k) {
    const uint32_t bit = (j >> k) & 1u;
    off |= (uint64_t)bit << a.tq[k];
    const uint32_t code = a.tcode[k];
    if (code < 8) r |= bit << code;
    else c |= bit << (code - 8);
  }
}

typedef float f32x4m __attribute__((ext_vector_type(4)));

template <int K, bool NT>
__global__ __launch_bounds__(256) void k_blockk_lds(float2* __restrict__ st, BlockKTile<K> a,
                                                    const float2* __restrict__ U) {
  constexpr int R = 1 << K;                     // rows
  constexpr int SLOTS = (R * kTileGroups) / 2 / 256;   // float4 slots per thread
  __shared__ float2 X[R * kRowStride];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint64_t base = 0;
  const uint64_t b = blockIdx.x;
#pragma unroll
  for (int k = 0; k < QSIM_MAX_QUBITS; ++k)
    if (k < a.nh) base |= ((b >> k) & 1ull) << a.hp[k];
  float4* sp = reinterpret_cast<float4*>(st);
  const uint32_t c0 = a.tcode[0];
  {
    float4 v[SLOTS];
    uint32_t l0[SLOTS], l1[SLOTS];
#pragma unroll
    for (int k = 0; k < SLOTS; ++k) {
      uint32_t r, cc;
      uint64_t off;
      tilek_rc<K>(2u * (threadIdx.x + 256 * k), a, r, cc, off);
      v[k] = ld4<NT>(sp + ((base | off) >> 1));
      const uint32_t r1 = c0 < 8 ? (r | (1u << c0)) : r;
      const uint32_t cc1 = c0 < 8 ? cc : (cc | (1u << (c0 - 8)));
      l0[k] = r * kRowStride + cc;
      l1[k] = r1 * kRowStride + cc1;
    }
#pragma unroll
    for (int k = 0; k < SLOTS; ++k) {
      X[l0[k]] = lo2(v[k]);
      X[l1[k]] = hi2(v[k]);
    }
  }
  if constexpr (K == 6) {
    const int h = lane >> 5, c = lane & 31;
    float ur[2][32], ui[2][32];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int s = 0; s < 32; ++s) {
        const float2 u = U[(32 * t + c) * 64 + 2 * s + h];
        ur[t][s] = u.x;
        ui[t][s] = u.y;
      }
    __syncthreads();
    const int col = wv * 32 + c;
    f32x16 ar0 = {0.f}, ai0 = {0.f}, ar1 = {0.f}, ai1 = {0.f};
#pragma unroll
    for (int s = 0; s < 32; ++s) {
      const float2 x = X[(2 * s + h) * kRowStride + col];
      ar0 = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[0][s], x.x, ar0, 0, 0, 0);
      ar0 = __builtin_amdgcn_mfma_f32_32x32x2f32(-ui[0][s], x.y, ar0, 0, 0, 0);
      ai0 = __builtin_amdgcn_mfma_f32_32x32x2f32(ui[0][s], x.x, ai0, 0, 0, 0);
      ai0 = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[0][s], x.y, ai0, 0, 0, 0);
      ar1 = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[1][s], x.x, ar1, 0, 0, 0);
      ar1 = __builtin_amdgcn_mfma_f32_32x32x2f32(-ui[1][s], x.y, ar1, 0, 0, 0);
      ai1 = __builtin_amdgcn_mfma_f32_32x32x2f32(ui[1][s], x.x, ai1, 0, 0, 0);
      ai1 = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[1][s], x.y, ai1, 0, 0, 0);
    }
    // a wave reads and then overwrites only its own 32 columns, and LDS
    // operations of one wave complete in order, so no barrier is needed here
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = 8 * (r >> 2) + 4 * h + (r & 3);
      X[i * kRowStride + col] = make_float2(ar0[r], ai0[r]);
      X[(32 + i) * kRowStride + col] = make_float2(ar1[r], ai1[r]);
    }
  } else {
    static_assert(K == 4, "k_blockk_lds handles K = 4 and K = 6");
    const int kq = lane >> 4, m = lane & 15;     // K index within a slice, row / column
    float ur[4], ui[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const float2 u = U[m * 16 + 4 * s + kq];
      ur[s] = u.x;
      ui[s] = u.y;
    }
    __syncthreads();
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      const int col = wv * 32 + 16 * ct + m;
      f32x4m accr = {0.f, 0.f, 0.f, 0.f}, acci = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const float2 x = X[(4 * s + kq) * kRowStride + col];
        accr = __builtin_amdgcn_mfma_f32_16x16x4f32(ur[s], x.x, accr, 0, 0, 0);
        accr = __builtin_amdgcn_mfma_f32_16x16x4f32(-ui[s], x.y, accr, 0, 0, 0);
        acci = __builtin_amdgcn_mfma_f32_16x16x4f32(ui[s], x.x, acci, 0, 0, 0);
        acci = __builtin_amdgcn_mfma_f32_16x16x4f32(ur[s], x.y, acci, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) X[(4 * kq + r) * kRowStride + col] = make_float2(accr[r], acci[r]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SLOTS; ++k) {
    uint32_t r, cc;
    uint64_t off;
    tilek_rc<K>(2u * (threadIdx.x + 256 * k), a, r, cc, off);
    const uint32_t r1 = c0 < 8 ? (r | (1u << c0)) : r;
    const uint32_t cc1 = c0 < 8 ? cc : (cc | (1u << (c0 - 8)));
    st4<NT>(sp + ((base | off) >> 1), cat4(X[r * kRowStride + cc], X[r1 * kRowStride + cc1]));
  }
}

// Tile of the K block qubits qs[0..K) and the 7 lowest other qubits.
template <int K>
static BlockKTile<K> make_tilek(int n, const int* qs) {
  BlockKTile<K> tl;
  memset(&tl, 0, sizeof tl);
  tl.n = n;
  uint64_t bm = 0;
  for (int j = 0; j < K; ++j) bm |= 1ull << qs[j];
  int low[7], nl = 0;
  for (int q = 0; q < n && nl < 7; ++q)
    if (!(bm >> q & 1)) low[nl++] = q;
  uint64_t tmask = bm;
  for (int k = 0; k < 7; ++k) tmask |= 1ull << low[k];
  int nt = 0;
  for (int q = 0; q < n; ++q) {
    if (!(tmask >> q & 1)) continue;
    tl.tq[nt] = (uint8_t)q;
    int code = -1;
    for (int j = 0; j < K; ++j)
      if (qs[j] == q) code = j;
    if (code < 0)
      for (int k = 0; k < 7; ++k)
        if (low[k] == q) code = 8 + k;
    tl.tcode[nt++] = (uint8_t)code;
  }
  tl.nh = 0;
  for (int q = 0; q < n; ++q)
    if (!(tmask >> q & 1)) tl.hp[tl.nh++] = (uint8_t)q;
  return tl;
}

template <int K>
static int launch_blockk(hipStream_t s, void* state, int n, const void* block_dev, const int* qs) {
  const BlockKTile<K> tl = make_tilek<K>(n, qs);
  const dim3 grid((unsigned)(1ull << tl.nh));
  if (nt_for_qubits(n))
    hipLaunchKernelGGL((k_blockk_lds<K, true>), grid, dim3(256), 0, s, (float2*)state, tl,
                       (const float2*)block_dev);
  else
    hipLaunchKernelGGL((k_blockk_lds<K, false>), grid, dim3(256), 0, s, (float2*)state, tl,
                       (const float2*)block_dev);
  return launch_status("apply_block_mfma");
}

}  // namespace qsim

using namespace qsim;

extern "C" {

int qsim_apply_block_mfma(void* stream, void* state, int n_qubits, const void* block_dev,
                          const int* qubits, int k) {
  if (int rc = check_n(n_qubits)) return rc;
  if (!state || !block_dev || !qubits) return fail(QSIM_EINVAL, "NULL pointer argument");
  if (k < 4 || k > 6) return fail(QSIM_ERANGE, "MFMA blocks support k = 4, 5, 6 (got %d)", k);
  if (n_qubits < k + 7 && k != 5)
    return fail(QSIM_ERANGE, "MFMA blocks of %d qubits need n_qubits >= %d (got %d)", k, k + 7, n_qubits);
  if (n_qubits < 10) return fail(QSIM_ERANGE, "MFMA blocks need n_qubits >= 10 (got %d)", n_qubits);
  if (k != 5) {
    uint64_t seen = 0;
    for (int j = 0; j < k; ++j) {
      if (int rc = check_q(qubits[j], n_qubits, "block qubit")) return rc;
      if (seen >> qubits[j] & 1) return fail(QSIM_EINVAL, "block qubits must differ");
      seen |= 1ull << qubits[j];
    }
    return k == 4 ? launch_blockk<4>(as_stream(stream), state, n_qubits, block_dev, qubits)
                  : launch_blockk<6>(as_stream(stream), state, n_qubits, block_dev, qubits);
  }
  BlockArgs a;
  memset(&a, 0, sizeof a);
  a.n = n_qubits;
  uint64_t seen = 0;
  for (int j = 0; j < 5; ++j) {
    if (int rc = check_q(qubits[j], n_qubits, "block qubit")) return rc;
    if (seen >> qubits[j] & 1) return fail(QSIM_EINVAL, "block qubits must differ");
    seen |= 1ull << qubits[j];
    a.q[j] = qubits[j];
  }
  a.nc = 0;
  for (int q = 0; q < n_qubits; ++q)
    if (!(seen >> q & 1)) a.cp[a.nc++] = (uint8_t)q;
  // Blocks with three or more qubits among 0..4 make the direct loads stride
  // across many cache lines per instruction; they go through the LDS tile.
  int low = 0;
  for (int j = 0; j < 5; ++j) low += a.q[j] < 5;
  int qmin = a.q[0];
  for (int j = 1; j < 5; ++j) qmin = std::min(qmin, a.q[j]);
  // Wide kernel (16-byte lanes); variants chosen from 30-qubit measurements:
  //   no block qubit in 0..3: prefetch, two workgroups per CU, nontemporal
  //   above the MALL size (cached below it, so the state stays resident);
  //   lowest block qubit in 1..3: prefetch, cached accesses, one per CU;
  //   qubit 0 in the block: cached accesses; above the MALL size prefetch
  //   and one per CU (3.45 against 3.61 ms on five patterns at 30 qubits),
  //   below it no prefetch and two per CU.
  if (n_qubits >= 11 && n_qubits <= 32 && low <= 2 && (qmin != 0 || a.q[0] == 0)) {
    const uint64_t n_sets = (1ull << (n_qubits - 5)) / 64;
    hipStream_t s = as_stream(stream);
    auto grid = [&](int per_cu) { return dim3((unsigned)std::min<uint64_t>(ceil_div(n_sets, 4), 256u * per_cu)); };
    // window order for streamed states whose lowest block qubit is below 12
    const bool il = nt_for_qubits(n_qubits) && qmin < 12;
    const float2* U = (const float2*)block_dev;
    float2* sv = (float2*)state;
    if (qmin == 0) {
      if (il) hipLaunchKernelGGL((k_block5w<1, 0, 1, 1>), grid(1), dim3(256), 0, s, sv, n_sets, a, U);
      else hipLaunchKernelGGL((k_block5w<0, 0, 1, 0>), grid(2), dim3(256), 0, s, sv, n_sets, a, U);
    } else if (qmin < 4) {
      if (il) hipLaunchKernelGGL((k_block5w<1, 0, 0, 1>), grid(1), dim3(256), 0, s, sv, n_sets, a, U);
      else hipLaunchKernelGGL((k_block5w<1, 0, 0, 0>), grid(1), dim3(256), 0, s, sv, n_sets, a, U);
    } else if (nt_for_qubits(n_qubits)) {
      if (il) hipLaunchKernelGGL((k_block5w<1, 1, 0, 1>), grid(2), dim3(256), 0, s, sv, n_sets, a, U);
      else hipLaunchKernelGGL((k_block5w<1, 1, 0, 0>), grid(2), dim3(256), 0, s, sv, n_sets, a, U);
    } else {
      hipLaunchKernelGGL((k_block5w<1, 0, 0, 0>), grid(2), dim3(256), 0, s, sv, n_sets, a, U);
    }
    return launch_status("apply_block_mfma");
  }
  if (n_qubits >= 12 && (low >= 3 || n_qubits > 32)) {
    Block5Tile tl;
    memset(&tl, 0, sizeof tl);
    tl.n = n_qubits;
    uint64_t tmask = seen;
    for (int k = 0; k < 7; ++k) tmask |= 1ull << a.cp[k];   // 7 lowest non-block qubits
    int nt = 0;
    for (int q = 0; q < n_qubits; ++q) {
      if (!(tmask >> q & 1)) continue;
      tl.tq[nt] = (uint8_t)q;
      int code = -1;
      for (int j = 0; j < 5; ++j)
        if (a.q[j] == q) code = j;
      if (code < 0)
        for (int k = 0; k < 7; ++k)
          if (a.cp[k] == q) code = 8 + k;
      tl.tcode[nt++] = (uint8_t)code;
    }
    tl.nh = 0;
    for (int k = 7; k < a.nc; ++k) tl.hp[tl.nh++] = a.cp[k];
    if (nt_for_qubits(n_qubits))
      hipLaunchKernelGGL(k_block5_lds<true>, dim3(1ull << tl.nh), dim3(256), 0, as_stream(stream),
                         (float2*)state, tl, (const float2*)block_dev);
    else
      hipLaunchKernelGGL(k_block5_lds<false>, dim3(1ull << tl.nh), dim3(256), 0, as_stream(stream),
                         (float2*)state, tl, (const float2*)block_dev);
    return launch_status("apply_block_mfma");
  }
  const uint64_t n_sets = (1ull << (n_qubits - 5)) / 32;
  constexpr int W = 4;
  // persistent grid: two workgroups of 4 waves per CU (the kernel holds ~180
  // registers per lane, two waves per SIMD), never more than there are sets
  const uint64_t blocks = std::min<uint64_t>(ceil_div(n_sets, W), 256 * 2);
  hipLaunchKernelGGL((k_block5<W>), dim3(blocks), dim3(64 * W), 0, as_stream(stream),
                     (float2*)state, n_sets, a, (const float2*)block_dev);
  return launch_status("apply_block_mfma");
}

int qsim_apply_block_batch(void* stream, const qsim_block_rec* recs_host, const void* recs_dev,
                           int64_t n_recs) {
  clear_error();
  if (n_recs < 0 || n_recs > (1 << 24)) return fail(QSIM_ERANGE, "n_recs = %lld out of range", (long long)n_recs);
  if (n_recs == 0) return QSIM_OK;
  if (!recs_host || !recs_dev) return fail(QSIM_EINVAL, "NULL pointer argument");
  uint64_t total = 0;
  for (int64_t i = 0; i < n_recs; ++i) {
    const qsim_block_rec& r = recs_host[i];
    if (r.n_qubits < 10 || r.n_qubits > 32)
      return fail(QSIM_ERANGE, "record %lld: batched blocks need 10 <= n_qubits <= 32 (got %d)",
                  (long long)i, r.n_qubits);
    if (!r.state || !r.mat) return fail(QSIM_EINVAL, "record %lld: NULL state or matrix", (long long)i);
    uint64_t seen = 0;
    for (int j = 0; j < 5; ++j) {
      if (r.q[j] < 0 || r.q[j] >= r.n_qubits)
        return fail(QSIM_EINVAL, "record %lld: block qubit %d out of range", (long long)i, r.q[j]);
      if (seen >> r.q[j] & 1) return fail(QSIM_EINVAL, "record %lld: block qubits must differ", (long long)i);
      seen |= 1ull << r.q[j];
    }
    if (r.set_begin != total)
      return fail(QSIM_EINVAL, "record %lld: set_begin %llu, expected %llu (prefix of 2^(n-10))",
                  (long long)i, (unsigned long long)r.set_begin, (unsigned long long)total);
    total += 1ull << (r.n_qubits - 10);
  }
  const int spw = total < 2048 ? 1 : (total < 16384 ? 2 : 4);
  const uint64_t waves = ceil_div(total, spw);
  if (ceil_div(waves, 4) > 0x7fffffffull) return fail(QSIM_ERANGE, "batch too large (%llu sets)",
                                                     (unsigned long long)total);
  hipLaunchKernelGGL(k_block5_batch, dim3((unsigned)ceil_div(waves, 4)), dim3(256), 0, as_stream(stream),
                     (const qsim_block_rec*)recs_dev, (int)n_recs, total, spw);
  return launch_status("apply_block_batch");
}

// Block-plan launcher: one MFMA block launch per entry, without a round
// trip through the caller's language per block.
int qsim_apply_blocks(void* stream, void* state, int n_qubits, const void* blocks_dev,
                      const int32_t* block_qubits, int n_blocks) {
  if (n_blocks < 0) return fail(QSIM_EINVAL, "n_blocks < 0");
  if (n_blocks > 0 && (!blocks_dev || !block_qubits)) return fail(QSIM_EINVAL, "NULL pointer argument");
  const float2* b = (const float2*)blocks_dev;
  for (int i = 0; i < n_blocks; ++i) {
    int qs[5];
    for (int j = 0; j < 5; ++j) qs[j] = block_qubits[i * 8 + j];
    if (int rc = qsim_apply_block_mfma(stream, state, n_qubits, b + (size_t)i * 1024, qs, 5))
      return rc;
  }
  return QSIM_OK;
}

}  // extern "C"
