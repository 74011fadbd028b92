// Host-side runtime of libqsim_gfx950.so: the dense-block planner.  Pure
// C++ (no HIP headers, no kernels), compiled by the host compiler, so it
// also builds under the host sanitizers (tests/test_planner_sanitizers.py).
//
// Planner.  Ops arrive in circuit order, each with up to 8 qubits and a
// dense FP64 matrix on them (row bit j <-> qubit j of the op).  A block is
// a set S of at most k qubits together with the ops it absorbs: scanning the
// remaining ops in order, an op joins when its qubits lie in S and none of
// them was touched by an earlier op left out of the block (so every op still
// follows its predecessors).  A seed is any remaining op that no earlier
// remaining op touches; up to kSeeds of them are tried and the block that
// absorbs the most ops per unit of kernel time (block_score) is kept (ties go
// to the earliest).
//  * effort 0: S grows greedily from the seed's qubits, an op joining while
//    the union stays <= k; a depth-28 random SU(4) circuit on 28 qubits
//    takes 100 blocks (120 with the first op as the only seed).
//  * effort 1: in addition, every S made of the seed's qubits and up to
//    k - |seed| of the qubits that the next kSearchWindow ops connect to the
//    seed is scored.  The depth-30 SU(4) circuit on 30 qubits then took 108
//    blocks instead of 119 (28 qubits, depth 28: 92 instead of 100) when
//    scored by count, for about 5 ms of host time.  A 40-op window gave 110 and 93 blocks in twice the
//    time, a 20-op window 112 and 96.
// The block matrix is the FP64 product of the members in circuit order, each
// applied to the rows of the running 2^k x 2^k matrix.  Blocks are padded with
// the lowest unused qubits from kPadFrom up (then below it) to exactly
// min(k, n) qubits.  An op wider than k
// becomes a passthrough entry for the caller to run on its own kernel.
#include "qsim_gfx950.h"
#include "host_error.hpp"
#include <stdint.h>
#include <algorithm>
#include <string.h>
#include <vector>

namespace qsim {

constexpr int kSeeds = 32;
constexpr int kSearchWindow = 30;     // ops after the seed whose qubits may join S (effort 1)
constexpr int kSearchQubits = 16;     // at most this many candidate qubits per seed
// First padding qubit.  A padded qubit only selects the block kernel: a low one
// can put qubit 0 or a third qubit of 0..4 into the block, whose kernels are
// slower, and a lowest block qubit of 12 streams well.  The 30-qubit SU(4)
// step ran 363-364 ms against 369 ms with padding from qubit 0 (108 blocks
// either way; 368 ms from 16, 20 or 24).
constexpr int kPadFrom = 12;
// Blocks are compared by ops absorbed per unit of kernel time, in thousandths
// of a block with no qubit in 0..3: qubit 0 in the block 1.10, three or more
// block qubits in 0..4 (the LDS-staged kernel) 1.31, the ratios of their mean
// launch times in the 30-qubit SU(4) trace.  Scored by count alone, that
// circuit took 108 blocks (20 with qubit 0, 3 LDS-staged) and 367 ms per
// fused step; scored per cost, 107 blocks (15 and 2) and 361 ms.  The ratios
// hold for states above the 256 MiB MALL (n > 25); smaller states and the
// other widths (one kernel for every qubit set) are scored by count.
constexpr long long kCostQ0 = 1100, kCostLow3 = 1310;

static inline long long block_score(int n_ops, uint64_t S, bool weigh) {
  long long c = 1000;
  if (weigh) c = (S & 1ull) ? kCostQ0 : (__builtin_popcountll(S & 31ull) >= 3 ? kCostLow3 : 1000);
  return (long long)n_ops * 1000000 / c;
}

// Ops of rem (in order) that the qubit set S absorbs: an op joins when its
// qubits lie in S and no earlier op left out touched any of them.  With
// `out` null only the count is returned, and the scan stops once every
// qubit of S is blocked.
static int absorb_set(const std::vector<int>& rem, const std::vector<uint64_t>& mask, uint64_t S,
                      std::vector<int>* out) {
  uint64_t blocked = 0;
  int cnt = 0;
  for (int i : rem) {
    const uint64_t m = mask[i];
    if ((m & blocked) == 0 && (m & ~S) == 0) {
      ++cnt;
      if (out) out->push_back(i);
    } else {
      blocked |= m;
      if (!out && (blocked & S) == S) break;
    }
  }
  return cnt;
}

// Block matrix kept as separate real and imaginary D x D arrays so that the
// row updates below run as contiguous loops over the columns.
//
// rows of (re, im) (row bit b <-> block bit b) <- u (on block bits pos[0..kq))
// times those rows; u is row-major complex (re, im interleaved).
// One- and two-qubit ops (the common case) go through local row buffers, so
// the column loops vectorise without alias checks (with an AVX2 clone where
// the host has it); the operation order is that of the general form below
// (each product rounded, then summed over t from 0.0), so both give the same
// bits.
template <int KQ>
__attribute__((target_clones("avx2", "default")))
static void apply_rows_small(double* re, double* im, int D, const double* u, const int* pos) {
  constexpr int d = 1 << KQ;
  constexpr int kMaxD = 1 << QSIM_MAX_BLOCK_QUBITS;
  double ur[d][d], ui[d][d];
  for (int s = 0; s < d; ++s)
    for (int t = 0; t < d; ++t) {
      ur[s][t] = u[2 * (s * d + t)];
      ui[s][t] = u[2 * (s * d + t) + 1];
    }
  uint32_t pm = 0;
  for (int j = 0; j < KQ; ++j) pm |= 1u << pos[j];
  double xr[d][kMaxD], xi[d][kMaxD], yr[d][kMaxD], yi[d][kMaxD];
  for (int base = 0; base < D; ++base) {
    if (base & pm) continue;
    double* pr[d];
    double* pi[d];
    for (int s = 0; s < d; ++s) {
      int r = base;
      for (int j = 0; j < KQ; ++j)
        if (s >> j & 1) r |= 1 << pos[j];
      pr[s] = re + (size_t)r * D;
      pi[s] = im + (size_t)r * D;
      memcpy(xr[s], pr[s], sizeof(double) * D);
      memcpy(xi[s], pi[s], sizeof(double) * D);
    }
    for (int s = 0; s < d; ++s)
      for (int c = 0; c < D; ++c) {
        double ar = 0.0, ai = 0.0;
        for (int t = 0; t < d; ++t) {
          ar += ur[s][t] * xr[t][c] - ui[s][t] * xi[t][c];
          ai += ur[s][t] * xi[t][c] + ui[s][t] * xr[t][c];
        }
        yr[s][c] = ar;
        yi[s][c] = ai;
      }
    for (int s = 0; s < d; ++s) {
      memcpy(pr[s], yr[s], sizeof(double) * D);
      memcpy(pi[s], yi[s], sizeof(double) * D);
    }
  }
}

static void apply_rows(double* re, double* im, int D, const double* u, const int* pos, int kq,
                       double* tre, double* tim) {
  if (kq == 1) return apply_rows_small<1>(re, im, D, u, pos);
  if (kq == 2) return apply_rows_small<2>(re, im, D, u, pos);
  const int d = 1 << kq;
  uint32_t pm = 0;
  for (int j = 0; j < kq; ++j) pm |= 1u << pos[j];
  int idx[256];
  for (int base = 0; base < D; ++base) {
    if (base & pm) continue;
    for (int s = 0; s < d; ++s) {
      int r = base;
      for (int j = 0; j < kq; ++j)
        if (s >> j & 1) r |= 1 << pos[j];
      idx[s] = r;
    }
    for (int t = 0; t < d; ++t) {
      memcpy(tre + (size_t)t * D, re + (size_t)idx[t] * D, sizeof(double) * D);
      memcpy(tim + (size_t)t * D, im + (size_t)idx[t] * D, sizeof(double) * D);
    }
    for (int s = 0; s < d; ++s) {
      double* ore = re + (size_t)idx[s] * D;
      double* oim = im + (size_t)idx[s] * D;
      const double* us = u + 2 * (size_t)s * d;
      for (int c = 0; c < D; ++c) ore[c] = oim[c] = 0.0;
      for (int t = 0; t < d; ++t) {
        const double ur = us[2 * t], ui = us[2 * t + 1];
        const double* xr = tre + (size_t)t * D;
        const double* xi = tim + (size_t)t * D;
        for (int c = 0; c < D; ++c) {
          ore[c] += ur * xr[c] - ui * xi[c];
          oim[c] += ur * xi[c] + ui * xr[c];
        }
      }
    }
  }
}

}  // namespace qsim

using namespace qsim;

extern "C" {

int qsim_plan_blocks_search(int n_qubits, int k, int effort, int n_ops, const int32_t* op_nq,
                            const int32_t* op_qubits, const double* op_mats, int32_t* out_ref,
                            int32_t* block_qubits, double* block_mats, int* n_out) {
  if (effort < 0 || effort > 1) return fail(QSIM_EINVAL, "effort %d outside 0..1", effort);
  if (n_qubits < 1 || n_qubits > QSIM_MAX_QUBITS) return fail(QSIM_ERANGE, "n_qubits=%d", n_qubits);
  if (k < 1 || k > QSIM_MAX_BLOCK_QUBITS) return fail(QSIM_ERANGE, "block size k=%d outside 1..%d", k,
                                                      QSIM_MAX_BLOCK_QUBITS);
  if (n_ops < 0) return fail(QSIM_EINVAL, "n_ops < 0");
  if (!n_out || (n_ops > 0 && (!op_nq || !op_qubits || !op_mats || !out_ref || !block_qubits ||
                               !block_mats)))
    return fail(QSIM_EINVAL, "NULL pointer argument");
  const int kk = k < n_qubits ? k : n_qubits;
  const int D = 1 << kk;
  const bool weigh = kk == 5 && n_qubits > 25;
  std::vector<uint64_t> mask(n_ops);
  std::vector<size_t> moff(n_ops);
  size_t off = 0;
  for (int i = 0; i < n_ops; ++i) {
    const int nq = op_nq[i];
    if (nq < 1 || nq > 8) return fail(QSIM_EINVAL, "op %d has %d qubits (1..8)", i, nq);
    uint64_t m = 0;
    for (int j = 0; j < nq; ++j) {
      const int q = op_qubits[i * 8 + j];
      if (q < 0 || q >= n_qubits) return fail(QSIM_EINVAL, "op %d: qubit %d out of range", i, q);
      if (m >> q & 1) return fail(QSIM_EINVAL, "op %d: repeated qubit %d", i, q);
      m |= 1ull << q;
    }
    mask[i] = m;
    moff[i] = off;
    off += (size_t)1 << (2 * nq);
  }
  std::vector<int> rem(n_ops), keep, best;
  for (int i = 0; i < n_ops; ++i) rem[i] = i;
  std::vector<int> chosen;
  int n_entries = 0, n_blocks = 0;
  std::vector<double> mre((size_t)D * D), mim((size_t)D * D), tre((size_t)256 * D),
      tim((size_t)256 * D);
  uint64_t best_bq = 0;
  while (!rem.empty()) {
    // try up to kSeeds eligible seeds; keep the block that absorbs the most ops
    best.clear();
    best_bq = 0;
    uint64_t before = 0;   // qubits of the remaining ops ahead of the candidate seed
    int tried = 0;
    for (size_t sc = 0; sc < rem.size() && tried < kSeeds; before |= mask[rem[sc]], ++sc) {
      const uint64_t ms = mask[rem[sc]];
      if (ms & before) continue;                          // not movable to the front
      if (__builtin_popcountll(ms) > kk) continue;        // wider than a block
      ++tried;
      uint64_t bq = ms, blocked = 0;
      chosen.clear();
      for (size_t t = 0; t < rem.size(); ++t) {
        const int i = rem[t];
        const uint64_t m = mask[i];
        if (t == sc) {
          chosen.push_back(i);
          continue;
        }
        if (m & blocked) {
          blocked |= m;
          continue;
        }
        const uint64_t un = bq | m;
        if (__builtin_popcountll(un) <= kk) {
          bq = un;
          chosen.push_back(i);
        } else {
          blocked |= m;
        }
        // a full block whose qubits are all blocked can take nothing more
        if (t > sc && __builtin_popcountll(bq) == kk && (blocked & bq) == bq) break;
      }
      if (best.empty() || block_score((int)chosen.size(), bq, weigh) > block_score((int)best.size(), best_bq, weigh)) {
        best = chosen;
        best_bq = bq;
      }
      if (effort >= 1) {
        // qubits the next ops connect to the seed, then every S of the
        // seed's qubits and up to kk - |seed| of them
        uint64_t conn = ms;
        const size_t wend = std::min(rem.size(), sc + (size_t)kSearchWindow);
        for (size_t t = sc; t < wend; ++t)
          if (mask[rem[t]] & conn) conn |= mask[rem[t]];
        int extra[kSearchQubits], ne = 0;
        for (int q = 0; q < n_qubits && ne < kSearchQubits; ++q)
          if ((conn >> q & 1) && !(ms >> q & 1)) extra[ne++] = q;
        const int room = kk - __builtin_popcountll(ms);
        // subsets of extra with 1..room members (Gosper's next-combination)
        long long best_sc = block_score((int)best.size(), best_bq, weigh);
        uint64_t best_s = 0;
        for (int r = 1; r <= room && r <= ne; ++r) {
          for (uint32_t sub = (1u << r) - 1; sub < (1u << ne);) {
            uint64_t S = ms;
            for (int e = 0; e < ne; ++e)
              if (sub >> e & 1) S |= 1ull << extra[e];
            const long long sc = block_score(absorb_set(rem, mask, S, nullptr), S, weigh);
            if (sc > best_sc) {
              best_sc = sc;
              best_s = S;
            }
            const uint32_t c = sub & (0u - sub), nx = sub + c;
            sub = (((nx ^ sub) >> 2) / c) | nx;
          }
        }
        if (best_s) {
          best.clear();
          absorb_set(rem, mask, best_s, &best);
          uint64_t u = 0;
          for (int i : best) u |= mask[i];
          best_bq = u;
        }
      }
    }
    if (best.empty()) {  // every candidate is wider than k: the first op passes through
      out_ref[n_entries++] = -1 - rem[0];
      rem.erase(rem.begin());
      continue;
    }
    chosen = best;
    uint64_t bq = best_bq;
    keep.clear();
    {
      size_t c = 0;
      for (int i : rem) {
        if (c < chosen.size() && chosen[c] == i) ++c;
        else keep.push_back(i);
      }
    }
    rem.swap(keep);
    int bqs[8], nb = 0;
    for (int q = 0; q < n_qubits && nb < 8; ++q)
      if (bq >> q & 1) bqs[nb++] = q;
    // padding qubits: the lowest unused from kPadFrom up, then below it
    for (int pass = 0; pass < 2; ++pass)
      for (int q = pass ? 0 : std::min(kPadFrom, n_qubits); q < n_qubits && nb < kk; ++q)
        if (!(bq >> q & 1)) { bqs[nb++] = q; bq |= 1ull << q; }
    // ascending order
    for (int a = 0; a < nb; ++a)
      for (int b = a + 1; b < nb; ++b)
        if (bqs[b] < bqs[a]) { int t = bqs[a]; bqs[a] = bqs[b]; bqs[b] = t; }
    std::fill(mre.begin(), mre.end(), 0.0);
    std::fill(mim.begin(), mim.end(), 0.0);
    for (int r = 0; r < D; ++r) mre[(size_t)r * D + r] = 1.0;
    for (int i : chosen) {
      const int nq = op_nq[i];
      int pos[8];
      for (int j = 0; j < nq; ++j) {
        const int q = op_qubits[i * 8 + j];
        for (int b = 0; b < nb; ++b)
          if (bqs[b] == q) pos[j] = b;
      }
      apply_rows(mre.data(), mim.data(), D, op_mats + 2 * moff[i], pos, nq, tre.data(), tim.data());
    }
    for (int b = 0; b < 8; ++b) block_qubits[n_blocks * 8 + b] = b < nb ? bqs[b] : -1;
    double* dst = block_mats + (size_t)n_blocks * D * D * 2;
    for (size_t e = 0; e < (size_t)D * D; ++e) {
      dst[2 * e] = mre[e];
      dst[2 * e + 1] = mim[e];
    }
    out_ref[n_entries++] = n_blocks++;
  }
  *n_out = n_entries;
  return QSIM_OK;
}

int qsim_plan_blocks(int n_qubits, int k, int n_ops, const int32_t* op_nq, const int32_t* op_qubits,
                     const double* op_mats, int32_t* out_ref, int32_t* block_qubits,
                     double* block_mats, int* n_out) {
  return qsim_plan_blocks_search(n_qubits, k, 0, n_ops, op_nq, op_qubits, op_mats, out_ref,
                                 block_qubits, block_mats, n_out);
}

}  // extern "C"
