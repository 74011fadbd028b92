// Host driver for qsim_plan_blocks and qsim_plan_blocks_search
// (csrc/host_plan.cpp) built with the host
// sanitizers by tests/test_planner_sanitizers.py.  For random circuits of
// 1..3-qubit ops (some wider than the block size, so they pass through) it
// checks that the plan, replayed in FP64 on a random state (each block as a
// 2^k x 2^k matrix on its ascending qubits, each passthrough op on its own
// qubits), equals the ops applied one by one; and that malformed arguments
// are rejected.  Exit status 0 on success.
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "qsim_gfx950.h"

namespace qsim {
static char g_msg[256];
int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof g_msg, fmt, ap);
  va_end(ap);
  return code;
}
void clear_error() { g_msg[0] = 0; }
}  // namespace qsim

using cd = std::complex<double>;

// psi <- u on qubits q[0..kq) (row bit j <-> q[j]); u row-major (re, im)
static void apply(std::vector<cd>& psi, int n, const double* u, const int* q, int kq) {
  const int d = 1 << kq;
  uint64_t m = 0;
  for (int j = 0; j < kq; ++j) m |= 1ull << q[j];
  std::vector<cd> in(d), out(d);
  std::vector<uint64_t> idx(d);
  for (uint64_t b = 0; b < (1ull << n); ++b) {
    if (b & m) continue;
    for (int s = 0; s < d; ++s) {
      uint64_t i = b;
      for (int j = 0; j < kq; ++j)
        if (s >> j & 1) i |= 1ull << q[j];
      idx[s] = i;
      in[s] = psi[i];
    }
    for (int r = 0; r < d; ++r) {
      cd acc = 0;
      for (int c = 0; c < d; ++c) acc += cd(u[2 * (r * d + c)], u[2 * (r * d + c) + 1]) * in[c];
      out[r] = acc;
    }
    for (int s = 0; s < d; ++s) psi[idx[s]] = out[s];
  }
}

int main() {
  std::mt19937_64 rng(12345);
  std::normal_distribution<double> nd;
  int cases = 0, passthrough = 0;
  for (int trial = 0; trial < 400; ++trial) {
    const int n = 2 + (int)(rng() % 9);       // 2..10 qubits
    const int k = 1 + (int)(rng() % 6);       // 1..6
    const int m = (int)(rng() % 40);
    std::vector<int32_t> nq(m), qs(8 * (size_t)m, -1);
    std::vector<double> mats;
    std::vector<size_t> moff(m);
    for (int i = 0; i < m; ++i) {
      int w = 1 + (int)(rng() % 3);
      if (w > n) w = n;
      nq[i] = w;
      std::vector<int> perm(n);
      for (int a = 0; a < n; ++a) perm[a] = a;
      for (int a = 0; a < w; ++a) std::swap(perm[a], perm[a + rng() % (n - a)]);
      for (int a = 0; a < w; ++a) qs[8 * i + a] = perm[a];
      moff[i] = mats.size();
      for (int e = 0; e < (1 << (2 * w)); ++e) {
        mats.push_back(nd(rng));
        mats.push_back(nd(rng));
      }
    }
    const int kk = k < n ? k : n, D = 1 << kk;
    std::vector<int32_t> ref(m > 0 ? m : 1), bq(8 * (size_t)(m > 0 ? m : 1));
    std::vector<double> bm((size_t)(m > 0 ? m : 1) * D * D * 2);
    int n_out = -1;
    // odd trials plan with the qubit-set search (effort 1), even ones greedily
    const int rc = trial & 1
                       ? qsim_plan_blocks_search(n, k, 1, m, nq.data(), qs.data(),
                                                 mats.empty() ? nullptr : mats.data(), ref.data(),
                                                 bq.data(), bm.data(), &n_out)
                       : qsim_plan_blocks(n, k, m, nq.data(), qs.data(), mats.empty() ? nullptr : mats.data(),
                                          ref.data(), bq.data(), bm.data(), &n_out);
    if (rc != QSIM_OK || n_out < 0 || n_out > m) {
      std::printf("trial %d: plan failed rc=%d n_out=%d (%s)\n", trial, rc, n_out, qsim::g_msg);
      return 1;
    }
    std::vector<cd> a(1ull << n), b;
    for (auto& x : a) x = cd(nd(rng), nd(rng));
    b = a;
    for (int i = 0; i < m; ++i) apply(a, n, mats.data() + moff[i], &qs[8 * i], nq[i]);
    int seen_ops = 0;
    for (int e = 0; e < n_out; ++e) {
      if (ref[e] >= 0) {
        const int blk = ref[e];
        for (int j = 1; j < kk; ++j)
          if (bq[8 * blk + j] <= bq[8 * blk + j - 1]) {
            std::printf("trial %d: block %d qubits not ascending\n", trial, blk);
            return 1;
          }
        apply(b, n, bm.data() + (size_t)blk * D * D * 2, &bq[8 * blk], kk);
      } else {
        const int i = -1 - ref[e];
        if (i < 0 || i >= m || nq[i] <= kk) {
          std::printf("trial %d: bad passthrough %d\n", trial, i);
          return 1;
        }
        apply(b, n, mats.data() + moff[i], &qs[8 * i], nq[i]);
        ++passthrough;
      }
      ++seen_ops;
    }
    double err = 0, scale = 0;
    for (size_t i = 0; i < a.size(); ++i) {
      err = std::max(err, std::abs(a[i] - b[i]));
      scale = std::max(scale, std::abs(a[i]));
    }
    if (err > 1e-9 * (scale + 1)) {
      std::printf("trial %d: n=%d k=%d m=%d plan differs by %.3e (scale %.3e)\n", trial, n, k, m, err, scale);
      return 1;
    }
    ++cases;
  }
  // malformed arguments
  int32_t nq1[1] = {1}, q1[8] = {0, -1, -1, -1, -1, -1, -1, -1}, out[4], bqs[32];
  double u1[8] = {1, 0, 0, 0, 0, 0, 1, 0}, bm1[64 * 64 * 2];
  int no = 0;
  struct { int n, k, nq, q0; } bad[] = {{0, 5, 1, 0}, {4, 0, 1, 0}, {4, 7, 1, 0}, {4, 5, 9, 0},
                                        {4, 5, 1, 4}, {4, 5, 1, -1}, {35, 5, 1, 0}};
  for (auto& c : bad) {
    nq1[0] = c.nq;
    q1[0] = c.q0;
    if (qsim_plan_blocks(c.n, c.k, 1, nq1, q1, u1, out, bqs, bm1, &no) == QSIM_OK) {
      std::printf("malformed case n=%d k=%d nq=%d q0=%d accepted\n", c.n, c.k, c.nq, c.q0);
      return 1;
    }
  }
  int32_t nq2[1] = {2}, q2[8] = {1, 1, -1, -1, -1, -1, -1, -1};
  double u2[32] = {0};
  if (qsim_plan_blocks(4, 5, 1, nq2, q2, u2, out, bqs, bm1, &no) == QSIM_OK) {
    std::printf("repeated qubit accepted\n");
    return 1;
  }
  if (qsim_plan_blocks(4, 5, 1, nq2, q2, u2, out, bqs, bm1, nullptr) == QSIM_OK) {
    std::printf("NULL n_out accepted\n");
    return 1;
  }
  int32_t nq3[1] = {1}, q3[8] = {0, -1, -1, -1, -1, -1, -1, -1};
  if (qsim_plan_blocks_search(4, 5, 2, 1, nq3, q3, u2, out, bqs, bm1, &no) == QSIM_OK) {
    std::printf("effort 2 accepted\n");
    return 1;
  }
  std::printf("planner ok: %d random plans replayed exactly (%d passthrough entries), malformed input rejected\n",
              cases, passthrough);
  return 0;
}
