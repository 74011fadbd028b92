/* FP64 statevector restatement in C with OpenMP — TEST INFRASTRUCTURE ONLY.
 *
 * Used by tests/ (cross-check of oracle/statevector.py) and by bench.py's
 * cpu_baseline leg (the timed host-core baseline).  Same conventions as the
 * numpy oracle: amplitude index bit q = qubit q; a 4x4 gate on (q0, q1) has
 * row bit 0 <-> q0 and row bit 1 <-> q1 (pass q0 = min to get the 2Q kernel
 * convention of bwk/csrc/apply_gate2q_sm120.cu:11-20); pair enumeration by
 * inserting zero bits (bwk/csrc/apply_gate_sm120.cu:100-104).
 */
#include <complex.h>
#include <stdint.h>
#include <omp.h>

typedef double complex cd;

static inline uint64_t ins0(uint64_t x, int q) {
  uint64_t lo = x & ((1ull << q) - 1);
  return ((x ^ lo) << 1) | lo;
}

int sv_set_threads(int t) {
  if (t > 0) omp_set_num_threads(t);
  return omp_get_max_threads();
}

void sv_init(cd* psi, int n) {
  uint64_t N = 1ull << n;
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < N; ++i) psi[i] = 0;
  psi[0] = 1;
}

/* u8: re00 im00 re01 im01 re10 im10 re11 im11 */
void sv_apply_1q(cd* psi, int n, const double* u8, int t) {
  const cd g00 = u8[0] + I * u8[1], g01 = u8[2] + I * u8[3];
  const cd g10 = u8[4] + I * u8[5], g11 = u8[6] + I * u8[7];
  const uint64_t pairs = 1ull << (n - 1), m = 1ull << t;
#pragma omp parallel for schedule(static)
  for (uint64_t p = 0; p < pairs; ++p) {
    const uint64_t i0 = ins0(p, t), i1 = i0 | m;
    const cd a0 = psi[i0], a1 = psi[i1];
    psi[i0] = g00 * a0 + g01 * a1;
    psi[i1] = g10 * a0 + g11 * a1;
  }
}

/* u32: row-major 4x4 (re, im); row bit 0 <-> q0, row bit 1 <-> q1 */
void sv_apply_2q(cd* psi, int n, const double* u32, int q0, int q1) {
  cd g[16];
  for (int k = 0; k < 16; ++k) g[k] = u32[2 * k] + I * u32[2 * k + 1];
  const int lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;
  const uint64_t groups = 1ull << (n - 2), m0 = 1ull << q0, m1 = 1ull << q1;
#pragma omp parallel for schedule(static)
  for (uint64_t p = 0; p < groups; ++p) {
    const uint64_t b = ins0(ins0(p, lo), hi);
    const uint64_t idx[4] = {b, b | m0, b | m1, b | m0 | m1};
    cd a[4], r[4];
    for (int k = 0; k < 4; ++k) a[k] = psi[idx[k]];
    for (int i = 0; i < 4; ++i) r[i] = g[4 * i] * a[0] + g[4 * i + 1] * a[1] + g[4 * i + 2] * a[2] + g[4 * i + 3] * a[3];
    for (int k = 0; k < 4; ++k) psi[idx[k]] = r[k];
  }
}

void sv_apply_mcx(cd* psi, int n, const int* controls, int nc, int t) {
  uint64_t cm = 0;
  for (int j = 0; j < nc; ++j) cm |= 1ull << controls[j];
  const uint64_t N = 1ull << n, m = 1ull << t;
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < N; ++i) {
    if ((i & m) || (i & cm) != cm) continue;
    const cd a = psi[i];
    psi[i] = psi[i | m];
    psi[i | m] = a;
  }
}

double sv_norm2(const cd* psi, int n) {
  const uint64_t N = 1ull << n;
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (uint64_t i = 0; i < N; ++i) s += creal(psi[i]) * creal(psi[i]) + cimag(psi[i]) * cimag(psi[i]);
  return s;
}
