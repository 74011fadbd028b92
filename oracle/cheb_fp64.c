/* FP64 Chebyshev evolution of a Pauli-sum Hamiltonian, in C with OpenMP —
 * TEST INFRASTRUCTURE ONLY (tests/ and bench.py's CPU leg check the GPU
 * result against it; no product path calls it).
 *
 * Operator: H = sum_t c_t P_t with P_t = i^popcount(x_t & z_t) X^x_t Z^z_t
 * (qubit q carries X if only bit q of x_t is set, Z if only bit q of z_t, Y
 * if both; Y = i X Z).  From Z^z |j> = (-1)^popcount(j & z) |j> and
 * X^x |j> = |j ^ x>:
 *     (P v)[i] = i^popcount(x & z) (-1)^popcount((i ^ x) & z) v[i ^ x].
 * The matrix elements are computed from the strings on the fly, so this is
 * independent of the CSR the device builds (csrc/k_pauli.hip) and of the
 * numpy CSR of oracle/hamiltonian.py.  Terms sharing an x mask are summed per
 * row before the gathered row is used (one gather per distinct x).
 *
 * Evolution: the reference's three-term recurrence
 * (blackwell_backend/chebyshev.py:111-156) in complex128:
 *     T0 = psi, T1 = s H T0, T_{k+1} = 2 s H T_k - T_{k-1},
 *     out = sum_k c_k T_k,  s = 1 / spectral bound,
 * with the coefficients c_k supplied by the caller (oracle/hamiltonian.py
 * chebyshev_coefficients, pinned by the Jacobi-Anger identity).
 *
 * Vectors hold r right-hand sides RHS-contiguous: element (row, j) at
 * row * r + j.
 */
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef double complex cd;

/* groups: ng distinct x masks xm[g]; group g owns terms [toff[g], toff[g+1])
 * with z masks zm[t] and complex coefficients coef[2t], coef[2t+1] (the
 * i^popcount(x & z) factor is applied here, not by the caller). */
typedef struct {
  int n, ng, r;
  const uint64_t* xm;
  const int32_t* toff;
  const uint64_t* zm;
  const double* coef;
} pauli_op;

static inline cd ipow(int k) {
  switch (k & 3) {
    case 0: return 1.0;
    case 1: return I;
    case 2: return -1.0;
    default: return -I;
  }
}

/* y = scale * H v  (recur = 0)  or  y = 2 * scale * H v - y  (recur = 1).
 * Rows are processed in blocks of kRowBlock per thread, group by group, so
 * the rows gathered for one x mask form one contiguous run of the input (the
 * block with its low bits permuted) instead of 64-byte pieces from every
 * group in turn. */
enum { kRowBlock = 2048 };
static void matvec(const pauli_op* op, const cd* v, cd* y, double scale, int recur) {
  const uint64_t N = 1ull << op->n;
  const int r = op->r;
  const int nt = op->toff[op->ng];
  /* c_t * i^popcount(x & z), once per term */
  cd* ce = (cd*)malloc(sizeof(cd) * (size_t)(nt ? nt : 1));
  for (int g = 0; g < op->ng; ++g)
    for (int t = op->toff[g]; t < op->toff[g + 1]; ++t)
      ce[t] = (op->coef[2 * t] + I * op->coef[2 * t + 1]) *
              ipow(__builtin_popcountll(op->xm[g] & op->zm[t]));
  const uint64_t nblk = (N + kRowBlock - 1) / kRowBlock;
#pragma omp parallel
  {
    cd* acc = (cd*)malloc(sizeof(cd) * (size_t)r * kRowBlock);
#pragma omp for schedule(static)
    for (uint64_t b = 0; b < nblk; ++b) {
      const uint64_t i0 = b * kRowBlock;
      const uint64_t i1 = i0 + kRowBlock < N ? i0 + kRowBlock : N;
      memset(acc, 0, sizeof(cd) * (size_t)r * (i1 - i0));
      for (int g = 0; g < op->ng; ++g) {
        const uint64_t x = op->xm[g];
        for (uint64_t i = i0; i < i1; ++i) {
          const uint64_t src = i ^ x;
          cd d = 0;
          for (int t = op->toff[g]; t < op->toff[g + 1]; ++t)
            d += (__builtin_popcountll(src & op->zm[t]) & 1) ? -ce[t] : ce[t];
          if (d == 0) continue;
          const cd* vs = v + src * (uint64_t)r;
          cd* a = acc + (i - i0) * (uint64_t)r;
          for (int j = 0; j < r; ++j) a[j] += d * vs[j];
        }
      }
      for (uint64_t i = i0; i < i1; ++i) {
        cd* yi = y + i * (uint64_t)r;
        const cd* a = acc + (i - i0) * (uint64_t)r;
        if (recur)
          for (int j = 0; j < r; ++j) yi[j] = 2.0 * scale * a[j] - yi[j];
        else
          for (int j = 0; j < r; ++j) yi[j] = scale * a[j];
      }
    }
    free(acc);
  }
  free(ce);
}

int ph_set_threads(int t) {
  if (t > 0) omp_set_num_threads(t);
  return omp_get_max_threads();
}

/* y = scale * H v for r RHS-contiguous vectors of 2^n rows. */
void ph_matvec(int n, int ng, const uint64_t* xm, const int32_t* toff, const uint64_t* zm,
               const double* coef, int r, const cd* v, cd* y, double scale) {
  const pauli_op op = {n, ng, r, xm, toff, zm, coef};
  matvec(&op, v, y, scale, 0);
}

/* out = sum_{k=0..K} c_k T_k(s H) psi; c holds 2 (K + 1) doubles.
 * Returns 0, or -1 when the work vectors cannot be allocated. */
int ph_cheb_evolve(int n, int ng, const uint64_t* xm, const int32_t* toff, const uint64_t* zm,
                   const double* coef, int r, double scale, int K, const double* c,
                   const cd* psi, cd* out) {
  const pauli_op op = {n, ng, r, xm, toff, zm, coef};
  const uint64_t M = (1ull << n) * (uint64_t)r;
  cd* a = (cd*)malloc(sizeof(cd) * M);
  cd* b = (cd*)malloc(sizeof(cd) * M);
  if (!a || !b) {
    free(a);
    free(b);
    return -1;
  }
  const cd c0 = c[0] + I * c[1];
  memcpy(a, psi, sizeof(cd) * M);                 /* a = T0 */
#pragma omp parallel for schedule(static)
  for (uint64_t e = 0; e < M; ++e) out[e] = c0 * a[e];
  if (K >= 1) {
    matvec(&op, a, b, scale, 0);                  /* b = T1 */
    const cd c1 = c[2] + I * c[3];
#pragma omp parallel for schedule(static)
    for (uint64_t e = 0; e < M; ++e) out[e] += c1 * b[e];
  }
  /* (prev, cur) = (a, b); T_{k} overwrites prev */
  cd* prev = a;
  cd* cur = b;
  for (int k = 2; k <= K; ++k) {
    matvec(&op, cur, prev, scale, 1);             /* prev = 2 s H cur - prev */
    const cd ck = c[2 * k] + I * c[2 * k + 1];
#pragma omp parallel for schedule(static)
    for (uint64_t e = 0; e < M; ++e) out[e] += ck * prev[e];
    cd* t = prev;
    prev = cur;
    cur = t;
  }
  free(a);
  free(b);
  return 0;
}
