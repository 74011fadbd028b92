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

template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_block5(float2* __restrict__ st, uint64_t n_waves,
                                                       BlockArgs a, const float2* __restrict__ U) {
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5, c = lane & 31;
  const uint64_t wave = (uint64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
  // A fragments: U[c][2s+h], s = 0..15
  float ur[16], ui[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const float2 u = U[c * 32 + 2 * s + h];
    ur[s] = u.x;
    ui[s] = u.y;
  }
  if (wave >= n_waves) return;
  const uint64_t base = group_base(wave * 32 + c, a);
  float xr[16], xi[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const float2 x = st[base | row_off(2 * s + h, a)];
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
    st[base | row_off(i, a)] = make_float2(accr[r], acci[r]);
  }
}

}  // namespace qsim

using namespace qsim;

extern "C" {

int qsim_apply_block_mfma(void* stream, void* state, int n_qubits, const void* block_dev,
                          const int* qubits, int k) {
  if (int rc = check_n(n_qubits)) return rc;
  if (!state || !block_dev || !qubits) return fail(QSIM_EINVAL, "NULL pointer argument");
  if (k != 5) return fail(QSIM_ERANGE, "MFMA blocks support k = 5 (got %d)", k);
  if (n_qubits < 10) return fail(QSIM_ERANGE, "MFMA blocks need n_qubits >= 10 (got %d)", n_qubits);
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
  const uint64_t n_waves = (1ull << (n_qubits - 5)) / 32;
  constexpr int W = 4;
  hipLaunchKernelGGL((k_block5<W>), dim3(ceil_div(n_waves, W)), dim3(64 * W), 0, as_stream(stream),
                     (float2*)state, n_waves, a, (const float2*)block_dev);
  return launch_status("apply_block_mfma");
}

}  // extern "C"
