/*
 * qsim_gfx950.h — C ABI of the MI355X (gfx950) statevector library.
 *
 * One shared library, libqsim_gfx950.so, exports every function below with
 * C linkage.  Every entry point:
 *   - takes the HIP stream it launches on as `void* stream` (NULL = the null
 *     stream), device pointers as `void*`, sizes as 64-bit integers and qubit
 *     indices as `int`;
 *   - returns an int status (QSIM_OK = 0).  On failure the message is kept in
 *     a per-thread buffer returned by qsim_last_error();
 *   - is asynchronous with respect to the host unless its comment says it
 *     synchronises the stream (measure_qubit and the "_to_host" reads do).
 *
 * Statevector layout: interleaved complex64 (re, im float pairs), amplitude
 * index i has bit q equal to the value of qubit q (qubit 0 = least
 * significant bit).  A state of n qubits holds 2^n amplitudes, n <= 34.
 *
 * Each declaration names the reference interface it replaces as
 * file:line in DarrellThomas/qiskit-blackwell.  Entry points marked
 * "(new)" have no counterpart in the reference.
 */
#ifndef QSIM_GFX950_H
#define QSIM_GFX950_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define QSIM_OK          0
#define QSIM_EINVAL      1   /* bad argument (qubit out of range, duplicate qubits, ...) */
#define QSIM_ERANGE      2   /* size beyond what the entry point supports              */
#define QSIM_EHIP        3   /* a HIP runtime call or kernel launch failed             */
#define QSIM_ENOMEM      4   /* device allocation failed                               */

#define QSIM_MAX_QUBITS      34
#define QSIM_MAX_CONTROLS    7    /* controls per multi-controlled gate  */
#define QSIM_MAX_BLOCK_QUBITS 6   /* largest dense block (2^6 x 2^6)     */

/* ---- library meta (new) ------------------------------------------------ */
const char* qsim_version(void);
int         qsim_abi_version(void);
const char* qsim_last_error(void);
/* Number of HIP devices visible (0 on a host without a GPU; makes no launch). */
int         qsim_device_count(int* count);

/* ---- state initialisation ---------------------------------------------- */
/* |0...0> into a caller-allocated state of 2^n amplitudes (one write pass).
 * Replaces state_init, bwk/csrc/apply_diagonal_sm120.cu:210-224
 * (cuquantum.py:213-219), which allocates and writes twice.              */
int qsim_state_init(void* stream, void* state, int n_qubits);
/* |0...0> in each of n_states contiguous states of 2^n amplitudes.
 * Replaces state_init_batched, bwk/csrc/batch_sm120.cu:260-279.          */
int qsim_state_init_batched(void* stream, void* states, int64_t n_states, int n_qubits);

/* ---- dense gates (compensated FP32) ------------------------------------ */
/* 2x2 gate on `target`.  gate_dev: device complex64[2][2], row-major.
 * new0 = g00 a0 + g01 a1, new1 = g10 a0 + g11 a1 over pairs (a0, a1) that
 * differ only in bit `target`.
 * Replaces apply_gate, bwk/csrc/apply_gate_sm120.cu:269-307 (cuquantum.py:40-51). */
int qsim_apply_gate(void* stream, void* state, int n_qubits, const void* gate_dev, int target);
/* Same, with the matrix passed by value from the host as 8 doubles
 * (re00, im00, re01, im01, re10, im10, re11, im11), rounded to FP32. (new) */
int qsim_apply_gate_h(void* stream, void* state, int n_qubits, const double* gate8, int target);

/* 4x4 gate on (q0, q1), q0 != q1.  Matrix row/column index
 * r = 2*bit(max(q0,q1)) + bit(min(q0,q1)) whatever the argument order.
 * Replaces apply_gate_2q, bwk/csrc/apply_gate2q_sm120.cu:232-280
 * (cuquantum.py:54-66).                                                  */
int qsim_apply_gate_2q(void* stream, void* state, int n_qubits, const void* gate_dev, int q0, int q1);
/* Same with the 4x4 matrix by value: 32 doubles, row-major (re, im). (new) */
int qsim_apply_gate_2q_h(void* stream, void* state, int n_qubits, const double* gate32, int q0, int q1);

/* ---- diagonal gates (TwoProduct-compensated complex multiply) ---------- */
/* state[i] *= diag[bit(i, target)], diag_dev: device complex64[2].
 * Replaces apply_diagonal, bwk/csrc/apply_diagonal_sm120.cu:118-148.     */
int qsim_apply_diagonal(void* stream, void* state, int n_qubits, const void* diag_dev, int target);
/* Same, diag by value: 4 doubles (re0, im0, re1, im1). (new)             */
int qsim_apply_diagonal_h(void* stream, void* state, int n_qubits, const double* diag4, int target);
/* Multiply only the amplitudes with bit(target) = 1 by phase.
 * Replaces apply_diagonal_phase, bwk/csrc/apply_diagonal_sm120.cu:150-175. */
int qsim_apply_diagonal_phase(void* stream, void* state, int n_qubits,
                              double phase_re, double phase_im, int target);
/* state[i] *= diag[2*bit(i,max(q0,q1)) + bit(i,min(q0,q1))], diag_dev: complex64[4].
 * Replaces apply_diagonal_2q, bwk/csrc/apply_diagonal_sm120.cu:177-208.  */
int qsim_apply_diagonal_2q(void* stream, void* state, int n_qubits, const void* diag_dev, int q0, int q1);

/* ---- multi-controlled gates -------------------------------------------- */
/* Controlled 2x2 gate: applied on `target` where every control bit is 1.
 * controls: host array of n_controls (0..7) distinct qubits != target.
 * Replaces apply_mcgate, bwk/csrc/apply_mcgate_sm120.cu:102-165
 * (cuquantum.py:297-317).                                                */
int qsim_apply_mcgate(void* stream, void* state, int n_qubits, const void* gate_dev,
                      const int64_t* controls, int n_controls, int target);
/* Controlled 4x4 gate on (t0, t1), index convention as qsim_apply_gate_2q.
 * Replaces apply_mc2qgate, bwk/csrc/apply_mc2qgate_sm120.cu:127-196
 * (cuquantum.py:320-337).                                                */
int qsim_apply_mc2qgate(void* stream, void* state, int n_qubits, const void* gate_dev,
                        const int64_t* controls, int n_controls, int t0, int t1);

/* ---- permutation gates (pure data movement, bit-exact) ------------------ */
/* Multi-controlled X (n_controls = 0: X, 1: CNOT, 2: Toffoli, ... 7).
 * Bit-exact form of the backend's apply_mcgate(state, X, ...) call,
 * blackwell_backend/simulator.py:456-458.                                */
int qsim_apply_mcx(void* stream, void* state, int n_qubits,
                   const int64_t* controls, int n_controls, int target);
/* Multi-controlled SWAP of q0 and q1 (n_controls = 0: SWAP, 1: Fredkin).
 * Bit-exact form of apply_mc2qgate(state, SWAP, ...),
 * blackwell_backend/simulator.py:459-461.                                */
int qsim_apply_mcswap(void* stream, void* state, int n_qubits,
                      const int64_t* controls, int n_controls, int q0, int q1);

/* ---- fused 1Q gate lists ------------------------------------------------ */
/* Apply n_gates 2x2 gates in order, gates_dev: device complex64[n_gates][2][2],
 * targets: HOST int32[n_gates].  One HBM pass per group of distinct targets
 * that fits an LDS tile.
 * Replaces apply_gates_fused, bwk/csrc/apply_gate_sm120.cu:311-350.      */
int qsim_apply_gates_fused(void* stream, void* state, int n_qubits, const void* gates_dev,
                           const int32_t* targets, int n_gates);
/* Register-resident variant for targets < 6, one HBM pass.  targets_dev:
 * DEVICE int32[n_gates] (as the reference).
 * Replaces apply_gates_reg_tiled, bwk/csrc/apply_gate_sm120.cu:238-265.  */
int qsim_apply_gates_reg_tiled(void* stream, void* state, int n_qubits, const void* gates_dev,
                               const void* targets_dev, int n_gates);

/* ---- tile programs: many ops per HBM pass (new) ------------------------- */
/* One op of a tile program.  Qubit fields are positions INSIDE the tile
 * (0..tile_bits-1).  coef is an offset, in complex64 units, into the
 * program's coefficient array.                                           */
typedef struct qsim_op {
  uint8_t  kind;      /* QSIM_OP_*                                      */
  uint8_t  nq;        /* number of target qubits                        */
  uint8_t  q[6];      /* target positions (matrix bit j <-> q[j])       */
  uint32_t cmask;     /* control mask inside the tile (0 = none)        */
  uint32_t coef;      /* coefficient offset                             */
} qsim_op;

#define QSIM_OP_M1     1   /* dense 2x2, 4 coefs                               */
#define QSIM_OP_M2     2   /* dense 4x4 on (q[0], q[1]): row bit0 <-> q[0], 16  */
#define QSIM_OP_D1     3   /* diagonal 2, 2 coefs                              */
#define QSIM_OP_D2     4   /* diagonal 4, index 2*b(q[1]) + b(q[0]), 4 coefs   */
#define QSIM_OP_PH     5   /* phase on bit(q[0]) = 1, 1 coef                   */
#define QSIM_OP_X      6   /* X on q[0] (controls via cmask), bit-exact        */
#define QSIM_OP_SWAP   7   /* SWAP q[0], q[1] (controls via cmask), bit-exact  */
#define QSIM_OP_MK     8   /* dense 2^nq x 2^nq block, nq <= 6, row bit j <-> q[j] */
#define QSIM_OP_DK     9   /* diagonal on nq qubits, 2^nq coefs               */

/* Apply a program to every tile of one state.  A tile is the 2^tile_bits
 * amplitudes whose index varies over the qubits set in tile_mask (ascending
 * order gives the positions 0..tile_bits-1) with all other bits fixed.
 * tile_bits = popcount(tile_mask) must be 6..13; ops_dev/coefs_dev in device
 * memory; n_ops >= 1.                                                     */
int qsim_tile_program(void* stream, void* state, int n_qubits, uint64_t tile_mask,
                      const void* ops_dev, int n_ops, const void* coefs_dev);

/* ---- dense k-qubit blocks on FP32 MFMA (new) ---------------------------- */
/* Apply a dense 2^k x 2^k complex block on qubits[0..k-1] (row bit j <->
 * qubits[j], any order, all distinct), k = 5, n_qubits >= 10.  block_dev:
 * device complex64[32][32] row-major.  The batched contraction runs on
 * v_mfma_f32_32x32x2_f32 with FP32 accumulation (one wave per 32 groups).
 * Smaller or larger blocks go through qsim_tile_program (QSIM_OP_MK).     */
int qsim_apply_block_mfma(void* stream, void* state, int n_qubits, const void* block_dev,
                          const int* qubits, int k);

/* ---- host runtime (new) --------------------------------------------------- */
/* Dense-block planner (host only, FP64).  Ops in circuit order: op i has
 * op_nq[i] (1..8) qubits op_qubits[8*i .. 8*i+op_nq[i]) and a row-major
 * complex matrix of 2^nq x 2^nq (re, im doubles) in op_mats, concatenated.
 * Packs them greedily into blocks of exactly min(k, n) qubits (k <= 6) such
 * that every op still follows the earlier ops it shares a qubit with.
 * Outputs, in execution order, *n_out entries out_ref[e]: b >= 0 names block
 * b (qubits block_qubits[8*b ..], ascending, -1 padded; a block short of
 * k qubits takes the lowest unused qubits from 12 up, then below; matrix block_mats +
 * b*2^(2k)*2 doubles, row bit j <-> j-th block qubit) and -1-i passes op i
 * through (wider than k).  Capacity: n_ops entries and blocks.            */
int qsim_plan_blocks(int n_qubits, int k, int n_ops, const int32_t* op_nq, const int32_t* op_qubits,
                     const double* op_mats, int32_t* out_ref, int32_t* block_qubits,
                     double* block_mats, int* n_out);
/* qsim_plan_blocks with a search effort: 0 is qsim_plan_blocks; 1 also
 * scores, for each seed op, every block qubit set made of the seed's qubits
 * and up to k - |seed| of the qubits that the next 30 ops connect to it, and
 * keeps the set that absorbs the most ops (fewer blocks for a few
 * milliseconds of host time on a few hundred ops).  Same outputs.  For
 * k = 5 and n_qubits > 25 both efforts count ops per unit of kernel time:
 * a block holding qubit 0 weighs 1.10, one with three qubits in 0..4
 * 1.31.  (new)                                                            */
int qsim_plan_blocks_search(int n_qubits, int k, int effort, int n_ops, const int32_t* op_nq,
                            const int32_t* op_qubits, const double* op_mats, int32_t* out_ref,
                            int32_t* block_qubits, double* block_mats, int* n_out);
/* Apply n_blocks 5-qubit blocks in order with qsim_apply_block_mfma:
 * blocks_dev complex64 [n_blocks][32][32], block_qubits host int32
 * [n_blocks][8] (first 5 used).  One call issues the whole plan.         */
int qsim_apply_blocks(void* stream, void* state, int n_qubits, const void* blocks_dev,
                      const int32_t* block_qubits, int n_blocks);

/* ---- reductions, measurement, sampling ---------------------------------- */
/* Size in bytes of the workspace the reduction entry points need for a
 * state of n_qubits (new; the reference keeps process-global scratch,
 * bwk/csrc/measure_sm120.cu:135-148).                                    */
int64_t qsim_reduce_workspace_bytes(int n_qubits);
/* ||psi||^2 in FP64, deterministic two-pass reduction; result written to
 * the device double *out_dev. (new)                                      */
int qsim_norm2(void* stream, const void* state, int n_qubits, void* workspace, void* out_dev);
/* probs[i] = |a_i|^2 as float32.  Replaces measure_probs,
 * bwk/csrc/measure_sm120.cu:186-207.                                     */
int qsim_measure_probs(void* stream, const void* state, int n_qubits, void* probs_dev);
/* Measure `target`: p0 reduced in FP64, outcome = (random_val < p0/norm) ? 0 : 1,
 * kept half scaled by 1/sqrt(p) computed in FP64, other half zeroed.
 * Synchronises the stream; outcome returned in *outcome.
 * Replaces measure_qubit, bwk/csrc/measure_sm120.cu:150-184.             */
int qsim_measure_qubit(void* stream, void* state, int n_qubits, int target,
                       double random_val, void* workspace, int* outcome);
/* Scale to unit norm (FP64 norm and FP64 scale).
 * Replaces renormalize, bwk/csrc/renorm_sm120.cu:75-108.                 */
int qsim_renormalize(void* stream, void* state, int n_qubits, void* workspace);
/* Sampling workspace (FP64 block totals and their prefix). (new)         */
int64_t qsim_sample_workspace_bytes(int n_qubits);
/* For every uniform u in [0,1) (device double[n_shots]) return the smallest
 * index i with cdf(i) > u * ||psi||^2, cdf in FP64 (device int64[n_shots]).
 * Replaces sample, bwk/csrc/sample_sm120.cu:73-132 (FP32 CDF there).     */
int qsim_sample(void* stream, const void* state, int n_qubits, const void* uniforms_dev,
                int64_t n_shots, void* workspace, void* out_dev);
/* qsim_sample without a workspace argument and without any allocation, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  No counterpart in
 * the reference.  The sampling scratch is a static device array of 4 MiB
 * per GPU (held until the process ends), so n_qubits <= 26 (QSIM_ERANGE
 * above).  Same results as qsim_sample, bit for bit.  Calls that overlap on
 * one GPU share that scratch and race: serialise them (one stream, or
 * events).                                                               */
int qsim_sample_noalloc(void* stream, const void* state, int n_qubits, const void* uniforms_dev,
                        int64_t n_shots, void* out_dev);
/* uniforms_dev[k] = (SplitMix64(seed + (k+1) * 0x9E3779B97F4A7C15) >> 11) * 2^-53,
 * k < count (device double[count]): the stream qsim_run_instances samples
 * from in-kernel, so qsim_sample can draw the same shots for a state in HBM.
 * (new)                                                                  */
int qsim_uniforms_splitmix(void* stream, uint64_t seed, int64_t count, void* uniforms_dev);

/* One state of a batched sampling call (new): .state / .out are device
 * addresses of a complex64 state of n_qubits (8..20) and of its int64[shots]
 * output, .seed its SplitMix64 seed (as qsim_uniforms_splitmix), and
 * .leaf_off / .shot_off the prefix sums of 2^(n_qubits - 8) and of shots
 * over the records before it.                                            */
typedef struct qsim_sample_rec {
  uint64_t state;
  uint64_t out;
  uint64_t seed;
  uint64_t leaf_off;
  uint64_t shot_off;
  int32_t  n_qubits;
  uint32_t shots;
} qsim_sample_rec;
int64_t qsim_sample_batch_workspace_bytes(int64_t n_leaves, int64_t n_states);
/* qsim_sample on each of n_states states with its seeded uniforms, in three
 * launches for all of them; every state's samples are bitwise those of
 * qsim_uniforms_splitmix + qsim_sample on it alone.  recs_host is validated
 * and must hold the bytes of the device copy recs_dev.  (new; the serving
 * HBM lane)                                                              */
int qsim_sample_batch(void* stream, const qsim_sample_rec* recs_host, const void* recs_dev,
                      int64_t n_states, void* workspace);

/* ---- Pauli strings, expectation values, noise channels ------------------ */
/* A Pauli string is (x_mask, z_mask): qubit q carries X if only x bit q is
 * set, Z if only z bit q, Y if both; P = i^popcount(x&z) X^x Z^z.
 * Expectations are <psi|P|psi> of the unnormalised state, summed in FP64 in
 * a fixed order (deterministic), written as doubles to out_dev.            */
int64_t qsim_expectation_workspace_bytes(int n_qubits);
/* out_dev[k] = <psi|P_k|psi> for n_terms strings.  Terms with the same
 * x mask share one pass over the state (up to 8 per pass).  (new; the
 * reference evaluates one Z string or one X/Y qubit per call)            */
int qsim_expectation_paulis(void* stream, const void* state, int n_qubits, int64_t n_terms,
                            const uint64_t* x_masks, const uint64_t* z_masks, void* workspace,
                            void* out_dev);
/* <Z on the z_mask qubits>.
 * Replaces expectation_z, bwk/csrc/expectation_sm120.cu:180-204.         */
int qsim_expectation_z(void* stream, const void* state, int n_qubits, uint64_t z_mask,
                       void* workspace, void* out_dev);
/* <(Z on z_mask) (op on xy_qubit)>, op 0 = I/Z string only, 1 = X, 2 = Y; a
 * z bit on xy_qubit is ignored.
 * Replaces expectation_pauli, bwk/csrc/expectation_sm120.cu:206-239.     */
int qsim_expectation_pauli(void* stream, const void* state, int n_qubits, uint64_t z_mask,
                           int xy_qubit, int op, void* workspace, void* out_dev);
/* psi <- P psi, bit-exact (components are moved and negated only).  (new;
 * the reference's Pauli kernels are internal to its depolarizing channel,
 * bwk/csrc/noise_sm120.cu:39-101)                                        */
int qsim_apply_pauli_string(void* stream, void* state, int n_qubits, uint64_t x_mask,
                            uint64_t z_mask);
/* Depolarizing channel, one trajectory step: *branch = 0 (identity) when
 * random_val < 1 - p, else 1/2/3 (X/Y/Z on target) over three intervals of
 * width p/3; the Pauli is applied bit-exactly.  Does not synchronise.
 * Replaces apply_depolarizing, bwk/csrc/noise_sm120.cu:238-281.          */
int qsim_apply_depolarizing(void* stream, void* state, int n_qubits, int target, double p,
                            double random_val, int* branch);
/* Amplitude damping, one trajectory step: branch 1 (decay, E1 = [[0,
 * sqrt(gamma)], [0, 0]]) when random_val < gamma * P(target = 1), else
 * branch 0 (E0 = diag(1, sqrt(1 - gamma))); the result is renormalised to
 * unit norm.  Probabilities and scales are FP64.  workspace: see
 * qsim_reduce_workspace_bytes.  Synchronises the stream.
 * Replaces apply_amplitude_damping, bwk/csrc/noise_sm120.cu:284-329.     */
int qsim_apply_amplitude_damping(void* stream, void* state, int n_qubits, int target,
                                 double gamma, double random_val, void* workspace, int* branch);
/* Dephasing, as amplitude damping with E1 = diag(0, sqrt(lambda)).
 * Replaces apply_dephasing, bwk/csrc/noise_sm120.cu:332-373.             */
int qsim_apply_dephasing(void* stream, void* state, int n_qubits, int target, double lam,
                         double random_val, void* workspace, int* branch);

/* ---- batched same-size states ------------------------------------------- */
/* Same 2x2 gate on each of n_states states of 2^n amplitudes.
 * Replaces apply_gate_batched, bwk/csrc/batch_sm120.cu:186-228.          */
int qsim_apply_gate_batched(void* stream, void* states, int64_t n_states, int n_qubits,
                            const void* gate_dev, int target);
/* Same gate list (targets < 6, DEVICE int32) on each state.
 * Replaces apply_gates_batched_fused, bwk/csrc/batch_sm120.cu:230-258.   */
int qsim_apply_gates_batched_fused(void* stream, void* states, int64_t n_states, int n_qubits,
                                   const void* gates_dev, const void* targets_dev, int n_gates);

/* ---- per-instance gate tables: whole circuits in one launch -------------- */
/* One circuit instance of a batch.  Its program is ops[op_off .. op_off+n_ops)
 * with coefficient offsets into the shared coefficient array.            */
typedef struct qsim_instance {
  uint64_t state_off;  /* complex64 offset of the instance's HBM state (used
                          when init == 0 or want & 1)                     */
  uint64_t seed;       /* seed of the in-kernel uniforms (SplitMix64)     */
  uint64_t shot_off;   /* int64 offset of this instance's samples         */
  uint32_t op_off;
  uint32_t n_ops;
  uint32_t shots;      /* samples drawn from the final state (0 = none)   */
  uint8_t  n_qubits;   /* 1..14: the whole state is held in LDS           */
  uint8_t  init;       /* 1: start from |0..0>; 0: load the HBM state     */
  uint8_t  want;       /* bit0: write the final amplitudes back to HBM    */
  uint8_t  pad;
} qsim_instance;
/* Run n_inst circuit instances of uneven size (n_qubits <= max_qubits <= 14),
 * one workgroup per instance, state resident in LDS for the whole program.
 * Sample k of an instance uses u_k = SplitMix64(seed + (k+1)*0x9E3779B97F4A7C15)
 * >> 11 scaled by 2^-53, and returns the smallest index i with
 * cdf(i) > u_k * ||psi||^2 (FP64).  inst_host / ops_host must hold the bytes
 * of the device copies; they are validated (n_qubits <= max_qubits, op,
 * state, shot and coefficient ranges against n_ops_total, n_state_slots
 * complex64 slots, n_shot_slots int64 slots and n_coefs) before anything
 * is launched.  Generalises qv8_simulate (bwk/csrc/qv8_sim_sm120.cu:121-151)
 * and qv4_simulate (bwk/csrc/qv4_sim_sm120.cu:139-169) to per-instance n
 * and programs.                                                          */
int qsim_run_instances(void* stream, void* states_base, uint64_t n_state_slots,
                       const qsim_instance* inst_host, const void* inst_dev, int n_inst,
                       const qsim_op* ops_host, const void* ops_dev, uint64_t n_ops_total,
                       const void* coefs_dev, uint64_t n_coefs, void* shots_dev,
                       uint64_t n_shot_slots, int max_qubits);

/* One (instance, block) of a batched block step (new): block_rec.state is the
 * device address of a complex64 state of n_qubits (10..32), .mat the device
 * address of its 32x32 complex64 block (row-major, row bit j <-> q[j]), and
 * .set_begin the prefix sum of 2^(n_qubits - 10) over the records before it
 * in the same step.                                                       */
typedef struct qsim_block_rec {
  uint64_t state;
  uint64_t mat;
  uint64_t set_begin;
  int32_t  n_qubits;
  int8_t   q[5];
  int8_t   pad[7];
} qsim_block_rec;
/* Apply one 5-qubit block to each of n_recs states in one launch (FP32 MFMA,
 * the same arithmetic as qsim_apply_block_mfma, so each state's result is
 * independent of the batch).  recs_host is validated on the host and must
 * hold the same bytes as the device copy recs_dev, which the kernel reads.
 * The states of one call must not overlap.  (new; the serving HBM lane)   */
int qsim_apply_block_batch(void* stream, const qsim_block_rec* recs_host, const void* recs_dev,
                           int64_t n_recs);
/* Fixed-n spec paths kept for API parity: probs out, float32.
 * qv8: gate_matrices f32[C][G][4][4][2], gate_qubits i32[C][G][2], row index
 * r = 2*bit(q1) + bit(q0) with q0 the first listed qubit (qv8_sim_sm120.cu:88-90).
 * qv4: gate_data f32[C][G][32] (16 re, 16 im), pair_ids i32[C][G] (qv4.py:14-21). */
int qsim_qv8_simulate(void* stream, const void* gate_matrices, const void* gate_qubits,
                      int64_t n_circuits, int n_gates, void* probs_out);
int qsim_qv4_simulate(void* stream, const void* gate_data, const void* pair_ids,
                      int64_t n_circuits, int n_gates, void* probs_out);

/* ---- Pauli sum -> CSR on the device --------------------------------------- */
/* Replaces SparsePauliOp.to_matrix(sparse=True).tocsr() plus the H2D copy in
 * HamiltonianCSR.from_sparse_pauli_op (blackwell_backend/hamiltonian.py:44-69).
 * Terms grouped by X mask: group g (n_groups of them, X masks xm[g] uint64
 * ascending) owns the terms t in [toff[g], toff[g+1]) (int32 toff[n_groups+1])
 * in input order; term t has the Y|Z mask sm[t] (uint64) and an FP64 complex
 * coefficient coef[t] (2 doubles) with (-i)^popcount(Y) folded in.  Row i of
 * group g holds sum_t coef[t] (-1)^popcount(i & sm[t]) (FP64, term order) at
 * column i ^ xm[g]; entries that are exactly zero are dropped; values are
 * rounded to complex64.  All inputs are device pointers.
 * Two calls with the same workspace: _count (synchronises `stream`, returns
 * the number of nonzeros, QSIM_ERANGE above INT32_MAX), then _fill into
 * int32 row_ptr[2^n+1], int32 col_idx[nnz], complex64 values[nnz].     (new) */
int64_t qsim_pauli_csr_workspace_bytes(int n_qubits);
int qsim_pauli_csr_count(void* stream, int n_qubits, int n_groups, const void* xm, const void* toff,
                         const void* sm, const void* coef, void* ws, int64_t* nnz_out);
int qsim_pauli_csr_fill(void* stream, int n_qubits, int n_groups, const void* xm, const void* toff,
                        const void* sm, const void* coef, const void* ws, int64_t nnz,
                        void* row_ptr, void* col_idx, void* values);

/* ---- Chebyshev evolution ------------------------------------------------- */
/* y = alpha*A*x + beta*y, CSR int32 row_ptr[n_rows+1], int32 col_idx[nnz],
 * complex64 values[nnz]; per-row sums TwoProduct + Kahan compensated.
 * Replaces spmv_csr, bwk/csrc/spmv_sm120.cu:108-142.                     */
int qsim_spmv_csr(void* stream, const void* row_ptr, const void* col_idx, const void* values,
                  int64_t n_rows, const void* x, void* y,
                  double alpha_re, double alpha_im, double beta_re, double beta_im);
/* t_next = 2*s - t_prev.  Replaces chebyshev_step, chebyshev_sm120.cu:96-114. */
int qsim_chebyshev_step(void* stream, const void* s, const void* t_prev, void* t_next, int64_t n);
/* accum += c * t.  Replaces chebyshev_accum, chebyshev_sm120.cu:116-134.     */
int qsim_chebyshev_accum(void* stream, void* accum, const void* t, int64_t n, double c_re, double c_im);
/* Both of the above fused.  Replaces chebyshev_step_accum,
 * chebyshev_sm120.cu:136-159.                                            */
int qsim_chebyshev_step_accum(void* stream, const void* s, const void* t_prev, void* t_next,
                              void* accum, int64_t n, double c_re, double c_im);
/* Batched recurrence over n_rhs right-hand sides stored RHS-contiguous
 * (element (row, j) at row*n_rhs + j).  One launch does
 *   w   = scale * A * t_cur                     (SpMM, per-row sums below)
 *   t_nx = first ? w : 2*w - t_prev              (written over t_prev)
 *   acc += c * t_nx                              (Kahan: acc + comp)
 * first != 0 computes T1 = scale*A*T0 (t_prev ignored).  Row sums:
 *   QSIM_CHEB_ACC_FP64  every FP32 x FP32 product (exact in FP64) summed
 *                       with FP64 FMAs, rounded once to complex64;
 *   QSIM_CHEB_ACC_FP32C TwoProduct + Neumaier compensated FP32 (the
 *                       reference's TwoProduct/Kahan scheme,
 *                       bwk/csrc/spmv_sm120.cu:26-54).
 * (new; the reference runs one vector: blackwell_backend/chebyshev.py:141-156) */
#define QSIM_CHEB_ACC_FP64  0
#define QSIM_CHEB_ACC_FP32C 1
int qsim_cheb_spmm_step(void* stream, const void* row_ptr, const void* col_idx, const void* values,
                        int64_t n_rows, int n_rhs, double scale,
                        const void* t_cur, void* t_prev_next, void* acc, void* comp,
                        double c_re, double c_im, int first, int accumulate);
/* Split of a CSR of n_rows = 2^n rows at bit split_bit (1..n-1): the
 * entries with (row ^ col) < 2^split_bit form the low CSR (rows in natural
 * order); the others form the high CSR, whose row p holds the entries of
 * row rot(p) = ((p & (2^(n-split_bit) - 1)) << split_bit) | (p >> (n-split_bit)).
 * Each row keeps its entry order.  Two calls with the same workspace:
 * _count (synchronises `stream`, returns both nonzero counts, QSIM_ERANGE
 * above INT32_MAX), then _fill into int32 row pointers [n_rows+1] and
 * int32 / complex64 entries of those counts.  (new)                       */
int64_t qsim_csr_split_workspace_bytes(int64_t n_rows);
int qsim_csr_split_count(void* stream, const void* row_ptr, const void* col_idx, int64_t n_rows,
                         int split_bit, void* workspace, int64_t* nnz_lo, int64_t* nnz_hi);
int qsim_csr_split_fill(void* stream, const void* row_ptr, const void* col_idx, const void* values,
                        int64_t n_rows, int split_bit, const void* workspace, void* rp_lo,
                        void* ci_lo, void* val_lo, void* rp_hi, void* ci_hi, void* val_hi);
/* qsim_cheb_spmm_step (QSIM_CHEB_ACC_FP64) over a split CSR, 32 RHS,
 * n_rows <= 2^24, in two launches: w1 = A_low * t_cur in natural row order
 * (rounded to complex64 into w1, a [n_rows][32] complex64 buffer), then
 * A_high * t_cur + w1 and the recurrence and accumulation of
 * qsim_cheb_spmm_step in the rotated order, so each pass gathers rows of
 * t_cur from a neighbourhood of 2^split_bit (2^(n-split_bit)) rows.  Same
 * arguments and results as qsim_cheb_spmm_step up to the rounding of w1.
 * Opt-in: on a 22-qubit chain at 32 RHS it takes 3.65 ms per term against
 * 3.8 ms for qsim_cheb_spmm_step, for a second CSR and one more vector.
 * (new; the reference runs one SpMV per term: spmv_sm120.cu:108-142)      */
int qsim_cheb_spmm_step_split(void* stream, const void* rp_lo, const void* ci_lo, const void* val_lo,
                              const void* rp_hi, const void* ci_hi, const void* val_hi, int64_t n_rows,
                              int split_bit, int n_rhs, double scale, const void* t_cur,
                              void* t_prev_next, void* acc, void* comp, void* w1, double c_re,
                              double c_im, int first);
/* qsim_cheb_spmm_step (QSIM_CHEB_ACC_FP64, one right-hand side) without a
 * CSR: H is the Pauli sum in the grouped tables of qsim_pauli_csr_count
 * (device pointers), n_rows = 2^n_qubits.  Row i's entries are the group
 * values of qsim_pauli_csr_fill (FP64 in term order, rounded to complex64,
 * exact zeros skipped) at columns i ^ xm[g], summed in group order in FP64;
 * the result equals the CSR term's up to the order of that FP64 sum.  Reads
 * no index or value stream, only the state vectors.  gconst (device, 2
 * doubles per group, or NULL) holds each group's complex64-rounded entry
 * when no term of the group has a Y|Z mask (the entry is then the same in
 * every row), else NaN in its real part.
 * (new; the reference builds the CSR and runs one SpMV per term:
 *  blackwell_backend/hamiltonian.py:44-69, bwk/csrc/spmv_sm120.cu:108-142) */
int qsim_cheb_pauli_step(void* stream, int n_qubits, int n_groups, const void* xm, const void* toff,
                         const void* sm, const void* coef, const void* gconst, double scale,
                         const void* t_cur, void* t_prev_next, void* acc, void* comp, double c_re,
                         double c_im, int first);
/* acc = c * t, comp = 0 (first Chebyshev term), n elements. (new)        */
int qsim_cheb_init_acc(void* stream, void* acc, void* comp, const void* t, int64_t n,
                       double c_re, double c_im);
/* out = acc + comp (fold the Kahan compensation), n elements. (new)      */
int qsim_cheb_finish(void* stream, void* out, const void* acc, const void* comp, int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* QSIM_GFX950_H */
