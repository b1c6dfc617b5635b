/*
 * desv.h -- C ABI of the MI355X-native dense statevector engine.
 *
 * Plain pointers and sizes only.  Every entry point returns an int status
 * (DESV_OK = 0); on failure a thread-local message is available from
 * desv_last_error().  Host buffers are owned by the caller; device memory is
 * owned by the handle.  Calls on one handle are serialised by the caller and
 * are asynchronous on the handle's HIP stream unless they return host data.
 *
 * Conventions (same as the reference engine): the state is a flat array of
 * 2^n interleaved (re, im) complex numbers, complex64 or complex128; qubit q
 * is bit (n-1-q) of the basis index (MSB-first).  An op table row is
 * [g_id, q1, q2, param] as float64 with the reference's gate ids:
 *   0 id, 1 h, 2 x, 3 y, 4 z, 5 s, 6 sdg, 7 t, 8 tdg, 9 rx, 10 ry, 11 rz,
 *   12 p/u1/phase, 13 sx, 14 gphase, 20 cx, 21 cz, 22 cp, 23 swap, 24 cy, 25 crz
 * (reference: dense_evolution/circuits/compiler.py:24-48, gates.py:53-63).
 * Pauli strings are passed as bit masks over basis-index bits:
 *   xmask = bits where the letter is X or Y, zmask = bits where it is Z or Y.
 */
#ifndef DESV_H
#define DESV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DESV_OK = 0,
    DESV_EINVAL = 1,      /* bad argument: maps to ValueError             */
    DESV_ENOMEM = 2,      /* device or host allocation failed: MemoryError */
    DESV_EDEVICE = 3,     /* HIP runtime error: RuntimeError               */
    DESV_EUNSUPPORTED = 4 /* valid request the library does not implement  */
};

enum { DESV_C64 = 0, DESV_C128 = 1 };

typedef struct desv_state *desv_handle;
typedef struct desv_prog *desv_program;
typedef struct desv_batch_s *desv_batch;

/* ---- library ---------------------------------------------------------- */

/* Message of the last failed call on this thread ("" if none). */
const char *desv_last_error(void);
/* ABI version number of this library. */
int desv_version(void);
/* Number of visible HIP devices. */
int desv_device_count(int *out);

/* ---- state handles ------------------------------------------------------
 * Replace DenseSVSimulator's state allocation and host transfers:
 * __init__/_reset_sv (statevector.py:86-107), set_initial_state/set_state
 * (:109-137), normalize (:141-147), get_statevector (:609-611). */

/* Allocate a 2^n state (1 <= n <= 34) of dtype DESV_C64/DESV_C128 on
 * `device`, initialised to |0...0>. */
int desv_create(int n_qubits, int dtype, int device, desv_handle *out);
int desv_destroy(desv_handle h);
/* Use an existing HIP stream (e.g. torch's current stream) for all work on
 * this handle; NULL restores the handle's own stream. */
int desv_set_stream(desv_handle h, void *hip_stream);
int desv_synchronize(desv_handle h);
/* Double buffering (on by default): single-gate streaming passes read the
 * state and write a second buffer of the same size, which then becomes the
 * state (the device address of the state changes; the handle stays valid).
 * This is the counterpart of the reference's donated jit call
 * (circuits/compiler.py:354, backends/statevector.py:467-471), which also
 * hands back a new buffer.  enable = 0 frees the second buffer and updates in
 * place; an allocation failure also falls back to in place. */
int desv_set_double_buffer(desv_handle h, int enable);
/* complex64 handles: enable = 1 applies the gate tables of desv_run_ops and
 * the matrices of desv_apply_matrix with compensated matrices (see
 * desv_compile_opts.compensated); a no-op for complex128.  No counterpart in
 * the reference (its complex64 path, use_float32, rounds each gate matrix
 * once: circuits/compiler.py:172-179). */
int desv_set_compensated(desv_handle h, int enable);
/* Device address of the state (2^n interleaved complex values).  Valid until
 * the next call that may swap the state buffer (a streaming gate pass with
 * double buffering) or destroys the handle; for zero-copy views such as
 * torch tensors.  No counterpart in the reference (JAX owns its buffers). */
int desv_state_device_ptr(desv_handle h, void **out);
/* |0...0> */
int desv_reset(desv_handle h);
/* Copy 2^n host amplitudes in the handle's dtype (no normalisation). */
int desv_set_state(desv_handle h, const void *host_amps);
int desv_get_state(desv_handle h, void *host_amps);
/* dst <- src (same n and device; a dtype change converts complex64 <->
 * complex128 on the device, e.g. to compare a complex64 run with a
 * complex128 one). */
int desv_copy(desv_handle dst, desv_handle src);
/* state <- (re + i im) * state */
int desv_scale(desv_handle h, double re, double im);

/* ---- reductions (FP64 accumulation, deterministic order) ----------------
 * Replace get_probabilities (statevector.py:599-607), the marginal of
 * measure (:318-334) and the vdot of statevector_fidelity
 * (utils/measurement.py:71-94). */
int desv_norm2(desv_handle h, double *out);
/* desv_norm2 for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops): sum |a_i|^2 in FP64 into *out_dev, a double in
 * device memory, asynchronously on the handle's stream.  No counterpart in
 * the reference.  Allocates nothing: the per-workgroup partials live in a
 * fixed 64 KiB device array (8192 partials, enough for every state size;
 * larger requests are DESV_EINVAL); the result is bit-identical to
 * desv_norm2's.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (other streams, other threads) race. */
int desv_norm2_device(desv_handle h, double *out_dev);
int desv_inner(desv_handle a, desv_handle b, double *re, double *im);
/* |a_i|^2 as float64 into a host buffer of 2^n doubles (raw, not clipped). */
int desv_probabilities(desv_handle h, double *out_host);
/* Marginal probabilities of one qubit (raw sums, not normalised). */
int desv_qubit_probs(desv_handle h, int qubit, double *p0, double *p1);
/* Shot sampling, replacing sample_counts' rng.choice(dim, n_shots, p=|a|^2 /
 * total) (utils/measurement.py:17-68): the caller draws the uniforms in
 * [0, 1) (numpy's choice draws rng.random(n_shots)); out[s] is the basis
 * index whose normalised cumulative probability first exceeds uniforms[s]
 * (searchsorted side='right').  The cumulative probabilities are exact
 * integer prefix sums of the probabilities quantised to 2^-62 (of the
 * total), so an amplitude of probability zero is never drawn; the device
 * reads the state once for chunk totals and once more for the chunks that
 * hold shots, without building a full CDF (states far from unit norm: a full
 * FP64 CDF).  DESV_EINVAL if the total probability is zero or not finite. */
int desv_sample(desv_handle h, const double *uniforms, int64_t n_shots, int64_t *out);
/* Zero the amplitudes whose `qubit` differs from `outcome`, scale the rest. */
int desv_collapse(desv_handle h, int qubit, int outcome, double scale);
/* In place 2x2 matrix m (row-major, interleaved re/im, 8 doubles) on `qubit`
 * in one streaming pass; if next_qubit >= 0 the same pass returns the FP64
 * populations (p0, p1) of next_qubit in the updated state (not normalised).
 * Serves the per-qubit Kraus steps of the noise trajectories
 * (noise/kraus_channels.py:104-235: amplitude damping draws against the next
 * qubit's |1> population, computed here without a separate pass). */
int desv_apply1q_probs(desv_handle h, int qubit, const double *m, int next_qubit, double *p0, double *p1);

/* One layer of single-qubit operators on k distinct qubits, applied in place
 * in one streaming pass: mats holds k row-major 2x2 complex matrices (8
 * doubles each, interleaved re/im), each with at most one nonzero entry per
 * row and its nonzero rows reading different columns -- Paulis, diagonal
 * matrices, the amplitude-damping jump diag/|0><1| forms and their products.
 * If pops is not null, the same pass returns the FP64 populations of the
 * result over the 2^n_pop basis states of pop_qubits (index bit j <->
 * pop_qubits[j]; n_pop <= 12; n_pop = 0: the squared norm); k = 0 only
 * reads the state.  DESV_EINVAL for any other matrix or a repeated qubit.
 * Serves the noise trajectories (noise/kraus_channels.py:104-240): the
 * per-qubit draws of a group of qubits only need the populations over that
 * group, so one pass applies a whole group's operators and returns the
 * populations the next group's draws see. */
int desv_apply_monomial_layer(desv_handle h, int k, const int *qubits, const double *mats, int n_pop,
                              const int *pop_qubits, double *pops);

/* ---- gate application ----------------------------------------------------
 * desv_run_ops replaces _compile_and_run_circuit_jit(sv, ops) and its
 * donated twin (compiler.py:318-355) as called by run_circuit_jit
 * (statevector.py:414-471): the whole table is compiled (fusion + cache
 * blocking) and applied in place.  Qubits are range-checked; an unknown gate
 * id or a 2-qubit row with q1 == q2 is DESV_EINVAL. */
int desv_run_ops(desv_handle h, const double *ops, int64_t n_ops);
/* Arbitrary k-qubit unitary (1 <= k <= 5), row-major 2^k x 2^k complex128
 * as interleaved doubles; qubits[0] is the most significant index bit.
 * Replaces apply_gate_1q / apply_gate_2q (statevector.py:167-203). */
int desv_apply_matrix(desv_handle h, const double *mat, const int *qubits, int k);

/* ---- compiled programs ---------------------------------------------------
 * A program is a gate list compiled once (host) and replayed many times.
 * Compilation options: */
typedef struct {
    int fuse_bits;  /* max qubits of a fused block: 1 (none), 2..5          */
    int tile_bits;  /* amplitudes per LDS tile = 2^tile_bits; 0 = default    */
    int low_bits;   /* low index bits always in a tile (coalescing); 0 = dflt*/
    int per_gate;   /* 1: one HBM pass per fused block (no cache blocking)   */
    int use_mfma;   /* 1: 3-5-qubit blocks run as MFMA GEMMs (complex64 only:
                       DESV_EINVAL for complex128, whose blocks run on VALU) */
    int dtype;      /* DESV_C64 / DESV_C128: matrix precision of the upload */
    int compensated;/* 1 (complex64): every matrix is uploaded as an FP32 pair
                       hi + lo (lo = the rounding error of hi) and applied lo
                       part first, so the FP32 rounding of a matrix does not
                       act coherently when the same gates repeat (long Trotter
                       programs); about twice the gate arithmetic.  No
                       counterpart in the reference, whose complex64 path
                       rounds each gate matrix once (circuits/compiler.py:
                       172-179).                                              */
    int reg_groups; /* 1 (complex64 with use_mfma): stages of 2^12-amplitude
                       tiles whose blocks are all uncontrolled dense/diagonal
                       blocks of at most 3 qubits keep the tile in registers
                       between LDS round trips (groups of blocks, lane and
                       register bits exchanged in registers, every block on
                       MFMA); other stages are unchanged.  2: eligible stages
                       of 2^11- or 2^12-amplitude tiles (tile_bits 11 or 12)
                       hold one whole tile per wave in registers from its HBM
                       load to its HBM store (the tile never passes through LDS).  Same
                       results as the LDS path to FP32 rounding.  No
                       counterpart in the reference.                          */
    int plan_search;/* randomised planner passes tried beside the greedy one
                       (the cheapest plan by estimated time is kept; fixed
                       seeds, so a program always gets the same plan): -1 =
                       default (reg_groups 2: 64 passes for programs of up
                       to 512 gates, fewer for longer ones, at least 4),
                       0 = greedy only.  No counterpart in the reference.   */
} desv_compile_opts;

/* Fill `opts` with the defaults for a dtype. */
int desv_default_opts(int dtype, desv_compile_opts *opts);
int desv_program_create(int n_qubits, desv_program *out);
int desv_program_destroy(desv_program p);
int desv_program_add_ops(desv_program p, const double *ops, int64_t n_ops);
int desv_program_add_matrix(desv_program p, const double *mat, const int *qubits, int k);
int desv_program_compile(desv_program p, const desv_compile_opts *opts);
/* gates = logical gates added, passes = HBM sweeps per run, blocks = device
 * ops after fusion, mfma_blocks = ops run on MFMA. */
int desv_program_stats(desv_program p, int64_t *gates, int64_t *passes,
                       int64_t *blocks, int64_t *mfma_blocks);
/* Register-group program of a compiled program (introspection, no device
 * needed; see desv_compile_opts.reg_groups): sizes, then per stage the offset
 * of its words (-1: the stage runs its blocks from LDS) and of its first
 * fragment block, the words and the fragment floats (256 per block: 64 lanes
 * x 4 stacked real A fragments; compensated programs hold hi then lo blocks).
 * No counterpart in the reference; used by the CPU replay of the layouts. */
int desv_program_reg_sizes(desv_program p, int compensated, int64_t *n_words, int64_t *n_frag_floats);
int desv_program_reg_dump(desv_program p, int compensated, int32_t *stage_word, int32_t *stage_frag,
                          int32_t *words, float *frags);
/* Relabelled tail of a wave-tile program (introspection, no device needed).
 * With reg_groups = 2 the compiler ends a program with a stage that writes
 * into the second state buffer with every qubit moved to its home bit (the
 * qubits outside the tile move with the tile addresses), instead of in-place
 * restore passes; the in-place form of that tail is kept and runs when the
 * handle has no second buffer (desv_set_double_buffer(h, 0)).  Outputs:
 * fb_from = first stage the in-place tail replaces (-1: none), n_fb = its
 * stages; per stage of the program store_masks[s] (0: in place, else the bits
 * its tile is written to) and omaps[64 s ..] (destination bit of each bit
 * outside the tile, lowest first, -1 padded); per in-place tail stage its
 * tile mask, (op_begin, op_end, sperm[16]) as desv_program_plan_dump, and
 * the offsets of its words and first fragment block in the arrays of
 * desv_program_reg_dump.  Any pointer may be null.  No counterpart in the
 * reference; used by the CPU replay. */
int desv_program_tail_dump(desv_program p, int compensated, int64_t *fb_from, int64_t *n_fb, uint64_t *store_masks,
                           int32_t *omaps, uint64_t *fb_tile_masks, int64_t *fb_stage_ops, int32_t *fb_stage_word,
                           int32_t *fb_stage_frag);
int desv_run_program(desv_handle h, desv_program p);
/* Same as desv_run_program, with a HIP event recorded on the handle's stream
 * before every stage and after the last; waits for the run and writes each
 * stage's duration (milliseconds) to stage_ms[0 .. stages).  No counterpart
 * in the reference (its jit call has no per-step timing); used to measure
 * each stage kernel against its roofline. */
int desv_run_program_timed(desv_handle h, desv_program p, float *stage_ms);
/* Introspection of a compiled plan (no device needed): sizes, then per stage
 * the tile bit mask and 18 int64 fields (op_begin, op_end, then the 16-entry
 * write-back permutation sperm: store-order tile bit j takes tile-local bit
 * sperm[j]; entries past the tile width are -1); per op 11 int64 fields
 * (kind, k, mat, slot, pos[0..4], cloc, cglob) and the matrix pool as
 * interleaved complex128.  Used by the host-compiler tests. */
int desv_program_plan_sizes(desv_program p, int64_t *n_stages, int64_t *n_ops, int64_t *n_mats);
int desv_program_plan_dump(desv_program p, uint64_t *stage_tile_mask, int64_t *stage_ops,
                           int64_t *ops, double *mats);

/* ---- Pauli sums ----------------------------------------------------------
 * desv_pauli_expect replaces pauli_expectation / pauli_sum_expectation
 * (physics/observables.py:97-186): out[t] = <psi|P_t|psi> (real), FP64
 * accumulation, one state sweep per distinct xmask.
 * desv_pauli_matvec replaces pauli_sum_matvec(_jax) and
 * PauliSumOperator.__matmul__ (observables.py:250-308, :344-450):
 * out <- sum_t (coeff_re[t] + i coeff_im[t]) P_t in. */
int desv_pauli_expect(desv_handle h, const uint64_t *xmask, const uint64_t *zmask,
                      int64_t n_terms, double *out);
int desv_pauli_matvec(desv_handle out, desv_handle in, const uint64_t *xmask,
                      const uint64_t *zmask, const double *coeff_re,
                      const double *coeff_im, int64_t n_terms);
/* *out = sum_t coeff[t] <psi|P_t|psi> for real coefficients, i.e.
 * Re<psi|H|psi>, in one pass of the tiled Pauli-sum kernel with FP64 partial
 * sums (no per-term values, no output state); the fast path of
 * pauli_sum_expectation (physics/observables.py:152-186). */
int desv_pauli_sum_expect(desv_handle h, const uint64_t *xmask, const uint64_t *zmask,
                          const double *coeff, int64_t n_terms, double *out);
/* out <- H in for a dense row-major complex128 H (2^n x 2^n, n <= 14); the
 * `h_matrix @ sv` of energy_fn (solvers/autodiff.py:166). */
int desv_dense_matvec(desv_handle out, desv_handle in, const double *h_matrix);

/* ---- Krylov bases (Lanczos ground states) --------------------------------
 * The reference finds ground states with scipy's eigsh over a LinearOperator
 * whose matvec is pauli_sum_matvec (tools/dashboard/core/hamiltonians.py:
 * 512-577), one host round trip per product.  These entry points keep the
 * whole Krylov basis on the device: handles of one n, dtype and device, at
 * most 64 per call, FP64 accumulation.
 * out[2j], out[2j+1] = Re, Im <vs[j]|w> for j < m. */
int desv_inner_many(desv_handle w, const desv_handle *vs, int m, double *out);
/* outs[l] <- sum_j (coef[2(j p + l)] + i coef[2(j p + l) + 1]) ins[j] for
 * l < p, j < m, element by element: outputs may be among the inputs (in-place
 * Gram-Schmidt subtraction and Ritz rotation). */
int desv_combine(const desv_handle *outs, int p, const desv_handle *ins, int m, const double *coef);
/* Reproducible pseudo-random amplitudes in [-1/2, 1/2) + i[-1/2, 1/2) from a
 * 64-bit seed (a Lanczos start vector; not normalised). */
int desv_fill_random(desv_handle h, uint64_t seed);

/* ---- adjoint gradient ----------------------------------------------------
 * Energy and the full gradient of E(theta) = Re<psi|H|psi>, psi = U(ops)|h>,
 * in one forward + one reverse sweep.  Replaces jax.value_and_grad over
 * energy_fn (solvers/autodiff.py:92-169, caller test_autodiff.py:56).
 * `ops` are the patched rows (angles filled in); param_of_op[i] is the theta
 * index the row's angle came from, or -1.  Supported theta-carrying ids:
 * rx, ry, rz, p, cp, crz.  The handle keeps the final state psi.
 * H is a Pauli sum (coeff real) or, if xmask == NULL, the dense complex128
 * matrix passed in `dense_h`. */
int desv_adjoint(desv_handle h, const double *ops, int64_t n_ops,
                 const int64_t *param_of_op, int64_t n_params,
                 const uint64_t *xmask, const uint64_t *zmask, const double *coeff,
                 int64_t n_terms, const double *dense_h, double *energy, double *grad);

/* ---- heterogeneous batches ------------------------------------------------
 * Many independent (circuit, Hamiltonian, theta) instances of uneven qubit
 * counts evaluated together; generalises run_batch_jit
 * (statevector.py:496-584), which vmaps one circuit over parameter rows.
 * Instances are added once; desv_batch_run evaluates energy and adjoint
 * gradient of every instance, replaying a captured hipGraph after the first
 * call. */
int desv_batch_create(int dtype, int device, desv_batch *out);
int desv_batch_destroy(desv_batch b);
/* Returns the instance index in *index. */
int desv_batch_add(desv_batch b, int n_qubits, const double *ops, int64_t n_ops,
                   const int64_t *param_of_op, int64_t n_params, const uint64_t *xmask,
                   const uint64_t *zmask, const double *coeff, int64_t n_terms,
                   int64_t *index);
/* Gradients for the last energies-only desv_batch_run (grads == NULL), from
 * the psi copy and lambda that run left on the device: the
 * reverse sweep alone, so an energies-only run followed by this call costs one
 * full run.  DESV_EINVAL if parameters, initial states or the launch shape
 * changed since that run.  Serves autograd bindings that compute the energy
 * in forward and the gradient only when backward runs (reference: the
 * caller's jax.value_and_grad over energy_fn, solvers/autodiff.py:92-169). */
int desv_batch_run_reverse(desv_batch b, double *grads);
/* Device address of instance i's final state (valid until the batch is
 * rebuilt), for zero-copy views such as torch tensors. */
int desv_batch_state_device_ptr(desv_batch b, int64_t index, void **out);
/* Update the theta-carrying angles of instance i (n_params values). */
int desv_batch_set_params(desv_batch b, int64_t index, const double *theta);
/* Initial state of instance i (2^n amplitudes in the batch dtype), or NULL
 * for |0...0> (the default). */
int desv_batch_set_initial_state(desv_batch b, int64_t index, const void *host_amps);
/* Final state psi of instance i after the last run. */
int desv_batch_get_state(desv_batch b, int64_t index, void *host_amps);
/* Every instance's final state, concatenated in instance order (sum over i of
 * 2^n_i amplitudes in the batch dtype), in one device-to-host copy. */
int desv_batch_get_states(desv_batch b, void *host_amps);
/* enable = 1: the batch only applies each instance's circuit (instances may be
 * added with n_terms = 0; desv_batch_run takes NULL energies and gradients;
 * no completion stamps are written).  Serves run_batch_jit
 * (statevector.py:496-584): one row per instance, instances of equal gate
 * structure share one compiled plan, all rows advance together one stage
 * launch at a time. */
int desv_batch_set_forward_only(desv_batch b, int enable);
/* Launch shape of the next build: tile_bits (0 = default: 11 for complex64,
 * 10 for complex128, capped by LDS), fuse_bits (0 = 2), use_mfma (forward
 * 3-5-qubit blocks on MFMA, complex64), use_graph (capture + replay).  The
 * Python autotuner (desv.tune) times candidate shapes and sets the fastest. */
int desv_batch_configure(desv_batch b, int tile_bits, int fuse_bits, int use_mfma, int use_graph);
/* complex64 batches: enable = 1 applies every gate of the forward and
 * reverse sweeps with compensated matrices (see desv_compile_opts.
 * compensated), which keeps energies and gradients within 1e-6 relative of
 * the FP64 result; the next run rebuilds.  No counterpart in the reference. */
int desv_batch_set_compensated(desv_batch b, int enable);
/* Evaluate everything; energies[i], grads[offset_i ...] in instance order
 * (grads packed by instance, n_params_i each).  grads == NULL runs the
 * forward pass and energies only. */
int desv_batch_run(desv_batch b, double *energies, double *grads);
/* Wall time of the last desv_batch_run in microseconds (host clock). */
int desv_batch_last_time_us(desv_batch b, double *us);
/* Per-instance completion times of the last run, us[i] in microseconds from
 * the run's first kernel to the moment instance i's last result (energy or
 * gradient entry) was reduced, read from the device wall clock.  The sums of
 * each instance run right after its own last reverse stage, so these times
 * spread with the instances' circuit depths.  No counterpart in the
 * reference (its vmapped run_batch_jit, statevector.py:496-584, returns all
 * rows at once). */
int desv_batch_completion_us(desv_batch b, double *us);

#ifdef __cplusplus
}
#endif

#endif /* DESV_H */
