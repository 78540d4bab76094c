/*
 * statecraft_hip.h -- C-ABI of the MI355X (gfx950) statevector engine.
 *
 * This is the drop-in boundary for StateCraft's dense `statevector` engine.
 * A host engine (the Java FFM adapter `HipStatevectorEngine`, the C++ mirror in
 * statecraft_amd/engine, or the Python ctypes mirror) lowers a QuantumCircuit
 * into a flat array of `sc_op` records plus one array of complex matrix
 * entries, and calls these functions.  Every function is synchronous from the
 * caller's point of view, takes plain pointers and sizes, and never retains a
 * host pointer past the call.
 *
 * Reference interfaces replaced (paths relative to the StateCraft tree):
 *   SimulatorEngine.simulate           core/src/main/java/com/omaarr90/statecraft/core/engine/SimulatorEngine.java:3-9
 *   StatevectorEngine.simulate op loop engines/src/main/java/com/omaarr90/statecraft/engines/statevector/StatevectorEngine.java:63-167
 *   StatevectorOps.apply*              engines/.../statevector/StatevectorOps.java:19-747
 *
 * Error convention.  Every function returns an int status.  The reference
 * throws Java exceptions; the host maps the status to the same class:
 *   SC_E_NULL        -> NullPointerException
 *   SC_E_ARG         -> IllegalArgumentException
 *   SC_E_UNSUPPORTED -> UnsupportedOperationException
 *   SC_E_STATE       -> IllegalStateException
 *   SC_E_OOM / SC_E_HIP -> UnsupportedOperationException / IllegalStateException
 * sc_last_error() returns the message of the last failure on the calling thread.
 *
 * State layout in HBM: 2^n complex128 amplitudes, interleaved (re, im), index
 * bit q = qubit q (little-endian), exactly the reference's double[] layout
 * (docs/core/statevector-layout.md:13-31), addressed with 64-bit indices so a
 * state may hold up to 34 qubits (256 GiB).
 */
#ifndef STATECRAFT_HIP_H
#define STATECRAFT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define SC_OK 0
#define SC_E_NULL 1
#define SC_E_ARG 2
#define SC_E_UNSUPPORTED 3
#define SC_E_STATE 4
#define SC_E_OOM 5
#define SC_E_HIP 6

/* ---- op kinds: one per sealed QuantumCircuit.Operation record ---------------
 * QuantumCircuit.java:362-574.  Field meaning per kind:
 *   SC_OP_SINGLE  SingleGateOperation(gate, qubit)        q0=qubit, mat -> 4 complex (row-major 2x2)
 *   SC_OP_CNOT    CnotOperation(control, target)          q0=control, q1=target
 *   SC_OP_TWO     TwoQubitGateOperation(matrix, first, second)
 *                                                         q0=first, q1=second, mat -> 16 complex;
 *                                                         local index l = 2*bit(first) + bit(second)
 *   SC_OP_DIAG2   TwoQubitDiagonalOperation(diag, first, second)
 *                                                         q0=first, q1=second, mat -> 4 complex
 *   SC_OP_SWAP    SwapOperation(first, second)            q0=first, q1=second
 *   SC_OP_MCU     MultiControlOperation(gate, target, controls)
 *                                                         q0=target, controls=bit mask, mat -> 4 complex
 *   SC_OP_MEASURE MeasureOperation(qubits)                controls=bit mask of measured qubits
 * Gradient metadata (no reference counterpart): `param` is the index of the
 * circuit parameter the gate depends on (-1 for none) and the gate equals
 * exp(-i * theta * coeff * G) for the Pauli generator `gen` on its qubits, so
 * dU/dtheta = -i * coeff * G * U.
 */
#define SC_OP_SINGLE 1
#define SC_OP_CNOT 2
#define SC_OP_TWO 3
#define SC_OP_DIAG2 4
#define SC_OP_SWAP 5
#define SC_OP_MCU 6
#define SC_OP_MEASURE 7

#define SC_GEN_NONE 0
#define SC_GEN_X 1  /* X on q0 */
#define SC_GEN_Y 2  /* Y on q0 */
#define SC_GEN_Z 3  /* Z on q0 */
#define SC_GEN_ZZ 4 /* Z(q0) Z(q1) */

typedef struct sc_op {
    int32_t kind;      /* SC_OP_* */
    int32_t q0;        /* first qubit field of the reference record */
    int32_t q1;        /* second qubit field, -1 when absent */
    int32_t mat;       /* offset, in complex entries, into the `mats` array; -1 when unused */
    uint64_t controls; /* SC_OP_MCU control mask / SC_OP_MEASURE qubit mask */
    int32_t param;     /* gradient parameter index, -1 for none */
    int32_t gen;       /* SC_GEN_* */
    double coeff;      /* generator coefficient */
} sc_op;

/* ---- execution options (sc_ctx_set_option) -------------------------------- */
#define SC_OPT_EXEC_MODE 1      /* SC_EXEC_*; default SC_EXEC_SWEEP */
#define SC_OPT_MAX_FUSED 2      /* max qubits of a dense fused block (0..5; sweeps fuse at most 4, 0 = no sweep
                                   fusion, fused mode at least 1), default 4 */
#define SC_OPT_PROFILE 3        /* 1: time every sweep/gate launch with HIP events */
#define SC_OPT_LOW_QUBITS 4     /* qubits always resident in a sweep chunk for coalescing (0..8); -1: chosen per
                                   circuit structure with SC_OPT_CHUNK_BITS (4 for the compiled sweep kernel),
                                   default -1 */
#define SC_OPT_TEST_DRIFT 5     /* test hook: adds value*1e-12 to psi[0] at every adjoint checkpoint, default 0 */
#define SC_OPT_REMAP 6          /* 1: sweep passes relabel qubits in place to keep coalesced low positions useful, default 0 */
#define SC_OPT_SWEEP_BLOCKS_PER_CU 7 /* resident sweep workgroups per CU (0: as many as fit), default 0 */
#define SC_OPT_JIT 8            /* 1: sweep passes run as kernels compiled at run time (hipRTC) with the pass
                                   constant-folded; compiled once per circuit structure and cached; 2: the same
                                   with the next chunk's HBM loads issued before the current chunk's stages;
                                   0: the precompiled sweep kernel; default 2 */
#define SC_OPT_SWEEP_PASS_CAP 9 /* most estimated FP64 operations per amplitude one sweep pass takes on (0: no
                                   limit; -1: the default of the sweep kernel kind), default -1 */
#define SC_OPT_CHUNK_BITS 10    /* local qubits per sweep chunk with runtime-compiled passes: 12 (64 KiB of LDS, two
                                   256-thread workgroups per CU) or 13 (128 KiB, one 512-thread workgroup per CU);
                                   0: chosen per circuit structure (with SC_OPT_LOW_QUBITS -1 also the low qubits):
                                   the circuit is planned under 12/4, 13/4 and 13/3 chunk/low qubits and the plan
                                   with the smallest pass time predicted by a model fitted to measured MI355X pass
                                   times is kept and cached (sc_last_plan_geometry); default 0.  The compiled
                                   sweep kernel always uses 12. */
#define SC_OPT_DIRECT_IO 11     /* 1: a runtime-compiled pass whose first (last) stage keeps the chunk's low positions
                                   on lane bits reads HBM straight into (writes it straight from) that stage's
                                   registers, skipping an LDS round trip; 0: always through LDS; default 1 */
#define SC_OPT_MFMA 12          /* fused mode: dense blocks run as complex GEMMs on the FP64 matrix cores
                                   (v_mfma_f64_16x16x4_f64) instead of FP64 VALU multiply-adds; 0: never; 1: 5-qubit
                                   blocks (2.5x faster than the VALU kernel on MI355X); 2: 4- and 5-qubit blocks;
                                   default 1 */
#define SC_OPT_PASS_SCHED 13    /* wave priority in runtime-compiled sweep passes (s_setprio): 0 none; 1 / 2: the
                                   second half / the odd workgroups of the grid raised; 3: raised during each
                                   stage's arithmetic; 4: during its LDS reads and arithmetic; 5: reads, arithmetic
                                   and LDS writes; 6: arithmetic and writes; default 4 (4-5 % faster passes than 0
                                   on the 30-qubit layered circuit) */
#define SC_OPT_PERMLANE 14      /* 1: runtime-compiled passes enter a stage whose register qubits differ from the
                                   previous stage's in at most two (found on lane bits 4 / 5) by exchanging them
                                   inside each wavefront (v_permlane32_swap / v_permlane16_swap) instead of through
                                   LDS; default 0 (measured neutral: the transitions it can take are the wave-local
                                   ones, whose LDS traffic already overlaps the other workgroup's arithmetic) */
#define SC_OPT_FUSED_TERMS 15   /* 1: sc_run_jobs sums a job's Z-string terms in its last sweep pass as the states are
                                   stored (one read of the batch fewer); results agree with the separate calls to
                                   rounding instead of bit for bit; 0: a term kernel after the circuit; default 1 */
#define SC_OPT_MFMA_TILE 16     /* tile shape of the matrix-core dense blocks: -1 (default) chosen by the library's
                                   autotuner, which times every shape on a scratch state the first time a context
                                   runs a block size; 0..5 forces one: 16-group tiles per wavefront step 1 (0-2) or 2
                                   (3-5), resident workgroups per CU 16, 8, 32 (0, 1, 2) and (3, 4, 5) */
#define SC_OPT_MFMA_PICK4 17    /* read-only: the tile shape the autotuner chose for 4-qubit blocks (-1: not tuned) */
#define SC_OPT_MFMA_PICK5 18    /* read-only: the tile shape the autotuner chose for 5-qubit blocks (-1: not tuned) */
#define SC_EXEC_DIRECT 0 /* one kernel launch per reference op (StatevectorEngine.java:90-151 order) */
#define SC_EXEC_FUSED 1  /* gate fusion into dense 2^k blocks, one HBM sweep per block */
#define SC_EXEC_SWEEP 2  /* fusion + LDS-blocked multi-gate sweeps (default) */

typedef struct sc_ctx sc_ctx;
typedef struct sc_state sc_state;

/* ---- library / context ---------------------------------------------------- */
const char *sc_last_error(void);
int sc_version(void);
/* StatevectorEngine(StatevectorExecutionConfig) StatevectorEngine.java:46-55: one context = one device + one stream. */
int sc_ctx_create(int device, sc_ctx **out);
int sc_ctx_destroy(sc_ctx *ctx);
int sc_ctx_set_option(sc_ctx *ctx, int option, int64_t value);
int sc_ctx_get_option(sc_ctx *ctx, int option, int64_t *value);
int sc_ctx_synchronize(sc_ctx *ctx);
/* Device facts for reporting: HBM bytes free/total, CU count. */
int sc_ctx_device_info(sc_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes, int *cu_count);

/* ---- state lifetime / init ------------------------------------------------ */
/* new double[2^(n+1)] StatevectorEngine.java:68-69; batch>1 allocates `batch` independent states. */
int sc_state_create(sc_ctx *ctx, int num_qubits, int batch, sc_state **out);
int sc_state_destroy(sc_state *st);
int sc_state_qubits(const sc_state *st, int *num_qubits, int *batch);
/* resetZeroState / resetBasisState StatevectorEngine.java:345-357 (index = OR of 1<<qubit), every batch element. */
int sc_state_set_basis(sc_state *st, uint64_t index);
/* populateFromState StatevectorEngine.java:340-343; n_complex = 2^n * batch (batch-major). */
int sc_state_upload(sc_state *st, const double *interleaved, uint64_t n_complex);
/* toStateVector StatevectorEngine.java:367-369. */
int sc_state_download(sc_state *st, double *interleaved, uint64_t n_complex);
/* Device-to-device copy of all amplitudes (same shape). */
int sc_state_copy(sc_state *dst, const sc_state *src);
/* Reads amplitudes [offset, offset + count) of batch element `elem` (StateVector.amplitude access without
 * copying a whole 2^n state to the host). */
int sc_state_read_range(sc_state *st, int elem, uint64_t offset, uint64_t count, double *interleaved);

/* ---- gate application ----------------------------------------------------- */
/* The op loop of StatevectorEngine.simulate :90-151 for every unitary op kind;
 * validates each op like Operation.validateTargets (QuantumCircuit.java:362-574);
 * a unitary op after SC_OP_MEASURE fails with SC_E_UNSUPPORTED (:92-95).
 * `mats` holds interleaved complex entries referenced by sc_op.mat.  With
 * batch > 1 the same circuit is applied to every batch element. */
int sc_apply_ops(sc_state *st, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex);
/* Batched circuits with per-element matrices: element b uses mats + b*mat_stride (complex entries). */
int sc_apply_ops_batched(sc_state *st, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t mat_stride);

/* ---- measurement ---------------------------------------------------------- */
/* computeMeasurementProbabilities StatevectorEngine.java:274-296: out[o] for o in [0, 2^m), outcome bit k <-
 * qubit measured[k]; `measured` sorted ascending (MeasurementInstruction.java:64-81); batch element `elem`. */
int sc_probabilities(sc_state *st, int elem, const int32_t *measured, int m, double *out);
/* performMeasurement :229-257 = normalize :298-312 + cumulative :314-323 + sampleOutcome :325-338 for each
 * uniform u (SplittableRandom.nextDouble() values drawn by the host in shot order).  outcomes[s] in [0, 2^m).
 * The marginal probabilities, the normalising sum and the cumulative array are the reference's sequential
 * sums bit for bit (computed segment-parallel, exactly), so the outcomes equal the reference's for the same
 * state and uniforms.  boundary_gap (may be NULL) receives min over shots of |u - cum[outcome]|,
 * |u - cum[outcome-1]|.  A state holding NaN or infinite amplitudes fails with SC_E_STATE. */
int sc_sample(sc_state *st, int elem, const int32_t *measured, int m, const double *uniforms, int64_t shots,
              int64_t *outcomes, double *boundary_gap);
/* Statistics of the last sc_sample on this context: out[0..2] segments of 4096 cumulative values summed as
 * exact integers / serially / all zero, out[3] gather round trips, out[4] 1 when normalize rescaled. */
int sc_last_sample_stats(sc_ctx *ctx, uint64_t out[5]);
/* Sum of |a_i|^2 (ComplexArrays.norm2Sq, ComplexArrays.java:74-143) for batch element `elem`. */
int sc_norm_sq(sc_state *st, int elem, double *out);
/* Probability checksum: out[0] = sum_i p_i, out[1] = sum_i p_i * w(i) with w(i) = splitmix64(i)*2^-64. */
int sc_probability_checksum(sc_state *st, int elem, double out[2]);
/* Largest |psi_i - r_i| over the 2^n amplitudes of batch element `elem`, where r is a two-level reference state:
 * r_index = a and r_i = b elsewhere (complex, interleaved).  A whole-state comparison without a host copy, e.g.
 * against |0...0> after a circuit and its inverse, or against a closed form such as Grover's
 * a|m> + b sum_{x != m} |x>.  No reference counterpart (StateVector comparisons copy to the host). */
int sc_state_max_deviation(sc_state *st, int elem, uint64_t index, const double a[2], const double b[2],
                           double *out);

/* ---- noise (Monte-Carlo Kraus unravelling) -------------------------------- */
/* computeBranchWeight StatevectorEngine.java:511-530 for k single-qubit Kraus operators (4 complex each) on
 * `target`, all in one sweep: weights[j] = || K_j psi ||^2. */
int sc_kraus_weights(sc_state *st, int elem, int target, const double *kraus, int k, double *weights);
/* applySingleQubitKraus + renormalizeState :426-439, :590-600: apply K (4 complex) on target, then rescale by
 * 1/sqrt(normSq) when |normSq-1| > 1e-12.  norm_sq_out (may be NULL) receives normSq. */
int sc_apply_kraus(sc_state *st, int elem, int target, const double *kraus_op, double *norm_sq_out);
/* One channel application of a noisy run: after op `after_op`, a channel of `n_kraus` single-qubit Kraus
 * operators (4 complex each, row-major, from complex offset `kraus` of the kraus array) acts on `target`. */
typedef struct sc_noise_event {
    uint64_t after_op;
    int32_t target;
    int32_t n_kraus; /* 1..4 */
    uint64_t kraus;
} sc_noise_event;
/* The first failing channel of a run (event -1: none).  kind 1: weight `branch` is negative or not finite;
 * 2: the weights' total is <= 1e-15 (MIN_KRAUS_WEIGHT) or not finite; 3: the chosen branch's 1/sqrt(weight)
 * overflows (the state would collapse). */
typedef struct sc_noise_failure {
    int64_t event;
    int32_t elem;
    int32_t kind;
    int32_t branch;
    int32_t pad;
    double weights[4];
    double norm_sq;
} sc_noise_failure;
/* A noisy circuit in one call: the ops in order, and after op events[e].after_op (events in circuit order) the
 * channel of event e with the reference's unravelling (applyNoiseAfterOperation / applyErrorChannel
 * StatevectorEngine.java:371-407, sampleKrausOperator :463-491, computeBranchWeight :511-530,
 * sampleWeightedIndex :579-588, renormalizeState :590-600), the branch chosen on the device from the live
 * weights: r = uniforms[b * n_events + e] * total (SplittableRandom.nextDouble(total) from the nextDouble()
 * value the caller drew, one per event in circuit order), the chosen K_j applied divided by sqrt(w_j) (the
 * norm the reference recomputes) when |w_j - 1| > 1e-12.  Every batch element runs its own branches.  Replaces
 * the host loop of sc_kraus_weights / sc_apply_kraus calls per channel: nothing returns to the host until the
 * circuit ends.  A failing channel (the reference's IllegalStateException) gives SC_E_STATE and fills
 * `failure` (may be NULL); the state is then unspecified. */
int sc_run_noisy(sc_state *st, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex,
                 const sc_noise_event *events, uint64_t n_events, const double *kraus, uint64_t n_kraus_complex,
                 const double *uniforms, sc_noise_failure *failure);
/* Batched trajectories: state `st` holds `batch` copies of one circuit's state; after every op i with
 * noise_after[i] != 0, a depolarising channel of probability p acts on each of the op's qubits
 * (q0, then q1) with the reference's unravelling (sampleKrausOperator :463-491 + sampleWeightedIndex
 * :579-588 + renormalizeState :590-600).  uniforms: [batch][n_draws] SplittableRandom.nextDouble() values,
 * one per scheduled (op, qubit) in circuit order; n_draws must equal that count. */
int sc_run_depolarizing_trajectories(sc_state *st, const sc_op *ops, uint64_t n_ops, const double *mats,
                                     uint64_t n_mat_complex, const uint8_t *noise_after, double p,
                                     const double *uniforms, uint64_t n_draws);
/* sc_run_depolarizing_trajectories, then out[b * n_terms + t] = <Z^zmask[t]>_b of every trajectory (no
 * reference counterpart: the trajectory-averaged observables of a noisy batch).  With runtime-compiled passes
 * and SC_OPT_FUSED_TERMS the circuit's last pass sums the terms as it stores the states; otherwise the term
 * kernel reads the batch afterwards. */
int sc_run_depolarizing_trajectories_z(sc_state *st, const sc_op *ops, uint64_t n_ops, const double *mats,
                                       uint64_t n_mat_complex, const uint8_t *noise_after, double p,
                                       const double *uniforms, uint64_t n_draws, const uint64_t *zmask, int n_terms,
                                       double *out);

/* ---- observables and gradients (no reference counterpart) ----------------- */
/* Pauli-sum observable H = sum_t coeff[t] * P_t with P_t given by an x-mask and z-mask per term (Y = X and Z
 * on the same qubit, with the i factor accounted).  out[b] = <psi_b|H|psi_b> for every batch element. */
int sc_expectation_pauli(sc_state *st, const uint64_t *xmask, const uint64_t *zmask, const double *coeff,
                         int n_terms, double *out);
/* Every term on its own: out[b * n_terms + t] = <psi_b|P_t|psi_b> (real: each P_t is the Hermitian Pauli
 * string of its masks).  Terms sharing an x-mask are evaluated together in one read of the state (up to
 * 16 per read), so a diagonal observable of many terms costs one pass per 16 terms. */
int sc_expectation_pauli_terms(sc_state *st, const uint64_t *xmask, const uint64_t *zmask, int n_terms,
                               double *out);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): sc_expectation_pauli_terms with its partial
 * sums in a fixed 32 MiB device array that the library holds on every GPU it runs on until the process ends,
 * instead of the context's growable scratch.  Calls no hipMalloc or hipFree.  The call fails with SC_E_ARG
 * when n_terms * batch exceeds 2^21 (the partial sums of one term row per element no longer fit); it uses
 * fewer partial sums per row than sc_expectation_pauli_terms when the default count does not fit, which can
 * change the last bits of a result.  All calls on one GPU share its array: calls in one process take it in
 * turn, each overwrites what the previous one left, and every value a call reads it wrote first. */
int sc_expectation_pauli_terms_noalloc(sc_state *st, const uint64_t *xmask, const uint64_t *zmask, int n_terms,
                                       double *out);
/* Adjoint gradient of E(theta) = <psi(theta)|H|psi(theta)>.  On entry `st` holds the initial states; the
 * call runs the circuit forward (leaving psi_b = U(theta_b)|init_b> in `st`), then sweeps the bra
 * lambda = H psi and the ket psi backwards through U^dagger with one inner-product reduction per
 * parameterised op.  mats holds n_mat_complex entries, or batch * n_mat_complex entries (element b at
 * b * n_mat_complex) when mats_per_element != 0.  grads[b*n_params + p] = dE/dtheta_p; energies (may be
 * NULL) receives E_b.  checkpoint_every > 0 stores the forward state every that many ops and, during the
 * reverse sweep, re-anchors psi from a checkpoint when max-abs drift exceeds 1e-10.  Parameterised ops:
 * SC_OP_SINGLE / SC_OP_MCU with gen X, Y or Z on q0, SC_OP_DIAG2 with gen ZZ. */
int sc_adjoint_gradient(sc_state *st, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex,
                        int mats_per_element, const uint64_t *xmask, const uint64_t *zmask, const double *coeff,
                        int n_terms, int n_params, int checkpoint_every, double *grads, double *energies);
/* ---- batched jobs in one graph launch (no reference counterpart) ---------- */
/* One job: every element of `state` starts from the basis state |basis>, runs the circuit (ops; element b's
 * matrices at mats + 2 * b * mat_stride doubles, as sc_apply_ops_batched), and out[b * n_terms + t] receives
 * <psi_b|P_t|psi_b> (as sc_expectation_pauli_terms). */
typedef struct sc_job {
    sc_state *state;
    uint64_t basis;
    const sc_op *ops;
    uint64_t n_ops;
    const double *mats;
    uint64_t mat_stride;
    const uint64_t *xmask;
    const uint64_t *zmask;
    int32_t n_terms;
    int32_t reserved;
    double *out;
} sc_job;
/* Runs the jobs in order with results identical to the separate calls (with SC_OPT_FUSED_TERMS, a job whose
 * terms are all Z strings sums them in its last pass: equal to rounding).  With runtime-compiled sweep passes
 * (SC_OPT_JIT > 0, every state of at least 12 qubits) the first call of a job structure (shapes, circuit
 * structure and matrix classes, bases, terms) records pool uploads, every pass launch, the term kernels and
 * the result copies as one HIP graph; every later call rebuilds only the matrix pools and launches that graph
 * once (*graph_launched = 1).  Otherwise the jobs run one after another (*graph_launched = 0).  The graph
 * keeps the states' device addresses: destroying a state drops nothing but makes its graph unusable, and a
 * new graph is recorded for new states.  The context keeps its 8 most recently used graphs. */
int sc_run_jobs(sc_ctx *ctx, const sc_job *jobs, int n_jobs, int *graph_launched);
/* Statistics of the last sc_adjoint_gradient on this context: re-anchors performed, max drift seen. */
int sc_adjoint_stats(sc_ctx *ctx, uint64_t *reanchors, double *max_drift);

/* ---- tracing -------------------------------------------------------------- */
/* With SC_OPT_PROFILE=1 every kernel launch is bracketed by HIP events on the context's stream.  Query per
 * kernel family: number of launches, total milliseconds, and the algorithmic HBM bytes credited. */
#define SC_KFAM_GATE 0     /* per-op kernels */
#define SC_KFAM_FUSED 1    /* dense fused-block kernel */
#define SC_KFAM_SWEEP 2    /* LDS-blocked multi-gate sweep kernel */
#define SC_KFAM_MEASURE 3  /* probability / sampling kernels */
#define SC_KFAM_NOISE 4    /* Kraus weight / renormalise kernels */
#define SC_KFAM_ADJOINT 5  /* adjoint backward kernels */
#define SC_KFAM_COPY 6     /* sc_state_copy's plain copy kernel (bandwidth reference) */
#define SC_KFAM_COUNT 7
int sc_profile_reset(sc_ctx *ctx);
int sc_profile_read(sc_ctx *ctx, int family, uint64_t *launches, double *total_ms, double *bytes);
/* Plan statistics of the last sc_apply_ops: number of full-state sweeps executed and ops credited. */
int sc_last_plan_stats(sc_ctx *ctx, uint64_t *sweeps, uint64_t *kernel_launches, uint64_t *fused_blocks);
/* Geometry of the last sweep plan the context used (planned or from its cache): local qubits per chunk, low
 * qubits, how many geometries were planned to choose it (1 when both options fix it) and its predicted pass
 * time in ns per amplitude summed over its passes.  Any output may be NULL.  No counterpart in the reference:
 * it reports the SC_OPT_CHUNK_BITS / SC_OPT_LOW_QUBITS autotuner's choice. */
int sc_last_plan_geometry(sc_ctx *ctx, int *chunk_bits, int *low_qubits, int *candidates, double *ns_per_amp);
/* Host-only plan inspection (no device needed): builds the SC_EXEC_SWEEP plan for an n-qubit circuit
 * (n >= 12) as a context with the given low-qubit, remap, jit, pass-cap and chunk-bit (12, or 13 with jit)
 * options would (jit 2: also with SC_OPT_PERMLANE 1), and reports its passes,
 * stages, fused 16x16 blocks and FP64 operations per amplitude.  When state_inout is non-NULL (2^n
 * interleaved complex values) the plan is also executed on the host with the sweep kernel's exact data
 * movement, so plans can be validated without a GPU. */
int sc_plan_sweeps_host(int num_qubits, const sc_op *ops, uint64_t n_ops, const double *mats,
                        uint64_t n_mat_complex, int low_qubits, int remap, int jit, int pass_cap, int chunk_bits,
                        uint64_t *passes, uint64_t *stages, uint64_t *fused_blocks, double *fp64_ops_per_amp,
                        double *state_inout);
/* Host-only (no device needed): the product-state input of runtime-compiled passes from a basis state.  The
 * circuit's leading single-qubit ops (uncontrolled, before any other op touches their qubit) are taken out
 * of the plan; the first pass generates prod_q v_q[i_q], v_q those ops applied to bit q of `basis`, and the
 * rest runs as planned (as a context does when sc_apply_ops starts from sc_state_set_basis).  state_out
 * (2^n interleaved complex values) receives the result with the kernels' data movement; absorbed (may be
 * NULL) the number of ops taken out.  No reference counterpart: it validates that path without a GPU. */
int sc_plan_product_host(int num_qubits, const sc_op *ops, uint64_t n_ops, const double *mats,
                         uint64_t n_mat_complex, uint64_t basis, int low_qubits, int chunk_bits, int *absorbed,
                         double *state_out);
/* Host-only (no device needed): the geometry a context with these options picks for the circuit (chunk_bits 0
 * and / or low_qubits -1 leave it to the autotuner; see SC_OPT_CHUNK_BITS), as sc_last_plan_geometry reports. */
int sc_plan_geometry_host(int num_qubits, const sc_op *ops, uint64_t n_ops, const double *mats,
                          uint64_t n_mat_complex, int low_qubits, int remap, int jit, int chunk_bits,
                          int *chunk_bits_out, int *low_qubits_out, int *candidates, double *ns_per_amp);
/* Host-only (no device needed): per pass of the same plan (at most max_passes entries), its stages, the stage
 * exchanges that need a workgroup barrier, its FP64 operations per amplitude and its direct-HBM flags (bit 0:
 * loads into the first stage, bit 1: stores from the last; SC_OPT_DIRECT_IO); n_passes receives the count. */
int sc_plan_pass_stats_host(int num_qubits, const sc_op *ops, uint64_t n_ops, const double *mats,
                            uint64_t n_mat_complex, int low_qubits, int remap, int jit, int chunk_bits, int max_passes,
                            int32_t *stages, int32_t *barriers, double *fp64_ops_per_amp, int32_t *direct,
                            int *n_passes);
/* Host-only (no device needed): the reverse sweep of an adjoint gradient as sweep passes (the plan the GPU's
 * runtime-compiled adjoint passes run), executed on the host: psi_inout holds the circuit's output state,
 * lam_inout H applied to it (2^n interleaved complex values each, one batch element); on return they hold
 * the states uncomputed to the circuit's input and grads[n_params] the gradient. */
int sc_adjoint_plan_host(int num_qubits, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex,
                         int n_params, int low_qubits, double *psi_inout, double *lam_inout, double *grads);
/* Host-only (no device needed): the exact segment-parallel running sums sc_sample uses, on the CPU: cum[i]
 * = ((values[0] + values[1]) + ...) + values[i] summed left to right, as `cumulative` (StatevectorEngine.java
 * :314-323) does (without forcing the last entry).  stats (may be NULL) receives the sc_last_sample_stats
 * counts out[0..3]. */
int sc_cumulative_host(const double *values, uint64_t count, double *cum, uint64_t *stats);
/* Host-only (no device needed): sc_sample's exact flow on a host state of num_qubits <= 30 qubits (2^n
 * interleaved complex values); stats (may be NULL) receives the 5 sc_last_sample_stats counts. */
int sc_sample_host(int num_qubits, const double *state, const int32_t *measured, int m, const double *uniforms,
                   int64_t shots, int64_t *outcomes, double *boundary_gap, uint64_t *stats);
/* Host-only (no device needed): generates the runtime-compiled kernels of every pass of that plan (the
 * SC_OPT_JIT path, with chunk_bits local qubits per chunk; jit = 3: the adjoint passes of the circuit's
 * reverse sweep, always 12; jit = 4: as 2 for the product-state plan from a basis state) and compiles them with hipRTC on `threads` threads.  Reports the kernel count, the
 * compile wall time and the code object size; when dump_dir is non-NULL each module's source (.hip) and
 * code object (.co) are written there. */
int sc_jit_compile_host(int num_qubits, const sc_op *ops, uint64_t n_ops, const double *mats,
                        uint64_t n_mat_complex, int low_qubits, int remap, int jit, int chunk_bits, int threads,
                        uint64_t *kernels, double *seconds, uint64_t *code_bytes, const char *dump_dir);

#ifdef __cplusplus
}
#endif
#endif /* STATECRAFT_HIP_H */
