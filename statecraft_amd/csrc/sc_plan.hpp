// sc_plan.hpp -- host-side circuit analysis, gate fusion and sweep planning.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "sc_internal.hpp"

namespace sc {

// One reference op after validation, with its qubit roles.
struct PlanOp {
    int kind = 0;           // SC_OP_*
    int q0 = -1, q1 = -1;
    uint64_t controls = 0;  // SC_OP_MCU
    int mat = -1;           // complex offset into the caller's matrix array
    uint64_t targets = 0;   // qubits acted on non-diagonally
    uint64_t diag = 0;      // qubits acted on diagonally (controls, diagonal gates)
    bool is_x = false;      // gate is exactly Pauli X (pure exchange)
    bool is_diag = false;   // whole op is diagonal
    bool is_real = false;   // every matrix entry has a zero imaginary part (all batch elements)
    int src = -1;           // index into the caller's op array
    int gen = 0;            // kPlanGrad: SC_GEN_* of the probe; mat holds its gradient slot
    // Uncontrolled 2x2 with a unit-pivot form for every batch element: row r of the matrix has a real pivot
    // entry in column bit r of `unit` (both pivots equal, nonzero, and the other entries at most 2^40 times
    // larger), so M = p * M' with M' unit on its pivots.  -1: no such form.  unit_form: 0 complex, 1 real,
    // 2 purely imaginary off-pivot entries.
    int unit = -1;
    int unit_form = 0;
    // Absorbed into a product-state input (absorb_leading_ops): the op keeps its place in the plan, so the
    // passes and stages are those of the full circuit, but emits no register op.
    bool skip = false;
};

// Adjoint passes: a probe that accumulates <lambda| G |psi> for the Pauli generator G on q0 (and q1 for ZZ).
// It leaves both states unchanged; its qubits must be register bits of its stage.
constexpr int kPlanGrad = 100;

// Host threads a planning step may use (pool builds, matrix classification of large batches).
constexpr uint64_t kHostThreads = 16;
// Calls f(thread, begin, end) over [0, n) in ranges of `grain` handed out from a shared counter, on up to
// kHostThreads threads (the calling thread included; one thread when n <= grain).  `thread` < kHostThreads
// names the calling thread's slot, for per-thread partial results.
void parallel_ranges(uint64_t n, uint64_t grain, const std::function<void(unsigned, uint64_t, uint64_t)> &f);

// Validate ops against n qubits (QuantumCircuit.Operation.validateTargets, QuantumCircuit.java:362-574) and
// classify them.  `mats_all` lists every batch element's matrix array (one entry when shared) so diagonal /
// exchange classification holds for all of them.  Measurement records are dropped; a unitary after one fails
// with SC_E_UNSUPPORTED (StatevectorEngine.java:92-95).
int analyze_ops(int n, const sc_op *ops, uint64_t n_ops, const std::vector<const double *> &mats_all,
                uint64_t n_mat_complex, std::vector<PlanOp> &out);
// Z-string terms summed by a circuit's last runtime-compiled pass (SC_OPT_FUSED_TERMS): vals[b * terms + t] =
// <Z_t>_b of the output states.  `done` reports whether the pass did it (else the caller evaluates the terms).
struct ZTermsRequest {
    std::vector<uint64_t> z;
    double *vals = nullptr;
    bool done = false;
};
// Apply analysed ops to a state (one matrix array per batch element with per_elem, else elems[0] for all),
// in the context's execution mode (sc_capi.hip).
int apply_planned(State *st, const std::vector<PlanOp> &plan_ops, const std::vector<const double *> &elems,
                  bool per_elem, ZTermsRequest *terms = nullptr);
// One kernel per op (SC_EXEC_DIRECT), matrices passed by value: enqueued on the context's stream, no
// synchronisation.
int exec_direct(Context *ctx, State *st, const std::vector<PlanOp> &ops, const double *mats);

bool ops_conflict(const PlanOp &a, const PlanOp &b);

// ---- dense 2^k block fusion (SC_EXEC_FUSED) --------------------------------
struct FusedBlock {
    std::vector<int> qubits;  // ascending; local bit j <-> qubits[j]
    std::vector<int> ops;     // indices into the PlanOp list, program order
    bool standalone = false;  // op too wide for a block: launched with its own kernel
};
void plan_blocks(int n, const std::vector<PlanOp> &ops, int max_k, std::vector<FusedBlock> &blocks);
// Row-major 2^k x 2^k complex matrix of a block for one matrix array.
void block_matrix(const FusedBlock &b, const std::vector<PlanOp> &ops, const double *mats, std::vector<double> &out);

// ---- LDS-blocked sweeps (SC_EXEC_SWEEP) ------------------------------------
// A matrix-pool entry: how to build `count` complex values from the caller's matrices.
enum RecipeKind : int {
    RC_COPY = 0,      // `count` entries copied from mat
    RC_DIAG1 = 1,     // (g00, g00, g11, g11) of the diagonal 2x2 at mat
    RC_FUSED = 2,     // dense block of fused_ops on fused_bits
    RC_SWAP4 = 3,     // the 4x4 at mat with its two qubits exchanged
    RC_SWAPDIAG = 4,  // the 4-entry diagonal at mat with its two qubits exchanged
    RC_DIAG16 = 5,    // diagonal of the fused (diagonal) fused_ops over the 16 register indices
    RC_UNIT = 6,      // (m[0][1-c0] / m[0][c0], m[1][1-c1] / m[1][c1]) of the 2x2 at mat, pivot columns aux
    RC_SCALE = 7,     // one entry 1 (then scaled)
};

struct PoolRecipe {
    int kind;        // RecipeKind
    int mat;         // source offset in the caller's matrices (RC_COPY, RC_DIAG1, RC_SWAP4, RC_SWAPDIAG)
    int count;
    std::vector<SweepOp> fused_ops;  // RC_FUSED / RC_DIAG16: register ops in order (.mat = recipe index)
    std::vector<int> fused_srcs;     // RC_FUSED / RC_DIAG16: per fused op, recipe index of its matrix (-1 none)
    std::vector<int> fused_bits;     // RC_FUSED: register bits of the block, most significant index bit first
    int aux = 0;                     // RC_UNIT: pivot column bits
    // Every entry is multiplied by the product of the real pivots m[0][c] of these (mat, c) sources: the
    // common factors of a pass's unit-pivot ops, applied once in the pass.
    std::vector<std::pair<int, int>> scale_srcs;
};

struct SweepPlan {
    int chunk_bits = kChunkBits;  // local qubits per chunk: 12, or 13 (runtime-compiled passes only)
    std::vector<SweepPass> passes;
    std::vector<SweepOp> ops;      // all passes' ops (stage ranges index this)
    std::vector<PoolRecipe> pool;  // recipes; SweepOp.mat = complex offset of the recipe's output
    uint64_t pool_complex = 0;
    std::vector<int> standalone;   // unused for now (every op kind fits a sweep)
    int fused_blocks = 0;
    double fp64_ops = 0.0;  // FP64 operations per amplitude of the register ops over all passes
    // Adjoint plans: per gradient probe (slot), the pivots (caller matrix offset, column) of the unit-pivot
    // ops its pass ran before it.  Both states then carry the factor 1 / prod(pivots) (the pass's scaling op
    // comes last), so the probe's sum is prod(pivots)^2 times too small.
    std::vector<std::pair<uint64_t, std::vector<std::pair<int, int>>>> probe_pivots;
    // Product-state input (runtime-compiled passes from a basis state): the circuit's leading single-qubit ops
    // were taken out of the plan (absorb_leading_ops), and the first pass generates amp(i) = prod_q v_q[i_q]
    // with v_q those ops applied to the basis state's qubit q.  The pool then ends in the tables the first
    // pass multiplies (finalize_product): R[16] over its register index, TL / TH over the low / high
    // thread-index bits, CL / CH over the low / high chunk-index bits.
    bool product = false;
    std::vector<std::pair<int, int>> product_ops;  // absorbed (qubit, caller matrix offset), circuit order
    uint64_t prod_off = 0;                          // pool offset of R
    int prod_tl = 0, prod_h = 0;                    // thread-index bits in TL, chunk-index bits in CL
};
// prod(pivots)^2 of every probe of a plan for one matrix array: (slot, factor) pairs.
void probe_factors(const SweepPlan &plan, const double *mats, std::vector<std::pair<uint64_t, double>> &out);

// FP64 operations per amplitude of one register op (an FMA or a MUL counts one).
double sweep_op_fp64(const SweepOp &op);

// Cost model of the planner's fusion choices, in FP64 operations per amplitude.
struct SweepCostModel {
    double op_overhead;  // fixed cost of every register op
    bool const_ops;      // ops are compile-time constants (runtime-compiled pass kernels)
    double pass_cap;     // most estimated FP64 operations per amplitude a pass takes on (0: no limit)
    bool unit_1q;        // single-qubit ops with a unit-pivot form use it (SW_U1U) and fold their pivots
    bool direct_io;      // passes may load into their first stage's and store from their last stage's registers
    bool permlane;       // stages may be entered by in-register lane exchanges (SweepStage::perm)
};
// The compiled sweep kernel: on the 30-qubit layered circuit an overhead of 1 beat 4 and 8 (measured).
constexpr SweepCostModel kInterpretedCost{1.0, false, 0.0, false, false, false};
constexpr SweepCostModel kCompiledPassCost{0.5, true, 0.0, true, true, false};
// Adjoint passes: as the runtime-compiled passes; a pass with gradient probes applies its unit-pivot factor
// by a scaling op at its end, and each probe's sum is corrected on the host (SweepPlan::probe_pivots).  Their
// kernels always go through LDS at both ends.
constexpr SweepCostModel kAdjointPassCost{0.5, true, 0.0, true, false, false};

// max_fuse_bits: register bits a fused dense block may span (0: no fusion beyond what a diagonal on two
// register bits needs).
int plan_sweeps(int n, const std::vector<PlanOp> &ops, int low_qubits, bool remap, SweepPlan &plan,
                int max_fuse_bits = kRegBits, const SweepCostModel &cm = kInterpretedCost,
                int chunk_bits = kChunkBits);

// Predicted time of a runtime-compiled pass (and of a whole plan) on MI355X, in ns per amplitude: the larger
// of the chunk size's HBM floor and a linear cost in the pass's FP64 operations and LDS round trips, fitted to
// measured pass times.
double predicted_pass_ns_per_amp(const SweepPlan &plan, const SweepPass &pass);
double predicted_plan_ns_per_amp(const SweepPlan &plan);

// Low qubits of a sweep plan when the option leaves them to the library and no prediction applies.
constexpr int kDefaultLowQubits = 4;

struct GeometryPick {
    int chunk_bits = kChunkBits;
    int low_qubits = kDefaultLowQubits;
    double ns_per_amp = 0.0;  // predicted plan time per amplitude
    int candidates = 0;       // geometries planned
};
// The geometry autotuner.  chunk_bits 0 and / or low_qubits -1 leave that choice to the library: with
// runtime-compiled passes (compiled) the circuit is planned under each measured geometry that fits the fixed
// choices ((12, 4), (13, 4), (13, 3) chunk / low qubits; 13-bit chunks from 13 qubits) and the plan with the
// smallest predicted time is kept, where a later candidate must be predicted 3% faster than an earlier one.  The compiled sweep kernel always uses (12, 4) or the
// fixed low count.  Deterministic: the pick depends only on the circuit structure and the options.
int plan_sweeps_geometry(int n, const std::vector<PlanOp> &ops, int low_qubits, int chunk_bits, bool remap,
                         bool compiled, SweepPlan &plan, int max_fuse_bits, const SweepCostModel &cm,
                         GeometryPick *pick);
// Plans of recently executed circuits.  A plan depends on the op structure (kinds, qubits, controls,
// matrix offsets and the diagonal / real / exchange classification of every matrix) and the options that
// shape it, not on matrix values, so circuits that differ only in parameters share one entry.
struct PlanCache {
    struct Entry {
        std::string key;
        SweepPlan plan;
        std::vector<hipFunction_t> fns;  // runtime-compiled pass kernels (SC_OPT_JIT), else empty
        GeometryPick pick;               // the geometry the plan was made with
        uint64_t used = 0;
    };
    static constexpr size_t kMaxEntries = 16;
    std::vector<std::unique_ptr<Entry>> entries;
    uint64_t tick = 0;
    Entry *find(const std::string &key);
    Entry *insert(std::unique_ptr<Entry> e);  // evicts the least recently used entry when full
};
std::string plan_cache_key(int n, const std::vector<PlanOp> &ops, int low_qubits, bool remap, int max_fuse_bits,
                           int jit, double pass_cap, int chunk_bits = kChunkBits);

// The plan (and runtime-compiled pass kernels) of an analysed circuit on n qubits under the context's options,
// from the context's plan cache (planned and compiled on a miss).  The entry stays valid until 16 other
// structures have been planned after it.
// basis_input: the state is a pending basis state, so with runtime-compiled passes the leading single-qubit
// ops become a product-state input generated by the first pass (SweepPlan::product).
int sweep_plan_entry(Context *ctx, int n, const std::vector<PlanOp> &ops, PlanCache::Entry **out,
                     bool basis_input = false);

// Build the matrix pool of a plan for one caller matrix array into out (pool_complex complex values).
// basis: the basis-state input of a product plan (ignored otherwise).
void build_pool(const SweepPlan &plan, const double *mats, double *out, long long basis = -1);

// Leading single-qubit ops of a circuit that starts from a basis state: every uncontrolled single-qubit op
// before any other op touches its qubit.  They are marked `skip` and listed as (qubit, matrix offset).
void absorb_leading_ops(int n, std::vector<PlanOp> &ops, std::vector<std::pair<int, int>> &absorbed);
// Makes a plan of the remaining ops a product plan: records the absorbed ops and appends the table space to
// its pool.  The plan must have a pass.
void finalize_product(int n, SweepPlan &plan, const std::vector<std::pair<int, int>> &absorbed);
// The first pass's generated input of a product plan on the host, with the kernel's table products
// (pool built by build_pool with the basis): 2^n interleaved complex values.
void product_input_host(int n, const SweepPlan &plan, const double *pool, double *state);
// Execute a sweep plan on a host array with exactly the kernel's data movement (load phase, swizzled LDS
// slots, per-thread register sets, stage ops, store phase).  Used to validate plans on the CPU.  With lam, an
// adjoint plan runs every op on both states and adds each probe's <lambda| G |psi> to grads[2 * slot].
void execute_plan_host(int n, const SweepPlan &plan, const double *pool, double *state, double *lam = nullptr,
                       double *grads = nullptr);

}  // namespace sc
