// sc_internal.hpp -- shared host/device definitions of the gfx950 statevector engine.
//
// Layout: a state of n qubits and `batch` elements is one HBM allocation of
// batch * 2^n double2 values (re, im); element b starts at b << n.  All
// indices are 64-bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/statecraft_hip.h"
#include "sc_ops.hpp"

namespace sc {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define SC_HIP_CHECK(expr)                                                                            \
    do {                                                                                              \
        hipError_t _e = (expr);                                                                       \
        if (_e != hipSuccess)                                                                         \
            return ::sc::fail(SC_E_HIP, std::string(#expr) + " failed: " + hipGetErrorString(_e));    \
    } while (0)

// ---------------------------------------------------------------- profiling
struct Profiler {
    bool enabled = false;
    uint64_t launches[SC_KFAM_COUNT] = {};
    double ms[SC_KFAM_COUNT] = {};
    double bytes[SC_KFAM_COUNT] = {};
    struct Pending {
        hipEvent_t a, b;
        int fam;
        double bytes;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
};

// Wraps one kernel launch: records events around it when profiling is on.
struct ProfScope {
    Profiler *p;
    hipStream_t s;
    int fam;
    double bytes;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(Profiler *prof, hipStream_t stream, int family, double algo_bytes);
    ~ProfScope();
};

struct PlanCache;      // sc_plan.hpp
struct JobGraphCache;  // sc_graph.hip

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    int exec_mode = SC_EXEC_SWEEP;
    int max_fused = 4;
    int low_qubits = -1;  // SC_OPT_LOW_QUBITS; -1: chosen per circuit structure (plan_sweeps_geometry)
    int remap = 0;
    int sweep_blocks_per_cu = 0;
    int jit = 2;
    int pass_cap = -1;
    int chunk_bits = 0;   // SC_OPT_CHUNK_BITS; 0: chosen per circuit structure
    int direct_io = 1;    // SC_OPT_DIRECT_IO
    int mfma = 1;         // SC_OPT_MFMA
    int mfma_tile = -1;   // SC_OPT_MFMA_TILE; -1: autotuned
    int mfma_pick[2] = {-1, -1};  // the autotuner's tile shape for 4- and 5-qubit blocks (-1: not tuned yet)
    int permlane = 0;     // SC_OPT_PERMLANE
    int fused_terms = 1;  // SC_OPT_FUSED_TERMS
    int pass_sched = 4;   // SC_OPT_PASS_SCHED
    int cu_count = 256;
    Profiler prof;
    // scratch (grown on demand, freed with the context)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // sc_adjoint_gradient's checkpoints, lambda and copy of psi, kept for the next call (allocating and freeing
    // ~100 GiB per call has cost seconds); released by sc_ctx_destroy or when a state allocation needs the HBM
    void *adj_ws = nullptr;
    size_t adj_ws_bytes = 0;
    void *scratch_host = nullptr;  // pinned
    size_t scratch_host_bytes = 0;
    // device copy of a sweep plan's records and matrix pool (used by sweep execution only)
    void *plan_dev = nullptr;
    size_t plan_dev_bytes = 0;
    PlanCache *plan_cache = nullptr;  // owned
    // plan statistics of the last sc_apply_ops
    uint64_t last_sweeps = 0, last_launches = 0, last_blocks = 0;
    // geometry of the last sweep plan acquired (sc_last_plan_geometry)
    int last_chunk_bits = 0, last_low_qubits = 0, last_candidates = 0;
    double last_ns_per_amp = 0.0;
    // exact-sum statistics of the last sc_sample: fast / serial / zero segments, gather rounds, rescaled
    uint64_t last_sample[5] = {};
    JobGraphCache *job_graphs = nullptr;  // recorded sc_run_jobs graphs (owned)
};

struct State {
    Context *ctx = nullptr;
    int n = 0;
    int batch = 1;
    double2 *amp = nullptr;  // batch << n amplitudes
    // >= 0: every element is the basis state |pending_basis>, not yet written to amp.  sc_state_set_basis only
    // records it; the next runtime-compiled sweep generates it in its first pass, anything else writes it
    // first (materialize_basis).
    int64_t pending_basis = -1;
    uint64_t dim() const { return 1ULL << n; }
};
// Writes a pending basis state to amp (no-op otherwise).
int materialize_basis(State *st);

int ensure_scratch(Context *ctx, size_t bytes);
// sc_run_jobs (sc_graph.hip); destroy_job_graphs releases a context's recorded graphs.
int run_jobs(Context *ctx, const sc_job *jobs, int n_jobs, int *used_graph);
void destroy_job_graphs(Context *ctx);
// Device address of sc_expectation_pauli_terms_noalloc's fixed scratch array on `device` (current), once per device.
int init_noalloc_scratch(int device);
int ensure_host_scratch(Context *ctx, size_t bytes);
int drain_profiler(Context *ctx);

// ---------------------------------------------------------------- device-side op encodings
// Dense per-op gates (DIRECT mode) take their matrices by value.
// Tile shapes of the matrix-core dense-block kernel (SC_OPT_MFMA_TILE): 16-group tiles per wavefront step 1
// (shapes 0-2) or 2 (3-5), resident workgroups per CU 16, 8, 32.
constexpr int kMfmaTiles = 6;

struct Mat2 {
    double v[8];  // row-major 2x2 complex (re, im)
};
struct Mat4 {
    double v[32];  // row-major 4x4 complex
};

// ---- sweep program (LDS-blocked multi-gate pass) -------------------------
// One chunk = 2^12 amplitudes: 12 "local" qubits resident in LDS; every
// other qubit is fixed per workgroup.  A pass runs stages; in a stage each of
// 256 threads holds 16 amplitudes that differ in 4 "register" local bits.
constexpr int kMaxStages = 24;
constexpr int kSwzBits = 4;  // LDS slot bits permuted by the XOR swizzle


// 32 bytes: the kernel reads a whole record with one scalar load.
struct SweepOp {
    int8_t kind;
    int8_t k0, k1;   // register bit positions (-1: none)
    uint8_t thi;     // SW_DIAGL with 512 threads: bit 8 of ta (bit 0) and of tb (bit 1)
    int32_t mat;     // complex offset of the op's matrix / factor table in the pass pool
    uint16_t rctrl;  // register-index control mask
    uint16_t tctrl;  // thread-index control mask
    uint8_t ta, tb;  // SW_DIAGL: thread-index masks of the diagonal's bit sources
    int8_t oa, ob;   // SW_DIAGL: outer (global) qubit positions of the diagonal's bit sources, -1: none
    uint64_t octrl;  // outer (chunk-fixed) control mask, global qubit positions
    uint64_t slot;   // SW_GRAD: gradient slot the inner product accumulates into (ta = generator)
};
static_assert(sizeof(SweepOp) == 32, "sweep op records are 32 bytes");

// Slot tables are 32-bit so that the kernel reads them with scalar dword loads.
struct SweepStage {
    int32_t op_begin, op_end;
    // 1: the exchange into this stage stays inside each wavefront (the wave-index bits, thread bits 6 and 7,
    // sit on the same local positions before and after), so it needs no workgroup barrier
    int32_t wave_local;
    // Runtime-compiled passes only: nonzero when this stage is entered from the previous one inside the
    // registers (bit 6), its layout being the previous layout with register bit j5 exchanged with lane bit 5
    // (bits 0-2: j5 + 1, 0 for none) and register bit j4 with lane bit 4 (bits 3-5: j4 + 1); the LDS tables
    // still describe the layout, so other executors may take the exchange through LDS
    int32_t perm;
    uint32_t lds_off[kRegs];            // swizzled LDS slot offset per register index
    uint32_t thread_slot[kMaxThreadBits];  // swizzled slot contribution of each thread-index bit
};

struct SweepPass {
    uint64_t outer_mask;            // global qubits outside the chunk
    int32_t nstages;
    // bit 0: the last stage -> store phase exchange stays inside each wavefront; bit 1: so does the store
    // phase -> next chunk's load phase (the workgroup may overwrite its LDS without a barrier)
    int32_t wave_local;
    uint64_t glob_thread[kMaxThreadBits];  // load/store phase: global offset of thread-index bit i
    uint64_t glob_reg[kRegs];       // load/store phase: global offset of register index r
    uint32_t io_thread_slot[kMaxThreadBits];  // load phase: swizzled slot of thread-index bit i
    uint32_t io_reg_slot[kRegs];
    // Store phase: thread bit i / register bit j still address chunk positions i / 8+j in HBM, but read the
    // LDS slot of the local qubit the pass moves there.  Equal to io_* when the pass does not remap qubits.
    uint32_t st_thread_slot[kMaxThreadBits];
    uint32_t st_reg_slot[kRegs];
    // Runtime-compiled passes only.  bit 0: the HBM loads go straight into the first stage's registers (its
    // layout keeps the chunk's low positions on lane bits 0-3, so the loads stay coalesced) and the load
    // phase's LDS round trip is skipped; bit 1: the last stage's registers go straight to HBM and the store
    // phase's round trip is skipped.  Global offsets of thread-index bit i / register index r in the first
    // stage's layout (ld_*) and in the last stage's layout (sd_*).
    int32_t direct;
    int32_t pad2;
    uint64_t ld_thread[kMaxThreadBits];
    uint64_t ld_reg[kRegs];
    uint64_t sd_thread[kMaxThreadBits];
    uint64_t sd_reg[kRegs];
    SweepStage stages[kMaxStages];
};

// ---------------------------------------------------------------- host-side plan
struct HostOp {
    int kind;  // SC_OP_*
    int q0, q1;
    uint64_t controls;
    double m[32];  // up to 4x4 complex
    bool diagonal;  // gate matrix is diagonal (1q) -> diagonal role on its qubit
};

// Launchers (implemented in the .hip files)
int launch_single(Context *ctx, State *st, int target, const Mat2 &g, uint64_t controls);
int launch_x(Context *ctx, State *st, int target, uint64_t controls);
int launch_two(Context *ctx, State *st, int first, int second, const Mat4 &m);
int launch_diag2(Context *ctx, State *st, int first, int second, const double *d);
int launch_swap(Context *ctx, State *st, int first, int second);
int launch_dense(Context *ctx, State *st, int k, const int *qubits, const double *mat_dev);
int launch_copy(Context *ctx, const double2 *src, double2 *dst, uint64_t count);
int launch_sweep(Context *ctx, State *st, const SweepPass *pass_dev, const SweepOp *ops_dev,
                 const double *mats_dev, int elem_begin, int elem_count, size_t mat_stride);

int launch_probabilities_full(Context *ctx, State *st, int elem, double *out_dev);
int launch_probabilities_marginal(Context *ctx, State *st, int elem, const int32_t *measured, int m,
                                  double *out_dev);
int launch_norm_sq(Context *ctx, const double2 *amp, uint64_t count, double *result_host);
int launch_checksum(Context *ctx, const double2 *amp, uint64_t count, double *out_host);
int launch_max_deviation(Context *ctx, const double2 *amp, uint64_t count, uint64_t index, double2 a, double2 b,
                         double *out_host);
int launch_kraus_weights(Context *ctx, const double2 *amp, int n, int target, const double *kraus, int k,
                         double *weights_host);
int launch_scale_state(Context *ctx, double2 *amp, uint64_t count, double factor);

// Per-term Pauli expectations of every batch element (sc_adjoint.hip), in two halves so that the device part
// can be captured into a graph: pauli_terms_enqueue launches the term kernels into `scratch` (partial sums,
// term_scratch_bytes of them for term_parts partials per row) and leaves the raw row sums in out_dev
// ([batch][n_terms]); pauli_terms_finish applies the host-side factors to those rows once copied back.
uint64_t term_parts(const Context *ctx, uint64_t B, int n, uint64_t rows);
size_t term_scratch_bytes(uint64_t B, uint64_t rows, uint64_t parts);
int pauli_terms_enqueue(Context *ctx, State *st, const double2 *amp, const uint64_t *xm, const uint64_t *zm,
                        int n_terms, void *scratch, size_t scratch_bytes, double *out_dev);
void pauli_terms_finish(const uint64_t *xm, const uint64_t *zm, int n_terms, uint64_t B, double *vals);



// op application shared by the C-ABI entry points (sc_capi.hip)
int apply_ops_impl(State *st, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex,
                   uint64_t mat_stride, bool per_elem);
// The same with the states read from src and left in dst (either may be st->amp); st->amp is the work
// buffer in between.  Runtime-compiled sweep passes read and write the two buffers directly, other
// execution modes copy.
int apply_ops_between(State *st, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex,
                      uint64_t mat_stride, bool per_elem, const double2 *src, double2 *dst);

// A device buffer owned for the duration of one API call.
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes);
};

}  // namespace sc

struct sc_ctx : sc::Context {
    // adjoint statistics / test hook
    uint64_t adj_reanchors = 0;
    double adj_max_drift = 0.0;
    int64_t test_drift = 0;
};
struct sc_state : sc::State {};
