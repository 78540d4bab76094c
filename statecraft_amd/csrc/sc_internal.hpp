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

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    int exec_mode = SC_EXEC_SWEEP;
    int max_fused = 4;
    int low_qubits = 4;
    int remap = 1;
    int cu_count = 256;
    Profiler prof;
    // scratch (grown on demand, freed with the context)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    void *scratch_host = nullptr;  // pinned
    size_t scratch_host_bytes = 0;
    // plan statistics of the last sc_apply_ops
    uint64_t last_sweeps = 0, last_launches = 0, last_blocks = 0;
};

struct State {
    Context *ctx = nullptr;
    int n = 0;
    int batch = 1;
    double2 *amp = nullptr;  // batch << n amplitudes
    uint64_t dim() const { return 1ULL << n; }
};

int ensure_scratch(Context *ctx, size_t bytes);
int ensure_host_scratch(Context *ctx, size_t bytes);
int drain_profiler(Context *ctx);

// ---------------------------------------------------------------- device-side op encodings
// Dense per-op gates (DIRECT mode) take their matrices by value.
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
constexpr int kChunkBits = 12;
constexpr int kChunk = 1 << kChunkBits;
constexpr int kRegBits = 4;
constexpr int kRegs = 1 << kRegBits;
constexpr int kThreadBits = kChunkBits - kRegBits;  // 8
constexpr int kThreads = 1 << kThreadBits;          // 256
constexpr int kMaxStages = 24;
constexpr int kSwzBits = 4;  // LDS slot bits permuted by the XOR swizzle

enum SweepOpKind : int32_t {
    SW_U1 = 1,     // (controlled) 2x2 on register bit k0
    SW_X = 2,      // (controlled) exact swap on register bit k0 (CNOT / MC-X)
    SW_U2 = 3,     // dense 4x4 on register bits (k0 = first = MSB, k1 = second)
    SW_DIAG = 4,   // 2-entry or 4-entry diagonal on up to two qubits anywhere
    SW_SWAP = 5,   // exact exchange of register bits k0, k1
    SW_DENSE4 = 6, // dense 16x16 on all four register bits (fused block)
    SW_U1R = 7,    // SW_U1 with a real matrix (imaginary parts zero): half the FP64 work
    SW_U2R = 8,    // SW_U2 with a real matrix
    SW_DENSE4R = 9,// SW_DENSE4 with a real matrix
};

// Where a qubit's bit of an amplitude index comes from inside a stage.
struct BitSrc {
    uint32_t rmask;  // register-index mask (one bit) or 0
    uint32_t tmask;  // thread-index mask (one bit) or 0
    uint32_t outer;  // 1 if the bit is an outer qubit that is set in this chunk (resolved in-kernel)
    uint32_t omask_lo, omask_hi;  // outer qubit mask (global) when the qubit is outer, else 0
};

struct SweepOp {
    int32_t kind;
    int32_t k0, k1;     // register bit positions
    int32_t mat;        // complex offset into the pass matrix pool
    // controls (SW_U1 / SW_X): register part, thread part, outer part
    uint32_t rctrl;     // register-index mask
    uint32_t tctrl;     // thread-index mask
    uint64_t octrl;     // global mask of outer controls
    // SW_DIAG: qubit a (MSB of the 2-bit diagonal index) and b (LSB); nd = 1 or 2 entries bits
    int32_t nd;
    int32_t pad;
    BitSrc da, db;
};

struct SweepStage {
    int32_t op_begin, op_end;
    uint16_t lds_off[kRegs];      // swizzled LDS slot offset per register index
    uint16_t thread_slot[kThreadBits];  // swizzled slot contribution of each thread-index bit
};

struct SweepPass {
    uint64_t outer_mask;            // global qubits outside the chunk
    int32_t nstages;
    int32_t pad;
    uint64_t glob_thread[kThreadBits];  // load/store phase: global offset of thread-index bit i
    uint64_t glob_reg[kRegs];       // load/store phase: global offset of register index r
    uint16_t io_thread_slot[kThreadBits];  // load phase: swizzled slot of thread-index bit i
    uint16_t io_reg_slot[kRegs];
    // Store phase: thread bit i / register bit j still address chunk positions i / 8+j in HBM, but read the
    // LDS slot of the local qubit the pass moves there.  Equal to io_* when the pass does not remap qubits.
    uint16_t st_thread_slot[kThreadBits];
    uint16_t st_reg_slot[kRegs];
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
int device_sum(Context *ctx, const double *x, uint64_t count, double *result_host);
int launch_scale(Context *ctx, double *x, uint64_t count, double factor);
int launch_cumulative(Context *ctx, double *p, uint64_t count);
int launch_search(Context *ctx, const double *cum, uint64_t count, const double *u_dev, int64_t shots,
                  int64_t *out_dev, double *gap_dev);
int launch_norm_sq(Context *ctx, const double2 *amp, uint64_t count, double *result_host);
int launch_checksum(Context *ctx, const double2 *amp, uint64_t count, double *out_host);
int launch_kraus_weights(Context *ctx, const double2 *amp, int n, int target, const double *kraus, int k,
                         double *weights_host);
int launch_scale_state(Context *ctx, double2 *amp, uint64_t count, double factor);

// batched Kraus channels (sc_noise.hip)
size_t noise_work_bytes(int batch);
int launch_channel_batched(Context *ctx, State *st, int target, const double *kraus, int K, const double *u_dev,
                           uint64_t u_stride, uint64_t draw, void *work);
int *noise_err_flag(void *work, int batch);

// op application shared by the C-ABI entry points (sc_capi.hip)
int apply_ops_impl(State *st, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex,
                   uint64_t mat_stride, bool per_elem);

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
