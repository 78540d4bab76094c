// sc_gates.hip -- one-kernel-per-op gate application (SC_EXEC_DIRECT) and the
// dense fused-block kernel (SC_EXEC_FUSED).
//
// Each kernel is a grid-stride loop over compacted index groups: a group index
// g is expanded by inserting zero bits at the gate's qubits (and, for
// controlled gates, OR-ing the control mask in), so a controlled gate with c
// controls touches only 2^(n-1-c) pairs instead of sweeping every pair with a
// mask test as StatevectorOps.applyMultiControlledSingleGate does
// (StatevectorOps.java:786-828).  Consecutive lanes take consecutive groups, so
// a wavefront's two loads cover 2 KiB of contiguous amplitudes whenever the
// lowest touched qubit is >= 6, and whole cache lines otherwise.
#include "sc_internal.hpp"

namespace sc {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint64_t insert_zero(uint64_t x, int bit) {
    uint64_t lo = x & ((1ULL << bit) - 1ULL);
    return ((x ^ lo) << 1) | lo;
}

// Insert zero bits at every set position of `mask`, ascending.
__device__ __forceinline__ uint64_t expand_mask(uint64_t x, uint64_t mask) {
    while (mask) {
        int p = __builtin_ctzll(mask);
        x = insert_zero(x, p);
        mask &= mask - 1ULL;
    }
    return x;
}

__device__ __forceinline__ double2 cmul(double ar, double ai, double2 x) {
    return make_double2(ar * x.x - ai * x.y, ar * x.y + ai * x.x);
}

__device__ __forceinline__ double2 cmad(double ar, double ai, double2 x, double2 acc) {
    acc.x = fma(ar, x.x, acc.x);
    acc.x = fma(-ai, x.y, acc.x);
    acc.y = fma(ar, x.y, acc.y);
    acc.y = fma(ai, x.x, acc.y);
    return acc;
}

// (Controlled) 2x2 gate: StatevectorOps.java:19-126 (uncontrolled) and :631-828 (controlled).
__global__ void __launch_bounds__(kBlock) k_single(double2 *__restrict__ amp, int n, int target, uint64_t fixed_mask,
                                                   uint64_t controls, int log_groups, uint64_t total, Mat2 g) {
    const uint64_t stride = 1ULL << target;
    const uint64_t gmask = (1ULL << log_groups) - 1ULL;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < total; i += (uint64_t)gridDim.x * kBlock) {
        uint64_t elem = i >> log_groups;
        uint64_t idx0 = expand_mask(i & gmask, fixed_mask) | controls;
        double2 *base = amp + (elem << n);
        double2 a0 = base[idx0];
        double2 a1 = base[idx0 | stride];
        double2 n0 = cmul(g.v[0], g.v[1], a0);
        n0 = cmad(g.v[2], g.v[3], a1, n0);
        double2 n1 = cmul(g.v[4], g.v[5], a0);
        n1 = cmad(g.v[6], g.v[7], a1, n1);
        base[idx0] = n0;
        base[idx0 | stride] = n1;
    }
}

// (Controlled) X: exact exchange, StatevectorOps.java:128-172 (applyCnot).
__global__ void __launch_bounds__(kBlock) k_x(double2 *__restrict__ amp, int n, int target, uint64_t fixed_mask,
                                              uint64_t controls, int log_groups, uint64_t total) {
    const uint64_t stride = 1ULL << target;
    const uint64_t gmask = (1ULL << log_groups) - 1ULL;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < total; i += (uint64_t)gridDim.x * kBlock) {
        uint64_t elem = i >> log_groups;
        uint64_t idx0 = expand_mask(i & gmask, fixed_mask) | controls;
        double2 *base = amp + (elem << n);
        double2 a0 = base[idx0];
        double2 a1 = base[idx0 | stride];
        base[idx0] = a1;
        base[idx0 | stride] = a0;
    }
}

// Dense 4x4 on (first, second), local index l = 2*bit(first) + bit(second): StatevectorOps.java:174-407, :851-929.
__global__ void __launch_bounds__(kBlock) k_two(double2 *__restrict__ amp, int n, int first, int second,
                                                int log_groups, uint64_t total, Mat4 m) {
    const uint64_t fm = 1ULL << first, sm = 1ULL << second;
    const uint64_t fixed = fm | sm;
    const uint64_t gmask = (1ULL << log_groups) - 1ULL;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < total; i += (uint64_t)gridDim.x * kBlock) {
        uint64_t elem = i >> log_groups;
        uint64_t b = expand_mask(i & gmask, fixed);
        double2 *base = amp + (elem << n);
        uint64_t off[4] = {b, b | sm, b | fm, b | fm | sm};
        double2 a[4];
#pragma unroll
        for (int c = 0; c < 4; c++) a[c] = base[off[c]];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            double2 acc = cmul(m.v[8 * r], m.v[8 * r + 1], a[0]);
#pragma unroll
            for (int c = 1; c < 4; c++) acc = cmad(m.v[8 * r + 2 * c], m.v[8 * r + 2 * c + 1], a[c], acc);
            base[off[r]] = acc;
        }
    }
}

// Two-qubit diagonal: StatevectorOps.java:409-541, :931-972.
__global__ void __launch_bounds__(kBlock) k_diag2(double2 *__restrict__ amp, int first, int second, uint64_t total,
                                                  double d0r, double d0i, double d1r, double d1i, double d2r,
                                                  double d2i, double d3r, double d3i) {
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < total; i += (uint64_t)gridDim.x * kBlock) {
        // batch elements are contiguous and qubit bits are below n, so bit tests on i are per element.
        int bf = (int)((i >> first) & 1ULL), bs = (int)((i >> second) & 1ULL);
        double dr = bf ? (bs ? d3r : d2r) : (bs ? d1r : d0r);
        double di = bf ? (bs ? d3i : d2i) : (bs ? d1i : d0i);
        amp[i] = cmul(dr, di, amp[i]);
    }
}

// SWAP: StatevectorOps.java:543-629.
__global__ void __launch_bounds__(kBlock) k_swap(double2 *__restrict__ amp, int n, int first, int second,
                                                 int log_groups, uint64_t total) {
    const uint64_t fm = 1ULL << first, sm = 1ULL << second;
    const uint64_t gmask = (1ULL << log_groups) - 1ULL;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < total; i += (uint64_t)gridDim.x * kBlock) {
        uint64_t elem = i >> log_groups;
        uint64_t b = expand_mask(i & gmask, fm | sm);
        double2 *base = amp + (elem << n);
        double2 x = base[b | fm];
        double2 y = base[b | sm];
        base[b | fm] = y;
        base[b | sm] = x;
    }
}

// Dense 2^K x 2^K fused block on qubits q[0..K-1] (local bit i <-> q[i]).  The matrix is uniform across the
// grid; row-major complex entries are read with scalar loads.
template <int K>
__global__ void __launch_bounds__(kBlock) k_dense(double2 *__restrict__ amp, int n, uint64_t fixed,
                                                  uint64_t qoff0, uint64_t qoff1, uint64_t qoff2, uint64_t qoff3,
                                                  uint64_t qoff4, int log_groups, uint64_t total,
                                                  const double2 *__restrict__ mat) {
    constexpr int D = 1 << K;
    const uint64_t gmask = (1ULL << log_groups) - 1ULL;
    const uint64_t qo[5] = {qoff0, qoff1, qoff2, qoff3, qoff4};
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < total; i += (uint64_t)gridDim.x * kBlock) {
        uint64_t elem = i >> log_groups;
        uint64_t b = expand_mask(i & gmask, fixed);
        double2 *base = amp + (elem << n);
        double2 x[D];
#pragma unroll
        for (int c = 0; c < D; c++) {
            uint64_t o = b;
#pragma unroll
            for (int j = 0; j < K; j++)
                if ((c >> j) & 1) o |= qo[j];
            x[c] = base[o];
        }
#pragma unroll
        for (int r = 0; r < D; r++) {
            double2 m0 = mat[r * D];
            double2 acc = cmul(m0.x, m0.y, x[0]);
#pragma unroll
            for (int c = 1; c < D; c++) {
                double2 mc = mat[r * D + c];
                acc = cmad(mc.x, mc.y, x[c], acc);
            }
            uint64_t o = b;
#pragma unroll
            for (int j = 0; j < K; j++)
                if ((r >> j) & 1) o |= qo[j];
            base[o] = acc;
        }
    }
}

// The same dense 2^K x 2^K block (K = 4, 5) as a complex GEMM on the FP64 matrix cores: Y = U X, where column
// j of X is the 2^K amplitudes of index group j.  A wavefront takes 16 groups at a time and forms their outputs
// with v_mfma_f64_16x16x4_f64 (D[16x16] += A[16x4] B[4x16]) as four real products per complex product:
// Yr = Ur Xr + Ui (-Xi), Yi = Ur Xi + Ui Xr.  Operand layouts of that instruction on gfx950 (measured with a
// one-wavefront probe): lane L supplies A[L % 16][L / 16] and B[L / 16][L % 16] (one double each) and holds
// D[4 v + L / 16][L % 16] in accumulator element v.  So lane L reads block rows k = 4 s + L / 16 of group L % 16
// straight from HBM (16 lanes = 16 consecutive groups, 256-byte runs when no block qubit is below position 4)
// and writes output rows 4 v + L / 16 (+16 per row tile): no
// LDS and no cross-lane exchange.  U's operands stay in registers for the kernel's life (2^K / 4 k-steps x
// 2^(K-4) row tiles x 2 doubles per lane).  Every group is read and written by the four lanes of one wavefront
// that own it, after all of the wavefront's loads, so the block is applied in place.
typedef double v4d __attribute__((ext_vector_type(4)));

template <int K, int T>
__global__ void __launch_bounds__(kBlock) k_dense_mfma(double2 *__restrict__ amp, int n, uint64_t fixed,
                                                       uint64_t qoff0, uint64_t qoff1, uint64_t qoff2, uint64_t qoff3,
                                                       uint64_t qoff4, int log_groups, uint64_t total,
                                                       const double2 *__restrict__ mat) {
    static_assert(K == 4 || K == 5, "MFMA dense blocks span 4 or 5 qubits");
    static_assert(T == 1 || T == 2, "one or two 16-group tiles per wavefront step");
    constexpr int D = 1 << K;
    constexpr int KS = D / 4;   // k-steps of 4
    constexpr int MT = D / 16;  // row tiles of 16
    const uint64_t qo[5] = {qoff0, qoff1, qoff2, qoff3, qoff4};
    const uint32_t lane = threadIdx.x & 63u, col = lane & 15u, kq = lane >> 4;
    auto block_off = [&](uint32_t c) -> uint64_t {
        uint64_t o = 0;
#pragma unroll
        for (int j = 0; j < K; j++)
            if ((c >> j) & 1u) o |= qo[j];
        return o;
    };
    double ur[MT][KS], ui[MT][KS];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const double2 u = mat[(16 * mt + col) * D + 4 * s + kq];
            ur[mt][s] = u.x;
            ui[mt][s] = u.y;
        }
    uint64_t koff[KS], moff[MT][4];
#pragma unroll
    for (int s = 0; s < KS; s++) koff[s] = block_off(4u * s + kq);
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int v = 0; v < 4; v++) moff[mt][v] = block_off(16u * mt + 4u * v + kq);
    const uint64_t gmask = (1ULL << log_groups) - 1ULL;
    const uint64_t steps = (total >> 4) / T;  // the launcher passes a tile count divisible by T
    const uint64_t waves = (uint64_t)gridDim.x * (kBlock / 64);
    for (uint64_t w = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); w < steps; w += waves) {
        // every tile's loads first (T independent 16-group reads in flight), then its products and stores
        double2 *base[T];
        double xr[T][KS], xi[T][KS], nxi[T][KS];
#pragma unroll
        for (int tt = 0; tt < T; tt++) {
            const uint64_t g = ((w * T + tt) << 4) | col;
            base[tt] = amp + ((g >> log_groups) << n) + expand_mask(g & gmask, fixed);
#pragma unroll
            for (int s = 0; s < KS; s++) {
                const double2 x = base[tt][koff[s]];
                xr[tt][s] = x.x;
                xi[tt][s] = x.y;
                nxi[tt][s] = -x.y;
            }
        }
#pragma unroll
        for (int tt = 0; tt < T; tt++)
#pragma unroll
            for (int mt = 0; mt < MT; mt++) {
                v4d yr = {0.0, 0.0, 0.0, 0.0}, yi = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s = 0; s < KS; s++) {
                    yr = __builtin_amdgcn_mfma_f64_16x16x4f64(ur[mt][s], xr[tt][s], yr, 0, 0, 0);
                    yr = __builtin_amdgcn_mfma_f64_16x16x4f64(ui[mt][s], nxi[tt][s], yr, 0, 0, 0);
                    yi = __builtin_amdgcn_mfma_f64_16x16x4f64(ur[mt][s], xi[tt][s], yi, 0, 0, 0);
                    yi = __builtin_amdgcn_mfma_f64_16x16x4f64(ui[mt][s], xr[tt][s], yi, 0, 0, 0);
                }
#pragma unroll
                for (int v = 0; v < 4; v++) base[tt][moff[mt][v]] = make_double2(yr[v], yi[v]);
            }
    }
}

// Plain device copy, 16 B per lane per iteration: the bandwidth reference for the sweep kernel.
// Each block moves 4 x 256 consecutive values (4 independent 16 B loads in flight per lane).
__global__ void __launch_bounds__(kBlock) k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                 uint64_t count) {
    const uint64_t base = (uint64_t)blockIdx.x * (4 * kBlock) + threadIdx.x;
    double2 v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        uint64_t i = base + (uint64_t)j * kBlock;
        if (i < count) v[j] = src[i];
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        uint64_t i = base + (uint64_t)j * kBlock;
        if (i < count) dst[i] = v[j];
    }
}

unsigned grid_for(Context *ctx, uint64_t total) {
    uint64_t blocks = (total + kBlock - 1) / kBlock;
    uint64_t cap = (uint64_t)ctx->cu_count * 16;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    return (unsigned)blocks;
}

}  // namespace

int launch_single(Context *ctx, State *st, int target, const Mat2 &g, uint64_t controls) {
    int c = __builtin_popcountll(controls);
    int log_groups = st->n - 1 - c;
    uint64_t total = (uint64_t)st->batch << log_groups;
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_GATE, 32.0 * (double)total * 2.0);
    hipLaunchKernelGGL(k_single, dim3(grid_for(ctx, total)), dim3(kBlock), 0, ctx->stream, st->amp, st->n, target,
                       controls | (1ULL << target), controls, log_groups, total, g);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

int launch_x(Context *ctx, State *st, int target, uint64_t controls) {
    int c = __builtin_popcountll(controls);
    int log_groups = st->n - 1 - c;
    uint64_t total = (uint64_t)st->batch << log_groups;
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_GATE, 32.0 * (double)total * 2.0);
    hipLaunchKernelGGL(k_x, dim3(grid_for(ctx, total)), dim3(kBlock), 0, ctx->stream, st->amp, st->n, target,
                       controls | (1ULL << target), controls, log_groups, total);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

int launch_two(Context *ctx, State *st, int first, int second, const Mat4 &m) {
    int log_groups = st->n - 2;
    uint64_t total = (uint64_t)st->batch << log_groups;
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_GATE, 32.0 * (double)total * 4.0);
    hipLaunchKernelGGL(k_two, dim3(grid_for(ctx, total)), dim3(kBlock), 0, ctx->stream, st->amp, st->n, first, second,
                       log_groups, total, m);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

int launch_diag2(Context *ctx, State *st, int first, int second, const double *d) {
    uint64_t total = (uint64_t)st->batch << st->n;
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_GATE, 32.0 * (double)total);
    hipLaunchKernelGGL(k_diag2, dim3(grid_for(ctx, total)), dim3(kBlock), 0, ctx->stream, st->amp, first, second, total,
                       d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7]);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

int launch_swap(Context *ctx, State *st, int first, int second) {
    int log_groups = st->n - 2;
    uint64_t total = (uint64_t)st->batch << log_groups;
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_GATE, 32.0 * (double)total * 2.0);
    hipLaunchKernelGGL(k_swap, dim3(grid_for(ctx, total)), dim3(kBlock), 0, ctx->stream, st->amp, st->n, first,
                       second, log_groups, total);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

int launch_copy(Context *ctx, const double2 *src, double2 *dst, uint64_t count) {
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_COPY, 32.0 * (double)count);
    const uint64_t blocks = (count + 4 * kBlock - 1) / (4 * kBlock);
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(kBlock), 0, ctx->stream, src, dst, count);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

namespace {

constexpr int kMfmaTileSteps[kMfmaTiles] = {1, 1, 1, 2, 2, 2};  // 16-group tiles per wavefront step
constexpr int kMfmaPerCu[kMfmaTiles] = {16, 8, 32, 16, 8, 32};   // resident workgroups per CU (grid cap)

template <int K>
void launch_mfma_variant(Context *ctx, double2 *amp, int n, int steps_per_wave, uint64_t grid, uint64_t fixed,
                         const uint64_t *qo, int log_groups, uint64_t total, const double2 *m) {
    if (steps_per_wave == 2)
        hipLaunchKernelGGL((k_dense_mfma<K, 2>), dim3((unsigned)grid), dim3(kBlock), 0, ctx->stream, amp, n, fixed,
                           qo[0], qo[1], qo[2], qo[3], qo[4], log_groups, total, m);
    else
        hipLaunchKernelGGL((k_dense_mfma<K, 1>), dim3((unsigned)grid), dim3(kBlock), 0, ctx->stream, amp, n, fixed,
                           qo[0], qo[1], qo[2], qo[3], qo[4], log_groups, total, m);
}

// One k-qubit block (k = 4, 5) on the matrix cores with tile shape `shape`.  A two-tile shape needs an even
// tile count; otherwise its one-tile counterpart runs.
int launch_dense_mfma(Context *ctx, double2 *amp, int n, int k, int shape, uint64_t fixed, const uint64_t *qo,
                      int log_groups, uint64_t total, const double2 *m) {
    const uint64_t tiles = total >> 4;
    const int T = (kMfmaTileSteps[shape] == 2 && tiles % 2 == 0) ? 2 : 1;
    const uint64_t steps = tiles / T;
    uint64_t blocks = (steps + kBlock / 64 - 1) / (kBlock / 64);
    const uint64_t cap = (uint64_t)ctx->cu_count * kMfmaPerCu[shape];
    if (blocks > cap) blocks = cap;
    if (k == 4)
        launch_mfma_variant<4>(ctx, amp, n, T, blocks, fixed, qo, log_groups, total, m);
    else
        launch_mfma_variant<5>(ctx, amp, n, T, blocks, fixed, qo, log_groups, total, m);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

// The tile-shape autotuner: every shape applies a k-qubit block (qubits 1, 6, 11, 16[, 21] -- block qubits
// below and above the 4 coalesced positions) to a scratch state of 2^26 amplitudes (2^24 when that does not
// fit), one warm-up launch and three timed ones each; the fastest shape is kept for the context, where a shape
// replaces the earlier choice only when at least 3% faster (shape 0 is the untuned kernel).  Runs once per
// context and block size, before the first matrix-core block.
int tune_dense_mfma(Context *ctx, int k) {
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    DevBuf buf;
    int n = 26;
    const size_t mat_bytes = (size_t)16 << (2 * k);
    if (buf.alloc(((size_t)16 << n) + mat_bytes)) {
        n = 24;
        if (buf.alloc(((size_t)16 << n) + mat_bytes)) {
            ctx->mfma_pick[k - 4] = 0;  // no room for a scratch state: the untuned shape
            return SC_OK;
        }
    }
    double2 *amp = (double2 *)buf.p;
    const double2 *m = (const double2 *)((char *)buf.p + ((size_t)16 << n));
    SC_HIP_CHECK(hipMemsetAsync(buf.p, 0, ((size_t)16 << n) + mat_bytes, ctx->stream));
    const int qubits[5] = {1, 6, 11, 16, 21};
    uint64_t qo[5] = {0, 0, 0, 0, 0}, fixed = 0;
    for (int j = 0; j < k; j++) {
        qo[j] = 1ULL << qubits[j];
        fixed |= qo[j];
    }
    const int log_groups = n - k;
    const uint64_t total = 1ULL << log_groups;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SC_HIP_CHECK(hipEventCreate(&e0));
    SC_HIP_CHECK(hipEventCreate(&e1));
    int best = 0, rc = SC_OK;
    float best_ms = 0.0f;
    for (int shape = 0; shape < kMfmaTiles && rc == SC_OK; shape++) {
        rc = launch_dense_mfma(ctx, amp, n, k, shape, fixed, qo, log_groups, total, m);
        if (rc == SC_OK && hipEventRecord(e0, ctx->stream) != hipSuccess) rc = fail(SC_E_HIP, "hipEventRecord failed");
        for (int r = 0; r < 3 && rc == SC_OK; r++) rc = launch_dense_mfma(ctx, amp, n, k, shape, fixed, qo, log_groups, total, m);
        if (rc == SC_OK && hipEventRecord(e1, ctx->stream) != hipSuccess) rc = fail(SC_E_HIP, "hipEventRecord failed");
        float ms = 0.0f;
        if (rc == SC_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = fail(SC_E_HIP, "timing the dense-block tile shapes failed");
        if (rc == SC_OK && (shape == 0 || ms < 0.97f * best_ms)) {
            best = shape;
            best_ms = ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    ctx->mfma_pick[k - 4] = best;
    return SC_OK;
}

}  // namespace

int launch_dense(Context *ctx, State *st, int k, const int *qubits, const double *mat_dev) {
    uint64_t qo[5] = {0, 0, 0, 0, 0};
    uint64_t fixed = 0;
    for (int j = 0; j < k; j++) {
        qo[j] = 1ULL << qubits[j];
        fixed |= qo[j];
    }
    int log_groups = st->n - k;
    uint64_t total = (uint64_t)st->batch << log_groups;
    unsigned grid = grid_for(ctx, total);
    const double2 *m = reinterpret_cast<const double2 *>(mat_dev);
    // matrix-core variant (blocks with whole 16-group tiles, n - k >= 4): 5-qubit blocks by default, where the
    // VALU kernel is arithmetic-bound (17.6 ms vs 7.0 ms per 30-qubit sweep on MI355X); 4-qubit blocks with
    // SC_OPT_MFMA 2 (the VALU kernel is at HBM speed there: 6.8 ms vs 7.2 ms)
    const bool mfma = ((ctx->mfma >= 1 && k == 5) || (ctx->mfma >= 2 && k == 4)) && log_groups >= 4;
    int shape = ctx->mfma_tile;
    if (mfma && shape < 0) {
        if (ctx->mfma_pick[k - 4] < 0) {
            int rc = tune_dense_mfma(ctx, k);  // outside the profiled region
            if (rc) return rc;
        }
        shape = ctx->mfma_pick[k - 4];
    }
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_FUSED, 32.0 * (double)((uint64_t)st->batch << st->n));
    if (mfma) return launch_dense_mfma(ctx, st->amp, st->n, k, shape, fixed, qo, log_groups, total, m);
    switch (k) {
    case 1:
        hipLaunchKernelGGL(k_dense<1>, dim3(grid), dim3(kBlock), 0, ctx->stream, st->amp, st->n, fixed, qo[0], qo[1],
                           qo[2], qo[3], qo[4], log_groups, total, m);
        break;
    case 2:
        hipLaunchKernelGGL(k_dense<2>, dim3(grid), dim3(kBlock), 0, ctx->stream, st->amp, st->n, fixed, qo[0], qo[1],
                           qo[2], qo[3], qo[4], log_groups, total, m);
        break;
    case 3:
        hipLaunchKernelGGL(k_dense<3>, dim3(grid), dim3(kBlock), 0, ctx->stream, st->amp, st->n, fixed, qo[0], qo[1],
                           qo[2], qo[3], qo[4], log_groups, total, m);
        break;
    case 4:
        hipLaunchKernelGGL(k_dense<4>, dim3(grid), dim3(kBlock), 0, ctx->stream, st->amp, st->n, fixed, qo[0], qo[1],
                           qo[2], qo[3], qo[4], log_groups, total, m);
        break;
    case 5:
        hipLaunchKernelGGL(k_dense<5>, dim3(grid), dim3(kBlock), 0, ctx->stream, st->amp, st->n, fixed, qo[0], qo[1],
                           qo[2], qo[3], qo[4], log_groups, total, m);
        break;
    default: return fail(SC_E_ARG, "dense block size must be 1..5");
    }
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

}  // namespace sc
