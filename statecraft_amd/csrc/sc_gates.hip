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
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_FUSED, 32.0 * (double)((uint64_t)st->batch << st->n));
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
