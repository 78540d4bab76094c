// sc_measure.hip -- probabilities, deterministic reductions, norms, checksums and Kraus branch weights.
// (The exact cumulative sums and the shot search are in sc_seqsum.hip.)
//
// Every reduction is two-level with a fixed grid and a fixed LDS tree, so a
// result is bit-identical from run to run on one device (no atomics).
#include "sc_internal.hpp"

namespace sc {

namespace {

constexpr int kBlock = 256;
constexpr int kRedGrid = 1024;  // fixed number of partial sums per reduction

// p = re*re + im*im with separately rounded products (StatevectorEngine.java:282).
__device__ __forceinline__ double prob_nofma(double2 a) { return __dadd_rn(__dmul_rn(a.x, a.x), __dmul_rn(a.y, a.y)); }

template <int K>
__device__ __forceinline__ void block_reduce(double (&v)[K], double *out) {
    __shared__ double sh[K][kBlock];
#pragma unroll
    for (int k = 0; k < K; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int k = 0; k < K; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) out[k] = sh[k][0];
    }
}

// ---- probabilities ---------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_prob_full(const double2 *__restrict__ amp, double *__restrict__ out,
                                                      uint64_t count) {
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < count; i += (uint64_t)gridDim.x * kBlock)
        out[i] = prob_nofma(amp[i]);
}

__device__ __forceinline__ uint64_t pdep64(uint64_t x, uint64_t mask) {
    uint64_t r = 0;
    while (mask) {
        uint64_t low = mask & (~mask + 1ULL);
        if (x & 1ULL) r |= low;
        x >>= 1;
        mask &= mask - 1ULL;
    }
    return r;
}

// Marginal, many outcomes: one thread per outcome walks all complement indices in increasing order.
__global__ void __launch_bounds__(kBlock) k_prob_marginal_narrow(const double2 *__restrict__ amp, uint64_t mmask,
                                                                 uint64_t cmask, uint64_t outcomes,
                                                                 double *__restrict__ out) {
    for (uint64_t o = blockIdx.x * (uint64_t)kBlock + threadIdx.x; o < outcomes; o += (uint64_t)gridDim.x * kBlock) {
        uint64_t obase = pdep64(o, mmask);
        double acc = 0.0;
        uint64_t c = 0;
        do {
            acc += prob_nofma(amp[obase | c]);
            c = ((c | ~cmask) + 1ULL) & cmask;  // next submask of cmask in increasing order
        } while (c != 0);
        out[o] = acc;
    }
}

// ---- generic reductions -----------------------------------------------------
template <int K>
__global__ void __launch_bounds__(kBlock) k_sum_final(const double *__restrict__ partial, int n_partial,
                                                      double *__restrict__ out) {
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0.0;
    for (int i = threadIdx.x; i < n_partial; i += kBlock) {
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += partial[(size_t)i * K + k];
    }
    block_reduce<K>(acc, out);
}

__global__ void __launch_bounds__(kBlock) k_norm_stage1(const double2 *__restrict__ amp, uint64_t count,
                                                        double *__restrict__ partial) {
    double acc[1] = {0.0};
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < count; i += (uint64_t)gridDim.x * kBlock) {
        double2 a = amp[i];
        acc[0] = fma(a.x, a.x, acc[0]);
        acc[0] = fma(a.y, a.y, acc[0]);
    }
    block_reduce<1>(acc, partial + blockIdx.x);
}

__device__ __forceinline__ double checksum_weight(uint64_t i) {
    uint64_t z = (i + 1ULL) * 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * 0x1.0p-53;
}

__global__ void __launch_bounds__(kBlock) k_checksum_stage1(const double2 *__restrict__ amp, uint64_t count,
                                                            double *__restrict__ partial) {
    double acc[2] = {0.0, 0.0};
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < count; i += (uint64_t)gridDim.x * kBlock) {
        double p = prob_nofma(amp[i]);
        acc[0] += p;
        acc[1] = fma(p, checksum_weight(i), acc[1]);
    }
    double r[2];
    block_reduce<2>(acc, r);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = r[0];
        partial[2 * blockIdx.x + 1] = r[1];
    }
}

// Kraus branch weights (StatevectorEngine.java:511-577) for up to 4 single-qubit operators in one sweep.
template <int K>
__global__ void __launch_bounds__(kBlock) k_kraus_stage1(const double2 *__restrict__ amp, int target,
                                                         uint64_t pairs, Mat2 k0, Mat2 k1, Mat2 k2, Mat2 k3,
                                                         double *__restrict__ partial) {
    const Mat2 *ks[4] = {&k0, &k1, &k2, &k3};
    double acc[K];
#pragma unroll
    for (int j = 0; j < K; j++) acc[j] = 0.0;
    const uint64_t stride = 1ULL << target;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < pairs; i += (uint64_t)gridDim.x * kBlock) {
        uint64_t lo = i & (stride - 1ULL);
        uint64_t idx0 = ((i ^ lo) << 1) | lo;
        double2 a0 = amp[idx0], a1 = amp[idx0 | stride];
#pragma unroll
        for (int j = 0; j < K; j++) {
            const double *g = ks[j]->v;
            double r0 = g[0] * a0.x - g[1] * a0.y + g[2] * a1.x - g[3] * a1.y;
            double i0 = g[0] * a0.y + g[1] * a0.x + g[2] * a1.y + g[3] * a1.x;
            double r1 = g[4] * a0.x - g[5] * a0.y + g[6] * a1.x - g[7] * a1.y;
            double i1 = g[4] * a0.y + g[5] * a0.x + g[6] * a1.y + g[7] * a1.x;
            acc[j] += r0 * r0 + i0 * i0 + r1 * r1 + i1 * i1;
        }
    }
    double r[K];
    block_reduce<K>(acc, r);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < K; j++) partial[(size_t)blockIdx.x * K + j] = r[j];
    }
}

__global__ void __launch_bounds__(kBlock) k_scale_state(double2 *__restrict__ x, uint64_t count, double f) {
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < count; i += (uint64_t)gridDim.x * kBlock) {
        double2 a = x[i];
        x[i] = make_double2(f * a.x, f * a.y);
    }
}

// Largest |amp[i] - ref_i| with ref_i = a at i == index and b elsewhere (a two-level reference state), per block.
__global__ void __launch_bounds__(kBlock) k_maxdev_stage1(const double2 *__restrict__ amp, uint64_t count,
                                                          uint64_t index, double2 a, double2 b,
                                                          double *__restrict__ partial) {
    __shared__ double sh[kBlock];
    double m = 0.0;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < count; i += (uint64_t)gridDim.x * kBlock) {
        const double2 x = amp[i];
        const double2 r = i == index ? a : b;
        const double dr = x.x - r.x, di = x.y - r.y;
        const double d = sqrt(dr * dr + di * di);
        m = (d > m || d != d) ? d : m;  // NaN propagates
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double o = sh[threadIdx.x + s];
            if (o > sh[threadIdx.x] || o != o) sh[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kBlock) k_max_final(const double *__restrict__ partial, int n_partial,
                                                      double *__restrict__ out) {
    __shared__ double sh[kBlock];
    double m = 0.0;
    for (int i = threadIdx.x; i < n_partial; i += kBlock) m = (partial[i] > m || partial[i] != partial[i]) ? partial[i] : m;
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double o = sh[threadIdx.x + s];
            if (o > sh[threadIdx.x] || o != o) sh[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

unsigned grid_for(Context *ctx, uint64_t total) {
    uint64_t blocks = (total + kBlock - 1) / kBlock;
    uint64_t cap = (uint64_t)ctx->cu_count * 16;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    return (unsigned)blocks;
}

// Finish a K-wide reduction of `nblocks` partials into host memory.
template <int K>
int finish_reduce(Context *ctx, double *partial, int nblocks, double *result_host) {
    double *final_dev = partial + (size_t)kRedGrid * K;
    hipLaunchKernelGGL(k_sum_final<K>, dim3(1), dim3(kBlock), 0, ctx->stream, partial, nblocks, final_dev);
    SC_HIP_CHECK(hipGetLastError());
    SC_HIP_CHECK(hipMemcpyAsync(result_host, final_dev, K * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

}  // namespace

int launch_probabilities_full(Context *ctx, State *st, int elem, double *out_dev) {
    uint64_t count = st->dim();
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_MEASURE, 24.0 * (double)count);
    hipLaunchKernelGGL(k_prob_full, dim3(grid_for(ctx, count)), dim3(kBlock), 0, ctx->stream,
                       st->amp + ((uint64_t)elem << st->n), out_dev, count);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

int launch_probabilities_marginal(Context *ctx, State *st, int elem, const int32_t *measured, int m,
                                  double *out_dev) {
    uint64_t mmask = 0;
    for (int k = 0; k < m; k++) mmask |= 1ULL << measured[k];
    const uint64_t full = st->dim() - 1ULL;
    const uint64_t cmask = full & ~mmask;
    const double2 *amp = st->amp + ((uint64_t)elem << st->n);
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_MEASURE, 16.0 * (double)st->dim());
    hipLaunchKernelGGL(k_prob_marginal_narrow, dim3(grid_for(ctx, 1ULL << m)), dim3(kBlock), 0, ctx->stream, amp,
                       mmask, cmask, 1ULL << m, out_dev);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

int launch_scale_state(Context *ctx, double2 *amp, uint64_t count, double factor) {
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_NOISE, 32.0 * (double)count);
    hipLaunchKernelGGL(k_scale_state, dim3(grid_for(ctx, count)), dim3(kBlock), 0, ctx->stream, amp, count, factor);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

int launch_norm_sq(Context *ctx, const double2 *amp, uint64_t count, double *result_host) {
    int rc = ensure_scratch(ctx, (size_t)(kRedGrid + 1) * sizeof(double));
    if (rc) return rc;
    double *partial = (double *)ctx->scratch;
    {
        ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_NOISE, 16.0 * (double)count);
        hipLaunchKernelGGL(k_norm_stage1, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, amp, count, partial);
        SC_HIP_CHECK(hipGetLastError());
    }
    return finish_reduce<1>(ctx, partial, kRedGrid, result_host);
}

int launch_checksum(Context *ctx, const double2 *amp, uint64_t count, double *out_host) {
    int rc = ensure_scratch(ctx, (size_t)(kRedGrid + 1) * 2 * sizeof(double));
    if (rc) return rc;
    double *partial = (double *)ctx->scratch;
    {
        ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_MEASURE, 16.0 * (double)count);
        hipLaunchKernelGGL(k_checksum_stage1, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, amp, count, partial);
        SC_HIP_CHECK(hipGetLastError());
    }
    return finish_reduce<2>(ctx, partial, kRedGrid, out_host);
}

int launch_kraus_weights(Context *ctx, const double2 *amp, int n, int target, const double *kraus, int k,
                         double *weights_host) {
    if (k < 1 || k > 4) return fail(SC_E_ARG, "kraus operator count must be 1..4");
    Mat2 m[4];
    for (int j = 0; j < 4; j++)
        for (int e = 0; e < 8; e++) m[j].v[e] = j < k ? kraus[8 * j + e] : 0.0;
    uint64_t pairs = 1ULL << (n - 1);
    int rc = ensure_scratch(ctx, (size_t)(kRedGrid + 1) * 4 * sizeof(double));
    if (rc) return rc;
    double *partial = (double *)ctx->scratch;
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_NOISE, 16.0 * (double)(pairs * 2));
    switch (k) {
    case 1:
        hipLaunchKernelGGL(k_kraus_stage1<1>, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, amp, target, pairs, m[0],
                           m[1], m[2], m[3], partial);
        break;
    case 2:
        hipLaunchKernelGGL(k_kraus_stage1<2>, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, amp, target, pairs, m[0],
                           m[1], m[2], m[3], partial);
        break;
    case 3:
        hipLaunchKernelGGL(k_kraus_stage1<3>, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, amp, target, pairs, m[0],
                           m[1], m[2], m[3], partial);
        break;
    default:
        hipLaunchKernelGGL(k_kraus_stage1<4>, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, amp, target, pairs, m[0],
                           m[1], m[2], m[3], partial);
        break;
    }
    SC_HIP_CHECK(hipGetLastError());
    switch (k) {
    case 1: return finish_reduce<1>(ctx, partial, kRedGrid, weights_host);
    case 2: return finish_reduce<2>(ctx, partial, kRedGrid, weights_host);
    case 3: return finish_reduce<3>(ctx, partial, kRedGrid, weights_host);
    default: return finish_reduce<4>(ctx, partial, kRedGrid, weights_host);
    }
}

int launch_max_deviation(Context *ctx, const double2 *amp, uint64_t count, uint64_t index, double2 a, double2 b,
                         double *out_host) {
    int rc = ensure_scratch(ctx, (size_t)(kRedGrid + 1) * sizeof(double));
    if (rc) return rc;
    double *partial = (double *)ctx->scratch;
    {
        ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_MEASURE, 16.0 * (double)count);
        hipLaunchKernelGGL(k_maxdev_stage1, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, amp, count, index, a, b,
                           partial);
        SC_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(kBlock), 0, ctx->stream, partial, kRedGrid, partial + kRedGrid);
    SC_HIP_CHECK(hipGetLastError());
    SC_HIP_CHECK(hipMemcpyAsync(out_host, partial + kRedGrid, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

}  // namespace sc
