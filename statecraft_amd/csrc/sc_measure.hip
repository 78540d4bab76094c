// sc_measure.hip -- probabilities, deterministic reductions, inclusive scan,
// inverse-CDF shot search, norms, checksums and Kraus branch weights.
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

// Marginal, few outcomes: one block per (outcome, slice); lanes walk the complement index.
// partial[slice * outcomes + o] = sum over the slice's complement indices.
__global__ void __launch_bounds__(kBlock) k_prob_marginal_wide(const double2 *__restrict__ amp, uint64_t mmask,
                                                               uint64_t cmask, int cbits, int log_slices, int m,
                                                               double *__restrict__ partial) {
    const uint64_t o = blockIdx.x;
    const uint64_t slice = blockIdx.y;
    const uint64_t per_slice = 1ULL << (cbits - log_slices);
    const uint64_t obase = pdep64(o, mmask);
    double acc[1] = {0.0};
    for (uint64_t c = threadIdx.x; c < per_slice; c += kBlock) {
        uint64_t cc = slice * per_slice + c;
        acc[0] += prob_nofma(amp[obase | pdep64(cc, cmask)]);
    }
    double r[1];
    block_reduce<1>(acc, r);
    if (threadIdx.x == 0) partial[(slice << m) + o] = r[0];
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

__global__ void __launch_bounds__(kBlock) k_sum_slices(const double *__restrict__ partial, int log_slices, int m,
                                                       double *__restrict__ out) {
    const uint64_t outcomes = 1ULL << m;
    for (uint64_t o = blockIdx.x * (uint64_t)kBlock + threadIdx.x; o < outcomes; o += (uint64_t)gridDim.x * kBlock) {
        double acc = 0.0;
        for (uint64_t s = 0; s < (1ULL << log_slices); s++) acc += partial[(s << m) + o];
        out[o] = acc;
    }
}

// ---- generic reductions -----------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_sum_stage1(const double *__restrict__ x, uint64_t count,
                                                       double *__restrict__ partial) {
    double acc[1] = {0.0};
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < count; i += (uint64_t)gridDim.x * kBlock)
        acc[0] += x[i];
    block_reduce<1>(acc, partial + blockIdx.x);
}

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

__global__ void __launch_bounds__(kBlock) k_scale(double *__restrict__ x, uint64_t count, double f) {
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < count; i += (uint64_t)gridDim.x * kBlock)
        x[i] *= f;
}

__global__ void __launch_bounds__(kBlock) k_scale_state(double2 *__restrict__ x, uint64_t count, double f) {
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < count; i += (uint64_t)gridDim.x * kBlock) {
        double2 a = x[i];
        x[i] = make_double2(f * a.x, f * a.y);
    }
}

// ---- inclusive scan: tiles of kScanTile values ------------------------------
constexpr int kScanPer = 16;
constexpr int kScanTile = kBlock * kScanPer;  // 4096

// Sum of each tile.
__global__ void __launch_bounds__(kBlock) k_tile_sums(const double *__restrict__ x, uint64_t count,
                                                      double *__restrict__ sums) {
    uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    double acc[1] = {0.0};
    // coalesced: lane t sums elements t, t + 256, ... of the tile (fixed order)
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
        uint64_t i = base + threadIdx.x + (uint64_t)j * kBlock;
        if (i < count) acc[0] += x[i];
    }
    block_reduce<1>(acc, sums + blockIdx.x);
}

// In-place inclusive scan of each tile, plus offsets[blockIdx] (exclusive prefix of tile sums) when non-null.
// The tile is loaded and stored coalesced through LDS (one pad slot per 16 values keeps the per-thread
// runs of 16 on distinct banks); each thread scans 16 consecutive values, then the thread totals are scanned.
__global__ void __launch_bounds__(kBlock) k_tile_scan(double *__restrict__ x, uint64_t count,
                                                      const double *__restrict__ offsets) {
    __shared__ double sh[kBlock];
    __shared__ double tile[kScanTile + kScanTile / kScanPer];
    uint64_t base = (uint64_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
        int l = threadIdx.x + j * kBlock;
        uint64_t i = base + l;
        tile[l + l / kScanPer] = i < count ? x[i] : 0.0;
    }
    __syncthreads();
    double v[kScanPer];
    double run = 0.0;
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
        int l = threadIdx.x * kScanPer + j;
        run += tile[l + l / kScanPer];
        v[j] = run;
    }
    // Hillis-Steele scan of per-thread totals (fixed order).
    sh[threadIdx.x] = run;
    __syncthreads();
    for (int s = 1; s < kBlock; s <<= 1) {
        double add = (int)threadIdx.x >= s ? sh[threadIdx.x - s] : 0.0;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    double prefix = (threadIdx.x > 0 ? sh[threadIdx.x - 1] : 0.0);
    if (offsets) prefix += (blockIdx.x > 0 ? offsets[blockIdx.x - 1] : 0.0);
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
        int l = threadIdx.x * kScanPer + j;
        tile[l + l / kScanPer] = prefix + v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
        int l = threadIdx.x + j * kBlock;
        uint64_t i = base + l;
        if (i < count) x[i] = tile[l + l / kScanPer];
    }
}

__global__ void k_set_last_one(double *x, uint64_t count) { x[count - 1] = 1.0; }

// ---- full-register sampling without a probability array ------------------------------------------------
// k_tile_scan's in-tile order applied to the probabilities of tile t of the state (|a|^2, times scale):
// the tile goes to LDS (padded as there), each thread accumulates its 16 consecutive values into v[], and
// the thread totals are scanned in sh[].  Returns the thread's exclusive prefix within the tile; the tile
// total is sh[kBlock - 1].
__device__ double state_tile_scan(const double2 *__restrict__ amp, uint64_t base, uint64_t count, double scale,
                                  double *tile, double *sh, double (&v)[kScanPer]) {
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
        const int l = threadIdx.x + j * kBlock;
        const uint64_t i = base + l;
        double p = i < count ? prob_nofma(amp[i]) : 0.0;
        if (scale != 1.0) p *= scale;
        tile[l + l / kScanPer] = p;
    }
    __syncthreads();
    double run = 0.0;
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
        const int l = threadIdx.x * kScanPer + j;
        run += tile[l + l / kScanPer];
        v[j] = run;
    }
    sh[threadIdx.x] = run;
    __syncthreads();
    for (int s = 1; s < kBlock; s <<= 1) {
        const double add = (int)threadIdx.x >= s ? sh[threadIdx.x - s] : 0.0;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    return threadIdx.x > 0 ? sh[threadIdx.x - 1] : 0.0;
}

// totals[t]: the last in-tile cumulative value of tile t
__global__ void __launch_bounds__(kBlock) k_state_tile_totals(const double2 *__restrict__ amp, uint64_t count,
                                                              double scale, double *__restrict__ totals) {
    __shared__ double sh[kBlock];
    __shared__ double tile[kScanTile + kScanTile / kScanPer];
    double v[kScanPer];
    (void)state_tile_scan(amp, (uint64_t)blockIdx.x * kScanTile, count, scale, tile, sh, v);
    if (threadIdx.x == kBlock - 1) totals[blockIdx.x] = sh[kBlock - 1];
}

// One workgroup per shot: the tile whose cumulative end (ends = inclusive scan of the tile totals, last
// forced to 1.0) first reaches u, then inside it the smallest index whose cumulative value reaches u -- the
// result of java.util.Arrays.binarySearch + StatevectorEngine.sampleOutcome (:325-338) on the cumulative
// array sc_sample otherwise materialises.
__global__ void __launch_bounds__(kBlock) k_search_state_tiles(const double2 *__restrict__ amp, uint64_t count,
                                                               double scale, const double *__restrict__ ends,
                                                               uint64_t ntiles, const double *__restrict__ u,
                                                               int64_t shots, int64_t *__restrict__ out,
                                                               double *__restrict__ gap) {
    __shared__ double sh[kBlock];
    __shared__ double tile[kScanTile + kScanTile / kScanPer];
    __shared__ uint64_t best[kBlock];
    for (int64_t s = blockIdx.x; s < shots; s += gridDim.x) {
        const double value = u[s];
        uint64_t t = ntiles - 1;
        if (value < 1.0) {
            uint64_t lo = 0, hi = ntiles - 1;  // smallest t with ends[t] >= value (ends[ntiles - 1] = 1.0)
            while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                if (ends[mid] >= value) hi = mid;
                else lo = mid + 1;
            }
            t = lo;
        }
        const uint64_t base = t * (uint64_t)kScanTile;
        double v[kScanPer];
        const double prefix = state_tile_scan(amp, base, count, scale, tile, sh, v) + (t > 0 ? ends[t - 1] : 0.0);
        uint64_t cand = ~0ULL;
#pragma unroll
        for (int j = 0; j < kScanPer; j++) {
            const int l = threadIdx.x * kScanPer + j;
            const uint64_t i = base + l;
            const double c = i == count - 1 ? 1.0 : prefix + v[j];
            tile[l + l / kScanPer] = c;
            if (i < count && cand == ~0ULL && (value >= 1.0 ? i == count - 1 : c >= value)) cand = i;
        }
        best[threadIdx.x] = cand;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) best[threadIdx.x] = best[threadIdx.x] < best[threadIdx.x + w] ? best[threadIdx.x]
                                                                                                  : best[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            uint64_t r = best[0];
            if (r == ~0ULL) r = base + kScanTile < count ? base + kScanTile : count - 1;  // rounding at a tile end
            out[s] = (int64_t)r;
            if (gap) {
                const uint64_t l = r - base;
                double g = fabs(value - (r < base + kScanTile ? tile[l + l / kScanPer] : ends[t]));
                if (r > base && r < base + kScanTile) g = fmin(g, fabs(value - tile[(l - 1) + (l - 1) / kScanPer]));
                else if (r == base && t > 0) g = fmin(g, fabs(value - ends[t - 1]));
                gap[s] = g;
            }
        }
        __syncthreads();
    }
}

// ---- shot search: java.util.Arrays.binarySearch + StatevectorEngine.sampleOutcome (:325-338) -------------
__global__ void __launch_bounds__(kBlock) k_search(const double *__restrict__ cum, int64_t len,
                                                   const double *__restrict__ u, int64_t shots,
                                                   int64_t *__restrict__ out, double *__restrict__ gap) {
    for (int64_t s = blockIdx.x * (int64_t)kBlock + threadIdx.x; s < shots; s += (int64_t)gridDim.x * kBlock) {
        double value = u[s];
        int64_t result;
        if (value >= 1.0) {
            result = len - 1;
        } else {
            int64_t low = 0, high = len - 1, found = -1;
            while (low <= high) {
                int64_t mid = (int64_t)(((uint64_t)low + (uint64_t)high) >> 1);
                double midVal = cum[mid];
                if (midVal < value) {
                    low = mid + 1;
                } else if (midVal > value) {
                    high = mid - 1;
                } else {
                    long long mb = __double_as_longlong(midVal), kb = __double_as_longlong(value);
                    if (mb == kb) {
                        found = mid;
                        break;
                    }
                    if (mb < kb) low = mid + 1;
                    else high = mid - 1;
                }
            }
            if (found >= 0) {
                while (found > 0 && cum[found - 1] >= value) found--;
                result = found;
            } else {
                result = low;
            }
        }
        out[s] = result;
        if (gap) {
            double g = fabs(value - cum[result]);
            if (result > 0) g = fmin(g, fabs(value - cum[result - 1]));
            gap[s] = g;
        }
    }
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
    const int cbits = st->n - m;
    const double2 *amp = st->amp + ((uint64_t)elem << st->n);
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_MEASURE, 16.0 * (double)st->dim());
    if (cbits >= 8 && m <= 12) {
        // slices so that outcomes * slices >= ~4096 blocks
        int log_slices = 0;
        while (m + log_slices < 12 && log_slices < cbits - 8) log_slices++;
        size_t need = ((size_t)1 << (m + log_slices)) * sizeof(double);
        int rc = ensure_scratch(ctx, need);
        if (rc) return rc;
        double *partial = (double *)ctx->scratch;
        hipLaunchKernelGGL(k_prob_marginal_wide, dim3(1u << m, 1u << log_slices), dim3(kBlock), 0, ctx->stream, amp,
                           mmask, cmask, cbits, log_slices, m, partial);
        SC_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_sum_slices, dim3(grid_for(ctx, 1ULL << m)), dim3(kBlock), 0, ctx->stream, partial,
                           log_slices, m, out_dev);
        SC_HIP_CHECK(hipGetLastError());
    } else {
        hipLaunchKernelGGL(k_prob_marginal_narrow, dim3(grid_for(ctx, 1ULL << m)), dim3(kBlock), 0, ctx->stream, amp,
                           mmask, cmask, 1ULL << m, out_dev);
        SC_HIP_CHECK(hipGetLastError());
    }
    return SC_OK;
}

int device_sum(Context *ctx, const double *x, uint64_t count, double *result_host) {
    int rc = ensure_scratch(ctx, (size_t)(kRedGrid + 1) * sizeof(double));
    if (rc) return rc;
    double *partial = (double *)ctx->scratch;
    hipLaunchKernelGGL(k_sum_stage1, dim3(kRedGrid), dim3(kBlock), 0, ctx->stream, x, count, partial);
    SC_HIP_CHECK(hipGetLastError());
    return finish_reduce<1>(ctx, partial, kRedGrid, result_host);
}

int launch_scale(Context *ctx, double *x, uint64_t count, double factor) {
    hipLaunchKernelGGL(k_scale, dim3(grid_for(ctx, count)), dim3(kBlock), 0, ctx->stream, x, count, factor);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

int launch_scale_state(Context *ctx, double2 *amp, uint64_t count, double factor) {
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_NOISE, 32.0 * (double)count);
    hipLaunchKernelGGL(k_scale_state, dim3(grid_for(ctx, count)), dim3(kBlock), 0, ctx->stream, amp, count, factor);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

// Recursive tiled inclusive scan; scratch holds the tile sums of every level.
static int scan_level(Context *ctx, double *x, uint64_t count, double *tmp) {
    if (count <= (uint64_t)kScanTile) {
        hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(kBlock), 0, ctx->stream, x, count, (const double *)nullptr);
        SC_HIP_CHECK(hipGetLastError());
        return SC_OK;
    }
    uint64_t tiles = (count + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)tiles), dim3(kBlock), 0, ctx->stream, x, count, tmp);
    SC_HIP_CHECK(hipGetLastError());
    int rc = scan_level(ctx, tmp, tiles, tmp + tiles);
    if (rc) return rc;
    hipLaunchKernelGGL(k_tile_scan, dim3((unsigned)tiles), dim3(kBlock), 0, ctx->stream, x, count, tmp);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

int launch_cumulative(Context *ctx, double *p, uint64_t count) {
    uint64_t need = 0, c = count;
    while (c > (uint64_t)kScanTile) {
        c = (c + kScanTile - 1) / kScanTile;
        need += c;
    }
    int rc = ensure_scratch(ctx, (size_t)(need + 1) * sizeof(double));
    if (rc) return rc;
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_MEASURE, 24.0 * (double)count);
    rc = scan_level(ctx, p, count, (double *)ctx->scratch);
    if (rc) return rc;
    hipLaunchKernelGGL(k_set_last_one, dim3(1), dim3(1), 0, ctx->stream, p, count);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

int launch_search(Context *ctx, const double *cum, uint64_t count, const double *u_dev, int64_t shots,
                  int64_t *out_dev, double *gap_dev) {
    hipLaunchKernelGGL(k_search, dim3(grid_for(ctx, (uint64_t)shots)), dim3(kBlock), 0, ctx->stream, cum,
                       (int64_t)count, u_dev, shots, out_dev, gap_dev);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

// Sampling over every qubit in index order: no 2^n probability array.  One read of the state gives the
// tile totals (k_tile_scan's order), whose sum is the normalisation (StatevectorEngine.normalize :298-312:
// rescaled only when off by more than 1e-12, then the totals are read again with the factor); the totals'
// inclusive scan, last forced to 1.0, gives the tile ends, and each shot rescans one tile.
int sample_full_state(Context *ctx, State *st, int elem, const double *u_dev, int64_t shots, int64_t *out_dev,
                      double *gap_dev) {
    const uint64_t count = st->dim();
    const uint64_t ntiles = (count + kScanTile - 1) / kScanTile;
    const double2 *amp = st->amp + ((uint64_t)elem << st->n);
    DevBuf ends_buf;
    int rc = ends_buf.alloc(ntiles * sizeof(double));
    if (rc) return rc;
    double *ends = (double *)ends_buf.p;
    {
        ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_MEASURE, 16.0 * (double)count);
        hipLaunchKernelGGL(k_state_tile_totals, dim3((unsigned)ntiles), dim3(kBlock), 0, ctx->stream, amp, count, 1.0,
                           ends);
        SC_HIP_CHECK(hipGetLastError());
    }
    double sum = 0.0;
    if ((rc = device_sum(ctx, ends, ntiles, &sum))) return rc;
    if (sum == 0.0) return fail(SC_E_STATE, "State collapsed to zero probability amplitudes");
    double scale = 1.0;
    if (std::fabs(sum - 1.0) > 1e-12) {
        scale = 1.0 / sum;
        ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_MEASURE, 16.0 * (double)count);
        hipLaunchKernelGGL(k_state_tile_totals, dim3((unsigned)ntiles), dim3(kBlock), 0, ctx->stream, amp, count,
                           scale, ends);
        SC_HIP_CHECK(hipGetLastError());
    }
    if ((rc = launch_cumulative(ctx, ends, ntiles))) return rc;
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_MEASURE, 0.0);
    const unsigned grid = (unsigned)std::min<uint64_t>((uint64_t)shots, (uint64_t)ctx->cu_count * 8);
    hipLaunchKernelGGL(k_search_state_tiles, dim3(grid), dim3(kBlock), 0, ctx->stream, amp, count, scale,
                       (const double *)ends, ntiles, u_dev, shots, out_dev, gap_dev);
    SC_HIP_CHECK(hipGetLastError());
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // ends_buf is released on return
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

}  // namespace sc
