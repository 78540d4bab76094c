// sc_noise.hip -- batched Monte-Carlo Kraus unravelling (many trajectories of one circuit).
//
// Per scheduled single-qubit channel the reference (StatevectorEngine.java:463-491, :426-439, :590-600)
// sweeps the state once per Kraus operator for the branch weights, once to apply the chosen operator
// and once more for the LASSQ norm.  Here a whole batch of trajectories advances together:
//   k_kraus_partial  one read sweep: all K branch weights of every trajectory (fixed-order partials)
//   k_kraus_select   one thread per trajectory: fixed-order total, r = u * total (nextDouble(bound)
//                    semantics), sampleWeightedIndex, renormalisation factor 1/sqrt(w_chosen)
//   k_kraus_apply    one read+write sweep: K_chosen * factor on every trajectory
// ||K_j psi||^2 is exactly the chosen branch weight, so the norm sweep of the reference is not repeated;
// the factor differs from the reference's LASSQ-based one only by rounding (<= a few ulp).
#include "sc_internal.hpp"

namespace sc {

namespace {

constexpr int kBlock = 256;
constexpr int kPartials = 64;  // partial sums per trajectory

struct Kraus4 {
    Mat2 k[4];
};

__device__ __forceinline__ void apply2(const double *g, double2 a0, double2 a1, double2 &o0, double2 &o1) {
    o0.x = g[0] * a0.x - g[1] * a0.y + g[2] * a1.x - g[3] * a1.y;
    o0.y = g[0] * a0.y + g[1] * a0.x + g[2] * a1.y + g[3] * a1.x;
    o1.x = g[4] * a0.x - g[5] * a0.y + g[6] * a1.x - g[7] * a1.y;
    o1.y = g[4] * a0.y + g[5] * a0.x + g[6] * a1.y + g[7] * a1.x;
}

// grid (kPartials, batch): partial[(elem * kPartials + block) * 4 + j]
__global__ void __launch_bounds__(kBlock) k_kraus_partial(const double2 *__restrict__ amp, int n, int target, int K,
                                                          Kraus4 kr, double *__restrict__ partial) {
    __shared__ double sh[4][kBlock];
    const uint64_t elem = blockIdx.y;
    const double2 *base = amp + (elem << n);
    const uint64_t pairs = 1ULL << (n - 1);
    const uint64_t stride = 1ULL << target;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < pairs; i += (uint64_t)kPartials * kBlock) {
        uint64_t lo = i & (stride - 1ULL);
        uint64_t idx0 = ((i ^ lo) << 1) | lo;
        double2 a0 = base[idx0], a1 = base[idx0 | stride];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (j < K) {
                double2 o0, o1;
                apply2(kr.k[j].v, a0, a1, o0, o1);
                acc[j] += o0.x * o0.x + o0.y * o0.y + o1.x * o1.x + o1.y * o1.y;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) sh[j][threadIdx.x] = acc[j];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
#pragma unroll
            for (int j = 0; j < 4; j++) sh[j][threadIdx.x] += sh[j][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 4) partial[(elem * kPartials + blockIdx.x) * 4 + threadIdx.x] = sh[threadIdx.x][0];
}

// One thread per trajectory.  u[elem * u_stride + draw] is the trajectory's SplittableRandom.nextDouble().
__global__ void k_kraus_select(const double *__restrict__ partial, int K, int batch, const double *__restrict__ u,
                               uint64_t u_stride, uint64_t draw, int *__restrict__ choice, double *__restrict__ scale,
                               int *__restrict__ err) {
    int elem = blockIdx.x * blockDim.x + threadIdx.x;
    if (elem >= batch) return;
    double w[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < kPartials; b++)
        for (int j = 0; j < K; j++) w[j] += partial[((uint64_t)elem * kPartials + b) * 4 + j];
    double total = 0.0;
    bool bad = false;
    for (int j = 0; j < K; j++) {
        if (!isfinite(w[j]) || w[j] < 0.0) bad = true;
        total += w[j];
    }
    if (!isfinite(total) || total <= 1e-15) bad = true;  // MIN_KRAUS_WEIGHT, StatevectorEngine.java:36
    if (bad) {
        err[0] = 1;
        choice[elem] = 0;
        scale[elem] = 1.0;
        return;
    }
    double r = u[(uint64_t)elem * u_stride + draw] * total;  // SplittableRandom.nextDouble(bound)
    if (r >= total) r = nextafter(total, 0.0);
    int idx = K - 1;
    double cum = 0.0;
    for (int j = 0; j < K; j++) {  // sampleWeightedIndex :579-588
        cum += w[j];
        if (r < cum) {
            idx = j;
            break;
        }
    }
    choice[elem] = idx;
    double ns = w[idx];
    if (!isfinite(ns) || ns == 0.0) {
        err[0] = 2;
        scale[elem] = 1.0;
        return;
    }
    scale[elem] = fabs(ns - 1.0) > 1e-12 ? 1.0 / sqrt(ns) : 1.0;  // renormalizeState :590-600
}

__global__ void __launch_bounds__(kBlock) k_kraus_apply(double2 *__restrict__ amp, int n, int target, Kraus4 kr,
                                                        const int *__restrict__ choice, const double *__restrict__ scale,
                                                        uint64_t total) {
    const uint64_t stride = 1ULL << target;
    const int lg = n - 1;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < total; i += (uint64_t)gridDim.x * kBlock) {
        uint64_t elem = i >> lg;
        uint64_t p = i & ((1ULL << lg) - 1ULL);
        uint64_t lo = p & (stride - 1ULL);
        uint64_t idx0 = ((p ^ lo) << 1) | lo;
        double2 *base = amp + (elem << n);
        int c = choice[elem];
        double f = scale[elem];
        const double *g = c == 0 ? kr.k[0].v : c == 1 ? kr.k[1].v : c == 2 ? kr.k[2].v : kr.k[3].v;
        double2 a0 = base[idx0], a1 = base[idx0 | stride];
        double2 o0, o1;
        apply2(g, a0, a1, o0, o1);
        base[idx0] = make_double2(o0.x * f, o0.y * f);
        base[idx0 | stride] = make_double2(o1.x * f, o1.y * f);
    }
}

}  // namespace

// Scratch layout expected in `work`: partial [batch*kPartials*4] doubles, choice [batch] ints, scale [batch]
// doubles, err [1] int.  Returns the size needed when work == nullptr.
size_t noise_work_bytes(int batch) {
    return (size_t)batch * kPartials * 4 * sizeof(double) + (size_t)batch * (sizeof(int) + sizeof(double)) + 64;
}

int launch_channel_batched(Context *ctx, State *st, int target, const double *kraus, int K, const double *u_dev,
                           uint64_t u_stride, uint64_t draw, void *work) {
    Kraus4 kr;
    for (int j = 0; j < 4; j++)
        for (int e = 0; e < 8; e++) kr.k[j].v[e] = j < K ? kraus[8 * j + e] : 0.0;
    double *partial = (double *)work;
    double *scale = partial + (size_t)st->batch * kPartials * 4;
    int *choice = (int *)(scale + st->batch);
    int *err = choice + st->batch;
    {
        ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_NOISE, 16.0 * (double)((uint64_t)st->batch << st->n));
        hipLaunchKernelGGL(k_kraus_partial, dim3(kPartials, st->batch), dim3(kBlock), 0, ctx->stream, st->amp, st->n,
                           target, K, kr, partial);
        SC_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_kraus_select, dim3((st->batch + 63) / 64), dim3(64), 0, ctx->stream, partial, K, st->batch,
                       u_dev, u_stride, draw, choice, scale, err);
    SC_HIP_CHECK(hipGetLastError());
    uint64_t total = (uint64_t)st->batch << (st->n - 1);
    uint64_t blocks = std::min<uint64_t>((total + kBlock - 1) / kBlock, (uint64_t)ctx->cu_count * 16);
    {
        ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_NOISE, 32.0 * (double)((uint64_t)st->batch << st->n));
        hipLaunchKernelGGL(k_kraus_apply, dim3((unsigned)blocks), dim3(kBlock), 0, ctx->stream, st->amp, st->n, target,
                           kr, choice, scale, total);
        SC_HIP_CHECK(hipGetLastError());
    }
    return SC_OK;
}

int *noise_err_flag(void *work, int batch) {
    double *partial = (double *)work;
    double *scale = partial + (size_t)batch * kPartials * 4;
    int *choice = (int *)(scale + batch);
    return choice + batch;
}

}  // namespace sc
