// sc_noise.hip -- batched Monte-Carlo Kraus unravelling (many trajectories of one circuit).
//
// Per scheduled single-qubit channel the reference (StatevectorEngine.java:463-491, :426-439, :590-600)
// sweeps the state once per Kraus operator for the branch weights, once to apply the chosen operator
// and once more for the LASSQ norm.  Here a whole batch of trajectories advances together:
//   k_kraus_partial  one read sweep: all K branch weights of every trajectory (fixed-order partials)
//   k_kraus_select   one thread per trajectory: fixed-order total, r = u * total (nextDouble(bound)
//                    semantics), sampleWeightedIndex, renormalisation factor 1/sqrt(w_chosen)
//   k_kraus_apply    one read+write sweep: K_chosen * factor on every trajectory, accumulating the norm
//                    of what it writes
// ||K_j psi||^2 is exactly the chosen branch weight, so the norm sweep of the reference is not repeated;
// the factor differs from the reference's LASSQ-based one only by rounding (<= a few ulp).  When every
// K_j is a multiple of a unitary (depolarising channels: sqrt(1-p) I, sqrt(p/3) X, Y, Z), the weights are
// c_j ||psi||^2 and the norm comes from the previous channel's apply sweep, so a channel after the first
// costs one read+write sweep of the batch.
#include <algorithm>

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
// Branch weights: with norm_parts, every K_j is proportional to a unitary (K_j^dag K_j = c_j I), so
// ||K_j psi||^2 = c_j ||psi||^2 with ||psi||^2 the sum the previous channel's apply sweep accumulated over
// the state it wrote (the unitary segment in between keeps the norm); otherwise the K weight partials of
// k_kraus_partial.  Writes the chosen operator times the renormalisation factor to kmat[elem].
__global__ void k_kraus_select(const double *__restrict__ partial, const double *__restrict__ norm_parts,
                               int nparts, Kraus4 kr, double4 cj, int K, int batch, const double *__restrict__ u,
                               uint64_t u_stride, uint64_t draw, Mat2 *__restrict__ kmat, int *__restrict__ err) {
    int elem = blockIdx.x * blockDim.x + threadIdx.x;
    if (elem >= batch) return;
    double w[4] = {0.0, 0.0, 0.0, 0.0};
    if (norm_parts) {
        double nrm = 0.0;
        for (int b = 0; b < nparts; b++) nrm += norm_parts[(uint64_t)elem * nparts + b];
        const double c[4] = {cj.x, cj.y, cj.z, cj.w};
        for (int j = 0; j < K; j++) w[j] = c[j] * nrm;
    } else {
        for (int b = 0; b < kPartials; b++)
            for (int j = 0; j < K; j++) w[j] += partial[((uint64_t)elem * kPartials + b) * 4 + j];
    }
    double total = 0.0;
    bool bad = false;
    for (int j = 0; j < K; j++) {
        if (!isfinite(w[j]) || w[j] < 0.0) bad = true;
        total += w[j];
    }
    if (!isfinite(total) || total <= 1e-15) bad = true;  // MIN_KRAUS_WEIGHT, StatevectorEngine.java:36
    int idx = K - 1;
    double f = 1.0;
    if (bad) {
        err[0] = 1;
        idx = 0;
    } else {
        double r = u[(uint64_t)elem * u_stride + draw] * total;  // SplittableRandom.nextDouble(bound)
        if (r >= total) r = nextafter(total, 0.0);
        double cum = 0.0;
        for (int j = 0; j < K; j++) {  // sampleWeightedIndex :579-588
            cum += w[j];
            if (r < cum) {
                idx = j;
                break;
            }
        }
        const double ns = w[idx];
        if (!isfinite(ns) || ns == 0.0) err[0] = 2;
        else f = fabs(ns - 1.0) > 1e-12 ? 1.0 / sqrt(ns) : 1.0;  // renormalizeState :590-600
    }
    const double *g = idx == 0 ? kr.k[0].v : idx == 1 ? kr.k[1].v : idx == 2 ? kr.k[2].v : kr.k[3].v;
    Mat2 m;
    for (int e = 0; e < 8; e++) m.v[e] = g[e] * f;
    kmat[elem] = m;
}

// grid (nparts, elements): the chosen operator of each trajectory on its pairs (the element is uniform per
// workgroup, so its matrix is a scalar load), and per workgroup the sum of |out|^2 into
// norm_parts[elem * nparts + block] for the next channel's weights.
__global__ void __launch_bounds__(kBlock) k_kraus_apply(double2 *__restrict__ amp, int n, int target, int nparts,
                                                        uint64_t e0, const Mat2 *__restrict__ kmat,
                                                        double *__restrict__ norm_parts) {
    const uint64_t elem = e0 + blockIdx.y;
    const uint64_t pairs = 1ULL << (n - 1);
    const uint64_t stride = 1ULL << target;
    double2 *base = amp + (elem << n);
    const Mat2 m = kmat[elem];
    double acc = 0.0;
    for (uint64_t p = blockIdx.x * (uint64_t)kBlock + threadIdx.x; p < pairs; p += (uint64_t)nparts * kBlock) {
        const uint64_t lo = p & (stride - 1ULL);
        const uint64_t idx0 = ((p ^ lo) << 1) | lo;
        const double2 a0 = base[idx0], a1 = base[idx0 | stride];
        double2 o0, o1;
        apply2(m.v, a0, a1, o0, o1);
        base[idx0] = o0;
        base[idx0 | stride] = o1;
        acc += o0.x * o0.x + o0.y * o0.y + o1.x * o1.x + o1.y * o1.y;
    }
    __shared__ double sh[kBlock];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) norm_parts[elem * (uint64_t)nparts + blockIdx.x] = sh[0];
}

}  // namespace

// Scratch layout in `work`: weight partials [batch*kPartials*4] doubles, norm partials [batch*kPartials]
// doubles, chosen operators [batch] Mat2, err [1] int.  noise_work_bytes gives the size.
size_t noise_work_bytes(int batch) {
    return (size_t)batch * kPartials * 5 * sizeof(double) + (size_t)batch * sizeof(Mat2) + 64;
}

// K_j^dag K_j = c_j I for every j (c_j returned), so branch weights follow from the norm.
static bool unitary_proportional(const double *kraus, int K, double *c) {
    for (int j = 0; j < K; j++) {
        const double *g = kraus + 8 * j;
        // columns (g00, g10) and (g01, g11): equal squared lengths, orthogonal
        const double n0 = g[0] * g[0] + g[1] * g[1] + g[4] * g[4] + g[5] * g[5];
        const double n1 = g[2] * g[2] + g[3] * g[3] + g[6] * g[6] + g[7] * g[7];
        const double dr = g[0] * g[2] + g[1] * g[3] + g[4] * g[6] + g[5] * g[7];
        const double di = g[0] * g[3] - g[1] * g[2] + g[4] * g[7] - g[5] * g[6];
        if (n0 != n1 || dr != 0.0 || di != 0.0) return false;
        c[j] = n0;
    }
    return true;
}

int launch_channel_batched(Context *ctx, State *st, int target, const double *kraus, int K, const double *u_dev,
                           uint64_t u_stride, uint64_t draw, void *work, bool norm_ready) {
    Kraus4 kr;
    for (int j = 0; j < 4; j++)
        for (int e = 0; e < 8; e++) kr.k[j].v[e] = j < K ? kraus[8 * j + e] : 0.0;
    const uint64_t B = (uint64_t)st->batch;
    double *partial = (double *)work;
    double *norm_parts = partial + B * kPartials * 4;
    Mat2 *kmat = (Mat2 *)(norm_parts + B * kPartials);
    int *err = (int *)(kmat + B);
    const uint64_t pairs = 1ULL << (st->n - 1);
    const int nparts = (int)std::max<uint64_t>(1, std::min<uint64_t>(kPartials, pairs / (4ULL * kBlock)));
    double c[4] = {0.0, 0.0, 0.0, 0.0};
    const bool from_norm = norm_ready && unitary_proportional(kraus, K, c);
    if (!from_norm) {
        ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_NOISE, 16.0 * (double)(B << st->n));
        hipLaunchKernelGGL(k_kraus_partial, dim3(kPartials, (unsigned)B), dim3(kBlock), 0, ctx->stream, st->amp,
                           st->n, target, K, kr, partial);
        SC_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_kraus_select, dim3((unsigned)((B + 63) / 64)), dim3(64), 0, ctx->stream, partial,
                       from_norm ? (const double *)norm_parts : nullptr, nparts, kr, make_double4(c[0], c[1], c[2], c[3]),
                       K, (int)B, u_dev, u_stride, draw, kmat, err);
    SC_HIP_CHECK(hipGetLastError());
    for (uint64_t e0 = 0; e0 < B; e0 += 65535) {
        const unsigned cnt = (unsigned)std::min<uint64_t>(65535, B - e0);
        ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_NOISE, 32.0 * (double)((uint64_t)cnt << st->n));
        hipLaunchKernelGGL(k_kraus_apply, dim3((unsigned)nparts, cnt), dim3(kBlock), 0, ctx->stream, st->amp, st->n,
                           target, nparts, e0, (const Mat2 *)kmat, norm_parts);
        SC_HIP_CHECK(hipGetLastError());
    }
    return SC_OK;
}

int *noise_err_flag(void *work, int batch) {
    double *partial = (double *)work;
    Mat2 *kmat = (Mat2 *)(partial + (size_t)batch * kPartials * 5);
    return (int *)(kmat + batch);
}

}  // namespace sc
