// sc_noise.hip -- noisy circuits with state-dependent Kraus channels in one call (sc_run_noisy).
//
// The reference unravels every channel on the host (StatevectorEngine.java:382-407, 463-600): branch weights
// w_j = ||K_j psi||^2 from the live state, r = rng.nextDouble(sum w), the first j with r < w_0 + .. + w_j,
// then K_j applied and the state renormalised.  The uniform behind r does not depend on the state (one
// nextDouble per channel and target, in circuit order), so the caller draws them all up front and the branch
// choice runs on the device: per channel and target, one read of the state for every weight
// (k_noise_weights), one workgroup per batch element to sum them, check them and pick the branch
// (k_noise_select), and one sweep that applies the chosen operator already divided by sqrt(w_j)
// (k_noise_apply; ||K_j psi||^2 is w_j, the norm the reference recomputes after the apply).  Channels that
// follow the same op share sweeps: the apply of one and the weights of the next are one read and one write
// (k_noise_apply_weights).  Nothing returns
// to the host until the circuit ends; the ops between channels run as per-op kernels when there are few of
// them and through the context's execution mode otherwise.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sc_internal.hpp"
#include "sc_plan.hpp"

namespace sc {

namespace {

constexpr int kBlock = 256;
constexpr int kRedGrid = 1024;     // partial sums per element with a batch of one (the sc_kraus_weights grid)
constexpr int kMaxKraus = 4;       // operators per channel
constexpr size_t kDirectOps = 4;   // segments of at most this many ops run as per-op kernels

// First failure of a run, written by the workgroup that claims `failed`.
struct NoiseFailDev {
    int32_t failed;
    int32_t kind;  // 1 a weight is negative or not finite, 2 the total is <= 1e-15 or not finite, 3 1/sqrt(w) overflows
    int32_t branch;
    int32_t elem;
    int64_t event;
    double weights[kMaxKraus];
    double norm_sq;
};

template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double (&out)[K]) {
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
#pragma unroll
    for (int k = 0; k < K; k++) out[k] = sh[k][0];
}

// Partial weights ||K_j psi||^2 (computeBranchWeight :511-530): element blockIdx.y, bpe workgroups per element.
template <int K>
__global__ void __launch_bounds__(kBlock) k_noise_weights(const double2 *__restrict__ amp, int n, int target,
                                                          Mat2 k0, Mat2 k1, Mat2 k2, Mat2 k3,
                                                          double *__restrict__ partial) {
    const Mat2 *ks[4] = {&k0, &k1, &k2, &k3};
    const double2 *base = amp + ((uint64_t)blockIdx.y << n);
    const uint64_t pairs = 1ULL << (n - 1), stride = 1ULL << target;
    double acc[K];
#pragma unroll
    for (int j = 0; j < K; j++) acc[j] = 0.0;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < pairs; i += (uint64_t)gridDim.x * kBlock) {
        const uint64_t lo = i & (stride - 1ULL);
        const uint64_t idx0 = ((i ^ lo) << 1) | lo;
        const double2 a0 = base[idx0], a1 = base[idx0 | stride];
#pragma unroll
        for (int j = 0; j < K; j++) {
            const double *g = ks[j]->v;
            const double r0 = g[0] * a0.x - g[1] * a0.y + g[2] * a1.x - g[3] * a1.y;
            const double i0 = g[0] * a0.y + g[1] * a0.x + g[2] * a1.y + g[3] * a1.x;
            const double r1 = g[4] * a0.x - g[5] * a0.y + g[6] * a1.x - g[7] * a1.y;
            const double i1 = g[4] * a0.y + g[5] * a0.x + g[6] * a1.y + g[7] * a1.x;
            acc[j] += r0 * r0 + i0 * i0 + r1 * r1 + i1 * i1;
        }
    }
    double r[K];
    block_sum<K>(acc, r);
    if (threadIdx.x == 0) {
        double *out = partial + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * K;
#pragma unroll
        for (int j = 0; j < K; j++) out[j] = r[j];
    }
}

// One workgroup per element: the weights, their checks (applyErrorChannel's), the branch
// (sampleKrausOperator :463-491, sampleWeightedIndex :579-588, SplittableRandom.nextDouble(bound)) and the
// operator to apply, K_j / sqrt(w_j) when |w_j - 1| > 1e-12 (renormalizeState :590-600), else K_j.
template <int K>
__global__ void __launch_bounds__(kBlock) k_noise_select(const double *__restrict__ partial, int bpe, Mat2 k0, Mat2 k1,
                                                         Mat2 k2, Mat2 k3, const double *__restrict__ uniforms,
                                                         uint64_t n_events, uint64_t event,
                                                         double *__restrict__ sel, NoiseFailDev *__restrict__ fail_rec) {
    const int elem = blockIdx.x;
    const double *p = partial + (size_t)elem * bpe * K;
    double acc[K];
#pragma unroll
    for (int j = 0; j < K; j++) acc[j] = 0.0;
    for (int i = threadIdx.x; i < bpe; i += kBlock) {
#pragma unroll
        for (int j = 0; j < K; j++) acc[j] += p[(size_t)i * K + j];
    }
    double w[K];
    block_sum<K>(acc, w);
    // Every lane evaluates the selection from the same sums, without branches (a single-lane, branchy form of
    // it was compiled wrongly here: its first branch was never taken); lane 0 writes the results.
    const Mat2 *ks[4] = {&k0, &k1, &k2, &k3};
    int kind = 0, branch = 0;
    double total = 0.0;
#pragma unroll
    for (int j = 0; j < K; j++) {
        const bool bad = !isfinite(w[j]) || w[j] < 0.0;
        branch = (kind == 0 && bad) ? j : branch;
        kind = (kind == 0 && bad) ? 1 : kind;
        total += w[j];
    }
    kind = (kind == 0 && (!isfinite(total) || total <= 1e-15)) ? 2 : kind;
    double r = __dmul_rn(uniforms[(uint64_t)elem * n_events + event], total);
    r = r >= total ? nextafter(total, -INFINITY) : r;
    // the first j with r < w_0 + .. + w_j (K - 1 when there is none): with nonnegative weights the partial
    // sums do not decrease, so it is the number of partial sums before the last that r is not below
    int idx = 0;
    double cum = 0.0;
#pragma unroll
    for (int j = 0; j < K - 1; j++) {
        cum += w[j];
        idx += (r < cum) ? 0 : 1;
    }
    double wsel = w[0];
#pragma unroll
    for (int j = 1; j < K; j++) wsel = j == idx ? w[j] : wsel;
    const bool rescale = fabs(wsel - 1.0) > 1e-12;
    const double f = rescale ? 1.0 / sqrt(wsel) : 1.0;
    kind = (kind == 0 && !isfinite(f)) ? 3 : kind;
    if (threadIdx.x != 0) return;
    double *s = sel + 8 * (size_t)elem;
    if (kind) {
        // the run has failed: leave this element's state as it is
        const double one[8] = {1, 0, 0, 0, 0, 0, 1, 0};
#pragma unroll
        for (int e = 0; e < 8; e++) s[e] = one[e];
        if (atomicCAS(&fail_rec->failed, 0, 1) == 0) {
            fail_rec->kind = kind;
            fail_rec->branch = branch;
            fail_rec->elem = elem;
            fail_rec->event = (int64_t)event;
#pragma unroll
            for (int j = 0; j < kMaxKraus; j++) fail_rec->weights[j] = j < K ? w[j] : 0.0;
            fail_rec->norm_sq = wsel;
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < 8; e++) {
        double v = ks[0]->v[e];
#pragma unroll
        for (int j = 1; j < K; j++) v = j == idx ? ks[j]->v[e] : v;
        s[e] = v * f;
    }
}

// The chosen operator of each element on `target` (every pair of the element's state).
__global__ void __launch_bounds__(kBlock) k_noise_apply(double2 *__restrict__ amp, int n, int target, uint64_t total,
                                                        const double *__restrict__ sel) {
    const uint64_t stride = 1ULL << target;
    const int log_pairs = n - 1;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < total; i += (uint64_t)gridDim.x * kBlock) {
        const uint64_t elem = i >> log_pairs, g = i & ((1ULL << log_pairs) - 1ULL);
        const double *m = sel + 8 * elem;
        const uint64_t lo = g & (stride - 1ULL);
        const uint64_t idx0 = ((g ^ lo) << 1) | lo;
        double2 *base = amp + (elem << n);
        const double2 a0 = base[idx0], a1 = base[idx0 | stride];
        double2 n0, n1;
        n0.x = m[0] * a0.x - m[1] * a0.y;
        n0.y = m[0] * a0.y + m[1] * a0.x;
        n0.x = fma(m[2], a1.x, n0.x);
        n0.x = fma(-m[3], a1.y, n0.x);
        n0.y = fma(m[2], a1.y, n0.y);
        n0.y = fma(m[3], a1.x, n0.y);
        n1.x = m[4] * a0.x - m[5] * a0.y;
        n1.y = m[4] * a0.y + m[5] * a0.x;
        n1.x = fma(m[6], a1.x, n1.x);
        n1.x = fma(-m[7], a1.y, n1.x);
        n1.y = fma(m[6], a1.y, n1.y);
        n1.y = fma(m[7], a1.x, n1.y);
        base[idx0] = n0;
        base[idx0 | stride] = n1;
    }
}

// The previous event's chosen operator on `ta` and, from the updated amplitudes, the partial weights of the
// next event's operators on `tb` (the two events follow the same op): one read and one write of the state
// instead of a write pass and a read pass.  With ta == tb a thread takes the pair the two share; otherwise
// the four amplitudes of a (ta, tb) quad, whose two ta-pairs are updated before its two tb-pairs are weighed.
// Same arithmetic per pair as k_noise_apply and k_noise_weights; the partials are laid out as
// k_noise_weights writes them (the sums run in another order).
template <int K>
__global__ void __launch_bounds__(kBlock) k_noise_apply_weights(double2 *__restrict__ amp, int n, int ta,
                                                                const double *__restrict__ sel, int tb, Mat2 k0,
                                                                Mat2 k1, Mat2 k2, Mat2 k3,
                                                                double *__restrict__ partial) {
    const Mat2 *ks[4] = {&k0, &k1, &k2, &k3};
    double2 *base = amp + ((uint64_t)blockIdx.y << n);
    const double *m = sel + 8 * (size_t)blockIdx.y;
    double s[8];
#pragma unroll
    for (int x = 0; x < 8; x++) s[x] = m[x];
    double acc[K];
#pragma unroll
    for (int j = 0; j < K; j++) acc[j] = 0.0;
    auto apply = [&](double2 &a0, double2 &a1) {
        double2 n0, n1;
        n0.x = s[0] * a0.x - s[1] * a0.y;
        n0.y = s[0] * a0.y + s[1] * a0.x;
        n0.x = fma(s[2], a1.x, n0.x);
        n0.x = fma(-s[3], a1.y, n0.x);
        n0.y = fma(s[2], a1.y, n0.y);
        n0.y = fma(s[3], a1.x, n0.y);
        n1.x = s[4] * a0.x - s[5] * a0.y;
        n1.y = s[4] * a0.y + s[5] * a0.x;
        n1.x = fma(s[6], a1.x, n1.x);
        n1.x = fma(-s[7], a1.y, n1.x);
        n1.y = fma(s[6], a1.y, n1.y);
        n1.y = fma(s[7], a1.x, n1.y);
        a0 = n0;
        a1 = n1;
    };
    auto weigh = [&](const double2 a0, const double2 a1) {
#pragma unroll
        for (int j = 0; j < K; j++) {
            const double *g = ks[j]->v;
            const double r0 = g[0] * a0.x - g[1] * a0.y + g[2] * a1.x - g[3] * a1.y;
            const double i0 = g[0] * a0.y + g[1] * a0.x + g[2] * a1.y + g[3] * a1.x;
            const double r1 = g[4] * a0.x - g[5] * a0.y + g[6] * a1.x - g[7] * a1.y;
            const double i1 = g[4] * a0.y + g[5] * a0.x + g[6] * a1.y + g[7] * a1.x;
            acc[j] += r0 * r0 + i0 * i0 + r1 * r1 + i1 * i1;
        }
    };
    const uint64_t sa = 1ULL << ta, sb = 1ULL << tb;
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    if (ta == tb) {
        const uint64_t pairs = 1ULL << (n - 1);
        for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < pairs; i += step) {
            const uint64_t lo = i & (sa - 1ULL);
            const uint64_t i0 = ((i ^ lo) << 1) | lo;
            double2 a0 = base[i0], a1 = base[i0 | sa];
            apply(a0, a1);
            base[i0] = a0;
            base[i0 | sa] = a1;
            weigh(a0, a1);
        }
    } else {
        const int lo_bit = ta < tb ? ta : tb, hi_bit = ta < tb ? tb : ta;
        const uint64_t quads = 1ULL << (n - 2);
        for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < quads; i += step) {
            uint64_t q = i;
            uint64_t lo = q & ((1ULL << lo_bit) - 1ULL);
            q = ((q ^ lo) << 1) | lo;
            lo = q & ((1ULL << hi_bit) - 1ULL);
            const uint64_t b = ((q ^ lo) << 1) | lo;
            double2 x00 = base[b], x10 = base[b | sa], x01 = base[b | sb], x11 = base[b | sa | sb];
            apply(x00, x10);  // pairs along ta
            apply(x01, x11);
            base[b] = x00;
            base[b | sa] = x10;
            base[b | sb] = x01;
            base[b | sa | sb] = x11;
            weigh(x00, x01);  // pairs along tb
            weigh(x10, x11);
        }
    }
    double r[K];
    block_sum<K>(acc, r);
    if (threadIdx.x == 0) {
        double *out = partial + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * K;
#pragma unroll
        for (int j = 0; j < K; j++) out[j] = r[j];
    }
}

unsigned grid_for(Context *ctx, uint64_t total) {
    uint64_t blocks = (total + kBlock - 1) / kBlock;
    const uint64_t cap = (uint64_t)ctx->cu_count * 16;
    return (unsigned)std::max<uint64_t>(1, std::min(blocks, cap));
}

void event_ops(const sc_noise_event &ev, const double *kraus, Mat2 (&m)[kMaxKraus]) {
    for (int j = 0; j < kMaxKraus; j++)
        for (int x = 0; x < 8; x++) m[j].v[x] = j < ev.n_kraus ? kraus[2 * ev.kraus + 8 * j + x] : 0.0;
}

// Partial weights of event `ev` into d_partial: alone (prev < 0), or fused with the apply of the chosen
// operator of the previous event on qubit `prev`.
int launch_weights(Context *ctx, State *st, const sc_noise_event &ev, const double *kraus, int prev, int bpe,
                   double *d_partial, const double *d_sel) {
    Mat2 m[kMaxKraus];
    event_ops(ev, kraus, m);
    const double amps = (double)((uint64_t)st->batch << st->n);
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_NOISE, (prev < 0 ? 16.0 : 32.0) * amps);
    const dim3 grid((unsigned)bpe, (unsigned)st->batch);
#define SC_NOISE_WEIGHTS(K)                                                                                       \
    if (prev < 0)                                                                                                 \
        hipLaunchKernelGGL(k_noise_weights<K>, grid, dim3(kBlock), 0, ctx->stream, st->amp, st->n, ev.target, m[0], \
                           m[1], m[2], m[3], d_partial);                                                          \
    else                                                                                                          \
        hipLaunchKernelGGL(k_noise_apply_weights<K>, grid, dim3(kBlock), 0, ctx->stream, st->amp, st->n, prev,      \
                           d_sel, ev.target, m[0], m[1], m[2], m[3], d_partial);
    switch (ev.n_kraus) {
    case 1: SC_NOISE_WEIGHTS(1) break;
    case 2: SC_NOISE_WEIGHTS(2) break;
    case 3: SC_NOISE_WEIGHTS(3) break;
    default: SC_NOISE_WEIGHTS(4) break;
    }
#undef SC_NOISE_WEIGHTS
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

int launch_select(Context *ctx, State *st, const sc_noise_event &ev, const double *kraus, uint64_t e,
                  uint64_t n_events, int bpe, const double *d_uniforms, const double *d_partial, double *d_sel,
                  NoiseFailDev *d_fail) {
    Mat2 m[kMaxKraus];
    event_ops(ev, kraus, m);
    const dim3 sgrid((unsigned)st->batch);
    switch (ev.n_kraus) {
    case 1:
        hipLaunchKernelGGL(k_noise_select<1>, sgrid, dim3(kBlock), 0, ctx->stream, d_partial, bpe, m[0], m[1], m[2],
                           m[3], d_uniforms, n_events, e, d_sel, d_fail);
        break;
    case 2:
        hipLaunchKernelGGL(k_noise_select<2>, sgrid, dim3(kBlock), 0, ctx->stream, d_partial, bpe, m[0], m[1], m[2],
                           m[3], d_uniforms, n_events, e, d_sel, d_fail);
        break;
    case 3:
        hipLaunchKernelGGL(k_noise_select<3>, sgrid, dim3(kBlock), 0, ctx->stream, d_partial, bpe, m[0], m[1], m[2],
                           m[3], d_uniforms, n_events, e, d_sel, d_fail);
        break;
    default:
        hipLaunchKernelGGL(k_noise_select<4>, sgrid, dim3(kBlock), 0, ctx->stream, d_partial, bpe, m[0], m[1], m[2],
                           m[3], d_uniforms, n_events, e, d_sel, d_fail);
        break;
    }
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

int launch_apply(Context *ctx, State *st, int target, const double *d_sel) {
    const uint64_t total = ((uint64_t)st->batch << st->n) / 2;
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_NOISE, 32.0 * 2.0 * (double)total);
    hipLaunchKernelGGL(k_noise_apply, dim3(grid_for(ctx, total)), dim3(kBlock), 0, ctx->stream, st->amp, st->n,
                       target, total, d_sel);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

std::string fmt(double x) {
    char b[40];
    std::snprintf(b, sizeof b, "%.17g", x);
    return b;
}

}  // namespace

int run_noisy(State *st, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex,
              const sc_noise_event *events, uint64_t n_events, const double *kraus, uint64_t n_kraus_complex,
              const double *uniforms, sc_noise_failure *failure) {
    Context *ctx = st->ctx;
    const int n = st->n;
    for (uint64_t e = 0; e < n_events; e++) {
        const sc_noise_event &ev = events[e];
        if (ev.after_op >= n_ops) return fail(SC_E_ARG, "noise event " + std::to_string(e) + " follows no op");
        if (e && ev.after_op < events[e - 1].after_op)
            return fail(SC_E_ARG, "noise events must follow their ops in circuit order");
        if (ev.target < 0 || ev.target >= n)
            return fail(SC_E_ARG, "noise channel qubit out of range: " + std::to_string(ev.target));
        if (ev.n_kraus < 1 || ev.n_kraus > kMaxKraus) return fail(SC_E_ARG, "kraus operator count must be 1..4");
        if (ev.kraus + 4ULL * (uint64_t)ev.n_kraus > n_kraus_complex)
            return fail(SC_E_ARG, "kraus offset out of range for noise event " + std::to_string(e));
    }
    if (st->batch > 65535) return fail(SC_E_UNSUPPORTED, "sc_run_noisy takes at most 65535 batch elements");
    for (uint64_t i = 0; i < (uint64_t)st->batch * n_events; i++)
        if (!(uniforms[i] >= 0.0 && uniforms[i] < 1.0)) return fail(SC_E_ARG, "uniforms must lie in [0, 1)");
    // every op validated up front, as the reference's circuit is
    std::vector<PlanOp> all;
    int rc = analyze_ops(n, ops, n_ops, {mats}, n_mat_complex, all);
    if (rc) return rc;
    SC_HIP_CHECK(hipSetDevice(ctx->device));
    if ((rc = materialize_basis(st))) return rc;
    const int bpe = st->batch == 1 ? kRedGrid : std::max(1, std::min(kRedGrid, 2048 / st->batch));
    const size_t n_uni = (size_t)st->batch * n_events;
    const size_t n_part = (size_t)st->batch * bpe * kMaxKraus, n_sel = (size_t)st->batch * 8;
    DevBuf buf;
    if ((rc = buf.alloc((n_uni + n_part + n_sel) * sizeof(double) + sizeof(NoiseFailDev)))) return rc;
    double *d_uni = (double *)buf.p, *d_part = d_uni + n_uni, *d_sel = d_part + n_part;
    NoiseFailDev *d_fail = (NoiseFailDev *)(d_sel + n_sel);
    if (n_uni)
        SC_HIP_CHECK(hipMemcpyAsync(d_uni, uniforms, n_uni * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    SC_HIP_CHECK(hipMemsetAsync(d_fail, 0, sizeof(NoiseFailDev), ctx->stream));
    // the ops of [begin, end) (measurement records carry no PlanOp), then the events that follow op end - 1
    size_t next = 0;
    auto segment = [&](uint64_t end) -> int {
        std::vector<PlanOp> seg;
        while (next < all.size() && (uint64_t)all[next].src < end) seg.push_back(all[next++]);
        if (seg.empty()) return SC_OK;
        if (seg.size() <= kDirectOps) return exec_direct(ctx, st, seg, mats);
        return apply_planned(st, seg, {mats}, false);
    };
    // per event: weights, choice, apply.  An event that follows the same op as the one before it takes its
    // weights in the same sweep as the previous event's apply.
    bool pending = false;  // the previous event's operator is chosen (d_sel) but not applied yet
    for (uint64_t e = 0; e < n_events; e++) {
        const sc_noise_event &ev = events[e];
        const bool fuse = pending && ev.after_op == events[e - 1].after_op;
        if (pending && !fuse && (rc = launch_apply(ctx, st, events[e - 1].target, d_sel))) return rc;
        if (!fuse && (rc = segment(ev.after_op + 1))) return rc;
        if ((rc = launch_weights(ctx, st, ev, kraus, fuse ? events[e - 1].target : -1, bpe, d_part, d_sel))) return rc;
        if ((rc = launch_select(ctx, st, ev, kraus, e, n_events, bpe, d_uni, d_part, d_sel, d_fail))) return rc;
        pending = true;
    }
    if (pending && (rc = launch_apply(ctx, st, events[n_events - 1].target, d_sel))) return rc;
    if ((rc = segment(n_ops))) return rc;
    NoiseFailDev rec;
    SC_HIP_CHECK(hipMemcpyAsync(&rec, d_fail, sizeof rec, hipMemcpyDeviceToHost, ctx->stream));
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (!rec.failed) return SC_OK;
    if (failure) {
        failure->event = rec.event;
        failure->elem = rec.elem;
        failure->kind = rec.kind;
        failure->branch = rec.branch;
        for (int j = 0; j < kMaxKraus; j++) failure->weights[j] = rec.weights[j];
        failure->norm_sq = rec.norm_sq;
    }
    const sc_noise_event &ev = events[rec.event];
    std::string w = "[";
    for (int j = 0; j < ev.n_kraus; j++) w += (j ? ", " : "") + fmt(rec.weights[j]);
    w += "]";
    const std::string where = "noise event " + std::to_string(rec.event) + " (target " + std::to_string(ev.target) +
                              ", batch element " + std::to_string(rec.elem) + ")";
    if (rec.kind == 1)
        return fail(SC_E_STATE, "invalid Kraus branch weight at " + where + ": weights=" + w +
                                    ", branch=" + fmt(rec.weights[rec.branch]));
    if (rec.kind == 2) {
        double total = 0.0;
        for (int j = 0; j < ev.n_kraus; j++) total += rec.weights[j];
        return fail(SC_E_STATE, "invalid Kraus branch weights at " + where + ": total=" + fmt(total) + ", weights=" + w);
    }
    return fail(SC_E_STATE, "state collapsed to invalid norm after Kraus application at " + where +
                                ": normSq=" + fmt(rec.norm_sq));
}

}  // namespace sc

using namespace sc;

extern "C" int sc_run_noisy(sc_state *st, const sc_op *ops, uint64_t n_ops, const double *mats,
                            uint64_t n_mat_complex, const sc_noise_event *events, uint64_t n_events,
                            const double *kraus, uint64_t n_kraus_complex, const double *uniforms,
                            sc_noise_failure *failure) {
    if (failure) {
        std::memset(failure, 0, sizeof *failure);
        failure->event = -1;
    }
    if (!st) return fail(SC_E_NULL, "state is null");
    if (n_ops && !ops) return fail(SC_E_NULL, "ops is null");
    if (!mats && n_mat_complex) return fail(SC_E_NULL, "mats is null");
    if (n_events && (!events || !kraus || !uniforms)) return fail(SC_E_NULL, "events, kraus and uniforms are required");
    return run_noisy(st, ops, n_ops, mats, n_mat_complex, events, n_events, kraus, n_kraus_complex, uniforms, failure);
}
