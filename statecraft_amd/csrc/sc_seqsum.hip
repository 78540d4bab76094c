// sc_seqsum.hip -- exact sequential running sums of nonnegative doubles, computed segment-parallel, and the
// measurement flow built on them (StatevectorEngine.performMeasurement, StatevectorEngine.java:229-338).
//
// The reference accumulates in basis order: each marginal probability is `probabilities[outcome] += p`
// (:274-296), `normalize` sums sequentially (:298-312) and `cumulative` is a running sum (:314-323).  A
// tree or tiled sum differs from those sums in the last bits, and a shot whose uniform falls between the
// two values of a CDF entry lands on a neighbouring outcome.  This file reproduces the sequential sums
// exactly, in parallel:
//
//   A running sum s in the binade [2^e, 2^(e+1)) is a multiple of ulp_e = 2^(e-52), and so is every double
//   up to 2^(e+1).  Adding p >= 0 without leaving the binade therefore gives s + ulp_e * rint(p / ulp_e)
//   exactly, unless p / ulp_e lies exactly halfway between two integers (then round-half-even looks at the
//   parity of s).  While the sum stays in one binade, each value contributes an integer that does not depend
//   on s, and integer-valued doubles below 2^53 add exactly in any order.
//
// Values are cut into segments of 4096.  Pass 1 sums every segment approximately; their approximate prefix
// gives each segment a candidate binade.  Pass 2 sums the segment's integers rint(v * 2^(52-e)) and counts
// ties.  A host chain then walks the segments in order with the exact running sum S: a segment whose
// candidate is the binade of S, with no tie and M + R <= 2^53 - 1 (M = S / ulp_e, R the integer sum) ends at
// exactly (M + R) * ulp_e; any other segment (the first nonzero one, binade crossings, ties) is summed
// serially on the host from its values.  The result is bit-identical to the sequential loop.
//
// The shot search uses the exact segment starts: the first segment whose end reaches u, then inside it the
// exact running sums (integer prefix sums for a single-binade segment, a serial loop otherwise) and the
// smallest index whose value reaches u, with the last entry forced to 1.0 -- Arrays.binarySearch plus the
// walk-back of sampleOutcome (:325-338) on the cumulative array.
//
// Compiled with -ffp-contract=off: every product and sum here is rounded on its own, like Java's.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "sc_internal.hpp"
#include "sc_seqsum.hpp"

namespace sc {
namespace seq {

namespace {

constexpr int kBlock = 256;
constexpr int kPer = (int)(kSeg / kBlock);  // 16 values per thread

__host__ __device__ inline uint64_t pdep_bits(uint64_t x, uint64_t mask) {
    uint64_t r = 0;
    while (mask) {
        const uint64_t low = mask & (~mask + 1ULL);
        if (x & 1ULL) r |= low;
        x >>= 1;
        mask &= mask - 1ULL;
    }
    return r;
}

// Exact 2^k for -1022 <= k <= 1023.
__host__ __device__ inline double pow2(int k) {
    const uint64_t bits = (uint64_t)(k + 1023) << 52;
    double d;
    __builtin_memcpy(&d, &bits, 8);
    return d;
}

// |a|^2 = (re * re) + (im * im) with both products rounded (StatevectorEngine.java:282).
__host__ __device__ inline double prob(double re, double im) {
    const double a = re * re;
    const double b = im * im;
    return a + b;
}

// Value j of sequence o.
__host__ __device__ inline double value_at(const Source &s, uint64_t o, uint64_t j) {
    double v;
    if (s.arr) {
        v = s.arr[o * s.len + j];
    } else {
        const uint64_t idx = s.contiguous ? j : (pdep_bits(o, s.mmask) | pdep_bits(j, s.cmask));
        const double *a = s.amp + 2 * idx;
        v = prob(a[0], a[1]);
    }
    if (s.scale != 1.0) v = v * s.scale;
    return v;
}

// rint(v * 2^(52-e)); tie: v * 2^(52-e) lies exactly halfway between two integers.
__host__ __device__ inline double grid_round(double v, double f, bool &tie) {
    const double x = v * f;
    const double fl = floor(x);
    tie = (x - fl) == 0.5;
    return rint(x);
}

__device__ inline double block_sum(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Value j (within the segment) of segment g, 0 beyond the sequence.
__device__ inline double seg_value(const Source &s, uint64_t g, int j) {
    const uint64_t o = g / s.segs_per_seq;
    const uint64_t pos = (g % s.segs_per_seq) * kSeg + (uint64_t)j;
    return pos < s.len ? value_at(s, o, pos) : 0.0;
}

// Pass 1: approximate sum of every segment (fixed order).
__global__ void __launch_bounds__(kBlock) k_seq_approx(Source s, double *__restrict__ approx) {
    __shared__ double sh[kBlock];
    const uint64_t g = blockIdx.x;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kPer; k++) acc += seg_value(s, g, (int)threadIdx.x + k * kBlock);
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) approx[g] = t;
}

// Pass 2: integer sum of the segment's values on the grid of its candidate binade, and its tie count.
__global__ void __launch_bounds__(kBlock) k_seq_int(Source s, const int32_t *__restrict__ cand,
                                                     double *__restrict__ isum, uint32_t *__restrict__ ties) {
    __shared__ double sh[kBlock];
    const uint64_t g = blockIdx.x;
    const int e = cand[g];
    if (e == kNoBinade) {
        if (threadIdx.x == 0) {
            isum[g] = -1.0;
            ties[g] = 0;
        }
        return;
    }
    const double f = pow2(52 - e);
    double acc = 0.0, nt = 0.0;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        bool tie;
        acc += grid_round(seg_value(s, g, (int)threadIdx.x + k * kBlock), f, tie);
        nt += tie ? 1.0 : 0.0;
    }
    const double r = block_sum(acc, sh);
    const double t = block_sum(nt, sh);
    if (threadIdx.x == 0) {
        isum[g] = r;
        ties[g] = (uint32_t)t;
    }
}

// The values of listed segments, kSeg per segment.
__global__ void __launch_bounds__(kBlock) k_seq_gather(Source s, const uint64_t *__restrict__ list,
                                                        double *__restrict__ out) {
    const uint64_t g = list[blockIdx.x];
    double *dst = out + (uint64_t)blockIdx.x * kSeg;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        const int j = (int)threadIdx.x + k * kBlock;
        dst[j] = seg_value(s, g, j);
    }
}

// One workgroup per segment that holds shots (grid-stride over them; the host found each shot's segment, the
// first whose exact end reaches u, and grouped the shots by segment): the segment's exact running sums once,
// then for each of its shots the smallest index whose running sum reaches u (the last index of the sequence
// counts as 1.0).  out[s] = -1 when no index qualifies, which the exact plan rules out; the host reports it as
// an internal error.
__global__ void __launch_bounds__(kBlock) k_seq_search(Source s, uint64_t nseg, const double *__restrict__ start,
                                                        const double *__restrict__ ends,
                                                        const int32_t *__restrict__ binade,
                                                        const uint8_t *__restrict__ mode,
                                                        const double *__restrict__ u,
                                                        const uint64_t *__restrict__ segs,
                                                        const uint64_t *__restrict__ seg_begin,
                                                        const int64_t *__restrict__ order, uint64_t nsegs,
                                                        int64_t *__restrict__ out, double *__restrict__ gap) {
    __shared__ double vals[kSeg];
    __shared__ double tsum[kBlock];
    __shared__ uint64_t best[kBlock];
    const uint64_t last = s.len - 1;
    for (uint64_t b = blockIdx.x; b < nsegs; b += gridDim.x) {
        const uint64_t g = segs[b];
        const uint64_t base = g * kSeg;
        const double S0 = start[g];
        const int md = mode[g];
        const int e = binade[g];
        // this thread's 16 consecutive values and their running sums within the segment
#pragma unroll
        for (int k = 0; k < kPer; k++) {
            const int j = (int)threadIdx.x + k * kBlock;
            vals[j] = seg_value(s, g, j);
        }
        __syncthreads();
        const int j0 = (int)threadIdx.x * kPer;
        double cum[kPer];
        if (md == kFast) {
            const double f = pow2(52 - e);
            double run = 0.0;
#pragma unroll
            for (int k = 0; k < kPer; k++) {
                bool tie;
                run += grid_round(vals[j0 + k], f, tie);
                cum[k] = run;
            }
            tsum[threadIdx.x] = run;
            __syncthreads();
            for (int w = 1; w < kBlock; w <<= 1) {  // inclusive scan of the thread totals (exact integers)
                const double add = (int)threadIdx.x >= w ? tsum[threadIdx.x - w] : 0.0;
                __syncthreads();
                tsum[threadIdx.x] += add;
                __syncthreads();
            }
            const double M = S0 * f;  // the start as an integer multiple of ulp_e
            const double pre = M + (threadIdx.x > 0 ? tsum[threadIdx.x - 1] : 0.0);
            const double ulp = pow2(e - 52);
#pragma unroll
            for (int k = 0; k < kPer; k++) cum[k] = (pre + cum[k]) * ulp;
        } else if (md == kSerial) {
            if (threadIdx.x == 0) {  // the sequential loop itself; running sums go back into vals
                double run = S0;
                for (int j = 0; j < (int)kSeg; j++) {
                    run = run + vals[j];
                    vals[j] = run;
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kPer; k++) cum[k] = vals[j0 + k];
        } else {
#pragma unroll
            for (int k = 0; k < kPer; k++) cum[k] = S0;
        }
#pragma unroll
        for (int k = 0; k < kPer; k++)
            if (base + (uint64_t)(j0 + k) == last) cum[k] = 1.0;
        __syncthreads();  // every thread has read vals
#pragma unroll
        for (int k = 0; k < kPer; k++) vals[j0 + k] = cum[k];
        __syncthreads();
        for (uint64_t t = seg_begin[b]; t < seg_begin[b + 1]; t++) {
            const int64_t shot = order[t];
            const double value = u[shot];
            uint64_t cand = ~0ULL;
#pragma unroll
            for (int k = 0; k < kPer; k++) {
                const uint64_t i = base + (uint64_t)(j0 + k);
                if (i <= last && cand == ~0ULL && (value >= 1.0 ? i == last : cum[k] >= value)) cand = i;
            }
            best[threadIdx.x] = cand;
            __syncthreads();
            for (int w = kBlock / 2; w > 0; w >>= 1) {
                if ((int)threadIdx.x < w && best[threadIdx.x + w] < best[threadIdx.x])
                    best[threadIdx.x] = best[threadIdx.x + w];
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                const uint64_t r = best[0];
                out[shot] = r == ~0ULL ? -1 : (int64_t)r;
                if (gap) {
                    double gv = 0.0;
                    if (r != ~0ULL) {
                        gv = fabs(value - vals[r - base]);
                        const double prev = r > base ? vals[r - 1 - base] : (g > 0 ? ends[g - 1] : -1.0);
                        if (prev >= 0.0) gv = fmin(gv, fabs(value - prev));
                    }
                    gap[shot] = gv;
                }
            }
            __syncthreads();
        }
    }
}

unsigned grid_blocks(uint64_t n) { return (unsigned)n; }

// ---------------------------------------------------------------- the exact chain (host)

// Serial running sum over one segment's values (the reference's loop itself).
double serial_run(double S, const double *v, uint64_t count) {
    for (uint64_t j = 0; j < count; j++) S = S + v[j];
    return S;
}

// The binade of x >= 0: its unbiased exponent field (ilogb for normal x; -1023 for zero and subnormals, which
// lie below every binade the chain uses).
inline int binade_of(double x) {
    uint64_t b;
    std::memcpy(&b, &x, 8);
    return (int)((b >> 52) & 0x7FF) - 1023;
}
constexpr uint64_t kMant = (1ULL << 52) - 1ULL;

}  // namespace

int plan_sequences(Provider &pv, const Source &src, Plan &plan) {
    const uint64_t nseq = src.nseq, sps = src.segs_per_seq, nseg = nseq * sps;
    std::vector<double> approx;
    int rc = pv.approx(approx);
    if (rc) return rc;
    for (uint64_t g = 0; g < nseg; g++)
        if (!std::isfinite(approx[g]))
            return fail(SC_E_STATE, "probabilities are not finite (the state holds NaN or infinite amplitudes)");
    // candidate binade of each segment from the approximate prefix of its sequence
    std::vector<int32_t> cand(nseg, kNoBinade);
    for (uint64_t o = 0; o < nseq; o++) {
        double P = 0.0;
        for (uint64_t k = 0; k < sps; k++) {
            const uint64_t g = o * sps + k;
            const double a = approx[g];
            if (a != 0.0 && P > 0.0) {
                const int e1 = binade_of(P), e2 = binade_of(P + a);
                if (e1 == e2 && e1 >= kMinBinade && e1 <= kMaxBinade) cand[g] = e1;
            }
            P += a;
        }
    }
    std::vector<double> isum;
    std::vector<uint32_t> ties;
    if ((rc = pv.int_sums(cand, isum, ties))) return rc;
    // values of the segments the chain will have to sum serially: gathered before it runs, and again for any
    // segment the chain finds it needs (a candidate binade off by one near a binade boundary)
    std::unordered_map<uint64_t, std::vector<double>> values;
    std::vector<uint64_t> want;
    for (uint64_t g = 0; g < nseg; g++)
        if (approx[g] != 0.0 && (cand[g] == kNoBinade || ties[g] != 0)) want.push_back(g);
    plan.start.assign(nseg, 0.0);
    plan.binade.assign(nseg, kNoBinade);
    plan.mode.assign(nseg, kZero);
    plan.total.assign(nseq, 0.0);
    plan.fast = plan.serial = plan.zero = 0;
    plan.gather_rounds = 0;
    const double two53m1 = 9007199254740991.0;
    for (int round = 0;; round++) {
        if (!want.empty()) {
            std::vector<double> vals;
            if ((rc = pv.gather(want, vals))) return rc;
            for (size_t i = 0; i < want.size(); i++)
                values[want[i]].assign(vals.begin() + i * kSeg, vals.begin() + (i + 1) * kSeg);
            plan.gather_rounds++;
            want.clear();
        }
        uint64_t fast = 0, serial = 0, zero = 0;
        for (uint64_t o = 0; o < nseq; o++) {
            double S = 0.0;
            for (uint64_t k = 0; k < sps; k++) {
                const uint64_t g = o * sps + k;
                plan.start[g] = S;
                if (approx[g] == 0.0) {  // a sum of nonnegative values is 0 only when every value is: S + 0 = S
                    plan.mode[g] = kZero;
                    plan.binade[g] = kNoBinade;
                    zero++;
                    continue;
                }
                const int e = cand[g];
                if (e != kNoBinade && ties[g] == 0 && S > 0.0 && binade_of(S) == e && isum[g] >= 0.0 &&
                    isum[g] <= two53m1) {
                    // S = M 2^(e-52) with the integer M in [2^52, 2^53): its significand
                    uint64_t bits;
                    std::memcpy(&bits, &S, 8);
                    const uint64_t MR = ((bits & kMant) | (1ULL << 52)) + (uint64_t)isum[g];
                    if (MR <= (uint64_t)two53m1) {
                        plan.mode[g] = kFast;
                        plan.binade[g] = e;
                        bits = ((uint64_t)(e + 1023) << 52) | (MR & kMant);  // (M + R) 2^(e-52), exact
                        std::memcpy(&S, &bits, 8);
                        fast++;
                        continue;
                    }
                }
                auto it = values.find(g);
                if (it == values.end()) {
                    // not gathered yet: this sequence resumes in the next round; on later rounds fetch the rest
                    // of the sequence at once so that the rounds stay few
                    want.push_back(g);
                    if (round >= 2)
                        for (uint64_t k2 = k + 1; k2 < sps; k2++)
                            if (approx[o * sps + k2] != 0.0 && !values.count(o * sps + k2)) want.push_back(o * sps + k2);
                    break;
                }
                const uint64_t pos = k * kSeg;
                const uint64_t cnt = std::min<uint64_t>(kSeg, src.len - pos);
                plan.mode[g] = kSerial;
                plan.binade[g] = kNoBinade;
                S = serial_run(S, it->second.data(), cnt);
                serial++;
            }
            plan.total[o] = S;
        }
        if (want.empty()) {
            plan.fast = fast;
            plan.serial = serial;
            plan.zero = zero;
            return SC_OK;
        }
        if (round > 64) return fail(SC_E_STATE, "exact cumulative sums did not converge");
    }
}

// ---------------------------------------------------------------- device provider

namespace {

struct DeviceProvider : Provider {
    Context *ctx;
    Source src;
    uint64_t nseg;
    // device workspace in the context's scratch (approx, candidates, integer sums, ties) and pinned staging in
    // its host scratch: no allocation per call, and transfers at pinned-memory rates
    double *d_approx = nullptr, *d_isum = nullptr;
    int32_t *d_cand = nullptr;
    uint32_t *d_ties = nullptr;
    char *h = nullptr;
    DeviceProvider(Context *c, const Source &s) : ctx(c), src(s), nseg(s.nseq * s.segs_per_seq) {}
    int init() {
        const size_t bytes = (size_t)nseg * kWorkBytesPerSeg;
        int rc = ensure_scratch(ctx, bytes);
        if (rc) return rc;
        if ((rc = ensure_host_scratch(ctx, bytes))) return rc;
        d_approx = (double *)ctx->scratch;
        d_isum = d_approx + nseg;
        d_cand = (int32_t *)(d_isum + nseg);
        d_ties = (uint32_t *)(d_cand + nseg);
        h = (char *)ctx->scratch_host;
        return SC_OK;
    }
    static constexpr size_t kWorkBytesPerSeg = 24;

    int approx(std::vector<double> &out) override {
        int rc = init();
        if (rc) return rc;
        {
            ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_MEASURE, 16.0 * (double)(src.nseq * src.len));
            hipLaunchKernelGGL(k_seq_approx, dim3(grid_blocks(nseg)), dim3(kBlock), 0, ctx->stream, src, d_approx);
            SC_HIP_CHECK(hipGetLastError());
        }
        SC_HIP_CHECK(hipMemcpyAsync(h, d_approx, nseg * 8, hipMemcpyDeviceToHost, ctx->stream));
        SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        out.assign((const double *)h, (const double *)h + nseg);
        return SC_OK;
    }
    int int_sums(const std::vector<int32_t> &cand, std::vector<double> &isum, std::vector<uint32_t> &ties) override {
        std::memcpy(h, cand.data(), nseg * 4);
        SC_HIP_CHECK(hipMemcpyAsync(d_cand, h, nseg * 4, hipMemcpyHostToDevice, ctx->stream));
        {
            ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_MEASURE, 16.0 * (double)(src.nseq * src.len));
            hipLaunchKernelGGL(k_seq_int, dim3(grid_blocks(nseg)), dim3(kBlock), 0, ctx->stream, src, d_cand, d_isum,
                               d_ties);
            SC_HIP_CHECK(hipGetLastError());
        }
        // results into the staging buffer after the candidates' region
        SC_HIP_CHECK(hipMemcpyAsync(h + nseg * 4, d_isum, nseg * 8, hipMemcpyDeviceToHost, ctx->stream));
        SC_HIP_CHECK(hipMemcpyAsync(h + nseg * 12, d_ties, nseg * 4, hipMemcpyDeviceToHost, ctx->stream));
        SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        isum.assign((const double *)(h + nseg * 4), (const double *)(h + nseg * 4) + nseg);
        ties.assign((const uint32_t *)(h + nseg * 12), (const uint32_t *)(h + nseg * 12) + nseg);
        return SC_OK;
    }
    int gather(const std::vector<uint64_t> &segs, std::vector<double> &vals) override {
        const uint64_t k = segs.size();
        DevBuf b;
        int rc = b.alloc(k * 8 + k * kSeg * 8);
        if (rc) return rc;
        uint64_t *l = (uint64_t *)b.p;
        double *v = (double *)((char *)b.p + k * 8);
        SC_HIP_CHECK(hipMemcpyAsync(l, segs.data(), k * 8, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_seq_gather, dim3((unsigned)k), dim3(kBlock), 0, ctx->stream, src, l, v);
        SC_HIP_CHECK(hipGetLastError());
        // through the pinned staging, after the provider's own region (its contents were copied out already)
        const size_t off = (size_t)nseg * kWorkBytesPerSeg;
        if ((rc = ensure_host_scratch(ctx, off + k * kSeg * 8))) return rc;
        h = (char *)ctx->scratch_host;
        SC_HIP_CHECK(hipMemcpyAsync(h + off, v, k * kSeg * 8, hipMemcpyDeviceToHost, ctx->stream));
        SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        vals.assign((const double *)(h + off), (const double *)(h + off) + k * kSeg);
        return SC_OK;
    }
};

// Host provider: the same values and sums computed on the CPU (test and inspection path).
struct HostProvider : Provider {
    Source src;
    uint64_t nseg;
    explicit HostProvider(const Source &s) : src(s), nseg(s.nseq * s.segs_per_seq) {}
    double val(uint64_t g, uint64_t j) const {
        const uint64_t o = g / src.segs_per_seq, pos = (g % src.segs_per_seq) * kSeg + j;
        return pos < src.len ? value_at(src, o, pos) : 0.0;
    }
    int approx(std::vector<double> &out) override {
        out.assign(nseg, 0.0);
        for (uint64_t g = 0; g < nseg; g++) {
            double acc = 0.0;
            for (uint64_t j = 0; j < kSeg; j++) acc += val(g, j);
            out[g] = acc;
        }
        return SC_OK;
    }
    int int_sums(const std::vector<int32_t> &cand, std::vector<double> &isum, std::vector<uint32_t> &ties) override {
        isum.assign(nseg, -1.0);
        ties.assign(nseg, 0);
        for (uint64_t g = 0; g < nseg; g++) {
            if (cand[g] == kNoBinade) continue;
            const double f = pow2(52 - cand[g]);
            double acc = 0.0;
            uint32_t nt = 0;
            for (uint64_t j = 0; j < kSeg; j++) {
                bool tie;
                acc += grid_round(val(g, j), f, tie);
                nt += tie;
            }
            isum[g] = acc;
            ties[g] = nt;
        }
        return SC_OK;
    }
    int gather(const std::vector<uint64_t> &segs, std::vector<double> &vals) override {
        vals.assign(segs.size() * kSeg, 0.0);
        for (size_t i = 0; i < segs.size(); i++)
            for (uint64_t j = 0; j < kSeg; j++) vals[i * kSeg + j] = val(segs[i], j);
        return SC_OK;
    }
};

// The shot search of k_seq_search, on the host.
int64_t host_search(const HostProvider &hp, const Plan &plan, const std::vector<double> &ends, double value,
                    double *gap) {
    const uint64_t nseg = plan.start.size(), last = hp.src.len - 1;
    uint64_t g = nseg - 1;
    if (value < 1.0) {
        uint64_t lo = 0, hi = nseg - 1;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (ends[mid] >= value) hi = mid;
            else lo = mid + 1;
        }
        g = lo;
    }
    std::vector<double> cum(kSeg);
    double run = plan.start[g], irun = 0.0;
    const int e = plan.binade[g];
    const double f = plan.mode[g] == kFast ? pow2(52 - e) : 0.0;
    for (uint64_t j = 0; j < kSeg; j++) {
        const double v = hp.val(g, j);
        if (plan.mode[g] == kFast) {
            bool tie;
            irun += grid_round(v, f, tie);
            cum[j] = (plan.start[g] * f + irun) * pow2(e - 52);
        } else if (plan.mode[g] == kSerial) {
            run = run + v;
            cum[j] = run;
        } else {
            cum[j] = plan.start[g];
        }
        if (g * kSeg + j == last) cum[j] = 1.0;
    }
    for (uint64_t j = 0; j < kSeg; j++) {
        const uint64_t i = g * kSeg + j;
        if (i > last) break;
        if (value >= 1.0 ? i == last : cum[j] >= value) {
            if (gap) {
                double gv = std::fabs(value - cum[j]);
                const double prev = j > 0 ? cum[j - 1] : (g > 0 ? ends[g - 1] : -1.0);
                if (prev >= 0.0) gv = std::fmin(gv, std::fabs(value - prev));
                *gap = gv;
            }
            return (int64_t)i;
        }
    }
    return -1;
}

Source make_state_source(const double *amp, int n, const int32_t *measured, int m, bool per_outcome) {
    Source s{};
    s.amp = amp;
    s.scale = 1.0;
    if (!per_outcome) {  // one sequence over the whole state in index order
        s.contiguous = 1;
        s.len = 1ULL << n;
        s.nseq = 1;
    } else {  // one sequence per outcome of the measured qubits, over the complement in increasing order
        uint64_t mm = 0;
        for (int k = 0; k < m; k++) mm |= 1ULL << measured[k];
        s.mmask = mm;
        s.cmask = ((n >= 64) ? ~0ULL : ((1ULL << n) - 1ULL)) & ~mm;
        s.len = 1ULL << (n - m);
        s.nseq = 1ULL << m;
    }
    s.segs_per_seq = (s.len + kSeg - 1) / kSeg;
    return s;
}

Source make_array_source(const double *arr, uint64_t len) {
    Source s{};
    s.arr = arr;
    s.scale = 1.0;
    s.len = len;
    s.nseq = 1;
    s.segs_per_seq = (len + kSeg - 1) / kSeg;
    return s;
}

bool identity_order(int n, const int32_t *measured, int m) {
    if (m != n) return false;
    for (int k = 0; k < m; k++)
        if (measured[k] != k) return false;
    return true;
}

// Exact plan of the CDF source, then StatevectorEngine.normalize (:298-312): the exact sequential total;
// zero -> IllegalStateException; |sum - 1| > 1e-12 -> every value times 1.0 / sum and the plan again.
template <typename MakeProvider>
int normalized_plan(Source &src, MakeProvider make, Plan &plan) {
    {
        auto pv = make(src);
        int rc = plan_sequences(*pv, src, plan);
        if (rc) return rc;
    }
    const double sum = plan.total[0];
    if (sum == 0.0) return fail(SC_E_STATE, "State collapsed to zero probability amplitudes");
    if (std::fabs(sum - 1.0) > 1e-12) {
        src.scale = 1.0 / sum;
        auto pv = make(src);
        int rc = plan_sequences(*pv, src, plan);
        if (rc) return rc;
    }
    return SC_OK;
}

std::vector<double> plan_ends(const Plan &plan, uint64_t nseg) {
    std::vector<double> ends(nseg);
    for (uint64_t g = 0; g + 1 < nseg; g++) ends[g] = plan.start[g + 1];
    ends[nseg - 1] = 1.0;  // cumulative[last] = 1.0 (:321)
    return ends;
}

}  // namespace

// ---------------------------------------------------------------- device flows

int exact_marginals(Context *ctx, const double2 *amp, int n, const int32_t *measured, int m, double *out_dev) {
    Source src = make_state_source((const double *)amp, n, measured, m, true);
    DeviceProvider pv(ctx, src);
    Plan plan;
    int rc = plan_sequences(pv, src, plan);
    if (rc) return rc;
    SC_HIP_CHECK(hipMemcpyAsync(out_dev, plan.total.data(), src.nseq * 8, hipMemcpyHostToDevice, ctx->stream));
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

int exact_sample(Context *ctx, const double2 *amp, int n, const int32_t *measured, int m, const double *marginals_dev,
                 const double *u_host, const double *u_dev, int64_t shots, int64_t *out_dev, double *gap_dev,
                 SampleStats *stats) {
    Source src = marginals_dev ? make_array_source(marginals_dev, 1ULL << m)
                               : make_state_source((const double *)amp, n, measured, m, false);
    Plan plan;
    int rc = normalized_plan(src, [&](const Source &s) { return std::make_unique<DeviceProvider>(ctx, s); }, plan);
    if (rc) return rc;
    const uint64_t nseg = src.segs_per_seq;
    const std::vector<double> ends = plan_ends(plan, nseg);
    // each shot's segment: the first whose exact end reaches u (binary search over the ends, the last one 1.0),
    // and the shots grouped by segment so that each segment's running sums are formed once
    std::vector<std::pair<uint64_t, int64_t>> by_seg((size_t)shots);
    for (int64_t t = 0; t < shots; t++) {
        const double value = u_host[t];
        uint64_t g = nseg - 1;
        if (value < 1.0) {
            uint64_t lo = 0, hi = nseg - 1;
            while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                if (ends[mid] >= value) hi = mid;
                else lo = mid + 1;
            }
            g = lo;
        }
        by_seg[(size_t)t] = {g, t};
    }
    std::sort(by_seg.begin(), by_seg.end());
    std::vector<uint64_t> segs, seg_begin;
    for (size_t t = 0; t < by_seg.size(); t++)
        if (t == 0 || by_seg[t].first != by_seg[t - 1].first) {
            segs.push_back(by_seg[t].first);
            seg_begin.push_back(t);
        }
    seg_begin.push_back(by_seg.size());
    const uint64_t nsegs = segs.size();
    // the search tables through the context's pinned staging into its device scratch (the providers are done
    // with both): start, ends, binade, mode, then the shot order, the segments and their first shots
    const size_t tbytes = ((size_t)nseg * 21 + 255) & ~(size_t)255;
    const size_t obytes = (size_t)shots * 8, sbytes = nsegs * 8, bbytes = (nsegs + 1) * 8;
    const size_t all = tbytes + obytes + sbytes + bbytes;
    if ((rc = ensure_scratch(ctx, all)) || (rc = ensure_host_scratch(ctx, all))) return rc;
    char *h = (char *)ctx->scratch_host;
    std::memcpy(h, plan.start.data(), nseg * 8);
    std::memcpy(h + nseg * 8, ends.data(), nseg * 8);
    std::memcpy(h + nseg * 16, plan.binade.data(), nseg * 4);
    std::memcpy(h + nseg * 20, plan.mode.data(), nseg);
    int64_t *h_order = (int64_t *)(h + tbytes);
    for (size_t t = 0; t < by_seg.size(); t++) h_order[t] = by_seg[t].second;
    std::memcpy(h + tbytes + obytes, segs.data(), sbytes);
    std::memcpy(h + tbytes + obytes + sbytes, seg_begin.data(), bbytes);
    char *d = (char *)ctx->scratch;
    double *d_start = (double *)d;
    double *d_ends = d_start + nseg;
    int32_t *d_bin = (int32_t *)(d_ends + nseg);
    uint8_t *d_mode = (uint8_t *)(d_bin + nseg);
    const int64_t *d_order = (const int64_t *)(d + tbytes);
    const uint64_t *d_segs = (const uint64_t *)(d + tbytes + obytes);
    const uint64_t *d_begin = (const uint64_t *)(d + tbytes + obytes + sbytes);
    SC_HIP_CHECK(hipMemcpyAsync(d, h, all, hipMemcpyHostToDevice, ctx->stream));
    {
        ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_MEASURE, 0.0);
        const unsigned grid = (unsigned)std::min<uint64_t>(nsegs, (uint64_t)ctx->cu_count * 8);
        hipLaunchKernelGGL(k_seq_search, dim3(grid), dim3(kBlock), 0, ctx->stream, src, nseg,
                           (const double *)d_start, (const double *)d_ends, (const int32_t *)d_bin,
                           (const uint8_t *)d_mode, u_dev, d_segs, d_begin, d_order, nsegs, out_dev, gap_dev);
        SC_HIP_CHECK(hipGetLastError());
    }
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // the tables live in the context's scratch
    if (stats) {
        stats->fast = plan.fast;
        stats->serial = plan.serial;
        stats->zero = plan.zero;
        stats->gather_rounds = plan.gather_rounds;
        stats->rescaled = src.scale != 1.0;
    }
    return SC_OK;
}

}  // namespace seq
}  // namespace sc

using namespace sc;
using namespace sc::seq;

extern "C" {

int sc_cumulative_host(const double *values, uint64_t count, double *cum, uint64_t *stats) {
    if (!values || !cum) return fail(SC_E_NULL, "values / cum is null");
    if (count == 0) return fail(SC_E_ARG, "count must be positive");
    Source src = make_array_source(values, count);
    HostProvider hp(src);
    Plan plan;
    int rc = plan_sequences(hp, src, plan);
    if (rc) return rc;
    // the running sums, segment by segment, from the exact segment starts
    for (uint64_t g = 0; g < src.segs_per_seq; g++) {
        const uint64_t base = g * kSeg, cnt = std::min<uint64_t>(kSeg, count - base);
        const int e = plan.binade[g];
        double run = plan.start[g], irun = 0.0;
        const double f = plan.mode[g] == kFast ? pow2(52 - e) : 0.0;
        for (uint64_t j = 0; j < cnt; j++) {
            if (plan.mode[g] == kFast) {
                bool tie;
                irun += grid_round(values[base + j], f, tie);
                cum[base + j] = (plan.start[g] * f + irun) * pow2(e - 52);
            } else if (plan.mode[g] == kSerial) {
                run = run + values[base + j];
                cum[base + j] = run;
            } else {
                cum[base + j] = plan.start[g];
            }
        }
    }
    if (stats) {
        stats[0] = plan.fast;
        stats[1] = plan.serial;
        stats[2] = plan.zero;
        stats[3] = plan.gather_rounds;
    }
    return SC_OK;
}

int sc_sample_host(int num_qubits, const double *state, const int32_t *measured, int m, const double *uniforms,
                   int64_t shots, int64_t *outcomes, double *boundary_gap, uint64_t *stats) {
    if (!state || !measured || !uniforms || !outcomes) return fail(SC_E_NULL, "argument is null");
    if (num_qubits <= 0 || num_qubits > 30) return fail(SC_E_ARG, "host sampling takes 1..30 qubits");
    if (m <= 0 || m > num_qubits) return fail(SC_E_ARG, "measured qubit count out of range");
    for (int k = 0; k < m; k++)
        if (measured[k] < 0 || measured[k] >= num_qubits || (k > 0 && measured[k] <= measured[k - 1]))
            return fail(SC_E_ARG, "measured qubits must be distinct, ascending and in range");
    if (shots <= 0) return fail(SC_E_ARG, "shots must be positive");
    std::vector<double> marg;
    const bool ident = identity_order(num_qubits, measured, m);
    if (!ident) {
        Source ms = make_state_source(state, num_qubits, measured, m, true);
        HostProvider hp(ms);
        Plan mp;
        int rc = plan_sequences(hp, ms, mp);
        if (rc) return rc;
        marg = mp.total;
    }
    Source src = ident ? make_state_source(state, num_qubits, measured, m, false)
                       : make_array_source(marg.data(), 1ULL << m);
    Plan plan;
    int rc = normalized_plan(src, [&](const Source &s) { return std::make_unique<HostProvider>(s); }, plan);
    if (rc) return rc;
    const std::vector<double> ends = plan_ends(plan, src.segs_per_seq);
    HostProvider hp(src);
    double gmin = INFINITY;
    for (int64_t s = 0; s < shots; s++) {
        double gv = 0.0;
        outcomes[s] = host_search(hp, plan, ends, uniforms[s], &gv);
        if (outcomes[s] < 0) return fail(SC_E_STATE, "exact cumulative search found no outcome");
        gmin = std::fmin(gmin, gv);
    }
    if (boundary_gap) *boundary_gap = gmin;
    if (stats) {
        stats[0] = plan.fast;
        stats[1] = plan.serial;
        stats[2] = plan.zero;
        stats[3] = plan.gather_rounds;
        stats[4] = src.scale != 1.0;
    }
    return SC_OK;
}

}  // extern "C"
