// sc_adjoint.hip -- Pauli-sum expectations, the adjoint-gradient reverse sweep, and batched noisy
// trajectories.  None of these has a counterpart in the reference engine; their inputs and outputs use
// the same op encoding and conventions as sc_apply_ops.
//
// Adjoint method (Jones & Gacon, "Efficient calculation of gradients in classical simulations of
// variational quantum algorithms", arXiv:2009.02823): with psi_g the state after op g and
// lambda_g = U_{g+1}^dag ... U_N^dag H psi_N, an op U_g = exp(-i theta c G) contributes
// dE/dtheta += 2 c Im <lambda_g | G | psi_g>.  The reverse sweep reads psi and lambda once per op, forms
// that inner product over the op's index pairs, applies U_g^dag to both and writes them back -- one
// fused kernel per op (k_adj_pair / k_adj_two / k_adj_diag2 / k_adj_swap).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <vector>

#include "sc_internal.hpp"
#include "sc_jit.hpp"
#include "sc_plan.hpp"

namespace sc {

namespace {

constexpr int kBlock = 256;
constexpr int kParts = 64;  // partial sums per batch element

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
// conj(a) * b accumulated
__device__ __forceinline__ void cdot_acc(double2 a, double2 b, double &re, double &im) {
    re = fma(a.x, b.x, re);
    re = fma(a.y, b.y, re);
    im = fma(a.x, b.y, im);
    im = fma(-a.y, b.x, im);
}

__device__ __forceinline__ void block_reduce2(double re, double im, double *out) {
    __shared__ double sr[kBlock], si[kBlock];
    sr[threadIdx.x] = re;
    si[threadIdx.x] = im;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sr[threadIdx.x] += sr[threadIdx.x + s];
            si[threadIdx.x] += si[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sr[0];
        out[1] = si[0];
    }
}

__device__ __forceinline__ uint64_t insert_zero(uint64_t x, int bit) {
    uint64_t lo = x & ((1ULL << bit) - 1ULL);
    return ((x ^ lo) << 1) | lo;
}
__device__ __forceinline__ uint64_t expand_mask(uint64_t x, uint64_t mask) {
    while (mask) {
        int p = __builtin_ctzll(mask);
        x = insert_zero(x, p);
        mask &= mask - 1ULL;
    }
    return x;
}

// Up to kTermsPerLaunch Pauli terms that share one x mask, evaluated in one read of the state.  With
// q = conj(psi_{i^x}) psi_i, the index pair {i, i^x} contributes s(i) q + s(i^x) conj(q) to term t's
// sum_i conj(psi_{i^x}) s(i) psi_i, s(i) = (-1)^{|i & z_t|}, and s(i^x) = s(i) (-1)^{|x & z_t|}: 2 s(i) Re q
// when |x & z_t| is even, 2i s(i) Im q when odd.  The kernel therefore visits each pair once (index bit
// ctz(x) clear) and accumulates s(i) Re q or s(i) Im q per term (bit t of `imag`); for x = 0 it visits every
// index with q = |psi_i|^2.  The host applies the factor 2 and the sign of i^{|x & z_t|}.
constexpr int kTermsPerLaunch = 32;  // a 32-qubit ring's ZZ terms in one read of the state
constexpr int kTermJ = 8;            // amplitudes a lane takes per iteration
struct TermGroup {
    uint64_t x;
    uint64_t z[kTermsPerLaunch];
    int32_t row[kTermsPerLaunch];  // output row of term t
    uint8_t jm[kTermsPerLaunch];   // term t's sign pattern over the kTermJ amplitudes of an iteration (below)
    uint32_t imag;
    int32_t nt;
};

// grid (parts, elements), parts a power of two: part[((e0 + elem) * rows + row_t) * parts + block].
// Let spread(h) be h with a zero inserted at bit ctz(x) (pairs; else h) and S = parts * kBlock.  A thread
// visits h = h0 + k0 + j S for its h0 < S, every k0 that is a multiple of 8 S, and j < 8: the index is
// i = spread(h0) | spread(k0) | spread(j S), three disjoint bit sets, so term t's sign is the product of a
// per-thread factor (applied once at the end), a factor uniform over the iteration (parity of
// spread(k0) & z_t) and (-1)^{<j, jm_t>}, where bit b of jm_t is the parity of spread(2^b S) & z_t.  Each
// iteration therefore loads 8 values, takes their Walsh-Hadamard transform W_m = sum_j (-1)^{<j, m>} q_j, and
// adds +-W_{jm_t} to term t: one add per term per 8 amplitudes, and the scalar unit computes one parity per
// term per iteration.  NT >= g.nt terms are unrolled.
__device__ __forceinline__ void wht8(double (&q)[kTermJ]) {
#pragma unroll
    for (int h = 1; h < kTermJ; h <<= 1)
#pragma unroll
        for (int j = 0; j < kTermJ; j++)
            if (!(j & h)) {
                const double a = q[j], b = q[j | h];
                q[j] = a + b;
                q[j | h] = a - b;
            }
}

// q[m] for a uniform m < 8 without indexing registers at run time (that would go through scratch memory):
// m = 0, the common case, takes a scalar branch; otherwise a select tree on m's three bits.
__device__ __forceinline__ double pick8(const double (&q)[kTermJ], uint32_t m) {
    if (m == 0) return q[0];
    const bool b0 = m & 1u, b1 = m & 2u, b2 = m & 4u;
    const double l0 = b0 ? q[1] : q[0], l1 = b0 ? q[3] : q[2], l2 = b0 ? q[5] : q[4], l3 = b0 ? q[7] : q[6];
    const double h0 = b1 ? l1 : l0, h1 = b1 ? l3 : l2;
    return b2 ? h1 : h0;
}

template <int NT, bool PAIRS>
__global__ void __launch_bounds__(kBlock) k_pauli_terms(const double2 *__restrict__ amp, int n, uint64_t e0,
                                                        TermGroup g, int rows, int parts, double *__restrict__ part) {
    const uint64_t elem = e0 + blockIdx.y;
    const double2 *base = amp + (elem << n);
    constexpr bool pairs = PAIRS;  // g.x != 0
    const uint64_t count = pairs ? (1ULL << (n - 1)) : (1ULL << n);
    const int pbit = pairs ? __builtin_ctzll(g.x) : 63;
    auto spread = [&](uint64_t h) {  // h with a zero inserted at bit pbit (pairs), else h
        const uint64_t lo = h & ((1ULL << pbit) - 1ULL);
        return pairs ? (((h ^ lo) << 1) | lo) : h;
    };
    const uint64_t S = (uint64_t)parts * kBlock;
    const uint64_t h0 = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
    const uint64_t c = spread(h0);
    double acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = 0.0;
    for (uint64_t k0 = 0; k0 < count; k0 += kTermJ * S) {
        double qr[kTermJ], qi[kTermJ] = {};
        const uint64_t ck = c | spread(k0);
#pragma unroll
        for (int j = 0; j < kTermJ; j++) {
            const uint64_t hk = k0 + (uint64_t)j * S;
            qr[j] = 0.0;
            qi[j] = 0.0;
            if (hk + h0 < count) {
                const uint64_t i = ck | spread((uint64_t)j * S);
                if (pairs) {
                    const double2 a = base[i], b = base[i ^ g.x];
                    qr[j] = fma(b.x, a.x, b.y * a.y);
                    qi[j] = fma(b.x, a.y, -b.y * a.x);
                } else {
                    const double2 a = base[i];
                    qr[j] = fma(a.x, a.x, a.y * a.y);
                }
            }
        }
        wht8(qr);
        if constexpr (PAIRS) wht8(qi);
        const uint64_t uk = spread(k0);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            uint32_t m = g.jm[t] & (kTermJ - 1);
            asm volatile("" : "+s"(m));  // derive the select masks here, not once per term before the loop
            const double w = (PAIRS && ((g.imag >> t) & 1u)) ? pick8(qi, m) : pick8(qr, m);
            acc[t] = (__popcll(uk & g.z[t]) & 1) ? acc[t] - w : acc[t] + w;
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++)
        if (__popcll(c & g.z[t]) & 1) acc[t] = -acc[t];
    // wavefront sums, then the workgroup's four wavefronts through LDS
    __shared__ double sh[kBlock / 64][NT];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        double v = acc[t];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) sh[wave][t] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < g.nt) {
        double v = 0.0;
        for (int w = 0; w < kBlock / 64; w++) v += sh[w][threadIdx.x];
        part[((elem * (uint64_t)rows) + (uint64_t)g.row[threadIdx.x]) * (uint64_t)parts + blockIdx.x] = v;
    }
}

template <int NT>
void launch_terms_nt(const TermGroup &g, dim3 grid, hipStream_t stream, const double2 *amp, int n, uint64_t e0,
                     int rows, int parts, double *part) {
    if (g.x != 0)
        hipLaunchKernelGGL((k_pauli_terms<NT, true>), grid, dim3(kBlock), 0, stream, amp, n, e0, g, rows, parts, part);
    else
        hipLaunchKernelGGL((k_pauli_terms<NT, false>), grid, dim3(kBlock), 0, stream, amp, n, e0, g, rows, parts, part);
}

// The k_pauli_terms instance for the group's term count and x mask (pairs, or one index per amplitude).
int launch_terms(const TermGroup &g, dim3 grid, hipStream_t stream, const double2 *amp, int n, uint64_t e0, int rows,
                 int parts, double *part) {
    if (g.nt <= 1) launch_terms_nt<1>(g, grid, stream, amp, n, e0, rows, parts, part);
    else if (g.nt <= 4) launch_terms_nt<4>(g, grid, stream, amp, n, e0, rows, parts, part);
    else if (g.nt <= 8) launch_terms_nt<8>(g, grid, stream, amp, n, e0, rows, parts, part);
    else if (g.nt <= 12) launch_terms_nt<12>(g, grid, stream, amp, n, e0, rows, parts, part);
    else if (g.nt <= 16) launch_terms_nt<16>(g, grid, stream, amp, n, e0, rows, parts, part);
    else if (g.nt <= 24) launch_terms_nt<24>(g, grid, stream, amp, n, e0, rows, parts, part);
    else launch_terms_nt<32>(g, grid, stream, amp, n, e0, rows, parts, part);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

// out[r] = sum of the parts partial sums of row r
__global__ void k_rows_final(const double *__restrict__ part, uint64_t rows, int parts, double *__restrict__ out) {
    const uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (r >= rows) return;
    double v = 0.0;
    for (int k = 0; k < parts; k++) v += part[r * (uint64_t)parts + k];
    out[r] = v;
}

// sum_i conj(a_i) b_i per element
__global__ void __launch_bounds__(kBlock) k_dot_partial(const double2 *__restrict__ a, const double2 *__restrict__ b,
                                                        int n, double *__restrict__ out) {
    const uint64_t elem = blockIdx.y;
    const uint64_t dim = 1ULL << n;
    const double2 *pa = a + (elem << n), *pb = b + (elem << n);
    double re = 0.0, im = 0.0;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < dim; i += (uint64_t)kParts * kBlock)
        cdot_acc(pa[i], pb[i], re, im);
    block_reduce2(re, im, out + (elem * kParts + blockIdx.x) * 2);
}

// lam = sum_t c_t i^ny_t (-1)^{|(i^x_t)&z_t|} psi_{i^x_t} in one pass: the terms are grouped by x (goff[k] ..
// goff[k+1] in tz / tc), so psi is gathered once per distinct x and lam written once.
__global__ void __launch_bounds__(kBlock) k_pauli_sum(const double2 *__restrict__ psi, double2 *__restrict__ lam,
                                                      int n, uint64_t total, const uint64_t *__restrict__ gx,
                                                      const int *__restrict__ goff, int ngroups,
                                                      const uint64_t *__restrict__ tz, const double *__restrict__ tc) {
    const uint64_t dmask = (1ULL << n) - 1ULL;
    for (uint64_t g = blockIdx.x * (uint64_t)kBlock + threadIdx.x; g < total; g += (uint64_t)gridDim.x * kBlock) {
        const uint64_t i = g & dmask;
        double ar = 0.0, ai = 0.0;
        for (int k = 0; k < ngroups; k++) {
            const uint64_t j = i ^ gx[k];
            const double2 v = psi[(g - i) + j];
            double wr = 0.0, wi = 0.0;
            for (int t = goff[k]; t < goff[k + 1]; t++) {
                const double sg = (__popcll(j & tz[t]) & 1) ? -1.0 : 1.0;
                wr = fma(sg, tc[2 * t], wr);
                wi = fma(sg, tc[2 * t + 1], wi);
            }
            ar += wr * v.x - wi * v.y;
            ai += wr * v.y + wi * v.x;
        }
        lam[g] = make_double2(ar, ai);
    }
}

// Max |a-b| over all components: grid-stride partial maxima.
__global__ void __launch_bounds__(kBlock) k_maxdiff(const double2 *__restrict__ a, const double2 *__restrict__ b,
                                                    uint64_t count, double *__restrict__ out) {
    __shared__ double sh[kBlock];
    double m = 0.0;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < count; i += (uint64_t)gridDim.x * kBlock) {
        double2 x = a[i], y = b[i];
        m = fmax(m, fmax(fabs(x.x - y.x), fabs(x.y - y.y)));
        if (x.x != x.x || x.y != x.y) m = INFINITY;
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

__global__ void k_add_to_first(double2 *amp, int n, int batch, double delta) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < batch) amp[(uint64_t)b << n].x += delta;
}

// ---- adjoint step kernels: grid (kParts, batch); inv = per-element inverse matrices ----------------
// Pair op on `target` with controls (compacted): U^dag 2x2 (or exact X), inner product with gen G on target.
__global__ void __launch_bounds__(kBlock) k_adj_pair(double2 *__restrict__ psi, double2 *__restrict__ lam, int n,
                                                     int target, uint64_t fixed, uint64_t controls, int log_groups,
                                                     const double *__restrict__ inv, uint64_t inv_stride, int is_x,
                                                     int gen, double *__restrict__ part) {
    const uint64_t elem = blockIdx.y;
    double2 *P = psi + (elem << n), *L = lam + (elem << n);
    const double *g = inv + elem * inv_stride;
    const double g0r = g[0], g0i = g[1], g1r = g[2], g1i = g[3], g2r = g[4], g2i = g[5], g3r = g[6], g3i = g[7];
    const uint64_t stride = 1ULL << target;
    const uint64_t groups = 1ULL << log_groups;
    double re = 0.0, im = 0.0;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < groups; i += (uint64_t)kParts * kBlock) {
        uint64_t i0 = expand_mask(i, fixed) | controls, i1 = i0 | stride;
        double2 p0 = P[i0], p1 = P[i1], l0 = L[i0], l1 = L[i1];
        if (gen == SC_GEN_X) {
            cdot_acc(l0, p1, re, im);
            cdot_acc(l1, p0, re, im);
        } else if (gen == SC_GEN_Y) {  // Y p: (-i p1, i p0)
            cdot_acc(l0, make_double2(p1.y, -p1.x), re, im);
            cdot_acc(l1, make_double2(-p0.y, p0.x), re, im);
        } else if (gen == SC_GEN_Z) {
            cdot_acc(l0, p0, re, im);
            cdot_acc(l1, make_double2(-p1.x, -p1.y), re, im);
        }
        if (is_x) {
            P[i0] = p1, P[i1] = p0, L[i0] = l1, L[i1] = l0;
        } else {
            P[i0] = cmad(g1r, g1i, p1, cmul(g0r, g0i, p0));
            P[i1] = cmad(g3r, g3i, p1, cmul(g2r, g2i, p0));
            L[i0] = cmad(g1r, g1i, l1, cmul(g0r, g0i, l0));
            L[i1] = cmad(g3r, g3i, l1, cmul(g2r, g2i, l0));
        }
    }
    if (part) block_reduce2(re, im, part + (elem * kParts + blockIdx.x) * 2);
}

__global__ void __launch_bounds__(kBlock) k_adj_two(double2 *__restrict__ psi, double2 *__restrict__ lam, int n,
                                                    int first, int second, const double *__restrict__ inv,
                                                    uint64_t inv_stride) {
    const uint64_t elem = blockIdx.y;
    double2 *S[2] = {psi + (elem << n), lam + (elem << n)};
    const double *m = inv + elem * inv_stride;
    const uint64_t fm = 1ULL << first, sm = 1ULL << second;
    const uint64_t groups = 1ULL << (n - 2);
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < groups; i += (uint64_t)kParts * kBlock) {
        uint64_t b = expand_mask(i, fm | sm);
        uint64_t off[4] = {b, b | sm, b | fm, b | fm | sm};
#pragma unroll
        for (int s = 0; s < 2; s++) {
            double2 a[4];
#pragma unroll
            for (int c = 0; c < 4; c++) a[c] = S[s][off[c]];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                double2 acc = cmul(m[8 * r], m[8 * r + 1], a[0]);
#pragma unroll
                for (int c = 1; c < 4; c++) acc = cmad(m[8 * r + 2 * c], m[8 * r + 2 * c + 1], a[c], acc);
                S[s][off[r]] = acc;
            }
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_adj_diag2(double2 *__restrict__ psi, double2 *__restrict__ lam, int n,
                                                      int first, int second, const double *__restrict__ inv,
                                                      uint64_t inv_stride, int gen, double *__restrict__ part) {
    const uint64_t elem = blockIdx.y;
    double2 *P = psi + (elem << n), *L = lam + (elem << n);
    const double *d = inv + elem * inv_stride;
    const uint64_t dim = 1ULL << n;
    double re = 0.0, im = 0.0;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < dim; i += (uint64_t)kParts * kBlock) {
        int bf = (int)((i >> first) & 1ULL), bs = (int)((i >> second) & 1ULL);
        int l = 2 * bf + bs;
        double2 p = P[i], q = L[i];
        if (gen == SC_GEN_ZZ) cdot_acc(q, (bf ^ bs) ? make_double2(-p.x, -p.y) : p, re, im);
        P[i] = cmul(d[2 * l], d[2 * l + 1], p);
        L[i] = cmul(d[2 * l], d[2 * l + 1], q);
    }
    if (part) block_reduce2(re, im, part + (elem * kParts + blockIdx.x) * 2);
}

__global__ void __launch_bounds__(kBlock) k_adj_swap(double2 *__restrict__ psi, double2 *__restrict__ lam, int n,
                                                     int first, int second) {
    const uint64_t elem = blockIdx.y;
    double2 *S[2] = {psi + (elem << n), lam + (elem << n)};
    const uint64_t fm = 1ULL << first, sm = 1ULL << second;
    const uint64_t groups = 1ULL << (n - 2);
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < groups; i += (uint64_t)kParts * kBlock) {
        uint64_t b = expand_mask(i, fm | sm);
#pragma unroll
        for (int s = 0; s < 2; s++) {
            double2 x = S[s][b | fm];
            S[s][b | fm] = S[s][b | sm];
            S[s][b | sm] = x;
        }
    }
}

// Final fixed-order reduction of kParts complex partials per (slot, elem).
__global__ void k_parts_final(const double *__restrict__ part, uint64_t count, double *__restrict__ out) {
    uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k >= count) return;
    double re = 0.0, im = 0.0;
    for (int b = 0; b < kParts; b++) {
        re += part[(k * kParts + b) * 2];
        im += part[(k * kParts + b) * 2 + 1];
    }
    out[2 * k] = re;
    out[2 * k + 1] = im;
}

unsigned grid_for(Context *ctx, uint64_t total) {
    uint64_t blocks = (total + kBlock - 1) / kBlock;
    uint64_t cap = (uint64_t)ctx->cu_count * 16;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    return (unsigned)blocks;
}

int set_device(Context *ctx) {
    SC_HIP_CHECK(hipSetDevice(ctx->device));
    return SC_OK;
}

// per-element complex results of kParts partials (count = slots*batch) -> host
int finish_parts(Context *ctx, double *part, uint64_t count, double *host) {
    double *out = part + count * kParts * 2;
    hipLaunchKernelGGL(k_parts_final, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, ctx->stream, part, count,
                       out);
    SC_HIP_CHECK(hipGetLastError());
    SC_HIP_CHECK(hipMemcpyAsync(host, out, count * 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

}  // namespace

// Partial sums per term row for k_pauli_terms: enough workgroups to fill the GPU, at least 8 amplitudes per
// thread, at most 32 Mi partials; a power of two (the kernel's index split).
uint64_t term_parts(const Context *ctx, uint64_t B, int n, uint64_t rows) {
    const uint64_t half = n > 0 ? (1ULL << (n - 1)) : 1ULL;
    uint64_t parts = std::max<uint64_t>(1, (8ULL * ctx->cu_count + B - 1) / B);
    parts = std::min<uint64_t>(parts, std::max<uint64_t>(1, half / (8ULL * kBlock)));
    parts = std::min<uint64_t>(parts, std::max<uint64_t>(1, (32ULL << 20) / (rows * B)));
    parts = std::min<uint64_t>(parts, 4096);
    return 1ULL << (63 - __builtin_clzll(parts));
}

// Device bytes pauli_terms_in needs for `parts` partials per row: the partials, then the row sums.
size_t term_scratch_bytes(uint64_t B, uint64_t rows, uint64_t parts) { return (size_t)(B * rows * (parts + 1)) * 8; }

// <P_t>_b for every term t and element b: vals[b * n_terms + t], each term a Hermitian Pauli string
// i^{|x & z|} X^x Z^z.  Terms are grouped by x mask (groups in order of first appearance); each launch
// evaluates up to kTermsPerLaunch terms of one group in one read of the state, and one synchronisation ends
// the call.  The partial sums live in `scratch` (scratch_bytes of device memory); fewer partials per row are
// used when the default count does not fit, and SC_E_ARG is returned when even one per row does not.
int pauli_terms_enqueue(Context *ctx, State *st, const double2 *amp, const uint64_t *xm, const uint64_t *zm,
                        int n_terms, void *scratch, size_t scratch_bytes, double *out) {
    const uint64_t B = (uint64_t)st->batch;
    const int n = st->n;
    const uint64_t rows = (uint64_t)n_terms;
    uint64_t parts = term_parts(ctx, B, n, rows);
    while (parts > 1 && term_scratch_bytes(B, rows, parts) > scratch_bytes) parts >>= 1;
    if (term_scratch_bytes(B, rows, parts) > scratch_bytes)
        return fail(SC_E_ARG, "expectation of " + std::to_string(n_terms) + " terms over " + std::to_string(B) +
                                  " states needs " + std::to_string(term_scratch_bytes(B, rows, 1)) +
                                  " bytes of scratch, more than the " + std::to_string(scratch_bytes) + " available");
    double *part = (double *)scratch;
    if (!out) out = (double *)((char *)scratch + (size_t)(B * rows * parts) * sizeof(double));
    for (int t0 = 0; t0 < n_terms; t0++) {
        bool seen = false;  // x mask of an earlier term: its group was done there
        for (int t = 0; t < t0 && !seen; t++) seen = xm[t] == xm[t0];
        if (seen) continue;
        size_t members = 0;
        for (int t = t0; t < n_terms; t++) members += xm[t] == xm[t0];
        // launches of equal size, at most kTermsPerLaunch terms each
        const size_t nl = (members + kTermsPerLaunch - 1) / kTermsPerLaunch;
        const size_t per = (members + nl - 1) / nl;
        int next = t0;  // next member term not yet launched
        for (size_t c0 = 0; c0 < members; c0 += per) {
            TermGroup g;
            std::memset(&g, 0, sizeof(g));
            g.x = xm[t0];
            g.nt = (int)std::min<size_t>(per, members - c0);
            // the index bits of spread(2^b S), b < 3, that the kernel's iterations step through
            const int pbit = g.x ? __builtin_ctzll(g.x) : 63;
            auto spread = [&](uint64_t h) {
                const uint64_t lo = pbit < 63 ? (h & ((1ULL << pbit) - 1ULL)) : h;
                return g.x ? (((h ^ lo) << 1) | lo) : h;
            };
            for (int j = 0; j < g.nt; j++) {
                while (xm[next] != xm[t0]) next++;
                const int t = next++;
                g.z[j] = zm[t];
                g.row[j] = t;
                if (__builtin_popcountll(xm[t] & zm[t]) & 1) g.imag |= 1u << j;
                for (int b = 0; b < 3; b++)
                    if (__builtin_popcountll(spread((uint64_t)parts * kBlock << b) & zm[t]) & 1) g.jm[j] |= (uint8_t)(1u << b);
            }
            for (uint64_t e0 = 0; e0 < B; e0 += 65535) {
                const unsigned cnt = (unsigned)std::min<uint64_t>(65535, B - e0);
                ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_ADJOINT, 16.0 * (double)((uint64_t)cnt << n));
                const dim3 grid((unsigned)parts, cnt);
                const int rc = launch_terms(g, grid, ctx->stream, amp, n, e0, (int)rows, (int)parts, part);
                if (rc) return rc;
                SC_HIP_CHECK(hipGetLastError());
            }
        }
    }
    const uint64_t total_rows = B * rows;
    hipLaunchKernelGGL(k_rows_final, dim3((unsigned)((total_rows + 255) / 256)), dim3(256), 0, ctx->stream, part,
                       total_rows, (int)parts, out);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

// Host side of pauli_terms_enqueue's rows: the factor 2 of x-paired terms and the sign of i^{|x & z|}.
void pauli_terms_finish(const uint64_t *xm, const uint64_t *zm, int n_terms, uint64_t B, double *vals) {
    for (int t = 0; t < n_terms; t++) {
        const int ny = __builtin_popcountll(xm[t] & zm[t]) & 3;
        const double f = (xm[t] != 0 ? 2.0 : 1.0) * ((ny == 1 || ny == 2) ? -1.0 : 1.0);
        for (uint64_t b = 0; b < B; b++) vals[b * (uint64_t)n_terms + t] *= f;
    }
}

namespace {

// Enqueue, copy the rows to the host, synchronise, finish.
int pauli_terms_in(Context *ctx, State *st, const double2 *amp, const uint64_t *xm, const uint64_t *zm, int n_terms,
                   double *vals, void *scratch, size_t scratch_bytes) {
    const uint64_t B = (uint64_t)st->batch, rows = (uint64_t)n_terms;
    uint64_t parts = term_parts(ctx, B, st->n, rows);
    while (parts > 1 && term_scratch_bytes(B, rows, parts) > scratch_bytes) parts >>= 1;
    double *out = (double *)((char *)scratch + (size_t)(B * rows * parts) * sizeof(double));
    int rc = pauli_terms_enqueue(ctx, st, amp, xm, zm, n_terms, scratch, scratch_bytes, out);
    if (rc) return rc;
    SC_HIP_CHECK(hipMemcpyAsync(vals, out, B * rows * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    pauli_terms_finish(xm, zm, n_terms, B, vals);
    return SC_OK;
}

// The same with the context's scratch, grown to the default number of partials.
int pauli_terms(Context *ctx, State *st, const double2 *amp, const uint64_t *xm, const uint64_t *zm, int n_terms,
                double *vals) {
    const uint64_t B = (uint64_t)st->batch, rows = (uint64_t)n_terms;
    int rc = ensure_scratch(ctx, term_scratch_bytes(B, rows, term_parts(ctx, B, st->n, rows)));
    if (rc) return rc;
    return pauli_terms_in(ctx, st, amp, xm, zm, n_terms, vals, ctx->scratch, ctx->scratch_bytes);
}

// ---- the fixed scratch array of sc_expectation_pauli_terms_noalloc ------------------------------------
// 32 MiB of partial sums that every GPU the library runs on holds until the process ends.  In host code the
// name `scratch` is only a stand-in for the symbol: its device address, one per GPU, comes from
// hipGetSymbolAddress with that GPU current (init_noalloc_scratch, called by sc_ctx_create) and is the only
// form in which host code passes the array on.
constexpr size_t kNoAllocScratchBytes = 32ULL << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

struct NoAllocSlot {
    std::mutex mu;
    void *ptr = nullptr;  // device address of `scratch` on this GPU
};
constexpr int kNoAllocDevices = 64;
NoAllocSlot g_noalloc[kNoAllocDevices];

// <psi_b|H|psi_b> for all b: out[b]
int expectation(Context *ctx, State *st, const double2 *amp, const uint64_t *xm, const uint64_t *zm, const double *c,
                int n_terms, double *out) {
    const uint64_t B = (uint64_t)st->batch;
    std::vector<double> vals(B * (uint64_t)n_terms);
    int rc = pauli_terms(ctx, st, amp, xm, zm, n_terms, vals.data());
    if (rc) return rc;
    for (uint64_t b = 0; b < B; b++) {
        double e = 0.0;
        for (int t = 0; t < n_terms; t++) e += c[t] * vals[b * (uint64_t)n_terms + t];
        out[b] = e;
    }
    return SC_OK;
}

int check_terms(const State *st, const uint64_t *xm, const uint64_t *zm, const double *c, int n_terms,
                bool need_coeff = true) {
    if (n_terms <= 0) return fail(SC_E_ARG, "observable must have at least one term");
    if (!xm || !zm || (need_coeff && !c)) return fail(SC_E_NULL, "observable arrays are null");
    for (int t = 0; t < n_terms; t++)
        if (st->n < 64 && (((xm[t] | zm[t]) >> st->n) != 0))
            return fail(SC_E_ARG, "observable term " + std::to_string(t) + " acts outside the register");
    return SC_OK;
}

// U^dag of an op's matrix (row-major complex), count entries
void inverse_matrix(int kind, const double *m, double *out) {
    if (kind == SC_OP_DIAG2) {
        for (int j = 0; j < 4; j++) out[2 * j] = m[2 * j], out[2 * j + 1] = -m[2 * j + 1];
        return;
    }
    int d = kind == SC_OP_TWO ? 4 : 2;
    for (int r = 0; r < d; r++)
        for (int c = 0; c < d; c++) {
            out[2 * (r * d + c)] = m[2 * (c * d + r)];
            out[2 * (r * d + c) + 1] = -m[2 * (c * d + r) + 1];
        }
}

// The reverse sweep over ops [begin, end) as one op list: the ops in reverse order with their inverse matrices
// (inv_elems, same offsets as the forward matrices), and before the inverse of every parameterized op a probe
// of its generator (gradient slot slot[g]).
int adjoint_sequence(int n, const sc_op *ops, uint64_t begin, uint64_t end, const std::vector<int> &slot,
                     const std::vector<const double *> &inv_elems, uint64_t n_mat_complex, std::vector<PlanOp> &seq) {
    std::vector<sc_op> rev(ops + begin, ops + end);
    std::reverse(rev.begin(), rev.end());
    std::vector<PlanOp> inv_ops;
    int rc = analyze_ops(n, rev.data(), rev.size(), inv_elems, n_mat_complex, inv_ops);
    if (rc) return rc;
    seq.clear();
    seq.reserve(2 * inv_ops.size());
    for (size_t j = 0; j < inv_ops.size(); j++) {
        const uint64_t g = end - 1 - j;
        if (slot[g] >= 0) {
            const sc_op &o = ops[g];
            PlanOp pr;
            pr.kind = kPlanGrad;
            pr.q0 = o.q0;
            pr.q1 = o.gen == SC_GEN_ZZ ? o.q1 : -1;
            pr.mat = slot[g];
            pr.gen = o.gen;
            pr.is_real = true;
            if (o.gen == SC_GEN_X || o.gen == SC_GEN_Y) {
                pr.targets = 1ULL << o.q0;
            } else {
                pr.is_diag = true;
                pr.diag = (1ULL << o.q0) | (pr.q1 >= 0 ? 1ULL << pr.q1 : 0ULL);
            }
            seq.push_back(pr);
        }
        seq.push_back(inv_ops[j]);
    }
    return SC_OK;
}

// Sum of gsub partials per (slot, element), in order: out[k] = sum_s part[k * gsub + s].
__global__ void k_sub_final(const double *__restrict__ part, uint64_t count, int gsub, double *__restrict__ out) {
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k >= count) return;
    double re = 0.0, im = 0.0;
    for (int s = 0; s < gsub; s++) {
        re += part[(k * gsub + s) * 2];
        im += part[(k * gsub + s) * 2 + 1];
    }
    out[2 * k] = re;
    out[2 * k + 1] = im;
}

// The reverse sweep as runtime-compiled adjoint passes.  Per checkpoint segment (last first) the reversed ops
// with inverse matrices, and a gradient probe before the inverse of every parameterized op, are planned like a
// forward circuit; each pass then reads and writes psi and lambda once for all the ops and probes it holds,
// instead of once per op.  res receives <lambda| G |psi> per (slot, element).
int fused_backward(Context *ctx, State *st, double2 *lam, const sc_op *ops, uint64_t n_ops, const std::vector<int> &slot,
                   int n_slots, const std::vector<const double *> &inv_elems, uint64_t n_mat_complex, uint64_t C,
                   uint64_t n_ckpt, const std::function<int(uint64_t)> &check_drift, double *res) {
    const int n = st->n;
    const uint64_t B = (uint64_t)st->batch;
    const int log_chunks = n - kChunkBits;
    const uint64_t nchunks = 1ULL << log_chunks;
    const uint64_t glimit = (uint64_t)std::max(1, sweep_grid_size(ctx));
    const uint64_t Bl = std::min<uint64_t>(B, glimit);
    const int gsub = (int)std::max<uint64_t>(1, std::min<uint64_t>(nchunks, glimit / Bl));
    const uint64_t count = (uint64_t)n_slots * B;
    DevBuf dpart;
    int rc = dpart.alloc(std::max<uint64_t>(count * (uint64_t)(gsub + 1), 1) * 2 * sizeof(double));
    if (rc) return rc;
    if (count > 0) SC_HIP_CHECK(hipMemsetAsync(dpart.p, 0, count * (uint64_t)gsub * 2 * sizeof(double), ctx->stream));
    double *part = (double *)dpart.p;
    std::vector<double> corr(count, 1.0);  // probe factors of unit-pivot ops, per (slot, element)
    if (!ctx->plan_cache) ctx->plan_cache = new PlanCache();
    for (int64_t k = (int64_t)n_ckpt; k >= 0; k--) {
        const uint64_t begin = C > 0 ? (uint64_t)k * C : 0;
        const uint64_t end = (C > 0 && (uint64_t)k < n_ckpt) ? ((uint64_t)k + 1) * C : n_ops;
        if ((uint64_t)k < n_ckpt && (rc = check_drift((uint64_t)k))) return rc;
        if (end <= begin) continue;
        std::vector<PlanOp> seq;
        if ((rc = adjoint_sequence(n, ops, begin, end, slot, inv_elems, n_mat_complex, seq))) return rc;
        // plan and kernels of this segment's structure, cached like forward plans
        const int max_fuse = std::min(ctx->max_fused, 2);  // both states' registers stay within the budget
        const int low = ctx->low_qubits < 0 ? kDefaultLowQubits : ctx->low_qubits;
        const std::string key = plan_cache_key(n, seq, low, false, max_fuse, ctx->jit, 0.0) + "adjoint" +
                                (ctx->pass_sched >= 3 ? "+prio" : "");
        PlanCache::Entry *ent = ctx->plan_cache->find(key);
        if (!ent) {
            std::unique_ptr<PlanCache::Entry> e(new PlanCache::Entry());
            e->key = key;
            if ((rc = plan_sweeps(n, seq, low, false, e->plan, max_fuse, kAdjointPassCost))) return rc;
            if (!e->plan.passes.empty() && (rc = jit_adjoint_kernels(ctx, e->plan, e->fns))) return rc;
            ent = ctx->plan_cache->insert(std::move(e));
        }
        ent->used = ++ctx->plan_cache->tick;
        const SweepPlan &plan = ent->plan;
        if (plan.passes.empty()) continue;
        // matrix pools of the inverse matrices, one per element when they differ
        const uint64_t E = inv_elems.size();
        for (uint64_t e = 0; e < E && !plan.probe_pivots.empty(); e++) {
            std::vector<std::pair<uint64_t, double>> pf;
            probe_factors(plan, inv_elems[e], pf);
            for (const auto &x : pf)
                for (uint64_t b = (E > 1 ? e : 0); b < (E > 1 ? e + 1 : B); b++) corr[x.first * B + b] = x.second;
        }
        std::vector<double> hpool(2 * std::max<uint64_t>(plan.pool_complex, 1) * E);
        for (uint64_t e = 0; e < E; e++) build_pool(plan, inv_elems[e], hpool.data() + 2 * plan.pool_complex * e);
        DevBuf dpool;
        if ((rc = dpool.alloc(hpool.size() * sizeof(double)))) return rc;
        SC_HIP_CHECK(hipMemcpyAsync(dpool.p, hpool.data(), hpool.size() * sizeof(double), hipMemcpyHostToDevice,
                                    ctx->stream));
        const v2d *mats = (const v2d *)dpool.p;
        uint64_t stride = E > 1 ? plan.pool_complex : 0;
        v2d *psi = (v2d *)st->amp, *lamv = (v2d *)lam;
        int nn = n, lc = log_chunks, Bi = (int)B, gs = gsub;
        for (size_t p = 0; p < plan.passes.size(); p++) {
            for (uint64_t e0 = 0; e0 < B; e0 += Bl) {
                int e0i = (int)e0, nb = (int)std::min<uint64_t>(Bl, B - e0);
                void *args[] = {&psi, &lamv, &nn, &lc, &mats, &stride, &e0i, &nb, &gs, &part, &Bi};
                ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_ADJOINT, 64.0 * (double)((uint64_t)nb << n));
                SC_HIP_CHECK(hipModuleLaunchKernel(ent->fns[p], (unsigned)((uint64_t)nb * (uint64_t)gsub), 1, 1,
                                                   kThreads, 1, 1, 0, ctx->stream, args, nullptr));
            }
        }
        SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // the pool buffer is released with this segment
    }
    if (count > 0) {
        double *out = part + count * (uint64_t)gsub * 2;
        hipLaunchKernelGGL(k_sub_final, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, ctx->stream, part, count,
                           gsub, out);
        SC_HIP_CHECK(hipGetLastError());
        SC_HIP_CHECK(hipMemcpyAsync(res, out, count * 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (uint64_t i = 0; i < count; i++) res[2 * i + 1] *= corr[i];
    }
    return SC_OK;
}

}  // namespace

// Takes the device address of `scratch` on `device` (which must be current) once per device.
int init_noalloc_scratch(int device) {
    if (device < 0 || device >= kNoAllocDevices)
        return fail(SC_E_UNSUPPORTED, "device index " + std::to_string(device) + " beyond the library's table");
    NoAllocSlot &slot = g_noalloc[device];
    std::lock_guard<std::mutex> lock(slot.mu);
    if (slot.ptr) return SC_OK;
    void *p = nullptr;
    SC_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    slot.ptr = p;
    return SC_OK;
}

int adjoint_plan_host(int n, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex,
                      int n_params, int low_qubits, SweepPlan &plan, std::vector<double> &inv, std::vector<int> &slot) {
    if (n_ops > 0 && !ops) return fail(SC_E_NULL, "ops is null");
    if (n_mat_complex > 0 && !mats) return fail(SC_E_NULL, "mats is null");
    if (n < kChunkBits || n > 30) return fail(SC_E_ARG, "adjoint sweep plans need 12..30 qubits");
    if (low_qubits < 0 || low_qubits > 8) return fail(SC_E_ARG, "low qubits must be 0..8");
    if (n_params < 0) return fail(SC_E_ARG, "n_params must be non-negative");
    slot.assign(n_ops, -1);
    int n_slots = 0;
    for (uint64_t i = 0; i < n_ops; i++) {
        const sc_op &o = ops[i];
        if (o.param < 0) continue;
        if (o.param >= n_params) return fail(SC_E_ARG, "op " + std::to_string(i) + " parameter index out of range");
        const bool ok = (o.kind == SC_OP_SINGLE && (o.gen == SC_GEN_X || o.gen == SC_GEN_Y || o.gen == SC_GEN_Z)) ||
                        (o.kind == SC_OP_DIAG2 && o.gen == SC_GEN_ZZ);
        if (!ok) return fail(SC_E_UNSUPPORTED, "op " + std::to_string(i) + ": unsupported generator for adjoint passes");
        slot[i] = n_slots++;
    }
    inv.assign(2 * std::max<uint64_t>(n_mat_complex, 1), 0.0);
    for (uint64_t i = 0; i < n_ops; i++)
        if (ops[i].mat >= 0 && ops[i].kind != SC_OP_CNOT && ops[i].kind != SC_OP_SWAP)
            inverse_matrix(ops[i].kind, mats + 2 * (int64_t)ops[i].mat, inv.data() + 2 * ops[i].mat);
    std::vector<const double *> inv_elems{inv.data()};
    std::vector<PlanOp> seq;
    int rc = adjoint_sequence(n, ops, 0, n_ops, slot, inv_elems, n_mat_complex, seq);
    if (rc) return rc;
    return plan_sweeps(n, seq, low_qubits, false, plan, 2, kAdjointPassCost);
}

}  // namespace sc

using namespace sc;

extern "C" {

int sc_expectation_pauli(sc_state *st, const uint64_t *xmask, const uint64_t *zmask, const double *coeff, int n_terms,
                         double *out) {
    if (!st) return fail(SC_E_NULL, "state is null");
    if (!out) return fail(SC_E_NULL, "out is null");
    int rc = check_terms(st, xmask, zmask, coeff, n_terms);
    if (rc) return rc;
    if ((rc = set_device(st->ctx))) return rc;
    if ((rc = materialize_basis(st))) return rc;
    return expectation(st->ctx, st, st->amp, xmask, zmask, coeff, n_terms, out);
}

int sc_expectation_pauli_terms(sc_state *st, const uint64_t *xmask, const uint64_t *zmask, int n_terms,
                               double *out) {
    if (!st) return fail(SC_E_NULL, "state is null");
    if (!out) return fail(SC_E_NULL, "out is null");
    int rc = check_terms(st, xmask, zmask, nullptr, n_terms, false);
    if (rc) return rc;
    if ((rc = set_device(st->ctx))) return rc;
    if ((rc = materialize_basis(st))) return rc;
    return pauli_terms(st->ctx, st, st->amp, xmask, zmask, n_terms, out);
}

// No counterpart in the reference: sc_expectation_pauli_terms for callers that cannot allocate device memory.
int sc_expectation_pauli_terms_noalloc(sc_state *st, const uint64_t *xmask, const uint64_t *zmask, int n_terms,
                                       double *out) {
    if (!st) return fail(SC_E_NULL, "state is null");
    if (!out) return fail(SC_E_NULL, "out is null");
    int rc = check_terms(st, xmask, zmask, nullptr, n_terms, false);
    if (rc) return rc;
    if ((rc = set_device(st->ctx))) return rc;
    if ((rc = init_noalloc_scratch(st->ctx->device))) return rc;  // done by sc_ctx_create already
    NoAllocSlot &slot = g_noalloc[st->ctx->device];
    std::lock_guard<std::mutex> lock(slot.mu);  // calls on one device take the array in turn
    if ((rc = materialize_basis(st))) return rc;
    return pauli_terms_in(st->ctx, st, st->amp, xmask, zmask, n_terms, out, slot.ptr, kNoAllocScratchBytes);
}

int sc_adjoint_plan_host(int num_qubits, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex,
                         int n_params, int low_qubits, double *psi_inout, double *lam_inout, double *grads) {
    if (!psi_inout || !lam_inout || !grads) return fail(SC_E_NULL, "state, lambda and grads are required");
    SweepPlan plan;
    std::vector<double> inv;
    std::vector<int> slot;
    int rc = adjoint_plan_host(num_qubits, ops, n_ops, mats, n_mat_complex, n_params, low_qubits, plan, inv, slot);
    if (rc) return rc;
    int n_slots = 0;
    for (int v : slot) n_slots = std::max(n_slots, v + 1);
    std::vector<double> pool(2 * std::max<uint64_t>(plan.pool_complex, 1));
    build_pool(plan, inv.data(), pool.data());
    std::vector<double> sums(2 * (size_t)std::max(n_slots, 1), 0.0);
    execute_plan_host(num_qubits, plan, pool.data(), psi_inout, lam_inout, sums.data());
    std::vector<std::pair<uint64_t, double>> pf;
    probe_factors(plan, inv.data(), pf);
    for (const auto &x : pf) sums[2 * x.first + 1] *= x.second;
    std::fill(grads, grads + n_params, 0.0);
    for (uint64_t i = 0; i < n_ops; i++)
        if (slot[i] >= 0) grads[ops[i].param] += 2.0 * ops[i].coeff * sums[2 * slot[i] + 1];
    return SC_OK;
}

int sc_adjoint_stats(sc_ctx *ctx, uint64_t *reanchors, double *max_drift) {
    if (!ctx) return fail(SC_E_NULL, "ctx is null");
    if (reanchors) *reanchors = ctx->adj_reanchors;
    if (max_drift) *max_drift = ctx->adj_max_drift;
    return SC_OK;
}

int sc_adjoint_gradient(sc_state *st, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex,
                        int mats_per_element, const uint64_t *xmask, const uint64_t *zmask, const double *coeff,
                        int n_terms, int n_params, int checkpoint_every, double *grads, double *energies) {
    if (!st) return fail(SC_E_NULL, "state is null");
    if (n_ops > 0 && !ops) return fail(SC_E_NULL, "ops is null");
    if (n_mat_complex > 0 && !mats) return fail(SC_E_NULL, "mats is null");
    if (!grads) return fail(SC_E_NULL, "grads is null");
    if (n_params < 0) return fail(SC_E_ARG, "n_params must be non-negative");
    if (checkpoint_every < 0) return fail(SC_E_ARG, "checkpoint_every must be non-negative");
    sc_ctx *ctx = static_cast<sc_ctx *>(st->ctx);
    int rc = check_terms(st, xmask, zmask, coeff, n_terms);
    if (rc) return rc;
    if ((rc = set_device(ctx))) return rc;
    if ((rc = materialize_basis(st))) return rc;
    const int n = st->n;
    const uint64_t B = (uint64_t)st->batch;
    std::vector<const double *> elems;
    if (mats_per_element)
        for (uint64_t b = 0; b < B; b++) elems.push_back(mats + 2 * n_mat_complex * b);
    else
        elems.push_back(mats);
    std::vector<PlanOp> pops;
    if ((rc = analyze_ops(n, ops, n_ops, elems, n_mat_complex, pops))) return rc;
    if (pops.size() != n_ops) return fail(SC_E_ARG, "adjoint circuits must not contain measurements");
    if (n < 2) return fail(SC_E_ARG, "adjoint gradients need at least 2 qubits");
    // parameter metadata
    std::vector<int> slot(n_ops, -1);
    int n_slots = 0;
    for (uint64_t i = 0; i < n_ops; i++) {
        const sc_op &o = ops[i];
        if (o.param < 0) continue;
        if (o.param >= n_params) return fail(SC_E_ARG, "op " + std::to_string(i) + " parameter index out of range");
        bool ok = ((o.kind == SC_OP_SINGLE || o.kind == SC_OP_MCU) &&
                   (o.gen == SC_GEN_X || o.gen == SC_GEN_Y || o.gen == SC_GEN_Z)) ||
                  (o.kind == SC_OP_DIAG2 && o.gen == SC_GEN_ZZ);
        if (!ok) return fail(SC_E_UNSUPPORTED, "op " + std::to_string(i) + ": unsupported generator for its kind");
        slot[i] = n_slots++;
    }
    // inverse matrices per element: inv[e * n_mat_complex + off]
    const uint64_t E = elems.size();
    std::vector<double> inv(2 * std::max<uint64_t>(n_mat_complex, 1) * E, 0.0);
    for (uint64_t e = 0; e < E; e++)
        for (uint64_t i = 0; i < n_ops; i++)
            if (ops[i].mat >= 0 && ops[i].kind != SC_OP_CNOT && ops[i].kind != SC_OP_SWAP)
                inverse_matrix(ops[i].kind, elems[e] + 2 * (int64_t)ops[i].mat,
                               inv.data() + 2 * (e * n_mat_complex + ops[i].mat));
    DevBuf dinv, dpart;
    if ((rc = dinv.alloc(inv.size() * sizeof(double)))) return rc;
    SC_HIP_CHECK(hipMemcpyAsync(dinv.p, inv.data(), inv.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const uint64_t inv_stride = mats_per_element ? 2 * n_mat_complex : 0;
    const size_t state_bytes = (size_t)(B << n) * sizeof(double2);
    // ---- forward with checkpoints
    const uint64_t C = (uint64_t)checkpoint_every;
    uint64_t n_ckpt = (C > 0 && n_ops > 0) ? (n_ops - 1) / C : 0;
    // the context's workspace: n_ckpt checkpoints, lambda, and (when it fits) a copy of the forward output
    bool saved = false;
    {
        const size_t need = state_bytes * (n_ckpt + 2), least = state_bytes * (n_ckpt + 1);
        if (ctx->adj_ws_bytes < need) {
            SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            if (ctx->adj_ws) (void)hipFree(ctx->adj_ws);
            ctx->adj_ws = nullptr;
            ctx->adj_ws_bytes = 0;
            for (size_t want : {need, least}) {
                if (hipMalloc(&ctx->adj_ws, want) == hipSuccess) {
                    ctx->adj_ws_bytes = want;
                    break;
                }
                (void)hipGetLastError();
                ctx->adj_ws = nullptr;
            }
            if (!ctx->adj_ws)
                return fail(SC_E_OOM, "device allocation of " + std::to_string(least) +
                                          " bytes for the adjoint checkpoints and lambda failed");
        }
        saved = ctx->adj_ws_bytes >= need;
    }
    char *ws = (char *)ctx->adj_ws;
    char *ckpt = ws;  // checkpoint k at ckpt + k * state_bytes
    {
        // segment k ends in checkpoint k, segment k + 1 starts from it: with runtime-compiled passes the
        // segment's last pass writes the checkpoint and the next segment's first pass reads it, no copies
        uint64_t begin = 0;
        const double2 *cur = st->amp;
        for (uint64_t k = 0; k <= n_ckpt; k++) {
            const uint64_t end = (k < n_ckpt) ? (k + 1) * C : n_ops;
            double2 *dst = k < n_ckpt ? (double2 *)(ckpt + state_bytes * k) : st->amp;
            if (end > begin) {
                rc = apply_ops_between(st, ops + begin, end - begin, mats, n_mat_complex,
                                       mats_per_element ? n_mat_complex : 0, mats_per_element != 0, cur, dst);
                if (rc) return rc;
            } else if (cur != dst) {
                SC_HIP_CHECK(hipMemcpyAsync(dst, cur, state_bytes, hipMemcpyDeviceToDevice, ctx->stream));
            }
            cur = dst;
            begin = end;
        }
    }
    // ---- lambda = H psi
    double2 *lam = (double2 *)(ws + state_bytes * n_ckpt);
    const uint64_t total = B << n;
    {
        // terms grouped by their X mask, coefficients with i^ny folded in
        std::vector<uint64_t> gx, tz;
        std::vector<int> goff;
        std::vector<double> tc;
        std::vector<int> order(n_terms);
        for (int t = 0; t < n_terms; t++) order[t] = t;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return xmask[a] < xmask[b]; });
        for (int t : order) {
            if (gx.empty() || gx.back() != xmask[t]) {
                gx.push_back(xmask[t]);
                goff.push_back((int)tz.size());
            }
            const int ny = __builtin_popcountll(xmask[t] & zmask[t]) & 3;
            const double pr = ny == 0 ? 1.0 : ny == 2 ? -1.0 : 0.0, pi = ny == 1 ? 1.0 : ny == 3 ? -1.0 : 0.0;
            tz.push_back(zmask[t]);
            tc.push_back(coeff[t] * pr);
            tc.push_back(coeff[t] * pi);
        }
        goff.push_back((int)tz.size());
        const size_t bytes = (gx.size() + tz.size()) * 8 + goff.size() * 4 + tc.size() * 8 + 64;
        DevBuf dterms;
        if ((rc = dterms.alloc(bytes))) return rc;
        char *d = (char *)dterms.p;
        const uint64_t *dgx = (const uint64_t *)d;
        const uint64_t *dtz = dgx + gx.size();
        const double *dtc = (const double *)(dtz + tz.size());
        const int *dgoff = (const int *)(dtc + tc.size());
        if (!gx.empty()) {
            SC_HIP_CHECK(hipMemcpyAsync((void *)dgx, gx.data(), gx.size() * 8, hipMemcpyHostToDevice, ctx->stream));
            SC_HIP_CHECK(hipMemcpyAsync((void *)dtz, tz.data(), tz.size() * 8, hipMemcpyHostToDevice, ctx->stream));
            SC_HIP_CHECK(hipMemcpyAsync((void *)dtc, tc.data(), tc.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        }
        SC_HIP_CHECK(hipMemcpyAsync((void *)dgoff, goff.data(), goff.size() * 4, hipMemcpyHostToDevice, ctx->stream));
        ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_ADJOINT, (16.0 * (double)gx.size() + 16.0) * (double)total);
        // a few elements per launch: the partner amplitudes psi_{i^x} of every X mask then come from a working
        // set the last-level cache can hold, instead of the whole batch's
        const uint64_t per = std::max<uint64_t>(1, (256ULL << 20) / (16ULL << n));
        for (uint64_t e0 = 0; e0 < B; e0 += per) {
            const uint64_t cnt = std::min<uint64_t>(per, B - e0);
            const uint64_t sub_total = cnt << n;
            hipLaunchKernelGGL(k_pauli_sum, dim3(grid_for(ctx, sub_total)), dim3(kBlock), 0, ctx->stream,
                               st->amp + (e0 << n), lam + (e0 << n), n, sub_total, dgx, dgoff, (int)gx.size(), dtz,
                               dtc);
            SC_HIP_CHECK(hipGetLastError());
        }
        SC_HIP_CHECK(hipGetLastError());
        SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // dterms is released at the end of this block
    }
    // ---- energies E_b = Re <psi|lambda>
    const uint64_t part_count = (uint64_t)std::max(n_slots, 1) * B;
    if ((rc = dpart.alloc((part_count * kParts + part_count) * 2 * sizeof(double) + B * (kParts + 1) * 16)))
        return rc;
    double *part = (double *)dpart.p;
    double *epart = part + (part_count * kParts + part_count) * 2;
    if (energies) {
        hipLaunchKernelGGL(k_dot_partial, dim3(kParts, (unsigned)B), dim3(kBlock), 0, ctx->stream, st->amp, lam, n,
                           epart);
        SC_HIP_CHECK(hipGetLastError());
        std::vector<double> e(2 * B);
        if ((rc = finish_parts(ctx, epart, B, e.data()))) return rc;
        for (uint64_t b = 0; b < B; b++) energies[b] = e[2 * b];
    }
    // ---- the forward output is what psi must hold when this call returns: the reverse sweep uncomputes a copy
    // of it (far cheaper than running the circuit again, which remains the fallback when the copy does not fit:
    // then the sweep runs on psi itself)
    double2 *dsave = (double2 *)(ws + state_bytes * (n_ckpt + 1));
    if (saved) SC_HIP_CHECK(hipMemcpyAsync(dsave, st->amp, state_bytes, hipMemcpyDeviceToDevice, ctx->stream));
    State work = *st;  // the state the reverse sweep uncomputes
    if (saved) work.amp = dsave;
    // ---- reverse sweep
    ctx->adj_reanchors = 0;
    ctx->adj_max_drift = 0.0;
    std::vector<double> mx;
    const unsigned dgrid = grid_for(ctx, total);
    // psi should equal the state the forward run checkpointed after its segment k: re-anchor when it drifted
    auto check_drift = [&](uint64_t k) -> int {
        double2 *ck = (double2 *)(ckpt + state_bytes * k);
        if (ctx->test_drift > 0) {
            hipLaunchKernelGGL(k_add_to_first, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, ctx->stream,
                               work.amp, n, (int)B, (double)ctx->test_drift * 1e-12);
            SC_HIP_CHECK(hipGetLastError());
        }
        int rc2 = ensure_scratch(ctx, (size_t)dgrid * sizeof(double));
        if (rc2) return rc2;
        hipLaunchKernelGGL(k_maxdiff, dim3(dgrid), dim3(kBlock), 0, ctx->stream, work.amp, ck, total,
                           (double *)ctx->scratch);
        SC_HIP_CHECK(hipGetLastError());
        mx.resize(dgrid);
        SC_HIP_CHECK(hipMemcpyAsync(mx.data(), ctx->scratch, dgrid * sizeof(double), hipMemcpyDeviceToHost,
                                    ctx->stream));
        SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        double drift = *std::max_element(mx.begin(), mx.end());
        ctx->adj_max_drift = std::max(ctx->adj_max_drift, drift);
        if (!(drift <= 1e-10)) {
            SC_HIP_CHECK(hipMemcpyAsync(work.amp, ck, state_bytes, hipMemcpyDeviceToDevice, ctx->stream));
            ctx->adj_reanchors++;
        }
        return SC_OK;
    };
    std::vector<double> res(2 * (uint64_t)n_slots * B, 0.0);
    bool fused = ctx->jit != 0 && ctx->exec_mode == SC_EXEC_SWEEP && n >= kChunkBits;
    for (uint64_t i = 0; i < n_ops && fused; i++)
        if (slot[i] >= 0 && ops[i].kind == SC_OP_MCU) fused = false;  // controlled generators: per-op path
    if (fused) {
        std::vector<const double *> inv_elems;
        for (uint64_t e = 0; e < E; e++) inv_elems.push_back(inv.data() + 2 * e * n_mat_complex);
        rc = fused_backward(ctx, &work, lam, ops, n_ops, slot, n_slots, inv_elems, n_mat_complex, C, n_ckpt,
                            check_drift, res.data());
        if (rc) return rc;
    } else {
        for (int64_t g = (int64_t)n_ops - 1; g >= 0; g--) {
            if (C > 0 && (uint64_t)(g + 1) % C == 0 && (uint64_t)(g + 1) / C >= 1 && (uint64_t)(g + 1) / C <= n_ckpt)
                if ((rc = check_drift((uint64_t)(g + 1) / C - 1))) return rc;
            const PlanOp &op = pops[g];
            const sc_op &o = ops[g];
            double *pslot = slot[g] >= 0 ? part + (uint64_t)slot[g] * B * kParts * 2 : nullptr;
            const double *dinvp = (const double *)dinv.p + 2 * (o.mat >= 0 ? o.mat : 0);
            ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_ADJOINT, 64.0 * (double)total);
            switch (op.kind) {
            case SC_OP_SINGLE:
            case SC_OP_MCU:
            case SC_OP_CNOT: {
                int target = op.kind == SC_OP_CNOT ? op.q1 : op.q0;
                uint64_t controls = op.controls;
                int lg = n - 1 - __builtin_popcountll(controls);
                hipLaunchKernelGGL(k_adj_pair, dim3(kParts, (unsigned)B), dim3(kBlock), 0, ctx->stream, work.amp, lam,
                                   n, target, controls | (1ULL << target), controls, lg, dinvp, inv_stride,
                                   op.is_x ? 1 : 0, slot[g] >= 0 ? o.gen : 0, pslot);
                break;
            }
            case SC_OP_TWO:
                hipLaunchKernelGGL(k_adj_two, dim3(kParts, (unsigned)B), dim3(kBlock), 0, ctx->stream, work.amp, lam,
                                   n, op.q0, op.q1, dinvp, inv_stride);
                break;
            case SC_OP_DIAG2:
                hipLaunchKernelGGL(k_adj_diag2, dim3(kParts, (unsigned)B), dim3(kBlock), 0, ctx->stream, work.amp, lam,
                                   n, op.q0, op.q1, dinvp, inv_stride, slot[g] >= 0 ? o.gen : 0, pslot);
                break;
            case SC_OP_SWAP:
                hipLaunchKernelGGL(k_adj_swap, dim3(kParts, (unsigned)B), dim3(kBlock), 0, ctx->stream, work.amp, lam,
                                   n, op.q0, op.q1);
                break;
            }
            SC_HIP_CHECK(hipGetLastError());
        }
        if (n_slots > 0 && (rc = finish_parts(ctx, part, (uint64_t)n_slots * B, res.data()))) return rc;
    }
    // ---- gradients
    std::fill(grads, grads + B * (uint64_t)n_params, 0.0);
    for (uint64_t i = 0; i < n_ops; i++) {
        if (slot[i] < 0) continue;
        for (uint64_t b = 0; b < B; b++)
            grads[b * n_params + ops[i].param] += 2.0 * ops[i].coeff * res[2 * ((uint64_t)slot[i] * B + b) + 1];
    }
    // psi holds the forward output unless the reverse sweep had to uncompute it in place
    if (!saved) {
        set_error("");
        rc = apply_ops_impl(st, ops, n_ops, mats, n_mat_complex, mats_per_element ? n_mat_complex : 0,
                            mats_per_element != 0);
        if (rc) return rc;
    }
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

namespace {

int run_trajectories(sc_state *st, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex,
                     const uint8_t *noise_after, double p, const double *uniforms, uint64_t n_draws,
                     ZTermsRequest *terms) {
    if (!st) return fail(SC_E_NULL, "state is null");
    if (n_ops > 0 && (!ops || !noise_after)) return fail(SC_E_NULL, "ops is null");
    if (!uniforms) return fail(SC_E_NULL, "uniforms is null");
    if (!(p >= 0.0 && p <= 1.0)) return fail(SC_E_ARG, "probability must be in [0,1], got " + std::to_string(p));
    Context *ctx = st->ctx;
    int rc = set_device(ctx);
    if (rc) return rc;
    {
        std::vector<const double *> elems{mats};
        std::vector<PlanOp> pops;
        if ((rc = analyze_ops(st->n, ops, n_ops, elems, n_mat_complex, pops))) return rc;
        if (pops.size() != n_ops) return fail(SC_E_ARG, "trajectory circuits must not contain measurements");
    }
    auto channel_qubits = [&](uint64_t i) {
        return (ops[i].q1 >= 0 && ops[i].kind != SC_OP_SINGLE && ops[i].kind != SC_OP_MCU) ? 2 : 1;
    };
    uint64_t draws = 0;
    for (uint64_t i = 0; i < n_ops; i++)
        if (noise_after[i]) draws += channel_qubits(i);
    if (draws != n_draws)
        return fail(SC_E_ARG, "expected " + std::to_string(draws) + " uniforms per trajectory, got " +
                                  std::to_string(n_draws));
    // DepolarizingChannel (core/.../noise/DepolarizingChannel.java:49-76): sqrt(1-p) I, sqrt(p/3) X, Y, Z
    const double a = std::sqrt(1.0 - p), s = std::sqrt(p / 3.0);
    const double kraus[32] = {a, 0, 0, 0, 0, 0, a, 0,  0, 0, s, 0,  s, 0, 0, 0,
                              0, 0, 0, -s, 0, s, 0, 0, s, 0, 0, 0, 0, 0, -s, 0};
    // Every Kraus operator is a multiple of a unitary, K_j^dag K_j = c_j I, so the reference's branch weight
    // ||K_j psi||^2 (StatevectorEngine.computeBranchWeight :511-530) is c_j ||psi||^2.  ||psi||^2 is read
    // from the live states once; unitary ops keep it and a channel that picks j and renormalises by f leaves
    // c_j ||psi||^2 f^2.  Each trajectory's branches therefore follow from its uniforms before any amplitude
    // moves, and the noisy circuit becomes one circuit per trajectory: the ops with, after every noisy op,
    // the chosen K_j times its renormalisation factor as a single-qubit op on each channel qubit.
    double c[4];
    for (int j = 0; j < 4; j++) {
        const double *g = kraus + 8 * j;
        c[j] = g[0] * g[0] + g[1] * g[1] + g[4] * g[4] + g[5] * g[5];
    }
    const uint64_t B = (uint64_t)st->batch;
    std::vector<double> nrm(B);
    {
        const uint64_t zero = 0;
        // a pending basis state has norm 1 exactly; the trajectory circuit's first pass then generates it
        if (st->pending_basis >= 0) std::fill(nrm.begin(), nrm.end(), 1.0);
        else if ((rc = pauli_terms(ctx, st, st->amp, &zero, &zero, 1, nrm.data()))) return rc;
    }
    const uint64_t stride = n_mat_complex + 4 * n_draws;  // complex entries per trajectory
    std::vector<double> emats(2 * stride * B);
    for (uint64_t b = 0; b < B; b++) {
        double *m = emats.data() + 2 * stride * b;
        if (n_mat_complex) std::memcpy(m, mats, 2 * n_mat_complex * sizeof(double));
        double N = nrm[b];
        for (uint64_t d = 0; d < n_draws; d++) {
            double w[4], total = 0.0;
            bool bad = false;
            for (int j = 0; j < 4; j++) {
                w[j] = c[j] * N;
                if (!std::isfinite(w[j]) || w[j] < 0.0) bad = true;
                total += w[j];
            }
            if (bad || !std::isfinite(total) || total <= 1e-15)  // MIN_KRAUS_WEIGHT, StatevectorEngine.java:36
                return fail(SC_E_STATE, "invalid Kraus branch weights in trajectory " + std::to_string(b));
            double r = uniforms[b * n_draws + d] * total;  // SplittableRandom.nextDouble(bound)
            if (r >= total) r = std::nextafter(total, 0.0);
            int idx = 3;
            double cum = 0.0;
            for (int j = 0; j < 4; j++) {  // sampleWeightedIndex :579-588
                cum += w[j];
                if (r < cum) {
                    idx = j;
                    break;
                }
            }
            const double ns = w[idx];
            if (!std::isfinite(ns) || ns == 0.0)
                return fail(SC_E_STATE, "invalid Kraus branch weights in trajectory " + std::to_string(b));
            const double f = std::fabs(ns - 1.0) > 1e-12 ? 1.0 / std::sqrt(ns) : 1.0;  // renormalizeState :590-600
            for (int e = 0; e < 8; e++) m[2 * (n_mat_complex + 4 * d) + e] = kraus[8 * idx + e] * f;
            N = ns * f * f;
        }
    }
    std::vector<sc_op> nops;
    std::vector<bool> inserted;
    uint64_t d = 0;
    for (uint64_t i = 0; i < n_ops; i++) {
        nops.push_back(ops[i]);
        inserted.push_back(false);
        if (!noise_after[i]) continue;
        const int qs[2] = {ops[i].q0, ops[i].q1};
        for (int j = 0; j < channel_qubits(i); j++) {
            sc_op o;
            std::memset(&o, 0, sizeof(o));
            o.kind = SC_OP_SINGLE;
            o.q0 = qs[j];
            o.q1 = -1;
            o.mat = (int32_t)(n_mat_complex + 4 * d++);
            o.param = -1;
            nops.push_back(o);
            inserted.push_back(true);
        }
    }
    std::vector<const double *> elems;
    for (uint64_t b = 0; b < B; b++) elems.push_back(emats.data() + 2 * stride * b);
    std::vector<PlanOp> pops;
    if ((rc = analyze_ops(st->n, nops.data(), nops.size(), elems, stride, pops))) return rc;
    // the channel ops' matrices change from call to call: plan them as general complex 2x2, so that every
    // call of one circuit shares its plan and pass kernels
    for (size_t k = 0; k < pops.size(); k++) {
        if (!inserted[k]) continue;
        PlanOp &op = pops[k];
        op.is_diag = op.is_x = op.is_real = false;
        op.unit = -1;
        op.targets = 1ULL << op.q0;
        op.diag = 0;
    }
    return apply_planned(st, pops, elems, true, terms);
}

}  // namespace

int sc_run_depolarizing_trajectories(sc_state *st, const sc_op *ops, uint64_t n_ops, const double *mats,
                                     uint64_t n_mat_complex, const uint8_t *noise_after, double p,
                                     const double *uniforms, uint64_t n_draws) {
    return run_trajectories(st, ops, n_ops, mats, n_mat_complex, noise_after, p, uniforms, n_draws, nullptr);
}

int sc_run_depolarizing_trajectories_z(sc_state *st, const sc_op *ops, uint64_t n_ops, const double *mats,
                                       uint64_t n_mat_complex, const uint8_t *noise_after, double p,
                                       const double *uniforms, uint64_t n_draws, const uint64_t *zmask, int n_terms,
                                       double *out) {
    if (!st) return fail(SC_E_NULL, "state is null");
    if (!zmask || !out) return fail(SC_E_NULL, "zmask and out are required");
    if (n_terms <= 0) return fail(SC_E_ARG, "n_terms must be positive");
    const uint64_t dim = 1ULL << st->n;
    for (int t = 0; t < n_terms; t++)
        if (zmask[t] >= dim) return fail(SC_E_ARG, "Pauli mask beyond the state's qubits");
    ZTermsRequest req;
    req.z.assign(zmask, zmask + n_terms);
    req.vals = out;
    int rc = run_trajectories(st, ops, n_ops, mats, n_mat_complex, noise_after, p, uniforms, n_draws, &req);
    if (rc || req.done) return rc;
    // the circuit ran without the term sums (conditions of the fused pass not met): evaluate them now
    const std::vector<uint64_t> x((size_t)n_terms, 0ULL);
    if ((rc = set_device(st->ctx))) return rc;
    if ((rc = materialize_basis(st))) return rc;
    return pauli_terms(st->ctx, st, st->amp, x.data(), zmask, n_terms, out);
}

}  // extern "C"
