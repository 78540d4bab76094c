// sc_sweep.hip -- the LDS-blocked multi-gate sweep kernel.
//
// One workgroup (256 threads, 4 wavefronts) owns one chunk of 2^12
// amplitudes: 12 "local" qubits vary inside the chunk and every other qubit is
// fixed by the chunk index.  The chunk is read from HBM once (lanes on the six
// lowest local qubits, so a wavefront load covers >= 256 B runs), staged in
// 64 KiB of LDS, and then run through a list of stages.  In a stage every
// thread holds 16 amplitudes in VGPRs that differ in 4 "register" local
// qubits; all ops of the stage act on those registers:
//   * 2x2 / 4x4 / fused 16x16 gates whose targets are register qubits,
//   * controls and diagonal factors on ANY qubit (register, thread-index or
//     chunk-fixed), resolved per lane without data exchange,
//   * CNOT / SWAP as exact register exchanges.
// Between stages the chunk goes back through LDS with an XOR swizzle chosen by
// the planner so that every 8- and 16-lane group of a ds_read/write_b128 hits
// distinct banks.  Finally the chunk is written back to HBM once.  A pass
// therefore costs one read and one write of the state regardless of how many
// gates it applies.
#include "sc_internal.hpp"

namespace sc {

namespace {

__device__ __forceinline__ double2 cmul(double2 m, double2 x) {
    return make_double2(m.x * x.x - m.y * x.y, m.x * x.y + m.y * x.x);
}
__device__ __forceinline__ double2 cmad(double2 m, double2 x, double2 acc) {
    acc.x = fma(m.x, x.x, acc.x);
    acc.x = fma(-m.y, x.y, acc.x);
    acc.y = fma(m.x, x.y, acc.y);
    acc.y = fma(m.y, x.x, acc.y);
    return acc;
}
__device__ __forceinline__ double2 sel(bool c, double2 a, double2 b) {
    return make_double2(c ? a.x : b.x, c ? a.y : b.y);
}

// ---- register-level ops ------------------------------------------------------
template <int K>
__device__ __forceinline__ void op_u1(double2 (&a)[kRegs], const double2 *__restrict__ m) {
    const double2 g00 = m[0], g01 = m[1], g10 = m[2], g11 = m[3];
#pragma unroll
    for (int r0 = 0; r0 < kRegs; r0++) {
        if (r0 & (1 << K)) continue;
        const int r1 = r0 | (1 << K);
        double2 x = a[r0], y = a[r1];
        a[r0] = cmad(g01, y, cmul(g00, x));
        a[r1] = cmad(g11, y, cmul(g10, x));
    }
}

// Real 2x2 (Ry, H, real products): two real multiply-adds per output component.
template <int K>
__device__ __forceinline__ void op_u1r(double2 (&a)[kRegs], const double2 *__restrict__ m) {
    const double g00 = m[0].x, g01 = m[1].x, g10 = m[2].x, g11 = m[3].x;
#pragma unroll
    for (int r0 = 0; r0 < kRegs; r0++) {
        if (r0 & (1 << K)) continue;
        const int r1 = r0 | (1 << K);
        double2 x = a[r0], y = a[r1];
        a[r0] = make_double2(fma(g01, y.x, g00 * x.x), fma(g01, y.y, g00 * x.y));
        a[r1] = make_double2(fma(g11, y.x, g10 * x.x), fma(g11, y.y, g10 * x.y));
    }
}

template <int K0, int K1>
__device__ __forceinline__ void op_u2r(double2 (&a)[kRegs], const double2 *__restrict__ m) {
#pragma unroll
    for (int b = 0; b < kRegs; b++) {
        if (b & ((1 << K0) | (1 << K1))) continue;
        const int idx[4] = {b, b | (1 << K1), b | (1 << K0), b | (1 << K0) | (1 << K1)};
        double2 x[4];
#pragma unroll
        for (int c = 0; c < 4; c++) x[c] = a[idx[c]];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            double re = m[4 * r].x * x[0].x, im = m[4 * r].x * x[0].y;
#pragma unroll
            for (int c = 1; c < 4; c++) {
                re = fma(m[4 * r + c].x, x[c].x, re);
                im = fma(m[4 * r + c].x, x[c].y, im);
            }
            a[idx[r]] = make_double2(re, im);
        }
    }
}

__device__ __forceinline__ void op_dense4r(double2 (&a)[kRegs], const double2 *__restrict__ m) {
    double2 y[kRegs];
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        double re = m[r * kRegs].x * a[0].x, im = m[r * kRegs].x * a[0].y;
#pragma unroll
        for (int c = 1; c < kRegs; c++) {
            re = fma(m[r * kRegs + c].x, a[c].x, re);
            im = fma(m[r * kRegs + c].x, a[c].y, im);
        }
        y[r] = make_double2(re, im);
    }
#pragma unroll
    for (int r = 0; r < kRegs; r++) a[r] = y[r];
}

template <int K>
__device__ __forceinline__ void op_cu1(double2 (&a)[kRegs], const double2 *__restrict__ m, uint32_t rctrl,
                                       bool tcond) {
    const double2 g00 = m[0], g01 = m[1], g10 = m[2], g11 = m[3];
#pragma unroll
    for (int r0 = 0; r0 < kRegs; r0++) {
        if (r0 & (1 << K)) continue;
        if (((uint32_t)r0 & rctrl) != rctrl) continue;  // uniform: skips pairs whose register controls are 0
        const int r1 = r0 | (1 << K);
        double2 x = a[r0], y = a[r1];
        double2 nx = cmad(g01, y, cmul(g00, x));
        double2 ny = cmad(g11, y, cmul(g10, x));
        a[r0] = sel(tcond, nx, x);
        a[r1] = sel(tcond, ny, y);
    }
}

template <int K>
__device__ __forceinline__ void op_x(double2 (&a)[kRegs], uint32_t rctrl, bool tcond) {
#pragma unroll
    for (int r0 = 0; r0 < kRegs; r0++) {
        if (r0 & (1 << K)) continue;
        if (((uint32_t)r0 & rctrl) != rctrl) continue;
        const int r1 = r0 | (1 << K);
        double2 x = a[r0], y = a[r1];
        a[r0] = sel(tcond, y, x);
        a[r1] = sel(tcond, x, y);
    }
}

// 4x4 on (first = register bit K0 -> MSB of the local index, second = K1 -> LSB).
template <int K0, int K1>
__device__ __forceinline__ void op_u2(double2 (&a)[kRegs], const double2 *__restrict__ m) {
#pragma unroll
    for (int b = 0; b < kRegs; b++) {
        if (b & ((1 << K0) | (1 << K1))) continue;
        const int idx[4] = {b, b | (1 << K1), b | (1 << K0), b | (1 << K0) | (1 << K1)};
        double2 x[4];
#pragma unroll
        for (int c = 0; c < 4; c++) x[c] = a[idx[c]];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            double2 acc = cmul(m[4 * r], x[0]);
#pragma unroll
            for (int c = 1; c < 4; c++) acc = cmad(m[4 * r + c], x[c], acc);
            a[idx[r]] = acc;
        }
    }
}

template <int K0, int K1>
__device__ __forceinline__ void op_swap(double2 (&a)[kRegs]) {
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        if ((r & (1 << K0)) && !(r & (1 << K1))) {
            const int s = r ^ (1 << K0) ^ (1 << K1);
            double2 t = a[r];
            a[r] = a[s];
            a[s] = t;
        }
    }
}

__device__ __forceinline__ void op_dense4(double2 (&a)[kRegs], const double2 *__restrict__ m) {
    double2 y[kRegs];
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        double2 acc = cmul(m[r * kRegs], a[0]);
#pragma unroll
        for (int c = 1; c < kRegs; c++) acc = cmad(m[r * kRegs + c], a[c], acc);
        y[r] = acc;
    }
#pragma unroll
    for (int r = 0; r < kRegs; r++) a[r] = y[r];
}

// Diagonal on up to two qubits: entry index l = 2*bit(a) + bit(b) (nd == 2) or bit(a) (nd == 1).
__device__ __forceinline__ void op_diag(double2 (&a)[kRegs], const double2 *__restrict__ m, const SweepOp &op,
                                        bool ta, bool tb) {
    const double2 d0 = m[0], d1 = m[1];
    if (op.nd == 1) {
#pragma unroll
        for (int r = 0; r < kRegs; r++) {
            bool ba = ((uint32_t)r & op.da.rmask) != 0u || ta;
            a[r] = cmul(sel(ba, d1, d0), a[r]);
        }
        return;
    }
    const double2 d2 = m[2], d3 = m[3];
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        bool ba = ((uint32_t)r & op.da.rmask) != 0u || ta;
        bool bb = ((uint32_t)r & op.db.rmask) != 0u || tb;
        double2 lo = sel(bb, d1, d0);
        double2 hi = sel(bb, d3, d2);
        a[r] = cmul(sel(ba, hi, lo), a[r]);
    }
}

// Controlled one-qubit diagonal (e.g. a multi-controlled Z): phase only where every control bit is set.
__device__ __forceinline__ void op_cdiag(double2 (&a)[kRegs], const double2 *__restrict__ m, const SweepOp &op,
                                         bool ta, bool tcond) {
    const double2 d0 = m[0], d1 = m[1];
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        if (((uint32_t)r & op.rctrl) != op.rctrl) continue;
        bool ba = ((uint32_t)r & op.da.rmask) != 0u || ta;
        a[r] = sel(tcond, cmul(sel(ba, d1, d0), a[r]), a[r]);
    }
}

__device__ __forceinline__ bool bit_from(const BitSrc &s, uint32_t tid, uint64_t chunk_base) {
    uint64_t om = ((uint64_t)s.omask_hi << 32) | s.omask_lo;
    return ((tid & s.tmask) != 0u) || ((chunk_base & om) != 0ULL);
}

template <int K>
__device__ __forceinline__ void dispatch_u1(double2 (&a)[kRegs], const double2 *m, const SweepOp &op, bool ctrl,
                                            bool tcond) {
    if (ctrl) op_cu1<K>(a, m, op.rctrl, tcond);
    else op_u1<K>(a, m);
}

__device__ __forceinline__ void run_op_fixed_impl(double2 (&a)[kRegs], const SweepOp &op,
                                                  const double2 *__restrict__ m);

// LANE = false: no control or diagonal source is a thread-index bit, so every condition is uniform across
// the workgroup (register bits are compile-time per element, chunk-fixed bits come from chunk_base) and the
// per-element selects fold away.  LANE = true only reaches the op kinds that can carry thread-bit sources.
template <bool LANE>
__device__ __forceinline__ void run_op_impl(double2 (&a)[kRegs], const SweepOp &op, const double2 *__restrict__ mats,
                                            uint32_t tid, uint64_t chunk_base) {
    const double2 *m = mats + op.mat;
    switch (op.kind) {
    case SW_U1:
    case SW_X: {
        const bool ctrl = (op.rctrl | op.tctrl) != 0u || op.octrl != 0ULL;
        const bool tcond = LANE ? (((tid & op.tctrl) == op.tctrl) && ((chunk_base & op.octrl) == op.octrl)) : true;
        if (op.kind == SW_U1) {
            switch (op.k0) {
            case 0: dispatch_u1<0>(a, m, op, ctrl, tcond); break;
            case 1: dispatch_u1<1>(a, m, op, ctrl, tcond); break;
            case 2: dispatch_u1<2>(a, m, op, ctrl, tcond); break;
            default: dispatch_u1<3>(a, m, op, ctrl, tcond); break;
            }
        } else {
            switch (op.k0) {
            case 0: op_x<0>(a, op.rctrl, tcond); break;
            case 1: op_x<1>(a, op.rctrl, tcond); break;
            case 2: op_x<2>(a, op.rctrl, tcond); break;
            default: op_x<3>(a, op.rctrl, tcond); break;
            }
        }
        break;
    }
    case SW_DIAG: {
        bool ta = LANE ? bit_from(op.da, tid, chunk_base) : bit_from(op.da, 0u, chunk_base);
        bool tb = LANE ? bit_from(op.db, tid, chunk_base) : bit_from(op.db, 0u, chunk_base);
        const bool ctrl = (op.rctrl | op.tctrl) != 0u || op.octrl != 0ULL;
        if (!ctrl) {
            op_diag(a, m, op, ta, tb);
        } else {
            const bool tcond = LANE ? (((tid & op.tctrl) == op.tctrl) && ((chunk_base & op.octrl) == op.octrl)) : true;
            op_cdiag(a, m, op, ta, tcond);
        }
        break;
    }
    default:
        if constexpr (!LANE) run_op_fixed_impl(a, op, m);
        break;
    }
}

__device__ __forceinline__ void run_op(double2 (&a)[kRegs], const SweepOp &op, const double2 *__restrict__ mats,
                                       uint32_t tid, uint64_t chunk_base) {
    if ((op.tctrl | op.da.tmask | op.db.tmask) != 0u) {
        run_op_impl<true>(a, op, mats, tid, chunk_base);
    } else {
        if ((chunk_base & op.octrl) != op.octrl) return;  // outer controls unset: identity on this chunk
        run_op_impl<false>(a, op, mats, tid, chunk_base);
    }
}

// Register-only op kinds (no controls, no diagonal sources): identical for every lane.
__device__ __forceinline__ void run_op_fixed_impl(double2 (&a)[kRegs], const SweepOp &op,
                                                  const double2 *__restrict__ m) {
    switch (op.kind) {
    case SW_U2: {
        switch (op.k0 * 4 + op.k1) {
        case 1: op_u2<0, 1>(a, m); break;
        case 2: op_u2<0, 2>(a, m); break;
        case 3: op_u2<0, 3>(a, m); break;
        case 4: op_u2<1, 0>(a, m); break;
        case 6: op_u2<1, 2>(a, m); break;
        case 7: op_u2<1, 3>(a, m); break;
        case 8: op_u2<2, 0>(a, m); break;
        case 9: op_u2<2, 1>(a, m); break;
        case 11: op_u2<2, 3>(a, m); break;
        case 12: op_u2<3, 0>(a, m); break;
        case 13: op_u2<3, 1>(a, m); break;
        default: op_u2<3, 2>(a, m); break;
        }
        break;
    }
    case SW_SWAP: {
        int lo = op.k0 < op.k1 ? op.k0 : op.k1;
        int hi = op.k0 < op.k1 ? op.k1 : op.k0;
        switch (lo * 4 + hi) {
        case 1: op_swap<1, 0>(a); break;
        case 2: op_swap<2, 0>(a); break;
        case 3: op_swap<3, 0>(a); break;
        case 6: op_swap<2, 1>(a); break;
        case 7: op_swap<3, 1>(a); break;
        default: op_swap<3, 2>(a); break;
        }
        break;
    }
    case SW_DENSE4: op_dense4(a, m); break;
    case SW_DENSE4R: op_dense4r(a, m); break;
    case SW_U1R:
        switch (op.k0) {
        case 0: op_u1r<0>(a, m); break;
        case 1: op_u1r<1>(a, m); break;
        case 2: op_u1r<2>(a, m); break;
        default: op_u1r<3>(a, m); break;
        }
        break;
    case SW_U2R: {
        switch (op.k0 * 4 + op.k1) {
        case 1: op_u2r<0, 1>(a, m); break;
        case 2: op_u2r<0, 2>(a, m); break;
        case 3: op_u2r<0, 3>(a, m); break;
        case 4: op_u2r<1, 0>(a, m); break;
        case 6: op_u2r<1, 2>(a, m); break;
        case 7: op_u2r<1, 3>(a, m); break;
        case 8: op_u2r<2, 0>(a, m); break;
        case 9: op_u2r<2, 1>(a, m); break;
        case 11: op_u2r<2, 3>(a, m); break;
        case 12: op_u2r<3, 0>(a, m); break;
        case 13: op_u2r<3, 1>(a, m); break;
        default: op_u2r<3, 2>(a, m); break;
        }
        break;
    }
    default: break;
    }
}

__global__ void __launch_bounds__(kThreads) k_sweep(double2 *__restrict__ amp, int n, int log_chunks,
                                                    const SweepPass *__restrict__ pass,
                                                    const SweepOp *__restrict__ ops,
                                                    const double2 *__restrict__ mats, uint64_t mat_stride,
                                                    int elem_begin) {
    __shared__ double2 lds[kChunk];
    const uint32_t tid = threadIdx.x;
    const uint64_t bid = blockIdx.x;
    const uint64_t elem = (uint64_t)elem_begin + (bid >> log_chunks);
    const uint64_t chunk = bid & ((1ULL << log_chunks) - 1ULL);

    // chunk index -> global offset of the fixed (outer) qubits
    uint64_t chunk_base = 0;
    {
        uint64_t om = pass->outer_mask, c = chunk;
        while (om) {
            uint64_t low = om & (~om + 1ULL);
            if (c & 1ULL) chunk_base |= low;
            c >>= 1;
            om &= om - 1ULL;
        }
    }
    double2 *base = amp + (elem << n) + chunk_base;
    const double2 *emats = mats + elem * mat_stride;

    // ---- load phase: thread bits -> local qubits 0..7, register bits -> local 8..11
    uint64_t gthread = 0;
    uint32_t sthread = 0;
#pragma unroll
    for (int i = 0; i < kThreadBits; i++) {
        if ((tid >> i) & 1u) {
            gthread += pass->glob_thread[i];
            sthread ^= pass->io_thread_slot[i];
        }
    }
    double2 a[kRegs];
#pragma unroll
    for (int r = 0; r < kRegs; r++) a[r] = base[gthread + pass->glob_reg[r]];
#pragma unroll
    for (int r = 0; r < kRegs; r++) lds[sthread ^ pass->io_reg_slot[r]] = a[r];
    __syncthreads();

    const int nstages = pass->nstages;
    for (int s = 0; s < nstages; s++) {
        const SweepStage &stg = pass->stages[s];
        uint32_t ts = 0;
#pragma unroll
        for (int i = 0; i < kThreadBits; i++)
            if ((tid >> i) & 1u) ts ^= stg.thread_slot[i];
#pragma unroll
        for (int r = 0; r < kRegs; r++) a[r] = lds[ts ^ stg.lds_off[r]];
        for (int o = stg.op_begin; o < stg.op_end; o++) run_op(a, ops[o], emats, tid, chunk_base);
#pragma unroll
        for (int r = 0; r < kRegs; r++) lds[ts ^ stg.lds_off[r]] = a[r];
        __syncthreads();
    }

    // ---- store phase (possibly remapping which local qubit lands on which chunk position)
    uint32_t sout = 0;
#pragma unroll
    for (int i = 0; i < kThreadBits; i++)
        if ((tid >> i) & 1u) sout ^= pass->st_thread_slot[i];
#pragma unroll
    for (int r = 0; r < kRegs; r++) a[r] = lds[sout ^ pass->st_reg_slot[r]];
#pragma unroll
    for (int r = 0; r < kRegs; r++) base[gthread + pass->glob_reg[r]] = a[r];
}

}  // namespace

int launch_sweep(Context *ctx, State *st, const SweepPass *pass_dev, const SweepOp *ops_dev, const double *mats_dev,
                 int elem_begin, int elem_count, size_t mat_stride) {
    const int log_chunks = st->n - kChunkBits;
    const uint64_t blocks = (uint64_t)elem_count << log_chunks;
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_SWEEP, 32.0 * (double)((uint64_t)elem_count << st->n));
    hipLaunchKernelGGL(k_sweep, dim3((unsigned)blocks), dim3(kThreads), 0, ctx->stream, st->amp, st->n, log_chunks,
                       pass_dev, ops_dev, reinterpret_cast<const double2 *>(mats_dev), (uint64_t)mat_stride / 1,
                       elem_begin);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

}  // namespace sc
