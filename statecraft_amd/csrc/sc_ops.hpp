// sc_ops.hpp -- the register-level ops of the sweep kernels.
//
// Shared by the compiled sweep kernel (sc_sweep.hip), which reads each op from a plan at run time, and by
// the per-pass kernels the library compiles at run time with hipRTC (sc_jit.cpp), in which every op is a
// compile-time constant.  The header therefore depends on nothing but the HIP device runtime.
//
// A stage holds 16 amplitudes per thread in registers that differ in 4 "register" qubits of the chunk;
// every op reads 16 source registers and writes all 16 destination registers.
#pragma once

#ifdef __HIPCC_RTC__
using __hip_internal::int32_t;
using __hip_internal::int8_t;
using __hip_internal::uint16_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
using __hip_internal::uint8_t;
#else
#include <hip/hip_runtime.h>

#include <cstdint>
#endif

namespace sc {

// ---- sweep geometry: one chunk = 2^12 amplitudes (12 local qubits) in LDS; a stage holds 16 amplitudes
// per thread in VGPRs that differ in 4 "register" local bits, one per thread-index combination.
constexpr int kChunkBits = 12;
constexpr int kChunk = 1 << kChunkBits;
constexpr int kRegBits = 4;
constexpr int kRegs = 1 << kRegBits;
constexpr int kThreadBits = kChunkBits - kRegBits;  // 8
constexpr int kThreads = 1 << kThreadBits;          // 256
// Runtime-compiled passes may also use 13-qubit chunks (128 KiB of LDS, 512 threads, one workgroup per CU);
// plan tables are sized for the larger geometry.
constexpr int kMaxChunkBits = 13;
constexpr int kMaxThreadBits = kMaxChunkBits - kRegBits;  // 9

// Register-level op kinds of a stage.  Every kind reads the 16 source registers and writes all 16
// destination registers (the kernel alternates between two register sets).  Any op whose outer
// (chunk-fixed) controls octrl are not all set in the chunk is the identity there.
enum SweepOpKind : int32_t {
    SW_U1 = 1,       // 2x2 complex on register bit k0
    SW_U1R = 2,      // 2x2 real on register bit k0
    SW_U1C = 3,      // 2x2 complex on bit k0 where the register controls rctrl and thread controls tctrl are set
    SW_X = 4,        // exchange on bit k0 where the register controls rctrl are set
    SW_XL = 5,       // SW_X that also needs the thread controls tctrl (per-lane select)
    SW_U2 = 6,       // 4x4 complex on register bits k0 > k1 (k0 is the MSB of the 2-bit local index)
    SW_U2R = 7,      // 4x4 real on register bits k0 > k1
    SW_SWAP = 8,     // exchange of register bits k0 > k1
    SW_D4 = 9,       // 16x16 complex on the four register bits (fused block)
    SW_D4R = 10,     // 16x16 real
    SW_DIAG16 = 11,  // diagonal over the register index: 16 complex factors
    SW_DIAGL = 12,   // 4-entry diagonal m[2A + B]; A is register bit k0 (k0 >= 0) or the thread / outer bit
                     // (ta, oa); B is the thread / outer bit (tb, ob) or 0; applied where rctrl and tctrl are set
    SW_CX = 13,      // exchange on bit k0 where register bit k1 is set (CNOT inside the registers): moves only
    SW_GRAD = 14,    // adjoint passes only: accumulate <lambda| G |psi> for the Pauli generator G (ta = SC_GEN_*)
    SW_U1U = 15,     // runtime-compiled passes only: 2x2 with unit pivots on bit k0 (row r reads its pivot from
                     // column bit r of ta; tb = 0 complex, 1 real, 2 imaginary off-pivot ratios m[0], m[1])
    SW_SCALE = 16,   // runtime-compiled passes only: every register times the real factor m[0].x
                     // on register bit k0 (and k1 for ZZ) into gradient slot pad1; the states are unchanged
};

// Amplitudes as a native two-element vector, so that register arrays of them stay in VGPRs.
typedef double v2d __attribute__((ext_vector_type(2)));
typedef v2d Regs[kRegs];

// Matrices are read through the constant address space: they never change while a kernel runs, and this lets
// the compiler use scalar loads for them although the kernel stores to the state (with plain global pointers
// it must assume those stores may alias and uses vector loads that wait for every outstanding HBM access).
typedef const __attribute__((address_space(4))) v2d CV2;

__device__ __forceinline__ v2d make_v2d(double x, double y) {
    v2d r;
    r.x = x;
    r.y = y;
    return r;
}

__device__ __forceinline__ v2d cmul(v2d m, v2d x) {
    return make_v2d(m.x * x.x - m.y * x.y, m.x * x.y + m.y * x.x);
}
__device__ __forceinline__ v2d cmad(v2d m, v2d x, v2d acc) {
    acc.x = fma(m.x, x.x, acc.x);
    // contracted to fma(-m.y, x.y, acc.x); an explicit negation of every matrix entry makes instruction
    // selection preparation quadratic in the size of straight-line pass kernels
    acc.x = acc.x - m.y * x.y;
    acc.y = fma(m.x, x.y, acc.y);
    acc.y = fma(m.y, x.x, acc.y);
    return acc;
}
__device__ __forceinline__ v2d sel(bool c, v2d a, v2d b) {
    return make_v2d(c ? a.x : b.x, c ? a.y : b.y);
}

// Returns v unchanged, but the compiler can no longer treat the result as loop-invariant or common to other
// uses of v: per-lane addresses derived from it are recomputed where they are used instead of being hoisted
// out of the chunk loop and kept live (16-32 VGPRs each) across the stages.
template <typename T>
__device__ __forceinline__ T opaque(T v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Orders this wavefront's LDS accesses before and after it: enough for an exchange in which every wavefront
// reads back only slots it wrote (a wavefront's LDS operations execute in order), where other exchanges
// need a workgroup barrier.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// In-register transition of a stage: register bit J exchanges its qubit with lane bit 4 (ROW16) or 5.  For every
// register pair (r, r | 2^J) the lanes of r whose lane bit is set trade places with the lanes of r | 2^J whose
// lane bit is clear; v_permlane16_swap / v_permlane32_swap (gfx950) swap exactly those halves of two VGPRs
// (odd 16-lane rows of the first with even rows of the second / upper 32 lanes of the first with lower 32 of
// the second), one instruction per dword pair.
__device__ __forceinline__ void lane_swap_u32(uint32_t &lo, uint32_t &hi, bool row16) {
    if (row16) {
        const auto r = __builtin_amdgcn_permlane16_swap(lo, hi, false, false);
        lo = r[0];
        hi = r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane32_swap(lo, hi, false, false);
        lo = r[0];
        hi = r[1];
    }
}
__device__ __forceinline__ void lane_swap_f64(double &lo, double &hi, bool row16) {
    const uint64_t a = __builtin_bit_cast(uint64_t, lo), b = __builtin_bit_cast(uint64_t, hi);
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    lane_swap_u32(a0, b0, row16);
    lane_swap_u32(a1, b1, row16);
    lo = __builtin_bit_cast(double, ((uint64_t)a1 << 32) | a0);
    hi = __builtin_bit_cast(double, ((uint64_t)b1 << 32) | b0);
}
template <int J, bool ROW16>
__device__ __forceinline__ void lane_exchange(Regs &x) {
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        if (r & (1 << J)) continue;
        double lx = x[r].x, ly = x[r].y, hx = x[r | (1 << J)].x, hy = x[r | (1 << J)].y;
        lane_swap_f64(lx, hx, ROW16);
        lane_swap_f64(ly, hy, ROW16);
        x[r] = make_v2d(lx, ly);
        x[r | (1 << J)] = make_v2d(hx, hy);
    }
}

__device__ __forceinline__ void copy_all(const Regs &s, Regs &d) {
#pragma unroll
    for (int r = 0; r < kRegs; r++) d[r] = s[r];
}

// ---- register ops: src -> dst ----------------------------------------------
// 2x2 complex on register bit K.  With CTRL only the pairs whose register index has every rctrl bit set change.
template <int K, bool CTRL>
__device__ __forceinline__ void op_u1(const Regs &s, Regs &d, v2d g00, v2d g01, v2d g10, v2d g11,
                                      uint32_t rctrl) {
#pragma unroll
    for (int r0 = 0; r0 < kRegs; r0++) {
        if (r0 & (1 << K)) continue;
        const int r1 = r0 | (1 << K);
        if (!CTRL || ((uint32_t)r0 & rctrl) == rctrl) {
            d[r0] = cmad(g01, s[r1], cmul(g00, s[r0]));
            d[r1] = cmad(g11, s[r1], cmul(g10, s[r0]));
        } else {
            d[r0] = s[r0];
            d[r1] = s[r1];
        }
    }
}

// Real 2x2 (Ry, H, real products): two real multiply-adds per output component.
template <int K>
__device__ __forceinline__ void op_u1r(const Regs &s, Regs &d, CV2 *m) {
    const double g00 = m[0].x, g01 = m[1].x, g10 = m[2].x, g11 = m[3].x;
#pragma unroll
    for (int r0 = 0; r0 < kRegs; r0++) {
        if (r0 & (1 << K)) continue;
        const int r1 = r0 | (1 << K);
        const v2d x = s[r0], y = s[r1];
        d[r0] = make_v2d(fma(g01, y.x, g00 * x.x), fma(g01, y.y, g00 * x.y));
        d[r1] = make_v2d(fma(g11, y.x, g10 * x.x), fma(g11, y.y, g10 * x.y));
    }
}

// 2x2 with unit pivots on register bit K: output row r is s[pivot] + x_r * s[other], where the pivot is
// column C0 (row 0) or C1 (row 1) of the pair and x_r = m[r].  The pivot entries of the original matrix are
// a common real factor that the pass applies elsewhere, once per amplitude.  FORM 1: real x_r (one FMA per
// component), FORM 2: purely imaginary x_r (one FMA per component), FORM 0: complex (a complex multiply-add).
template <int K, int C0, int C1, int FORM>
__device__ __forceinline__ void op_u1u(const Regs &s, Regs &d, CV2 *m) {
    const v2d x0 = m[0], x1 = m[1];
#pragma unroll
    for (int r0 = 0; r0 < kRegs; r0++) {
        if (r0 & (1 << K)) continue;
        const int r1 = r0 | (1 << K);
        const v2d p0 = C0 ? s[r1] : s[r0], o0 = C0 ? s[r0] : s[r1];
        const v2d p1 = C1 ? s[r1] : s[r0], o1 = C1 ? s[r0] : s[r1];
        if (FORM == 1) {
            d[r0] = make_v2d(fma(x0.x, o0.x, p0.x), fma(x0.x, o0.y, p0.y));
            d[r1] = make_v2d(fma(x1.x, o1.x, p1.x), fma(x1.x, o1.y, p1.y));
        } else if (FORM == 2) {
            d[r0] = make_v2d(p0.x - x0.y * o0.y, fma(x0.y, o0.x, p0.y));
            d[r1] = make_v2d(p1.x - x1.y * o1.y, fma(x1.y, o1.x, p1.y));
        } else {
            d[r0] = cmad(x0, o0, p0);
            d[r1] = cmad(x1, o1, p1);
        }
    }
}

// Every register times the real factor m[0].x.
__device__ __forceinline__ void op_scale(const Regs &s, Regs &d, CV2 *m) {
    const double f = m[0].x;
#pragma unroll
    for (int r = 0; r < kRegs; r++) d[r] = make_v2d(f * s[r].x, f * s[r].y);
}

// Exchange on register bit K in the pairs whose register index has every rctrl bit set; with LANE the exchange
// also needs the lane condition c.
template <int K, bool LANE>
__device__ __forceinline__ void op_x(const Regs &s, Regs &d, uint32_t rctrl, bool c) {
#pragma unroll
    for (int r0 = 0; r0 < kRegs; r0++) {
        if (r0 & (1 << K)) continue;
        const int r1 = r0 | (1 << K);
        if (((uint32_t)r0 & rctrl) == rctrl) {
            if (LANE) {
                d[r0] = sel(c, s[r1], s[r0]);
                d[r1] = sel(c, s[r0], s[r1]);
            } else {
                d[r0] = s[r1];
                d[r1] = s[r0];
            }
        } else {
            d[r0] = s[r0];
            d[r1] = s[r1];
        }
    }
}

// 4x4 on register bits K0 > K1; local index l = 2 * bit(K0) + bit(K1).
template <int K0, int K1, bool REAL>
__device__ __forceinline__ void op_u2(const Regs &s, Regs &d, CV2 *m) {
#pragma unroll
    for (int b = 0; b < kRegs; b++) {
        if (b & ((1 << K0) | (1 << K1))) continue;
        const int idx[4] = {b, b | (1 << K1), b | (1 << K0), b | (1 << K0) | (1 << K1)};
#pragma unroll
        for (int r = 0; r < 4; r++) {
            if (REAL) {
                double re = m[4 * r].x * s[idx[0]].x, im = m[4 * r].x * s[idx[0]].y;
#pragma unroll
                for (int c = 1; c < 4; c++) {
                    re = fma(m[4 * r + c].x, s[idx[c]].x, re);
                    im = fma(m[4 * r + c].x, s[idx[c]].y, im);
                }
                d[idx[r]] = make_v2d(re, im);
            } else {
                v2d acc = cmul(m[4 * r], s[idx[0]]);
#pragma unroll
                for (int c = 1; c < 4; c++) acc = cmad(m[4 * r + c], s[idx[c]], acc);
                d[idx[r]] = acc;
            }
        }
    }
}

template <int K0, int K1>
__device__ __forceinline__ void op_swap(const Regs &s, Regs &d) {
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        const int b0 = (r >> K0) & 1, b1 = (r >> K1) & 1;
        d[r] = s[(r & ~((1 << K0) | (1 << K1))) | (b1 << K0) | (b0 << K1)];
    }
}

template <bool REAL>
__device__ __forceinline__ void op_d4(const Regs &s, Regs &d, CV2 *m) {
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        if (REAL) {
            double re = m[r * kRegs].x * s[0].x, im = m[r * kRegs].x * s[0].y;
#pragma unroll
            for (int c = 1; c < kRegs; c++) {
                re = fma(m[r * kRegs + c].x, s[c].x, re);
                im = fma(m[r * kRegs + c].x, s[c].y, im);
            }
            d[r] = make_v2d(re, im);
        } else {
            v2d acc = cmul(m[r * kRegs], s[0]);
#pragma unroll
            for (int c = 1; c < kRegs; c++) acc = cmad(m[r * kRegs + c], s[c], acc);
            d[r] = acc;
        }
    }
}

__device__ __forceinline__ void op_diag16(const Regs &s, Regs &d, CV2 *m) {
#pragma unroll
    for (int r = 0; r < kRegs; r++) d[r] = cmul(m[r], s[r]);
}

// Factor hi where register bit K is set, lo elsewhere (K == kRegBits: lo everywhere), in the elements whose
// register index has every rctrl bit set.
template <int K>
__device__ __forceinline__ void op_diagl(const Regs &s, Regs &d, v2d lo, v2d hi, uint32_t rctrl) {
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        if (((uint32_t)r & rctrl) == rctrl) d[r] = cmul((K < kRegBits && ((r >> K) & 1)) ? hi : lo, s[r]);
        else d[r] = s[r];
    }
}

template <int K0, int K1>
__device__ __forceinline__ void two_bit(int kind, const Regs &s, Regs &d, CV2 *m) {
    if (kind == SW_U2) op_u2<K0, K1, false>(s, d, m);
    else if (kind == SW_U2R) op_u2<K0, K1, true>(s, d, m);
    else op_swap<K0, K1>(s, d);
}

// A 32-byte SweepOp record as eight dwords (one scalar load) and its decoded fields.

struct Op {
    int kind, k0, k1, mat;
    uint32_t rctrl, tctrl, ta, tb;
    int oa, ob;
    uint64_t octrl;
};


// Exchange on register bit K where register bit KC is set: a CNOT inside the registers, moves only.
// The empty asm names the case: cases that differ only in which registers they move would otherwise be
// merged by the compiler into pointer selects over the register array, which forces the array into scratch.
template <int K, int KC>
__device__ __forceinline__ void op_cx(const Regs &s, Regs &d) {
#pragma unroll
    for (int r = 0; r < kRegs; r++) d[r] = s[r ^ (((r >> KC) & 1) << K)];
    asm volatile("; cx %0 %1" ::"n"(K), "n"(KC));
}

__device__ __forceinline__ void run_cx(int k0, int k1, const Regs &s, Regs &d) {
    switch (k0 * 4 + k1) {
    case 1: op_cx<0, 1>(s, d); break;
    case 2: op_cx<0, 2>(s, d); break;
    case 3: op_cx<0, 3>(s, d); break;
    case 4: op_cx<1, 0>(s, d); break;
    case 6: op_cx<1, 2>(s, d); break;
    case 7: op_cx<1, 3>(s, d); break;
    case 8: op_cx<2, 0>(s, d); break;
    case 9: op_cx<2, 1>(s, d); break;
    case 11: op_cx<2, 3>(s, d); break;
    case 12: op_cx<3, 0>(s, d); break;
    case 13: op_cx<3, 1>(s, d); break;
    default: op_cx<3, 2>(s, d); break;
    }
}

// ---- one entry point per op kind with the register bits as template arguments: run_op dispatches to these,
// and the runtime-compiled pass kernels call them directly with every argument a constant.
template <int K>
__device__ __forceinline__ void op_u1m(const Regs &s, Regs &d, CV2 *m) {
    op_u1<K, false>(s, d, m[0], m[1], m[2], m[3], 0u);
}

// 2x2 on bit K where the register controls rctrl and the thread controls tctrl are set (other lanes: identity)
template <int K>
__device__ __forceinline__ void op_u1c(const Regs &s, Regs &d, CV2 *m, uint32_t rctrl, uint32_t tctrl,
                                       uint32_t tid) {
    const bool c = (tid & tctrl) == tctrl;
    const v2d one = make_v2d(1.0, 0.0), zero = make_v2d(0.0, 0.0);
    op_u1<K, true>(s, d, sel(c, m[0], one), sel(c, m[1], zero), sel(c, m[2], zero), sel(c, m[3], one), rctrl);
}

template <int K>
__device__ __forceinline__ void op_xl(const Regs &s, Regs &d, uint32_t rctrl, uint32_t tctrl, uint32_t tid) {
    op_x<K, true>(s, d, rctrl, (tid & tctrl) == tctrl);
}

// SW_DIAGL with A on register bit K (K == kRegBits: A is the thread / outer bit ta / oa).  The thread and
// chunk-fixed bits are resolved once per lane: lo / hi are the factors for A = 0 / 1.
template <int K>
__device__ __forceinline__ void op_diagl_sel(const Regs &s, Regs &d, CV2 *m, uint32_t rctrl, uint32_t tctrl,
                                             uint32_t ta, uint32_t tb, int oa, int ob, uint32_t tid,
                                             uint64_t cb) {
    const bool bb = ((tid & tb) != 0u) || (ob >= 0 && ((cb >> ob) & 1ULL));
    v2d lo = sel(bb, m[1], m[0]), hi = sel(bb, m[3], m[2]);
    if (K == kRegBits) {
        const bool ba = ((tid & ta) != 0u) || (oa >= 0 && ((cb >> oa) & 1ULL));
        lo = sel(ba, hi, lo);
    }
    const bool c = (tid & tctrl) == tctrl;
    const v2d one = make_v2d(1.0, 0.0);
    op_diagl<K>(s, d, sel(c, lo, one), sel(c, hi, one), rctrl);
}

__device__ __forceinline__ void run_op(const Regs &s, Regs &d, const Op &op, CV2 *mats, uint32_t tid, uint64_t cb) {
    if ((cb & op.octrl) != op.octrl) {  // outer controls unset: identity on this chunk
        copy_all(s, d);
        return;
    }
    CV2 *m = mats + op.mat;
    switch (op.kind) {
    case SW_CX: run_cx(op.k0, op.k1, s, d); break;
    case SW_U1:
        switch (op.k0) {
        case 0: op_u1m<0>(s, d, m); break;
        case 1: op_u1m<1>(s, d, m); break;
        case 2: op_u1m<2>(s, d, m); break;
        default: op_u1m<3>(s, d, m); break;
        }
        break;
    case SW_U1R:
        switch (op.k0) {
        case 0: op_u1r<0>(s, d, m); break;
        case 1: op_u1r<1>(s, d, m); break;
        case 2: op_u1r<2>(s, d, m); break;
        default: op_u1r<3>(s, d, m); break;
        }
        break;
    case SW_U1C: {
        // lanes whose thread controls are unset apply the identity
        const bool c = (tid & op.tctrl) == op.tctrl;
        const v2d one = make_v2d(1.0, 0.0), zero = make_v2d(0.0, 0.0);
        const v2d g00 = sel(c, m[0], one), g01 = sel(c, m[1], zero);
        const v2d g10 = sel(c, m[2], zero), g11 = sel(c, m[3], one);
        switch (op.k0) {
        case 0: op_u1<0, true>(s, d, g00, g01, g10, g11, op.rctrl); break;
        case 1: op_u1<1, true>(s, d, g00, g01, g10, g11, op.rctrl); break;
        case 2: op_u1<2, true>(s, d, g00, g01, g10, g11, op.rctrl); break;
        default: op_u1<3, true>(s, d, g00, g01, g10, g11, op.rctrl); break;
        }
        break;
    }
    case SW_X:
        switch (op.k0) {
        case 0: op_x<0, false>(s, d, op.rctrl, true); break;
        case 1: op_x<1, false>(s, d, op.rctrl, true); break;
        case 2: op_x<2, false>(s, d, op.rctrl, true); break;
        default: op_x<3, false>(s, d, op.rctrl, true); break;
        }
        break;
    case SW_XL: {
        const bool c = (tid & op.tctrl) == op.tctrl;
        switch (op.k0) {
        case 0: op_x<0, true>(s, d, op.rctrl, c); break;
        case 1: op_x<1, true>(s, d, op.rctrl, c); break;
        case 2: op_x<2, true>(s, d, op.rctrl, c); break;
        default: op_x<3, true>(s, d, op.rctrl, c); break;
        }
        break;
    }
    case SW_U2:
    case SW_U2R:
    case SW_SWAP:
        switch (op.k0 * 4 + op.k1) {
        case 4: two_bit<1, 0>(op.kind, s, d, m); break;
        case 8: two_bit<2, 0>(op.kind, s, d, m); break;
        case 9: two_bit<2, 1>(op.kind, s, d, m); break;
        case 12: two_bit<3, 0>(op.kind, s, d, m); break;
        case 13: two_bit<3, 1>(op.kind, s, d, m); break;
        default: two_bit<3, 2>(op.kind, s, d, m); break;
        }
        break;
    case SW_D4: op_d4<false>(s, d, m); break;
    case SW_D4R: op_d4<true>(s, d, m); break;
    case SW_DIAG16: op_diag16(s, d, m); break;
    case SW_DIAGL: {
        // resolve the thread / chunk-fixed bits once per lane: lo / hi are the factors for A = 0 / 1
        const bool bb = ((tid & op.tb) != 0u) || (op.ob >= 0 && ((cb >> op.ob) & 1ULL));
        v2d lo = sel(bb, m[1], m[0]), hi = sel(bb, m[3], m[2]);
        if (op.k0 < 0) {
            const bool ba = ((tid & op.ta) != 0u) || (op.oa >= 0 && ((cb >> op.oa) & 1ULL));
            lo = sel(ba, hi, lo);
        }
        const bool c = (tid & op.tctrl) == op.tctrl;
        const v2d one = make_v2d(1.0, 0.0);
        lo = sel(c, lo, one);
        hi = sel(c, hi, one);
        switch (op.k0) {
        case 0: op_diagl<0>(s, d, lo, hi, op.rctrl); break;
        case 1: op_diagl<1>(s, d, lo, hi, op.rctrl); break;
        case 2: op_diagl<2>(s, d, lo, hi, op.rctrl); break;
        case 3: op_diagl<3>(s, d, lo, hi, op.rctrl); break;
        default: op_diagl<kRegBits>(s, d, lo, hi, op.rctrl); break;
        }
        break;
    }
    default: copy_all(s, d); break;
    }
}

// ---- adjoint passes: <lambda| G |psi> over one thread's 16 register amplitudes, G a Pauli generator on
// register bit K0 (GEN 1 = X, 2 = Y, 3 = Z) or Z on K0 and K1 (GEN 4 = ZZ).  re/im accumulate conj(lambda) * (G psi).
__device__ __forceinline__ void cdot_acc(v2d a, v2d b, double &re, double &im) {
    re = fma(a.x, b.x, re);
    re = fma(a.y, b.y, re);
    im = fma(a.x, b.y, im);
    im = im - a.y * b.x;
}

// Im conj(a) * b accumulated
__device__ __forceinline__ void cdot_im_acc(v2d a, v2d b, double &im) {
    im = fma(a.x, b.y, im);
    im = im - a.y * b.x;
}

// Im <lambda| G |psi> over the registers (a gradient needs only the imaginary part; re is left alone).
template <int GEN, int K0, int K1>
__device__ __forceinline__ void grad_acc(const Regs &psi, const Regs &lam, double &re, double &im) {
    (void)re;
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        const int b0 = (r >> K0) & 1;
        if (GEN == 1) {
            cdot_im_acc(lam[r], psi[r ^ (1 << K0)], im);
        } else if (GEN == 2) {  // Y psi: (-i psi1, i psi0)
            const v2d p = psi[r ^ (1 << K0)];
            cdot_im_acc(lam[r], b0 ? make_v2d(-p.y, p.x) : make_v2d(p.y, -p.x), im);
        } else if (GEN == 3) {
            const v2d p = psi[r];
            cdot_im_acc(lam[r], b0 ? make_v2d(-p.x, -p.y) : p, im);
        } else {
            const v2d p = psi[r];
            cdot_im_acc(lam[r], (b0 ^ ((r >> K1) & 1)) ? make_v2d(-p.x, -p.y) : p, im);
        }
    }
}

// chunk index -> global offset of the outer qubits (bit j of the chunk index sits at the j-th set bit of om)
__device__ __forceinline__ uint64_t chunk_base_of(uint64_t chunk, uint64_t om) {
    uint64_t cb = 0;
    while (om) {
        const uint64_t low = om & (~om + 1ULL);
        if (chunk & 1ULL) cb |= low;
        chunk >>= 1;
        om &= om - 1ULL;
    }
    return cb;
}

}  // namespace sc
