// sc_sweep.hip -- the LDS-blocked multi-gate sweep kernel.
//
// A pass splits the qubits into 12 "local" qubits and the rest.  A chunk is
// the 2^12 amplitudes that share the values of all other ("outer") qubits.
// The kernel is persistent: each workgroup (256 threads, 4 wavefronts) walks
// the chunks c = blockIdx.x, blockIdx.x + gridDim.x, ... and for each one
//   * writes the chunk's amplitudes, already in registers, into 64 KiB of LDS
//     (lanes on the lowest local qubits, so every wavefront load covers runs
//     of >= 256 contiguous bytes in HBM),
//   * issues the HBM loads of its NEXT chunk into a register buffer, so they
//     are in flight while this chunk is computed,
//   * runs the pass's stages: in a stage each thread holds 16 amplitudes
//     that differ in 4 "register" local qubits and applies the stage's ops to
//     them -- 2x2 / 4x4 / fused 16x16 gates whose targets are register
//     qubits, diagonals and controls whose bits are register, thread-index or
//     chunk-fixed bits, and exact exchanges for X / CNOT / SWAP,
//   * writes the chunk back to HBM, optionally moving local qubits to other
//     positions of the chunk (in-place qubit remapping).
// A pass therefore reads and writes the state once, however many gates it
// applies.
//
// Register ops read 16 source registers and write all 16 destination
// registers; the stage loop alternates between two register sets, so no op
// needs temporaries or copies back into its inputs.  Between stages the chunk
// goes through LDS with an XOR swizzle chosen by the planner so that every 8-
// and 16-lane group of a ds_read/write_b128 hits distinct banks.
#include <algorithm>

#include "sc_internal.hpp"
#include "sc_jit.hpp"

namespace sc {

namespace {

// The pass and its ops are read through the constant address space (scalar loads; see CV2).
typedef const __attribute__((address_space(4))) SweepPass CPass;
typedef const __attribute__((address_space(4))) SweepStage CStage;

struct OpWords {
    uint32_t w[8];
};
typedef const __attribute__((address_space(4))) OpWords COpW;
static_assert(sizeof(OpWords) == sizeof(SweepOp), "op record size");

__device__ __forceinline__ OpWords load_op(COpW *p) {
    OpWords r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.w[i] = p->w[i];
    return r;
}


__device__ __forceinline__ Op decode(const OpWords &r) {
    Op o;
    o.kind = (int)(r.w[0] & 0xffu);
    o.k0 = (int)(int8_t)(r.w[0] >> 8);
    o.k1 = (int)(int8_t)(r.w[0] >> 16);
    o.mat = (int)r.w[1];
    o.rctrl = r.w[2] & 0xffffu;
    o.tctrl = r.w[2] >> 16;
    o.ta = r.w[3] & 0xffu;
    o.tb = (r.w[3] >> 8) & 0xffu;
    o.oa = (int)(int8_t)(r.w[3] >> 16);
    o.ob = (int)(int8_t)(r.w[3] >> 24);
    o.octrl = (uint64_t)r.w[4] | ((uint64_t)r.w[5] << 32);
    return o;
}


// Two workgroups per CU fit the LDS (64 KiB each), i.e. two waves per SIMD: the register budget is 256.
// A workgroup's next chunk is loaded (into the ping-pong register set, free once the stages are done)
// before its current chunk is stored.  Loads and stores share one in-order counter; with the loads issued
// first, waiting for them leaves the stores outstanding, and they drain while the next chunk is computed.
// (Issuing the loads before the stages instead needs a third 64-register set, which does not fit.)
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_sweep(v2d *__restrict__ amp, int n, int log_chunks, uint64_t total_chunks, const SweepPass *__restrict__ pass_arg,
        const SweepOp *__restrict__ ops_arg, const v2d *__restrict__ mats_arg, uint64_t mat_stride,
        int elem_begin) {
    __shared__ v2d lds[kChunk];
    CPass *pass = (CPass *)pass_arg;
    COpW *ops = (COpW *)ops_arg;
    CV2 *mats = (CV2 *)mats_arg;
    const uint32_t tid = threadIdx.x;
    uint64_t c = blockIdx.x;
    if (c >= total_chunks) return;
    const uint64_t outer_mask = pass->outer_mask;
    const uint64_t chunk_mask = (1ULL << log_chunks) - 1ULL;

    // load / store phase: thread bits -> local qubits 0..7, register bits -> local 8..11.  The pass's
    // tables are read unconditionally (uniform scalar loads) and selected per lane.
    uint64_t gthread = 0;
    uint32_t sthread = 0, sout = 0;
#pragma unroll
    for (int i = 0; i < kThreadBits; i++) {
        const bool on = (tid >> i) & 1u;
        gthread += on ? pass->glob_thread[i] : 0ULL;
        sthread ^= on ? (uint32_t)pass->io_thread_slot[i] : 0u;
        sout ^= on ? (uint32_t)pass->st_thread_slot[i] : 0u;
    }
    const int nstages = pass->nstages;
    Regs a, p;
    uint64_t cb = chunk_base_of(c & chunk_mask, outer_mask);
    {
        const v2d *src = amp + (((uint64_t)elem_begin + (c >> log_chunks)) << n) + cb + opaque(gthread);
#pragma unroll
        for (int r = 0; r < kRegs; r++) p[r] = src[pass->glob_reg[r]];
#pragma unroll
        for (int r = 0; r < kRegs; r++) lds[sthread ^ pass->io_reg_slot[r]] = p[r];
    }
    __syncthreads();
    // The chunk is in LDS at the top of every iteration; the LDS write of the next chunk sits after this
    // chunk's stores in the same block, so the wait for those loads leaves the stores outstanding.
    for (;;) {
        const uint64_t elem = (uint64_t)elem_begin + (c >> log_chunks);
        CV2 *emats = mats + elem * mat_stride;
        const uint64_t next = c + gridDim.x;
        const uint64_t ncb = next < total_chunks ? chunk_base_of(next & chunk_mask, outer_mask) : 0ULL;

        for (int st = 0; st < nstages; st++) {
            CStage &stg = pass->stages[st];
            uint32_t ts = 0;
#pragma unroll
            for (int i = 0; i < kThreadBits; i++) ts ^= ((tid >> i) & 1u) ? (uint32_t)stg.thread_slot[i] : 0u;
#pragma unroll
            for (int r = 0; r < kRegs; r++) a[r] = lds[ts ^ stg.lds_off[r]];
            const int oe = stg.op_end;
            int o = stg.op_begin;
            // each op's record arrives with one scalar load
            for (;;) {
                if (o == oe) {
                    const uint32_t tw = opaque(ts);
#pragma unroll
                    for (int r = 0; r < kRegs; r++) lds[tw ^ stg.lds_off[r]] = a[r];
                    break;
                }
                run_op(a, p, decode(load_op(ops + o)), emats, tid, cb);
                if (++o == oe) {
                    const uint32_t tw = opaque(ts);
#pragma unroll
                    for (int r = 0; r < kRegs; r++) lds[tw ^ stg.lds_off[r]] = p[r];
                    break;
                }
                run_op(p, a, decode(load_op(ops + o)), emats, tid, cb);
                ++o;
            }
            __syncthreads();
        }

        // store phase (possibly remapping which local qubit lands on which chunk position), after the next
        // chunk's loads have been issued
        {
            const uint32_t sl = opaque(sout);
#pragma unroll
            for (int r = 0; r < kRegs; r++) a[r] = lds[sl ^ pass->st_reg_slot[r]];
        }
        if (next < total_chunks) {
            const v2d *src = amp + (((uint64_t)elem_begin + (next >> log_chunks)) << n) + ncb + opaque(gthread);
#pragma unroll
            for (int r = 0; r < kRegs; r++) p[r] = src[pass->glob_reg[r]];
        }
        {
            v2d *dst = amp + (elem << n) + cb + opaque(gthread);
#pragma unroll
            for (int r = 0; r < kRegs; r++) dst[pass->glob_reg[r]] = a[r];
        }
        if (next >= total_chunks) break;
        c = next;
        cb = ncb;
        __syncthreads();  // every thread has read its store-phase slots before the next chunk lands in LDS
        {
            const uint32_t sl = opaque(sthread);
#pragma unroll
            for (int r = 0; r < kRegs; r++) lds[sl ^ pass->io_reg_slot[r]] = p[r];
        }
        __syncthreads();
    }
}

}  // namespace

// Runtime-compiled pass kernels have the same launch bounds and LDS footprint, hence the same occupancy.
int sweep_grid_size(Context *ctx) {
    static int per_cu = 0;
    if (per_cu == 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep, kThreads, 0) != hipSuccess || nb <= 0) {
            (void)hipGetLastError();
            nb = 1;
        }
        per_cu = nb;
    }
    const int want = ctx->sweep_blocks_per_cu > 0 ? std::min(ctx->sweep_blocks_per_cu, per_cu) : per_cu;
    return want * ctx->cu_count;
}

int launch_sweep(Context *ctx, State *st, const SweepPass *pass_dev, const SweepOp *ops_dev, const double *mats_dev,
                 int elem_begin, int elem_count, size_t mat_stride) {
    const int log_chunks = st->n - kChunkBits;
    const uint64_t total = (uint64_t)elem_count << log_chunks;
    const uint64_t grid = std::min<uint64_t>(total, (uint64_t)sweep_grid_size(ctx));
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_SWEEP, 32.0 * (double)((uint64_t)elem_count << st->n));
    hipLaunchKernelGGL(k_sweep, dim3((unsigned)grid), dim3(kThreads), 0, ctx->stream, reinterpret_cast<v2d *>(st->amp),
                       st->n, log_chunks, total, pass_dev, ops_dev, reinterpret_cast<const v2d *>(mats_dev),
                       (uint64_t)mat_stride, elem_begin);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

}  // namespace sc
