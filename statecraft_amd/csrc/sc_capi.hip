// sc_capi.hip -- the C-ABI of the gfx950 statevector engine (include/statecraft_hip.h).
//
// Host orchestration only: validation, planning (sc_plan.cpp), HBM ownership,
// uploads, launches and the reference's host-side rules (normalise threshold,
// zero-probability and invalid-norm errors).  Kernels live in sc_gates.hip, sc_sweep.hip, sc_measure.hip
// and sc_adjoint.hip (observables, adjoint gradients and batched trajectories).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "sc_internal.hpp"
#include "sc_jit.hpp"
#include "sc_plan.hpp"
#include "sc_seqsum.hpp"

namespace sc {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }
int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

// ---------------------------------------------------------------- profiler
ProfScope::ProfScope(Profiler *prof, hipStream_t stream, int family, double algo_bytes)
    : p(prof), s(stream), fam(family), bytes(algo_bytes) {
    if (!p->enabled) return;
    auto take = [&]() {
        hipEvent_t e;
        if (!p->pool.empty()) {
            e = p->pool.back();
            p->pool.pop_back();
        } else if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
        }
        return e;
    };
    a = take();
    b = take();
    if (a) (void)hipEventRecord(a, s);
}

ProfScope::~ProfScope() {
    if (!p->enabled || !a || !b) return;
    (void)hipEventRecord(b, s);
    p->pending.push_back({a, b, fam, bytes});
}

int drain_profiler(Context *ctx) {
    Profiler &p = ctx->prof;
    if (p.pending.empty()) return SC_OK;
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (auto &pe : p.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pe.a, pe.b) == hipSuccess) {
            p.launches[pe.fam] += 1;
            p.ms[pe.fam] += ms;
            p.bytes[pe.fam] += pe.bytes;
        }
        p.pool.push_back(pe.a);
        p.pool.push_back(pe.b);
    }
    p.pending.clear();
    return SC_OK;
}

int ensure_scratch(Context *ctx, size_t bytes) {
    if (bytes <= ctx->scratch_bytes) return SC_OK;
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (ctx->scratch) SC_HIP_CHECK(hipFree(ctx->scratch));
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
    size_t want = std::max(bytes, (size_t)1 << 20);
    if (hipMalloc(&ctx->scratch, want) != hipSuccess) {
        (void)hipGetLastError();
        return fail(SC_E_OOM, "device scratch allocation of " + std::to_string(want) + " bytes failed");
    }
    ctx->scratch_bytes = want;
    return SC_OK;
}

int ensure_host_scratch(Context *ctx, size_t bytes) {
    if (bytes <= ctx->scratch_host_bytes) return SC_OK;
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (ctx->scratch_host) SC_HIP_CHECK(hipHostFree(ctx->scratch_host));
    ctx->scratch_host = nullptr;
    ctx->scratch_host_bytes = 0;
    size_t want = std::max(bytes, (size_t)1 << 20);
    SC_HIP_CHECK(hipHostMalloc(&ctx->scratch_host, want, hipHostMallocDefault));
    ctx->scratch_host_bytes = want;
    return SC_OK;
}

int DevBuf::alloc(size_t bytes) {
    if (hipMalloc(&p, std::max(bytes, (size_t)16)) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        return fail(SC_E_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    return SC_OK;
}

namespace {

__global__ void k_set_basis(double2 *amp, int n, int batch, uint64_t index) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < batch) amp[((uint64_t)b << n) + index] = make_double2(1.0, 0.0);
}

int set_device(Context *ctx) {
    SC_HIP_CHECK(hipSetDevice(ctx->device));
    return SC_OK;
}

}  // namespace

// ---- execution of an analysed op list ----------------------------------------
int exec_direct(Context *ctx, State *st, const std::vector<PlanOp> &ops, const double *mats) {
    for (const PlanOp &op : ops) {
        const double *m = op.mat >= 0 ? mats + 2 * (int64_t)op.mat : nullptr;
        int rc = SC_OK;
        switch (op.kind) {
        case SC_OP_SINGLE:
        case SC_OP_MCU: {
            if (op.is_x) {
                rc = launch_x(ctx, st, op.q0, op.controls);
            } else {
                Mat2 g;
                std::memcpy(g.v, m, sizeof(g.v));
                rc = launch_single(ctx, st, op.q0, g, op.controls);
            }
            break;
        }
        case SC_OP_CNOT: rc = launch_x(ctx, st, op.q1, 1ULL << op.q0); break;
        case SC_OP_TWO: {
            Mat4 g;
            std::memcpy(g.v, m, sizeof(g.v));
            rc = launch_two(ctx, st, op.q0, op.q1, g);
            break;
        }
        case SC_OP_DIAG2: rc = launch_diag2(ctx, st, op.q0, op.q1, m); break;
        case SC_OP_SWAP: rc = launch_swap(ctx, st, op.q0, op.q1); break;
        }
        if (rc) return rc;
        ctx->last_launches++;
        ctx->last_sweeps++;
    }
    return SC_OK;
}

namespace {

int exec_fused(Context *ctx, State *st, const std::vector<PlanOp> &ops, const double *mats) {
    std::vector<FusedBlock> blocks;
    plan_blocks(st->n, ops, std::max(1, std::min(ctx->max_fused, 5)), blocks);
    // matrices of all blocks in one upload
    std::vector<double> pool;
    std::vector<size_t> offs(blocks.size(), 0);
    for (size_t b = 0; b < blocks.size(); b++) {
        if (blocks[b].standalone) continue;
        std::vector<double> m;
        block_matrix(blocks[b], ops, mats, m);
        offs[b] = pool.size();
        pool.insert(pool.end(), m.begin(), m.end());
    }
    DevBuf dpool;
    int rc = dpool.alloc(pool.size() * sizeof(double));
    if (rc) return rc;
    if (!pool.empty())
        SC_HIP_CHECK(hipMemcpyAsync(dpool.p, pool.data(), pool.size() * sizeof(double), hipMemcpyHostToDevice,
                                    ctx->stream));
    for (size_t b = 0; b < blocks.size(); b++) {
        const FusedBlock &fb = blocks[b];
        if (fb.standalone) {
            std::vector<PlanOp> one{ops[fb.ops[0]]};
            rc = exec_direct(ctx, st, one, mats);
        } else {
            rc = launch_dense(ctx, st, (int)fb.qubits.size(), fb.qubits.data(), (double *)dpool.p + offs[b]);
            ctx->last_launches++;
            ctx->last_sweeps++;
            ctx->last_blocks++;
        }
        if (rc) return rc;
    }
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // dpool lifetime
    return SC_OK;
}

}  // namespace

int sweep_plan_entry(Context *ctx, int n, const std::vector<PlanOp> &ops_in, PlanCache::Entry **out,
                     bool basis_input) {
    // ---- the plan (and its runtime-compiled kernels) of this circuit structure, cached per context
    const int max_fuse = std::min(ctx->max_fused, kRegBits);
    // from a basis state, runtime-compiled passes generate the product state of the leading single-qubit ops
    std::vector<PlanOp> rest;
    std::vector<std::pair<int, int>> absorbed;
    if (basis_input && ctx->jit) {
        rest = ops_in;
        absorb_leading_ops(n, rest, absorbed);
        if (absorbed.size() == rest.size()) absorbed.clear();  // nothing else: the plain plan is as good
    }
    const std::vector<PlanOp> &ops = absorbed.empty() ? ops_in : rest;
    if (!ctx->plan_cache) ctx->plan_cache = new PlanCache();
    SweepCostModel cm = ctx->jit ? kCompiledPassCost : kInterpretedCost;
    if (ctx->pass_cap >= 0) cm.pass_cap = (double)ctx->pass_cap;
    cm.direct_io = cm.direct_io && ctx->direct_io;
    cm.permlane = ctx->jit && ctx->permlane;
    // a fixed 13-bit chunk applies from 13 qubits with runtime-compiled passes; 0 leaves the choice to the planner
    int cb = ctx->chunk_bits;
    if (!ctx->jit || (cb == kMaxChunkBits && n < kMaxChunkBits)) cb = kChunkBits;
    std::string key = plan_cache_key(n, ops, ctx->low_qubits, ctx->remap != 0, max_fuse,
                                     ctx->jit | (cm.direct_io ? 8 : 0) | (ctx->pass_sched << 4) | (cm.permlane ? 256 : 0), cm.pass_cap, cb);
    if (!absorbed.empty()) {
        key += "|product";
        for (const auto &a : absorbed) key += " " + std::to_string(a.first) + ":" + std::to_string(a.second);
    }
    PlanCache::Entry *ent = ctx->plan_cache->find(key);
    if (!ent) {
        std::unique_ptr<PlanCache::Entry> e(new PlanCache::Entry());
        e->key = key;
        int rc = plan_sweeps_geometry(n, ops, ctx->low_qubits, cb, ctx->remap != 0, ctx->jit != 0, e->plan, max_fuse,
                                      cm, &e->pick);
        if (rc) return rc;
        if (!absorbed.empty()) {
            if (e->plan.passes.empty()) return sweep_plan_entry(ctx, n, ops_in, out, false);
            finalize_product(n, e->plan, absorbed);
        }
        if (ctx->jit && !e->plan.passes.empty() && (rc = jit_pass_kernels(ctx, e->plan, e->fns))) return rc;
        ent = ctx->plan_cache->insert(std::move(e));
    }
    ent->used = ++ctx->plan_cache->tick;
    ctx->last_chunk_bits = ent->pick.chunk_bits;
    ctx->last_low_qubits = ent->pick.low_qubits;
    ctx->last_candidates = ent->pick.candidates;
    ctx->last_ns_per_amp = ent->pick.ns_per_amp;
    *out = ent;
    return SC_OK;
}

namespace {

// mats_per_elem: one matrix array per batch element (size batch) or a single shared array.
// src / dst (runtime-compiled passes only): the first pass reads the states from src and the last one
// writes them to dst; passes in between work in st->amp.  nullptr means st->amp.
int exec_sweep(Context *ctx, State *st, const std::vector<PlanOp> &ops, const std::vector<const double *> &mats_elems,
               const double2 *src = nullptr, double2 *dst = nullptr, ZTermsRequest *terms = nullptr) {
    PlanCache::Entry *ent = nullptr;
    {
        int rc = sweep_plan_entry(ctx, st->n, ops, &ent, st->pending_basis >= 0);
        if (rc) return rc;
    }
    const SweepPlan &plan = ent->plan;
    const std::vector<hipFunction_t> &fns = ent->fns;
    if (!src) src = st->amp;
    if (!dst) dst = st->amp;
    if (plan.passes.empty() || fns.empty()) {
        int rc0 = materialize_basis(st);
        if (rc0) return rc0;
    }
    const long long basis = st->pending_basis;  // generated by the first pass (its src is not read)
    if (plan.passes.empty()) {
        if (src != dst)
            SC_HIP_CHECK(hipMemcpyAsync(dst, src, ((size_t)st->batch << st->n) * sizeof(double2),
                                        hipMemcpyDeviceToDevice, ctx->stream));
        return SC_OK;
    }
    if (fns.empty() && (src != st->amp || dst != st->amp))
        return fail(SC_E_STATE, "internal error: out-of-place sweeps need runtime-compiled passes");
    // ---- upload: pass and op records (compiled sweep kernel only) and every element's matrix pool
    const bool records = fns.empty();
    const size_t pass_bytes = records ? plan.passes.size() * sizeof(SweepPass) : 0;
    const size_t ops_bytes = records ? std::max<size_t>(plan.ops.size(), 1) * sizeof(SweepOp) : 0;
    const size_t nelem_pool = mats_elems.size();
    const size_t pool_bytes = std::max<uint64_t>(plan.pool_complex, 1) * 16 * nelem_pool;
    const size_t total = pass_bytes + ops_bytes + pool_bytes;
    int rc = ensure_host_scratch(ctx, total);
    if (rc) return rc;
    char *h = (char *)ctx->scratch_host;
    if (records) {
        std::memcpy(h, plan.passes.data(), pass_bytes);
        if (!plan.ops.empty()) std::memcpy(h + pass_bytes, plan.ops.data(), plan.ops.size() * sizeof(SweepOp));
    }
    double *hpool = (double *)(h + pass_bytes + ops_bytes);
    // per-element pools are independent: large batches build them on several host threads
    const size_t pool_work = nelem_pool * (plan.pool.size() + 1);
    const uint64_t grain = pool_work < 4096 ? nelem_pool : std::max<uint64_t>(1, nelem_pool / (4 * kHostThreads));
    parallel_ranges(nelem_pool, grain, [&](unsigned, uint64_t e0, uint64_t e1) {
        for (uint64_t e = e0; e < e1; e++) build_pool(plan, mats_elems[e], hpool + 2 * plan.pool_complex * e, basis);
    });
    if (total > ctx->plan_dev_bytes) {
        if (ctx->plan_dev) SC_HIP_CHECK(hipFree(ctx->plan_dev));
        ctx->plan_dev = nullptr;
        ctx->plan_dev_bytes = 0;
        const size_t want = std::max(total, (size_t)1 << 20);
        if (hipMalloc(&ctx->plan_dev, want) != hipSuccess) {
            (void)hipGetLastError();
            ctx->plan_dev = nullptr;
            return fail(SC_E_OOM, "device allocation of " + std::to_string(want) + " bytes for a sweep plan failed");
        }
        ctx->plan_dev_bytes = want;
    }
    char *d = (char *)ctx->plan_dev;
    SC_HIP_CHECK(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, ctx->stream));
    const SweepPass *dpass = (const SweepPass *)d;
    const SweepOp *dops = (const SweepOp *)(d + pass_bytes);
    const double *dpool = (const double *)(d + pass_bytes + ops_bytes);
    const size_t stride = nelem_pool > 1 ? plan.pool_complex : 0;
    // keep each launch's grid under ~2^20 blocks
    const int log_chunks = st->n - plan.chunk_bits;
    int per_launch = std::max(1, (int)std::min<uint64_t>((uint64_t)st->batch, (1ULL << 20) >> log_chunks));
    // Z-string terms summed by the last pass: per-element pools (contiguous chunk ranges per workgroup), every
    // element in one launch, in place, and a last pass that is not the first (which generates a basis input)
    hipFunction_t terms_fn = nullptr;
    int tparts = 0;
    double *tpart = nullptr, *trows = nullptr;
    if (terms && ctx->fused_terms && !records && stride != 0 && plan.passes.size() >= 2 && per_launch >= st->batch &&
        src == st->amp && dst == st->amp && !terms->z.empty() && terms->z.size() <= 32) {
        if ((rc = jit_last_pass_terms_kernel(ctx, plan, terms->z, terms_fn))) return rc;
        tparts = sweep_terms_parts(ctx, st->n, plan.chunk_bits, st->batch);
        const size_t rows = (size_t)st->batch * terms->z.size();
        if ((rc = ensure_scratch(ctx, (rows * (size_t)tparts + rows) * sizeof(double)))) return rc;
        tpart = (double *)ctx->scratch;
        trows = tpart + rows * (size_t)tparts;
    }
    for (size_t p = 0; p < plan.passes.size(); p++) {
        if (terms_fn && p + 1 == plan.passes.size()) {
            rc = launch_sweep_jit_terms(ctx, st, terms_fn, dpool, stride, plan.chunk_bits, (int)terms->z.size(), tpart,
                                        tparts, trows);
            if (rc) return rc;
            ctx->last_launches++;
            ctx->last_sweeps++;
            continue;
        }
        for (int e0 = 0; e0 < st->batch; e0 += per_launch) {
            int cnt = std::min(per_launch, st->batch - e0);
            const double2 *in = p == 0 ? src : st->amp;
            double2 *out = p + 1 == plan.passes.size() ? dst : st->amp;
            rc = records ? launch_sweep(ctx, st, dpass + p, dops, dpool, e0, cnt, stride)
                         : launch_sweep_jit(ctx, st, fns[p], dpool, e0, cnt, stride, plan.chunk_bits, in, out,
                                            p == 0 ? basis : -1);
            if (rc) return rc;
            ctx->last_launches++;
        }
        ctx->last_sweeps++;
        st->pending_basis = -1;
    }
    if (terms_fn) {
        SC_HIP_CHECK(hipMemcpyAsync(terms->vals, trows, (size_t)st->batch * terms->z.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, ctx->stream));
        terms->done = true;
    }
    ctx->last_blocks += plan.fused_blocks;
    // the pinned staging buffer and plan_dev are reused by the next call
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

}  // namespace

int materialize_basis(State *st) {
    if (st->pending_basis < 0) return SC_OK;
    Context *ctx = st->ctx;
    int rc = set_device(ctx);
    if (rc) return rc;
    SC_HIP_CHECK(hipMemsetAsync(st->amp, 0, ((size_t)st->batch << st->n) * sizeof(double2), ctx->stream));
    hipLaunchKernelGGL(k_set_basis, dim3((st->batch + 255) / 256), dim3(256), 0, ctx->stream, st->amp, st->n, st->batch,
                       (uint64_t)st->pending_basis);
    SC_HIP_CHECK(hipGetLastError());
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    st->pending_basis = -1;
    return SC_OK;
}

int apply_planned(State *st, const std::vector<PlanOp> &plan_ops, const std::vector<const double *> &elems,
                  bool per_elem, ZTermsRequest *terms) {
    Context *ctx = st->ctx;
    ctx->last_sweeps = ctx->last_launches = ctx->last_blocks = 0;
    if (plan_ops.empty()) return SC_OK;
    int mode = ctx->exec_mode;
    if (mode == SC_EXEC_SWEEP && st->n < kChunkBits) mode = SC_EXEC_FUSED;
    int rc = SC_OK;
    // only runtime-compiled sweeps generate a pending basis state themselves
    if (!(mode == SC_EXEC_SWEEP && ctx->jit) && (rc = materialize_basis(st))) return rc;
    if (per_elem && mode != SC_EXEC_SWEEP) {
        // element-by-element with a one-element view of the state
        for (int b = 0; b < st->batch; b++) {
            State view = *st;
            view.amp = st->amp + ((uint64_t)b << st->n);
            view.batch = 1;
            rc = mode == SC_EXEC_FUSED ? exec_fused(ctx, &view, plan_ops, elems[b])
                                       : exec_direct(ctx, &view, plan_ops, elems[b]);
            if (rc) return rc;
        }
    } else if (mode == SC_EXEC_DIRECT) {
        rc = exec_direct(ctx, st, plan_ops, elems[0]);
    } else if (mode == SC_EXEC_FUSED) {
        rc = exec_fused(ctx, st, plan_ops, elems[0]);
    } else {
        rc = exec_sweep(ctx, st, plan_ops, elems, nullptr, nullptr, terms);
    }
    if (rc) return rc;
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

int apply_ops_between(State *st, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex,
                      uint64_t mat_stride, bool per_elem, const double2 *src, double2 *dst) {
    Context *ctx = st->ctx;
    int rc = set_device(ctx);
    if (rc) return rc;
    const size_t bytes = ((size_t)st->batch << st->n) * sizeof(double2);
    if (!(ctx->jit && ctx->exec_mode == SC_EXEC_SWEEP && st->n >= kChunkBits)) {
        // in place, with copies at the ends
        if (src != st->amp) SC_HIP_CHECK(hipMemcpyAsync(st->amp, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        if ((rc = apply_ops_impl(st, ops, n_ops, mats, n_mat_complex, mat_stride, per_elem))) return rc;
        if (dst != st->amp) SC_HIP_CHECK(hipMemcpyAsync(dst, st->amp, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        return SC_OK;
    }
    std::vector<const double *> elems;
    if (per_elem) {
        for (int b = 0; b < st->batch; b++) elems.push_back(mats + 2 * mat_stride * (uint64_t)b);
    } else {
        elems.push_back(mats);
    }
    std::vector<PlanOp> plan_ops;
    if ((rc = analyze_ops(st->n, ops, n_ops, elems, per_elem ? mat_stride : n_mat_complex, plan_ops))) return rc;
    ctx->last_sweeps = ctx->last_launches = ctx->last_blocks = 0;
    if (plan_ops.empty()) {
        if (src != dst) SC_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        return SC_OK;
    }
    if ((rc = exec_sweep(ctx, st, plan_ops, elems, src, dst))) return rc;
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

int apply_ops_impl(State *st, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex,
                   uint64_t mat_stride, bool per_elem) {
    Context *ctx = st->ctx;
    int rc = set_device(ctx);
    if (rc) return rc;
    std::vector<const double *> elems;
    if (per_elem) {
        for (int b = 0; b < st->batch; b++) elems.push_back(mats + 2 * mat_stride * (uint64_t)b);
    } else {
        elems.push_back(mats);
    }
    std::vector<PlanOp> plan_ops;
    rc = analyze_ops(st->n, ops, n_ops, elems, per_elem ? mat_stride : n_mat_complex, plan_ops);
    if (rc) return rc;
    return apply_planned(st, plan_ops, elems, per_elem);
}

}  // namespace sc

using namespace sc;

#define SC_REQUIRE(ptr, name)                                                   \
    do {                                                                        \
        if (!(ptr)) return fail(SC_E_NULL, std::string(name) + " is null");     \
    } while (0)

extern "C" {

const char *sc_last_error(void) { return g_last_error.c_str(); }
int sc_version(void) { return 10000; }

int sc_ctx_create(int device, sc_ctx **out) {
    SC_REQUIRE(out, "out");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return fail(SC_E_HIP, "no HIP device is visible");
    }
    if (device < 0 || device >= count) return fail(SC_E_ARG, "device index out of range: " + std::to_string(device));
    SC_HIP_CHECK(hipSetDevice(device));
    std::unique_ptr<sc_ctx> ctx(new sc_ctx());
    ctx->device = device;
    SC_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        ctx->cu_count = cus;
    int rc = init_noalloc_scratch(device);
    if (rc) {
        (void)hipStreamDestroy(ctx->stream);
        return rc;
    }
    *out = ctx.release();
    return SC_OK;
}

int sc_ctx_destroy(sc_ctx *ctx) {
    if (!ctx) return SC_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &pe : ctx->prof.pending) {
        (void)hipEventDestroy(pe.a);
        (void)hipEventDestroy(pe.b);
    }
    for (auto e : ctx->prof.pool) (void)hipEventDestroy(e);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->scratch_host) (void)hipHostFree(ctx->scratch_host);
    if (ctx->plan_dev) (void)hipFree(ctx->plan_dev);
    if (ctx->adj_ws) (void)hipFree(ctx->adj_ws);
    destroy_job_graphs(ctx);
    delete ctx->plan_cache;
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SC_OK;
}

int sc_ctx_set_option(sc_ctx *ctx, int option, int64_t value) {
    SC_REQUIRE(ctx, "ctx");
    switch (option) {
    case SC_OPT_EXEC_MODE:
        if (value < SC_EXEC_DIRECT || value > SC_EXEC_SWEEP) return fail(SC_E_ARG, "unknown execution mode");
        ctx->exec_mode = (int)value;
        return SC_OK;
    case SC_OPT_MAX_FUSED:
        if (value < 0 || value > 5) return fail(SC_E_ARG, "max fused qubits must be 0..5");
        ctx->max_fused = (int)value;
        return SC_OK;
    case SC_OPT_PROFILE: {
        int rc = drain_profiler(ctx);
        if (rc) return rc;
        ctx->prof.enabled = value != 0;
        return SC_OK;
    }
    case SC_OPT_LOW_QUBITS:
        if (value < -1 || value > 8) return fail(SC_E_ARG, "low qubits must be 0..8, or -1 (chosen per circuit)");
        ctx->low_qubits = (int)value;
        return SC_OK;
    case SC_OPT_TEST_DRIFT:
        if (value < 0) return fail(SC_E_ARG, "test drift must be non-negative");
        ctx->test_drift = value;
        return SC_OK;
    case SC_OPT_REMAP:
        ctx->remap = value != 0;
        return SC_OK;
    case SC_OPT_SWEEP_BLOCKS_PER_CU:
        if (value < 0 || value > 8) return fail(SC_E_ARG, "sweep blocks per CU must be 0..8");
        ctx->sweep_blocks_per_cu = (int)value;
        return SC_OK;
    case SC_OPT_JIT:
        if (value < 0 || value > 2) return fail(SC_E_ARG, "jit must be 0, 1 or 2");
        ctx->jit = (int)value;
        return SC_OK;
    case SC_OPT_SWEEP_PASS_CAP:
        if (value < -1 || value > 100000) return fail(SC_E_ARG, "sweep pass cap must be -1..100000");
        ctx->pass_cap = (int)value;
        return SC_OK;
    case SC_OPT_CHUNK_BITS:
        if (value != 0 && value != kChunkBits && value != kMaxChunkBits)
            return fail(SC_E_ARG, "chunk bits must be 12 or 13, or 0 (chosen per circuit)");
        ctx->chunk_bits = (int)value;
        return SC_OK;
    case SC_OPT_DIRECT_IO:
        ctx->direct_io = value != 0;
        return SC_OK;
    case SC_OPT_MFMA:
        if (value < 0 || value > 2) return fail(SC_E_ARG, "MFMA option must be 0, 1 or 2");
        ctx->mfma = (int)value;
        return SC_OK;
    case SC_OPT_PASS_SCHED:
        if (value < 0 || value > 6) return fail(SC_E_ARG, "pass scheduling variant must be 0..6");
        ctx->pass_sched = (int)value;
        return SC_OK;
    case SC_OPT_PERMLANE:
        ctx->permlane = value != 0;
        return SC_OK;
    case SC_OPT_FUSED_TERMS:
        ctx->fused_terms = value != 0;
        return SC_OK;
    case SC_OPT_MFMA_TILE:
        if (value < -1 || value >= kMfmaTiles) return fail(SC_E_ARG, "SC_OPT_MFMA_TILE must be -1..5");
        ctx->mfma_tile = (int)value;
        return SC_OK;
    case SC_OPT_MFMA_PICK4:
    case SC_OPT_MFMA_PICK5: return fail(SC_E_ARG, "the autotuner's tile choice is read-only");
    default: return fail(SC_E_ARG, "unknown option " + std::to_string(option));
    }
}

int sc_ctx_get_option(sc_ctx *ctx, int option, int64_t *value) {
    SC_REQUIRE(ctx, "ctx");
    SC_REQUIRE(value, "value");
    switch (option) {
    case SC_OPT_EXEC_MODE: *value = ctx->exec_mode; return SC_OK;
    case SC_OPT_MAX_FUSED: *value = ctx->max_fused; return SC_OK;
    case SC_OPT_PROFILE: *value = ctx->prof.enabled ? 1 : 0; return SC_OK;
    case SC_OPT_LOW_QUBITS: *value = ctx->low_qubits; return SC_OK;
    case SC_OPT_TEST_DRIFT: *value = ctx->test_drift; return SC_OK;
    case SC_OPT_REMAP: *value = ctx->remap; return SC_OK;
    case SC_OPT_SWEEP_BLOCKS_PER_CU: *value = ctx->sweep_blocks_per_cu; return SC_OK;
    case SC_OPT_JIT: *value = ctx->jit; return SC_OK;
    case SC_OPT_SWEEP_PASS_CAP: *value = ctx->pass_cap; return SC_OK;
    case SC_OPT_CHUNK_BITS: *value = ctx->chunk_bits; return SC_OK;
    case SC_OPT_DIRECT_IO: *value = ctx->direct_io; return SC_OK;
    case SC_OPT_MFMA: *value = ctx->mfma; return SC_OK;
    case SC_OPT_PASS_SCHED: *value = ctx->pass_sched; return SC_OK;
    case SC_OPT_PERMLANE: *value = ctx->permlane; return SC_OK;
    case SC_OPT_FUSED_TERMS: *value = ctx->fused_terms; return SC_OK;
    case SC_OPT_MFMA_TILE: *value = ctx->mfma_tile; return SC_OK;
    case SC_OPT_MFMA_PICK4: *value = ctx->mfma_pick[0]; return SC_OK;
    case SC_OPT_MFMA_PICK5: *value = ctx->mfma_pick[1]; return SC_OK;
    default: return fail(SC_E_ARG, "unknown option " + std::to_string(option));
    }
}

int sc_ctx_synchronize(sc_ctx *ctx) {
    SC_REQUIRE(ctx, "ctx");
    int rc = set_device(ctx);
    if (rc) return rc;
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

int sc_ctx_device_info(sc_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes, int *cu_count) {
    SC_REQUIRE(ctx, "ctx");
    int rc = set_device(ctx);
    if (rc) return rc;
    size_t f = 0, t = 0;
    SC_HIP_CHECK(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    if (cu_count) *cu_count = ctx->cu_count;
    return SC_OK;
}

int sc_state_create(sc_ctx *ctx, int num_qubits, int batch, sc_state **out) {
    SC_REQUIRE(ctx, "ctx");
    SC_REQUIRE(out, "out");
    *out = nullptr;
    if (num_qubits <= 0) return fail(SC_E_ARG, "qubitCount must be positive");
    if (num_qubits > 34)
        return fail(SC_E_UNSUPPORTED,
                    "statevector-hip engine supports up to 34 qubits in dense mode, got " + std::to_string(num_qubits));
    if (batch <= 0) return fail(SC_E_ARG, "batch must be positive");
    int rc = set_device(ctx);
    if (rc) return rc;
    std::unique_ptr<sc_state> st(new sc_state());
    st->ctx = ctx;
    st->n = num_qubits;
    st->batch = batch;
    const size_t bytes = ((size_t)batch << num_qubits) * sizeof(double2);
    hipError_t e = hipMalloc((void **)&st->amp, bytes);
    if (e != hipSuccess && ctx->adj_ws) {
        // the adjoint workspace kept from an earlier call gives way to a state
        (void)hipGetLastError();
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(ctx->adj_ws);
        ctx->adj_ws = nullptr;
        ctx->adj_ws_bytes = 0;
        e = hipMalloc((void **)&st->amp, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(SC_E_OOM, "cannot allocate " + std::to_string(bytes) + " bytes of HBM for " +
                                  std::to_string(num_qubits) + " qubits x " + std::to_string(batch));
    }
    sc_state *raw = st.release();
    rc = sc_state_set_basis(raw, 0);
    if (rc) {
        (void)hipFree(raw->amp);
        delete raw;
        return rc;
    }
    *out = raw;
    return SC_OK;
}

int sc_state_destroy(sc_state *st) {
    if (!st) return SC_OK;
    (void)hipSetDevice(st->ctx->device);
    (void)hipStreamSynchronize(st->ctx->stream);
    if (st->amp) (void)hipFree(st->amp);
    delete st;
    return SC_OK;
}

int sc_state_qubits(const sc_state *st, int *num_qubits, int *batch) {
    SC_REQUIRE(st, "state");
    if (num_qubits) *num_qubits = st->n;
    if (batch) *batch = st->batch;
    return SC_OK;
}

int sc_state_set_basis(sc_state *st, uint64_t index) {
    SC_REQUIRE(st, "state");
    if (st->n < 64 && (index >> st->n) != 0) return fail(SC_E_ARG, "basis index out of range");
    st->pending_basis = (int64_t)index;  // written by the next operation that needs it
    return SC_OK;
}

int sc_state_upload(sc_state *st, const double *interleaved, uint64_t n_complex) {
    SC_REQUIRE(st, "state");
    SC_REQUIRE(interleaved, "data");
    if (n_complex != ((uint64_t)st->batch << st->n))
        return fail(SC_E_ARG, "Expected state vector of length " + std::to_string((uint64_t)st->batch << st->n) +
                                  ", got " + std::to_string(n_complex));
    int rc = set_device(st->ctx);
    if (rc) return rc;
    SC_HIP_CHECK(hipMemcpy(st->amp, interleaved, n_complex * sizeof(double2), hipMemcpyHostToDevice));
    st->pending_basis = -1;  // every amplitude was written
    return SC_OK;
}

int sc_state_download(sc_state *st, double *interleaved, uint64_t n_complex) {
    SC_REQUIRE(st, "state");
    SC_REQUIRE(interleaved, "data");
    if (n_complex != ((uint64_t)st->batch << st->n))
        return fail(SC_E_ARG, "Expected state vector of length " + std::to_string((uint64_t)st->batch << st->n) +
                                  ", got " + std::to_string(n_complex));
    int rc = set_device(st->ctx);
    if (rc) return rc;
    if ((rc = materialize_basis(st))) return rc;
    SC_HIP_CHECK(hipStreamSynchronize(st->ctx->stream));
    SC_HIP_CHECK(hipMemcpy(interleaved, st->amp, n_complex * sizeof(double2), hipMemcpyDeviceToHost));
    return SC_OK;
}

int sc_state_copy(sc_state *dst, const sc_state *src) {
    SC_REQUIRE(dst, "dst");
    SC_REQUIRE(src, "src");
    if (dst->n != src->n || dst->batch != src->batch) return fail(SC_E_ARG, "state shapes differ");
    int rc = set_device(dst->ctx);
    if (rc) return rc;
    if ((rc = materialize_basis(const_cast<sc_state *>(src)))) return rc;
    if ((rc = launch_copy(dst->ctx, src->amp, dst->amp, (uint64_t)src->batch << src->n))) return rc;
    dst->pending_basis = -1;
    SC_HIP_CHECK(hipStreamSynchronize(dst->ctx->stream));
    return SC_OK;
}

int sc_state_read_range(sc_state *st, int elem, uint64_t offset, uint64_t count, double *interleaved) {
    SC_REQUIRE(st, "state");
    if (count > 0) SC_REQUIRE(interleaved, "data");
    if (elem < 0 || elem >= st->batch) return fail(SC_E_ARG, "batch element out of range");
    if (offset > st->dim() || count > st->dim() - offset) return fail(SC_E_ARG, "amplitude range out of bounds");
    int rc = set_device(st->ctx);
    if (rc) return rc;
    if ((rc = materialize_basis(st))) return rc;
    SC_HIP_CHECK(hipStreamSynchronize(st->ctx->stream));
    if (count > 0)
        SC_HIP_CHECK(hipMemcpy(interleaved, st->amp + ((uint64_t)elem << st->n) + offset, count * sizeof(double2),
                               hipMemcpyDeviceToHost));
    return SC_OK;
}

int sc_apply_ops(sc_state *st, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex) {
    SC_REQUIRE(st, "state");
    if (n_ops > 0) SC_REQUIRE(ops, "ops");
    if (n_mat_complex > 0) SC_REQUIRE(mats, "mats");
    return apply_ops_impl(st, ops, n_ops, mats, n_mat_complex, 0, false);
}

int sc_apply_ops_batched(sc_state *st, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t mat_stride) {
    SC_REQUIRE(st, "state");
    if (n_ops > 0) SC_REQUIRE(ops, "ops");
    if (mat_stride > 0) SC_REQUIRE(mats, "mats");
    return apply_ops_impl(st, ops, n_ops, mats, mat_stride, mat_stride, true);
}

// ---- measurement ----------------------------------------------------------------
static int check_measured(const sc_state *st, int elem, const int32_t *measured, int m) {
    if (elem < 0 || elem >= st->batch) return fail(SC_E_ARG, "batch element out of range");
    if (m <= 0) return fail(SC_E_ARG, "measured qubit list must not be empty");
    if (m > st->n) return fail(SC_E_ARG, "more measured qubits than qubits");
    for (int k = 0; k < m; k++) {
        if (measured[k] < 0 || measured[k] >= st->n)
            return fail(SC_E_ARG, "measurement qubit out of range: " + std::to_string(measured[k]));
        if (k > 0 && measured[k] <= measured[k - 1])
            return fail(SC_E_ARG, "measured qubits must be sorted ascending and distinct");
    }
    return SC_OK;
}

// computeMeasurementProbabilities (:274-296), exactly: every qubit in order -> |a_i|^2; a few complement bits ->
// one thread per outcome running the reference's loop; otherwise the segment-parallel exact sums.
static int probabilities_dev(sc_state *st, int elem, const int32_t *measured, int m, double *out_dev) {
    bool identity = (m == st->n);
    for (int k = 0; identity && k < m; k++) identity = measured[k] == k;
    if (identity) return launch_probabilities_full(st->ctx, st, elem, out_dev);
    if (st->n - m <= seq::kSegBits) return launch_probabilities_marginal(st->ctx, st, elem, measured, m, out_dev);
    return seq::exact_marginals(st->ctx, st->amp + ((uint64_t)elem << st->n), st->n, measured, m, out_dev);
}

int sc_probabilities(sc_state *st, int elem, const int32_t *measured, int m, double *out) {
    SC_REQUIRE(st, "state");
    SC_REQUIRE(measured, "measured");
    SC_REQUIRE(out, "out");
    int rc = check_measured(st, elem, measured, m);
    if (rc) return rc;
    rc = set_device(st->ctx);
    if (rc) return rc;
    if ((rc = materialize_basis(st))) return rc;
    DevBuf p;
    const uint64_t count = 1ULL << m;
    if ((rc = p.alloc(count * sizeof(double)))) return rc;
    if ((rc = probabilities_dev(st, elem, measured, m, (double *)p.p))) return rc;
    SC_HIP_CHECK(hipStreamSynchronize(st->ctx->stream));
    SC_HIP_CHECK(hipMemcpy(out, p.p, count * sizeof(double), hipMemcpyDeviceToHost));
    return SC_OK;
}

// performMeasurement (:229-257) for given uniforms: exact marginals (unless every qubit is measured in order),
// then the exact normalize / cumulative / sampleOutcome of sc_seqsum.hip.
int sc_sample(sc_state *st, int elem, const int32_t *measured, int m, const double *uniforms, int64_t shots,
              int64_t *outcomes, double *boundary_gap) {
    SC_REQUIRE(st, "state");
    SC_REQUIRE(measured, "measured");
    if (shots <= 0) return fail(SC_E_ARG, "shots must be positive");
    SC_REQUIRE(uniforms, "uniforms");
    SC_REQUIRE(outcomes, "outcomes");
    int rc = check_measured(st, elem, measured, m);
    if (rc) return rc;
    Context *ctx = st->ctx;
    if ((rc = set_device(ctx))) return rc;
    if ((rc = materialize_basis(st))) return rc;
    bool identity = (m == st->n);
    for (int k = 0; identity && k < m; k++) identity = measured[k] == k;
    DevBuf p, io;
    if (!identity) {
        if ((rc = p.alloc((1ULL << m) * sizeof(double)))) return rc;
        if ((rc = probabilities_dev(st, elem, measured, m, (double *)p.p))) return rc;
    }
    if ((rc = io.alloc((size_t)shots * 24))) return rc;
    double *u_dev = (double *)io.p;
    int64_t *o_dev = (int64_t *)((char *)io.p + (size_t)shots * 8);
    double *g_dev = (double *)((char *)io.p + (size_t)shots * 16);
    SC_HIP_CHECK(hipMemcpyAsync(u_dev, uniforms, (size_t)shots * 8, hipMemcpyHostToDevice, ctx->stream));
    seq::SampleStats stats;
    if ((rc = seq::exact_sample(ctx, st->amp + ((uint64_t)elem << st->n), st->n, measured, m,
                                identity ? nullptr : (const double *)p.p, uniforms, u_dev, shots, o_dev,
                                boundary_gap ? g_dev : nullptr, &stats)))
        return rc;
    ctx->last_sample[0] = stats.fast;
    ctx->last_sample[1] = stats.serial;
    ctx->last_sample[2] = stats.zero;
    ctx->last_sample[3] = stats.gather_rounds;
    ctx->last_sample[4] = stats.rescaled ? 1 : 0;
    SC_HIP_CHECK(hipMemcpyAsync(outcomes, o_dev, (size_t)shots * 8, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<double> gaps;
    if (boundary_gap) {
        gaps.resize((size_t)shots);
        SC_HIP_CHECK(hipMemcpyAsync(gaps.data(), g_dev, (size_t)shots * 8, hipMemcpyDeviceToHost, ctx->stream));
    }
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (int64_t s = 0; s < shots; s++)
        if (outcomes[s] < 0) return fail(SC_E_STATE, "exact cumulative search found no outcome for a shot");
    if (boundary_gap) *boundary_gap = gaps.empty() ? 0.0 : *std::min_element(gaps.begin(), gaps.end());
    return SC_OK;
}

int sc_last_sample_stats(sc_ctx *ctx, uint64_t out[5]) {
    SC_REQUIRE(ctx, "ctx");
    SC_REQUIRE(out, "out");
    for (int i = 0; i < 5; i++) out[i] = ctx->last_sample[i];
    return SC_OK;
}

int sc_norm_sq(sc_state *st, int elem, double *out) {
    SC_REQUIRE(st, "state");
    SC_REQUIRE(out, "out");
    if (elem < 0 || elem >= st->batch) return fail(SC_E_ARG, "batch element out of range");
    int rc = set_device(st->ctx);
    if (rc) return rc;
    if ((rc = materialize_basis(st))) return rc;
    return launch_norm_sq(st->ctx, st->amp + ((uint64_t)elem << st->n), st->dim(), out);
}

int sc_probability_checksum(sc_state *st, int elem, double out[2]) {
    SC_REQUIRE(st, "state");
    SC_REQUIRE(out, "out");
    if (elem < 0 || elem >= st->batch) return fail(SC_E_ARG, "batch element out of range");
    int rc = set_device(st->ctx);
    if (rc) return rc;
    if ((rc = materialize_basis(st))) return rc;
    return launch_checksum(st->ctx, st->amp + ((uint64_t)elem << st->n), st->dim(), out);
}

int sc_state_max_deviation(sc_state *st, int elem, uint64_t index, const double a[2], const double b[2],
                           double *out) {
    SC_REQUIRE(st, "state");
    SC_REQUIRE(a, "a");
    SC_REQUIRE(b, "b");
    SC_REQUIRE(out, "out");
    if (elem < 0 || elem >= st->batch) return fail(SC_E_ARG, "batch element out of range");
    if (index >= st->dim()) return fail(SC_E_ARG, "basis index out of range");
    int rc = set_device(st->ctx);
    if (rc) return rc;
    if ((rc = materialize_basis(st))) return rc;
    return launch_max_deviation(st->ctx, st->amp + ((uint64_t)elem << st->n), st->dim(), index,
                                make_double2(a[0], a[1]), make_double2(b[0], b[1]), out);
}

// ---- noise ------------------------------------------------------------------------
int sc_kraus_weights(sc_state *st, int elem, int target, const double *kraus, int k, double *weights) {
    SC_REQUIRE(st, "state");
    SC_REQUIRE(kraus, "kraus");
    SC_REQUIRE(weights, "weights");
    if (elem < 0 || elem >= st->batch) return fail(SC_E_ARG, "batch element out of range");
    if (target < 0 || target >= st->n) return fail(SC_E_ARG, "noise channel qubit out of range: " + std::to_string(target));
    int rc = set_device(st->ctx);
    if (rc) return rc;
    if ((rc = materialize_basis(st))) return rc;
    return launch_kraus_weights(st->ctx, st->amp + ((uint64_t)elem << st->n), st->n, target, kraus, k, weights);
}

int sc_apply_kraus(sc_state *st, int elem, int target, const double *kraus_op, double *norm_sq_out) {
    SC_REQUIRE(st, "state");
    SC_REQUIRE(kraus_op, "kraus_op");
    if (elem < 0 || elem >= st->batch) return fail(SC_E_ARG, "batch element out of range");
    if (target < 0 || target >= st->n) return fail(SC_E_ARG, "noise channel qubit out of range: " + std::to_string(target));
    Context *ctx = st->ctx;
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = materialize_basis(st))) return rc;
    State view = *st;
    view.amp = st->amp + ((uint64_t)elem << st->n);
    view.batch = 1;
    Mat2 g;
    std::memcpy(g.v, kraus_op, sizeof(g.v));
    if ((rc = launch_single(ctx, &view, target, g, 0))) return rc;
    // renormalizeState (StatevectorEngine.java:590-600)
    double normSq = 0.0;
    if ((rc = launch_norm_sq(ctx, view.amp, st->dim(), &normSq))) return rc;
    if (norm_sq_out) *norm_sq_out = normSq;
    if (!std::isfinite(normSq) || normSq == 0.0)
        return fail(SC_E_STATE, "state collapsed to invalid norm after Kraus application: normSq=" + std::to_string(normSq));
    if (std::fabs(normSq - 1.0) > 1e-12) {
        if ((rc = launch_scale_state(ctx, view.amp, st->dim(), 1.0 / std::sqrt(normSq)))) return rc;
        SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    return SC_OK;
}

// ---- tracing ----------------------------------------------------------------------
int sc_profile_reset(sc_ctx *ctx) {
    SC_REQUIRE(ctx, "ctx");
    int rc = drain_profiler(ctx);
    if (rc) return rc;
    for (int f = 0; f < SC_KFAM_COUNT; f++) {
        ctx->prof.launches[f] = 0;
        ctx->prof.ms[f] = 0;
        ctx->prof.bytes[f] = 0;
    }
    return SC_OK;
}

int sc_profile_read(sc_ctx *ctx, int family, uint64_t *launches, double *total_ms, double *bytes) {
    SC_REQUIRE(ctx, "ctx");
    if (family < 0 || family >= SC_KFAM_COUNT) return fail(SC_E_ARG, "unknown kernel family");
    int rc = drain_profiler(ctx);
    if (rc) return rc;
    if (launches) *launches = ctx->prof.launches[family];
    if (total_ms) *total_ms = ctx->prof.ms[family];
    if (bytes) *bytes = ctx->prof.bytes[family];
    return SC_OK;
}

int sc_plan_sweeps_host(int num_qubits, const sc_op *ops, uint64_t n_ops, const double *mats,
                        uint64_t n_mat_complex, int low_qubits, int remap, int jit, int pass_cap, int chunk_bits,
                        uint64_t *passes, uint64_t *stages, uint64_t *fused_blocks, double *fp64_ops_per_amp,
                        double *state_inout) {
    if (n_ops > 0) SC_REQUIRE(ops, "ops");
    if (n_mat_complex > 0) SC_REQUIRE(mats, "mats");
    if (num_qubits < kChunkBits || num_qubits > 34) return fail(SC_E_ARG, "sweep plans need 12..34 qubits");
    if (chunk_bits != kChunkBits && !(chunk_bits == kMaxChunkBits && jit))
        return fail(SC_E_ARG, "chunk bits must be 12, or 13 with runtime-compiled passes");
    if (low_qubits < 0 || low_qubits > 8) return fail(SC_E_ARG, "low qubits must be 0..8");
    std::vector<const double *> elems{mats};
    std::vector<PlanOp> pops;
    int rc = analyze_ops(num_qubits, ops, n_ops, elems, n_mat_complex, pops);
    if (rc) return rc;
    SweepCostModel cm = jit ? kCompiledPassCost : kInterpretedCost;
    if (pass_cap >= 0) cm.pass_cap = (double)pass_cap;
    cm.permlane = jit == 2;  // in-register stage transitions (SC_OPT_PERMLANE)
    SweepPlan plan;
    if ((rc = plan_sweeps(num_qubits, pops, low_qubits, remap != 0, plan, kRegBits, cm, chunk_bits))) return rc;
    uint64_t ns = 0;
    for (const auto &p : plan.passes) ns += (uint64_t)p.nstages;
    if (passes) *passes = plan.passes.size();
    if (stages) *stages = ns;
    if (fused_blocks) *fused_blocks = (uint64_t)plan.fused_blocks;
    if (fp64_ops_per_amp) *fp64_ops_per_amp = plan.fp64_ops;
    if (state_inout) {
        std::vector<double> pool(2 * std::max<uint64_t>(plan.pool_complex, 1));
        build_pool(plan, mats, pool.data());
        execute_plan_host(num_qubits, plan, pool.data(), state_inout);
    }
    return SC_OK;
}

int sc_plan_pass_stats_host(int num_qubits, const sc_op *ops, uint64_t n_ops, const double *mats,
                            uint64_t n_mat_complex, int low_qubits, int remap, int jit, int chunk_bits, int max_passes,
                            int32_t *stages, int32_t *barriers, double *fp64_ops_per_amp, int32_t *direct,
                            int *n_passes) {
    if (n_ops > 0) SC_REQUIRE(ops, "ops");
    if (n_mat_complex > 0) SC_REQUIRE(mats, "mats");
    SC_REQUIRE(n_passes, "n_passes");
    if (num_qubits < kChunkBits || num_qubits > 34) return fail(SC_E_ARG, "sweep plans need 12..34 qubits");
    if (chunk_bits != kChunkBits && !(chunk_bits == kMaxChunkBits && jit))
        return fail(SC_E_ARG, "chunk bits must be 12, or 13 with runtime-compiled passes");
    if (low_qubits < 0 || low_qubits > 8) return fail(SC_E_ARG, "low qubits must be 0..8");
    std::vector<const double *> elems{mats};
    std::vector<PlanOp> pops;
    int rc = analyze_ops(num_qubits, ops, n_ops, elems, n_mat_complex, pops);
    if (rc) return rc;
    SweepPlan plan;
    if ((rc = plan_sweeps(num_qubits, pops, low_qubits, remap != 0, plan, kRegBits,
                          jit ? kCompiledPassCost : kInterpretedCost, chunk_bits)))
        return rc;
    *n_passes = (int)plan.passes.size();
    for (int p = 0; p < (int)plan.passes.size() && p < max_passes; p++) {
        const SweepPass &ps = plan.passes[p];
        int nb = 0;
        double f = 0.0;
        for (int st = 0; st < ps.nstages; st++) {
            nb += ps.stages[st].wave_local ? 0 : 1;
            for (int o = ps.stages[st].op_begin; o < ps.stages[st].op_end; o++) f += sweep_op_fp64(plan.ops[o]);
        }
        if (stages) stages[p] = ps.nstages;
        if (barriers) barriers[p] = nb;
        if (fp64_ops_per_amp) fp64_ops_per_amp[p] = f;
        if (direct) direct[p] = ps.direct;
    }
    return SC_OK;
}

int sc_jit_compile_host(int num_qubits, const sc_op *ops, uint64_t n_ops, const double *mats,
                        uint64_t n_mat_complex, int low_qubits, int remap, int jit, int chunk_bits, int threads,
                        uint64_t *kernels, double *seconds, uint64_t *code_bytes, const char *dump_dir) {
    if (n_ops > 0) SC_REQUIRE(ops, "ops");
    if (n_mat_complex > 0) SC_REQUIRE(mats, "mats");
    if (num_qubits < kChunkBits || num_qubits > 34) return fail(SC_E_ARG, "sweep plans need 12..34 qubits");
    if (low_qubits < 0 || low_qubits > 8) return fail(SC_E_ARG, "low qubits must be 0..8");
    if (threads < 1 || threads > 64) return fail(SC_E_ARG, "threads must be 1..64");
    if (jit < 1 || jit > 4) return fail(SC_E_ARG, "jit must be 1, 2, 3 or 4");
    if (chunk_bits != kChunkBits && !(chunk_bits == kMaxChunkBits && jit != 3))
        return fail(SC_E_ARG, "chunk bits must be 12, or 13 for forward passes");
    SweepPlan plan;
    int rc;
    if (jit == 3) {  // the circuit's adjoint reverse sweep (parameterized ops carry param / gen)
        std::vector<double> inv;
        std::vector<int> slot;
        int n_params = 0;
        for (uint64_t i = 0; i < n_ops; i++) n_params = std::max(n_params, ops[i].param + 1);
        if ((rc = adjoint_plan_host(num_qubits, ops, n_ops, mats, n_mat_complex, n_params, low_qubits, plan, inv,
                                    slot)))
            return rc;
    } else {
        std::vector<const double *> elems{mats};
        std::vector<PlanOp> pops;
        if ((rc = analyze_ops(num_qubits, ops, n_ops, elems, n_mat_complex, pops))) return rc;
        std::vector<std::pair<int, int>> absorbed;
        if (jit == 4) {  // the product-state plan from a basis state
            absorb_leading_ops(num_qubits, pops, absorbed);
            if (absorbed.empty() || absorbed.size() == pops.size())
                return fail(SC_E_ARG, "no leading single-qubit ops to absorb");
        }
        if ((rc = plan_sweeps(num_qubits, pops, low_qubits, remap != 0, plan, kRegBits, kCompiledPassCost,
                              chunk_bits)))
            return rc;
        if (jit == 4 && !plan.passes.empty()) finalize_product(num_qubits, plan, absorbed);
    }
    if (kernels) *kernels = plan.passes.size();
    return jit_compile_host(plan, jit == 4 ? 2 : jit, threads, seconds, code_bytes, dump_dir);
}

int sc_last_plan_stats(sc_ctx *ctx, uint64_t *sweeps, uint64_t *kernel_launches, uint64_t *fused_blocks) {
    SC_REQUIRE(ctx, "ctx");
    if (sweeps) *sweeps = ctx->last_sweeps;
    if (kernel_launches) *kernel_launches = ctx->last_launches;
    if (fused_blocks) *fused_blocks = ctx->last_blocks;
    return SC_OK;
}

int sc_last_plan_geometry(sc_ctx *ctx, int *chunk_bits, int *low_qubits, int *candidates, double *ns_per_amp) {
    SC_REQUIRE(ctx, "ctx");
    if (chunk_bits) *chunk_bits = ctx->last_chunk_bits;
    if (low_qubits) *low_qubits = ctx->last_low_qubits;
    if (candidates) *candidates = ctx->last_candidates;
    if (ns_per_amp) *ns_per_amp = ctx->last_ns_per_amp;
    return SC_OK;
}

int sc_plan_product_host(int num_qubits, const sc_op *ops, uint64_t n_ops, const double *mats,
                         uint64_t n_mat_complex, uint64_t basis, int low_qubits, int chunk_bits, int *absorbed,
                         double *state_out) {
    if (n_ops > 0) SC_REQUIRE(ops, "ops");
    if (n_mat_complex > 0) SC_REQUIRE(mats, "mats");
    SC_REQUIRE(state_out, "state_out");
    if (num_qubits < kChunkBits || num_qubits > 26) return fail(SC_E_ARG, "host product plans take 12..26 qubits");
    if (num_qubits < 64 && (basis >> num_qubits) != 0) return fail(SC_E_ARG, "basis index out of range");
    if (chunk_bits != 0 && chunk_bits != kChunkBits && chunk_bits != kMaxChunkBits)
        return fail(SC_E_ARG, "chunk bits must be 12 or 13, or 0 (chosen per circuit)");
    if (low_qubits < -1 || low_qubits > 8) return fail(SC_E_ARG, "low qubits must be 0..8, or -1 (chosen per circuit)");
    std::vector<const double *> elems{mats};
    std::vector<PlanOp> pops;
    int rc = analyze_ops(num_qubits, ops, n_ops, elems, n_mat_complex, pops);
    if (rc) return rc;
    std::vector<std::pair<int, int>> abs;
    absorb_leading_ops(num_qubits, pops, abs);
    if (absorbed) *absorbed = (int)abs.size();
    if (abs.empty() || abs.size() == pops.size())
        return fail(SC_E_ARG, "no leading single-qubit ops to absorb, or nothing after them");
    int cb = chunk_bits;
    if (cb == kMaxChunkBits && num_qubits < kMaxChunkBits) cb = kChunkBits;
    SweepPlan plan;
    GeometryPick pick;
    if ((rc = plan_sweeps_geometry(num_qubits, pops, low_qubits, cb, false, true, plan, kRegBits, kCompiledPassCost,
                                   &pick)))
        return rc;
    if (plan.passes.empty()) return fail(SC_E_ARG, "the remaining ops need no pass");
    finalize_product(num_qubits, plan, abs);
    std::vector<double> pool(2 * std::max<uint64_t>(plan.pool_complex, 1));
    build_pool(plan, mats, pool.data(), (long long)basis);
    product_input_host(num_qubits, plan, pool.data(), state_out);
    execute_plan_host(num_qubits, plan, pool.data(), state_out);
    return SC_OK;
}

int sc_plan_geometry_host(int num_qubits, const sc_op *ops, uint64_t n_ops, const double *mats,
                          uint64_t n_mat_complex, int low_qubits, int remap, int jit, int chunk_bits,
                          int *chunk_bits_out, int *low_qubits_out, int *candidates, double *ns_per_amp) {
    if (n_ops > 0) SC_REQUIRE(ops, "ops");
    if (n_mat_complex > 0) SC_REQUIRE(mats, "mats");
    if (num_qubits < kChunkBits || num_qubits > 34) return fail(SC_E_ARG, "sweep plans need 12..34 qubits");
    if (chunk_bits != 0 && chunk_bits != kChunkBits && chunk_bits != kMaxChunkBits)
        return fail(SC_E_ARG, "chunk bits must be 12 or 13, or 0 (chosen per circuit)");
    if (low_qubits < -1 || low_qubits > 8) return fail(SC_E_ARG, "low qubits must be 0..8, or -1 (chosen per circuit)");
    if (jit < 0 || jit > 2) return fail(SC_E_ARG, "jit must be 0, 1 or 2");
    std::vector<const double *> elems{mats};
    std::vector<PlanOp> pops;
    int rc = analyze_ops(num_qubits, ops, n_ops, elems, n_mat_complex, pops);
    if (rc) return rc;
    int cb = chunk_bits;
    if (!jit || (cb == kMaxChunkBits && num_qubits < kMaxChunkBits)) cb = kChunkBits;
    SweepPlan plan;
    GeometryPick pick;
    if ((rc = plan_sweeps_geometry(num_qubits, pops, low_qubits, cb, remap != 0, jit != 0, plan, kRegBits,
                                   jit ? kCompiledPassCost : kInterpretedCost, &pick)))
        return rc;
    if (chunk_bits_out) *chunk_bits_out = pick.chunk_bits;
    if (low_qubits_out) *low_qubits_out = pick.low_qubits;
    if (candidates) *candidates = pick.candidates;
    if (ns_per_amp) *ns_per_amp = pick.ns_per_amp;
    return SC_OK;
}

}  // extern "C"
