// sc_graph.hip -- many batched circuits and their per-term expectations in one HIP graph launch.
//
// sc_run_jobs takes a list of jobs (a batch of states, a circuit with one matrix array per element, Pauli
// terms) and does for each what sc_state_set_basis + sc_apply_ops_batched + sc_expectation_pauli_terms do,
// with the same plans and kernels.  When every term of a job is a Z string (SC_OPT_FUSED_TERMS), its last pass
// also sums the terms while it stores the states, instead of a term kernel reading the batch again; the
// results then agree with the separate calls to rounding (other summation order), otherwise bit for bit.  The first call of
// a job structure plans every circuit (runtime-compiled passes), allocates the matrix pools, the partial sums
// and pinned staging once, and records the whole sequence -- pool uploads, every pass launch of every job,
// the term kernels and the copies of the row sums back -- as one hipGraph on the context's stream.  Later
// calls with the same structure (new parameters, same circuit shapes) only rebuild the pools on host threads
// and launch that graph once: no planning, no per-launch host work, one synchronisation for all jobs.
//
// The graph holds no memset (each job's first pass generates its basis-state input) and no allocation.
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "sc_internal.hpp"
#include "sc_jit.hpp"
#include "sc_plan.hpp"

namespace sc {

namespace {

struct GraphJob {
    SweepPlan plan;                  // the job's plan (kept: the context's plan cache may evict its entry)
    std::vector<hipFunction_t> fns;  // its pass kernels (cached for the life of the process)
    // Z-string terms only (every x mask 0, at most 32 terms, a batch in one launch): the last pass sums the
    // terms as it stores the states (terms_fn, tparts partial slots per row) instead of a term kernel reading
    // the batch again
    hipFunction_t terms_fn = nullptr;
    int tparts = 0;
    size_t pool_off = 0, pool_bytes = 0;      // device and host offsets of the per-element pools
    size_t scratch_off = 0, scratch_bytes = 0;  // device: term partial sums and row sums
    size_t rows_off = 0, rows_bytes = 0;      // device row sums / host staging of them
    size_t host_rows_off = 0;
    uint64_t batch = 1;
};

struct JobGraph {
    std::string key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *dev = nullptr;
    void *host = nullptr;  // pinned
    std::vector<GraphJob> jobs;
    uint64_t used = 0;
    ~JobGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (dev) (void)hipFree(dev);
        if (host) (void)hipHostFree(host);
    }
};

constexpr size_t kMaxGraphs = 8;

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

void put(std::string &k, const void *p, size_t bytes) { k.append((const char *)p, bytes); }

// Builds every element's matrix pool of every job into the pinned staging, on host threads shared by all jobs.
void build_pools(const sc_job *jobs, const std::vector<GraphJob> &gjs, char *host) {
    std::vector<uint64_t> first(gjs.size() + 1, 0);  // global element index of each job's element 0
    uint64_t work = 0;
    for (size_t j = 0; j < gjs.size(); j++) {
        first[j + 1] = first[j] + gjs[j].batch;
        work += gjs[j].batch * (gjs[j].plan.pool.size() + 1);
    }
    const uint64_t total = first.back();
    const uint64_t grain = work < 4096 ? total : std::max<uint64_t>(1, total / (4 * kHostThreads));
    parallel_ranges(total, grain, [&](unsigned, uint64_t e0, uint64_t e1) {
        size_t j = std::upper_bound(first.begin(), first.end(), e0) - first.begin() - 1;
        for (uint64_t g = e0; g < e1; g++) {
            while (g >= first[j + 1]) j++;
            const sc_job &job = jobs[j];
            const GraphJob &gj = gjs[j];
            const uint64_t e = g - first[j];
            double *dst = (double *)(host + gj.pool_off) + 2 * gj.plan.pool_complex * e;
            build_pool(gj.plan, job.mats + 2 * job.mat_stride * e, dst, (long long)job.basis);
        }
    });
}

// The enqueue half of one job (recorded into the graph): pool upload, passes, term kernels, row copy.
int record_job(Context *ctx, const sc_job &job, GraphJob &gj, char *dev, char *host) {
    State *st = job.state;
    const double *dpool = (const double *)(dev + gj.pool_off);
    SC_HIP_CHECK(hipMemcpyAsync(dev + gj.pool_off, host + gj.pool_off, gj.pool_bytes, hipMemcpyHostToDevice,
                                ctx->stream));
    const size_t stride = gj.batch > 1 ? gj.plan.pool_complex : 0;
    const int log_chunks = st->n - gj.plan.chunk_bits;
    const int per_launch = std::max(1, (int)std::min<uint64_t>(gj.batch, (1ULL << 20) >> log_chunks));
    const size_t fused_last = gj.terms_fn ? gj.plan.passes.size() - 1 : gj.plan.passes.size();
    for (size_t p = 0; p < fused_last; p++)
        for (int e0 = 0; e0 < st->batch; e0 += per_launch) {
            const int cnt = std::min(per_launch, st->batch - e0);
            int rc = launch_sweep_jit(ctx, st, gj.fns[p], dpool, e0, cnt, stride, gj.plan.chunk_bits, st->amp, st->amp,
                                      p == 0 ? (long long)job.basis : -1);
            if (rc) return rc;
        }
    double *rows = (double *)(dev + gj.rows_off);
    int rc = gj.terms_fn ? launch_sweep_jit_terms(ctx, st, gj.terms_fn, dpool, stride, gj.plan.chunk_bits, job.n_terms,
                                                  (double *)(dev + gj.scratch_off), gj.tparts, rows)
                         : pauli_terms_enqueue(ctx, st, st->amp, job.xmask, job.zmask, job.n_terms, dev + gj.scratch_off,
                                               gj.scratch_bytes, rows);
    if (rc) return rc;
    SC_HIP_CHECK(hipMemcpyAsync(host + gj.host_rows_off, rows, gj.rows_bytes, hipMemcpyDeviceToHost, ctx->stream));
    return SC_OK;
}

int build_graph(Context *ctx, const sc_job *jobs, int n_jobs, std::vector<PlanCache::Entry *> &ents, JobGraph &g) {
    size_t dev_bytes = 0, host_bytes = 0;
    g.jobs.resize(n_jobs);
    for (int j = 0; j < n_jobs; j++) {
        GraphJob &gj = g.jobs[j];
        const State *st = jobs[j].state;
        gj.plan = ents[j]->plan;
        gj.fns = ents[j]->fns;
        gj.batch = (uint64_t)st->batch;
        gj.pool_bytes = (size_t)std::max<uint64_t>(gj.plan.pool_complex, 1) * 16 * gj.batch;
        gj.pool_off = dev_bytes;  // the same offsets in the pinned staging
        dev_bytes = align256(dev_bytes + gj.pool_bytes);
        host_bytes = dev_bytes;
    }
    for (int j = 0; j < n_jobs; j++) {
        GraphJob &gj = g.jobs[j];
        const sc_job &job = jobs[j];
        const uint64_t rows = (uint64_t)job.n_terms;
        const int n = job.state->n, log_chunks = n - gj.plan.chunk_bits;
        // (the first pass generates the basis input, so the term pass must be a later one)
        bool fuse = ctx->fused_terms && gj.batch > 1 && job.n_terms <= 32 && gj.plan.passes.size() >= 2 &&
                    (gj.batch << log_chunks) <= (1ULL << 20);
        for (int t = 0; t < job.n_terms && fuse; t++) fuse = job.xmask[t] == 0;
        if (fuse) {
            const std::vector<uint64_t> z(job.zmask, job.zmask + job.n_terms);
            int rc = jit_last_pass_terms_kernel(ctx, gj.plan, z, gj.terms_fn);
            if (rc) return rc;
            gj.tparts = sweep_terms_parts(ctx, n, gj.plan.chunk_bits, (int)gj.batch);
            gj.scratch_bytes = (size_t)(gj.batch * rows * (uint64_t)gj.tparts) * sizeof(double);
        } else {
            gj.scratch_bytes = term_scratch_bytes(gj.batch, rows, term_parts(ctx, gj.batch, n, rows));
        }
        gj.scratch_off = dev_bytes;
        dev_bytes = align256(dev_bytes + gj.scratch_bytes);
        gj.rows_bytes = (size_t)(gj.batch * rows) * sizeof(double);
        gj.rows_off = dev_bytes;
        dev_bytes = align256(dev_bytes + gj.rows_bytes);
        gj.host_rows_off = host_bytes;
        host_bytes = align256(host_bytes + gj.rows_bytes);
    }
    if (hipMalloc(&g.dev, std::max<size_t>(dev_bytes, 256)) != hipSuccess) {
        (void)hipGetLastError();
        g.dev = nullptr;
        return fail(SC_E_OOM, "device allocation of " + std::to_string(dev_bytes) + " bytes for a job graph failed");
    }
    SC_HIP_CHECK(hipHostMalloc(&g.host, std::max<size_t>(host_bytes, 256), hipHostMallocDefault));
    // record: profiling events stay out of the graph
    const bool prof = ctx->prof.enabled;
    ctx->prof.enabled = false;
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    SC_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = SC_OK;
    for (int j = 0; j < n_jobs && rc == SC_OK; j++) rc = record_job(ctx, jobs[j], g.jobs[j], (char *)g.dev, (char *)g.host);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
    ctx->prof.enabled = prof;
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return fail(SC_E_HIP, std::string("hipStreamEndCapture failed: ") + hipGetErrorString(ec));
    g.graph = graph;
    SC_HIP_CHECK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    return SC_OK;
}

// The jobs one after another through the ordinary entry points' machinery (no graph).
int run_sequential(const sc_job *jobs, int n_jobs) {
    for (int j = 0; j < n_jobs; j++) {
        const sc_job &job = jobs[j];
        State *st = job.state;
        st->pending_basis = (int64_t)job.basis;
        int rc = apply_ops_impl(st, job.ops, job.n_ops, job.mats, job.mat_stride, job.mat_stride, true);
        if (rc) return rc;
        if ((rc = materialize_basis(st))) return rc;
        Context *ctx = st->ctx;
        const uint64_t rows = (uint64_t)job.n_terms, B = (uint64_t)st->batch;
        const size_t need = term_scratch_bytes(B, rows, term_parts(ctx, B, st->n, rows));
        if ((rc = ensure_scratch(ctx, need))) return rc;
        double *out_dev = (double *)((char *)ctx->scratch + need - (size_t)(B * rows) * sizeof(double));
        if ((rc = pauli_terms_enqueue(ctx, st, st->amp, job.xmask, job.zmask, job.n_terms, ctx->scratch, need,
                                      out_dev)))
            return rc;
        SC_HIP_CHECK(hipMemcpyAsync(job.out, out_dev, B * rows * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        pauli_terms_finish(job.xmask, job.zmask, job.n_terms, B, job.out);
    }
    return SC_OK;
}

}  // namespace

struct JobGraphCache {
    std::vector<std::unique_ptr<JobGraph>> graphs;
    uint64_t tick = 0;
};

void destroy_job_graphs(Context *ctx) {
    delete ctx->job_graphs;
    ctx->job_graphs = nullptr;
}

int run_jobs(Context *ctx, const sc_job *jobs, int n_jobs, int *used_graph) {
    if (used_graph) *used_graph = 0;
    bool graphable = ctx->jit && ctx->exec_mode == SC_EXEC_SWEEP;
    for (int j = 0; j < n_jobs && graphable; j++) graphable = jobs[j].state->n >= kChunkBits;
    // analyse every circuit (validation and the matrix classes the plan depends on)
    std::vector<std::vector<PlanOp>> pops(n_jobs);
    for (int j = 0; j < n_jobs; j++) {
        const sc_job &job = jobs[j];
        std::vector<const double *> elems;
        for (int b = 0; b < job.state->batch; b++) elems.push_back(job.mats + 2 * job.mat_stride * (uint64_t)b);
        int rc = analyze_ops(job.state->n, job.ops, job.n_ops, elems, job.mat_stride, pops[j]);
        if (rc) return rc;
        graphable = graphable && !pops[j].empty();
    }
    if (!graphable) return run_sequential(jobs, n_jobs);
    // plans, and the graph key: every job's shape, plan, basis and terms, and the launch geometry
    std::vector<PlanCache::Entry *> ents(n_jobs);
    std::string key;
    const int32_t geo[3] = {ctx->cu_count, ctx->sweep_blocks_per_cu, ctx->fused_terms};
    put(key, geo, sizeof geo);
    for (int j = 0; j < n_jobs; j++) {
        int rc = sweep_plan_entry(ctx, jobs[j].state->n, pops[j], &ents[j], true);
        if (rc) return rc;
        if (ents[j]->fns.empty()) return run_sequential(jobs, n_jobs);
        const State *st = jobs[j].state;
        const uint64_t head[5] = {(uint64_t)(uintptr_t)st->amp, (uint64_t)st->n, (uint64_t)st->batch, jobs[j].basis,
                                  (uint64_t)jobs[j].n_terms};
        put(key, head, sizeof head);
        const uint64_t klen = ents[j]->key.size();
        put(key, &klen, sizeof klen);
        key += ents[j]->key;
        put(key, jobs[j].xmask, sizeof(uint64_t) * (size_t)jobs[j].n_terms);
        put(key, jobs[j].zmask, sizeof(uint64_t) * (size_t)jobs[j].n_terms);
    }
    if (!ctx->job_graphs) ctx->job_graphs = new JobGraphCache();
    JobGraphCache &cache = *ctx->job_graphs;
    JobGraph *g = nullptr;
    for (auto &e : cache.graphs)
        if (e->key == key) g = e.get();
    if (!g) {
        std::unique_ptr<JobGraph> ng(new JobGraph());
        ng->key = key;
        int rc = build_graph(ctx, jobs, n_jobs, ents, *ng);
        if (rc) return rc;
        if (cache.graphs.size() >= kMaxGraphs) {
            size_t lru = 0;
            for (size_t i = 1; i < cache.graphs.size(); i++)
                if (cache.graphs[i]->used < cache.graphs[lru]->used) lru = i;
            cache.graphs.erase(cache.graphs.begin() + (long)lru);
        }
        cache.graphs.push_back(std::move(ng));
        g = cache.graphs.back().get();
    }
    g->used = ++cache.tick;
    // replay: this call's matrices into the pools, one launch, one synchronisation
    build_pools(jobs, g->jobs, (char *)g->host);
    {
        ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_SWEEP, 0.0);
        SC_HIP_CHECK(hipGraphLaunch(g->exec, ctx->stream));
    }
    SC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (int j = 0; j < n_jobs; j++) {
        const GraphJob &gj = g->jobs[j];
        std::memcpy(jobs[j].out, (const char *)g->host + gj.host_rows_off, gj.rows_bytes);
        pauli_terms_finish(jobs[j].xmask, jobs[j].zmask, jobs[j].n_terms, gj.batch, jobs[j].out);
        jobs[j].state->pending_basis = -1;
    }
    ctx->last_sweeps = ctx->last_launches = 0;
    for (const auto &gj : g->jobs) ctx->last_sweeps += gj.plan.passes.size();
    ctx->last_launches = 1;
    if (used_graph) *used_graph = 1;
    return SC_OK;
}

}  // namespace sc

using namespace sc;

extern "C" int sc_run_jobs(sc_ctx *ctx, const sc_job *jobs, int n_jobs, int *graph_launched) {
    if (!ctx) return fail(SC_E_NULL, "ctx is null");
    if (!jobs) return fail(SC_E_NULL, "jobs is null");
    if (n_jobs <= 0) return fail(SC_E_ARG, "n_jobs must be positive");
    for (int j = 0; j < n_jobs; j++) {
        const sc_job &job = jobs[j];
        const std::string at = "job " + std::to_string(j) + ": ";
        if (!job.state) return fail(SC_E_NULL, at + "state is null");
        if (job.state->ctx != ctx) return fail(SC_E_ARG, at + "state belongs to another context");
        if (job.n_ops > 0 && !job.ops) return fail(SC_E_NULL, at + "ops is null");
        if (job.mat_stride > 0 && !job.mats) return fail(SC_E_NULL, at + "mats is null");
        if (job.n_terms <= 0) return fail(SC_E_ARG, at + "n_terms must be positive");
        if (!job.xmask || !job.zmask || !job.out) return fail(SC_E_NULL, at + "xmask, zmask and out are required");
        const uint64_t dim = 1ULL << job.state->n;
        if (job.basis >= dim) return fail(SC_E_ARG, at + "basis index out of range");
        for (int t = 0; t < job.n_terms; t++)
            if (job.xmask[t] >= dim || job.zmask[t] >= dim) return fail(SC_E_ARG, at + "Pauli mask beyond the state's qubits");
        for (int k = 0; k < j; k++)
            if (jobs[k].state == job.state) return fail(SC_E_ARG, at + "a state appears in two jobs");
    }
    SC_HIP_CHECK(hipSetDevice(ctx->device));
    int used = 0;
    const int rc = run_jobs(ctx, jobs, n_jobs, &used);
    if (graph_launched) *graph_launched = used;
    return rc;
}

