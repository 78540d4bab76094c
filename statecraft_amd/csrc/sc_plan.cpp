// sc_plan.cpp -- validation, gate fusion and sweep planning (host only).
//
// Dependency model.  Every op acts on each of its qubits in one of two roles:
// "target" (non-diagonal action) or "diagonal" (controls, phases).  Two ops
// commute when every qubit they share is diagonal for both.  Planners scan
// the remaining ops in program order and take an op into the current unit
// (fused block, sweep pass, or stage) only when no earlier op left behind
// conflicts with it; the units are therefore dependency-closed and executing
// them in order reproduces the reference's sequential op loop
// (StatevectorEngine.java:90-151) up to floating-point reassociation.
#include "sc_plan.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>

#include <unistd.h>

namespace sc {

namespace {

int popc(uint64_t x) { return __builtin_popcountll(x); }

// Thread-index masks of a SW_DIAGL's two bit sources (9 bits with 512-thread chunks).
uint32_t diag_ta(const SweepOp &op) { return op.ta | ((op.thi & 1u) << 8); }
uint32_t diag_tb(const SweepOp &op) { return op.tb | (((op.thi >> 1) & 1u) << 8); }

bool is_exact(double v, double want) { return v == want; }

// 2x2 complex g (8 doubles): diagonal when both off-diagonal entries are exactly zero.
bool mat_is_diag(const double *g) { return g[2] == 0.0 && g[3] == 0.0 && g[4] == 0.0 && g[5] == 0.0; }
bool mat_is_x(const double *g) {
    return is_exact(g[0], 0.0) && is_exact(g[1], 0.0) && is_exact(g[2], 1.0) && is_exact(g[3], 0.0) &&
           is_exact(g[4], 1.0) && is_exact(g[5], 0.0) && is_exact(g[6], 0.0) && is_exact(g[7], 0.0);
}

int mat_count(int kind) {
    switch (kind) {
    case SC_OP_SINGLE:
    case SC_OP_MCU:
    case SC_OP_DIAG2: return 4;
    case SC_OP_TWO: return 16;
    default: return 0;
    }
}

struct Blocker {
    uint64_t bt = 0, bd = 0;
    bool blocked(const PlanOp &op) const { return ((op.targets | op.diag) & bt) || (op.targets & bd); }
    void mark(const PlanOp &op) {
        bt |= op.targets;
        bd |= op.diag;
    }
};

}  // namespace

bool ops_conflict(const PlanOp &a, const PlanOp &b) {
    return (a.targets & (b.targets | b.diag)) || (b.targets & a.diag);
}

namespace {

// Matrix classes of one op's matrices in one batch element, as bits (AND over the elements gives the classes
// that hold for all of them).
constexpr uint32_t kClsReal = 1u;       // every entry has a zero imaginary part
constexpr uint32_t kClsDiag = 2u;       // 2x2 diagonal
constexpr uint32_t kClsX = 4u;          // 2x2 exactly Pauli X
constexpr int kUnitOrder[4] = {2, 0, 3, 1};  // unit-pivot column pairs in order of preference
uint32_t cls_unit_ok(int i) { return 8u << (3 * i); }     // pivot pair kUnitOrder[i] qualifies
uint32_t cls_unit_real(int i) { return 16u << (3 * i); }  // ... with real off-pivot entries
uint32_t cls_unit_imag(int i) { return 32u << (3 * i); }  // ... with purely imaginary off-pivot entries

uint32_t classify_matrix(int kind, const double *g, int cnt) {
    uint32_t c = 0;
    bool real = true;
    for (int e = 0; e < cnt; e++) real = real && g[2 * e + 1] == 0.0;
    if (real) c |= kClsReal;
    if (kind == SC_OP_SINGLE || kind == SC_OP_MCU) {
        if (mat_is_diag(g)) c |= kClsDiag;
        if (mat_is_x(g)) c |= kClsX;
    }
    if (kind == SC_OP_SINGLE) {
        auto bounded = [](const double *ov, const double *pv) {
            const double a = std::hypot(ov[0], ov[1]);
            return std::isfinite(a) && a <= std::ldexp(std::fabs(pv[0]), 40);
        };
        for (int i = 0; i < 4; i++) {
            const int cc = kUnitOrder[i], c0 = cc & 1, c1 = (cc >> 1) & 1;
            const double *p0 = g + 2 * c0, *o0 = g + 2 * (1 - c0);
            const double *p1 = g + 4 + 2 * c1, *o1 = g + 4 + 2 * (1 - c1);
            if (!(p0[1] == 0.0 && p1[1] == 0.0 && p0[0] == p1[0] && p0[0] != 0.0 && std::isfinite(p0[0]) &&
                  bounded(o0, p0) && bounded(o1, p1)))
                continue;
            c |= cls_unit_ok(i);
            if (o0[1] == 0.0 && o1[1] == 0.0) c |= cls_unit_real(i);
            if (o0[0] == 0.0 && o1[0] == 0.0) c |= cls_unit_imag(i);
        }
    }
    return c;
}

}  // namespace

namespace {

// Worker threads kept for the life of the process (starting threads for every parallel region costs about as
// much as the regions' work).  One region runs at a time; a caller that finds the pool busy runs its region
// on its own thread.  A forked child starts a pool of its own.
class HostPool {
  public:
    static HostPool *get() {
        static std::mutex m;
        static HostPool *pool = nullptr;
        std::lock_guard<std::mutex> lock(m);
        if (!pool || pool->pid_ != getpid()) pool = new HostPool();  // never destroyed: workers may still wait
        return pool;
    }
    unsigned size() const { return (unsigned)workers_ + 1; }
    // Runs task(slot) on the calling thread (slot 0) and on `extra` workers (slots 1..extra); false when another
    // region holds the pool.
    bool run(unsigned extra, const std::function<void(unsigned)> &task) {
        std::unique_lock<std::mutex> region(region_mu_, std::try_to_lock);
        if (!region.owns_lock()) return false;
        {
            std::lock_guard<std::mutex> lock(mu_);
            task_ = &task;
            want_ = extra;
            active_ = extra;
            gen_++;
        }
        cv_.notify_all();
        task(0);
        std::unique_lock<std::mutex> lock(mu_);
        done_.wait(lock, [&] { return active_ == 0; });
        task_ = nullptr;
        return true;
    }

  private:
    HostPool() : pid_(getpid()) {
        const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
        workers_ = (unsigned)std::min<uint64_t>(hc, kHostThreads) - 1;
        for (unsigned i = 1; i <= workers_; i++) std::thread([this, i] { loop(i); }).detach();
    }
    void loop(unsigned slot) {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(unsigned)> *task;
            {
                std::unique_lock<std::mutex> lock(mu_);
                cv_.wait(lock, [&] { return gen_ != seen; });
                seen = gen_;
                if (slot > want_) continue;
                task = task_;
            }
            (*task)(slot);
            std::lock_guard<std::mutex> lock(mu_);
            if (--active_ == 0) done_.notify_all();
        }
    }
    pid_t pid_;
    unsigned workers_ = 0;
    std::mutex region_mu_, mu_;
    std::condition_variable cv_, done_;
    const std::function<void(unsigned)> *task_ = nullptr;
    uint64_t gen_ = 0;
    unsigned want_ = 0, active_ = 0;
};

}  // namespace

void parallel_ranges(uint64_t n, uint64_t grain, const std::function<void(unsigned, uint64_t, uint64_t)> &f) {
    grain = std::max<uint64_t>(grain, 1);
    if (n <= grain) {
        if (n) f(0, 0, n);
        return;
    }
    HostPool *pool = HostPool::get();
    const unsigned threads = (unsigned)std::min<uint64_t>((n + grain - 1) / grain, pool->size());
    std::atomic<uint64_t> next{0};
    const std::function<void(unsigned)> work = [&](unsigned t) {
        for (;;) {
            const uint64_t b = next.fetch_add(grain);
            if (b >= n) return;
            f(t, b, std::min(n, b + grain));
        }
    };
    if (threads <= 1 || !pool->run(threads - 1, work)) work(0);
}

int analyze_ops(int n, const sc_op *ops, uint64_t n_ops, const std::vector<const double *> &mats_all,
                uint64_t n_mat_complex, std::vector<PlanOp> &out) {
    out.clear();
    out.reserve(n_ops);
    bool measured = false;
    auto range = [&](int q, const char *what) -> int {
        if (q >= n) return fail(SC_E_ARG, std::string(what) + " out of range: " + std::to_string(q));
        return SC_OK;
    };
    // validation and qubit roles
    for (uint64_t i = 0; i < n_ops; i++) {
        const sc_op &o = ops[i];
        PlanOp p;
        p.kind = o.kind;
        p.q0 = o.q0;
        p.q1 = o.q1;
        p.mat = o.mat;
        p.src = (int)i;
        if (o.kind == SC_OP_MEASURE) {
            // MeasureOperation.validateTargets (QuantumCircuit.java:538-573)
            if (o.controls == 0) return fail(SC_E_ARG, "at least one qubit must be measured");
            if (n < 64 && (o.controls >> n) != 0)
                return fail(SC_E_ARG, "measured qubit out of range: " + std::to_string(63 - __builtin_clzll(o.controls)));
            measured = true;
            continue;
        }
        if (o.kind < SC_OP_SINGLE || o.kind > SC_OP_MCU)
            return fail(SC_E_STATE, "Unsupported operation type: " + std::to_string(o.kind));
        if (measured)
            return fail(SC_E_UNSUPPORTED, "Unitary operations cannot follow measurement operations in the circuit");
        int rc = SC_OK;
        switch (o.kind) {
        case SC_OP_SINGLE:  // SingleGateOperation :368-383
            if (o.q0 < 0) return fail(SC_E_ARG, "qubit index must be non-negative");
            if ((rc = range(o.q0, "target qubit"))) return rc;
            break;
        case SC_OP_CNOT:  // CnotOperation :385-409
            if (o.q0 < 0) return fail(SC_E_ARG, "control qubit index must be non-negative");
            if (o.q1 < 0) return fail(SC_E_ARG, "target qubit index must be non-negative");
            if (o.q0 == o.q1) return fail(SC_E_ARG, "control and target qubits must differ");
            if ((rc = range(o.q0, "control qubit")) || (rc = range(o.q1, "target qubit"))) return rc;
            break;
        case SC_OP_TWO:    // TwoQubitGateOperation :411-438
        case SC_OP_DIAG2:  // TwoQubitDiagonalOperation :440-468
            if (o.q0 < 0 || o.q1 < 0) return fail(SC_E_ARG, "qubit indices must be non-negative");
            if (o.q0 == o.q1) return fail(SC_E_ARG, "qubits must be distinct");
            if ((rc = range(o.q0, "first qubit")) || (rc = range(o.q1, "second qubit"))) return rc;
            break;
        case SC_OP_SWAP:  // SwapOperation :470-490
            if (o.q0 < 0 || o.q1 < 0) return fail(SC_E_ARG, "qubit indices must be non-negative");
            if (o.q0 == o.q1) return fail(SC_E_ARG, "cannot swap qubit with itself");
            if ((rc = range(o.q0, "first qubit")) || (rc = range(o.q1, "second qubit"))) return rc;
            break;
        case SC_OP_MCU:  // MultiControlOperation :492-536
            if (o.q0 < 0) return fail(SC_E_ARG, "target qubit index must be non-negative");
            if (o.controls == 0) return fail(SC_E_ARG, "at least one control qubit is required");
            if ((o.controls >> o.q0) & 1ULL) return fail(SC_E_ARG, "control and target qubits must differ");
            if ((rc = range(o.q0, "target qubit"))) return rc;
            if (n < 64 && (o.controls >> n) != 0)
                return fail(SC_E_ARG, "control qubit out of range: " + std::to_string(63 - __builtin_clzll(o.controls)));
            p.controls = o.controls;
            break;
        }
        const int cnt = mat_count(o.kind);
        if (cnt > 0 && (o.mat < 0 || (uint64_t)o.mat + (uint64_t)cnt > n_mat_complex))
            return fail(SC_E_ARG, "matrix offset out of range for op " + std::to_string(i));
        switch (o.kind) {
        case SC_OP_CNOT:
            p.is_x = true;
            p.targets = 1ULL << o.q1;
            p.diag = 1ULL << o.q0;
            p.controls = 1ULL << o.q0;
            break;
        case SC_OP_TWO:
        case SC_OP_SWAP: p.targets = (1ULL << o.q0) | (1ULL << o.q1); break;
        case SC_OP_DIAG2:
            p.is_diag = true;
            p.diag = (1ULL << o.q0) | (1ULL << o.q1);
            break;
        default: break;  // single and multi-controlled 2x2: roles follow the matrix classes below
        }
        out.push_back(p);
    }
    // matrix classes that hold for every batch element; large batches classify on several host threads
    const size_t nops = out.size(), nel = mats_all.size();
    std::vector<uint32_t> cls(nops, ~0u);
    std::vector<std::vector<uint32_t>> part;
    const uint64_t grain = std::max<uint64_t>(1, 2048 / std::max<size_t>(nops, 1));
    part.assign(std::min<uint64_t>(kHostThreads, (nel + grain - 1) / grain), std::vector<uint32_t>(nops, ~0u));
    parallel_ranges(nel, grain, [&](unsigned t, uint64_t e0, uint64_t e1) {
        std::vector<uint32_t> &c = part[t];
        for (uint64_t e = e0; e < e1; e++)
            for (size_t k = 0; k < nops; k++) {
                const int cnt = mat_count(out[k].kind);
                if (cnt > 0) c[k] &= classify_matrix(out[k].kind, mats_all[e] + 2 * (int64_t)out[k].mat, cnt);
            }
    });
    for (const auto &c : part)
        for (size_t k = 0; k < nops; k++) cls[k] &= c[k];
    for (size_t k = 0; k < nops; k++) {
        PlanOp &p = out[k];
        const int cnt = mat_count(p.kind);
        p.is_real = cnt == 0 || (cls[k] & kClsReal);
        if (p.kind == SC_OP_SINGLE) {
            for (int i = 0; i < 4; i++)
                if (cls[k] & cls_unit_ok(i)) {
                    // unit-pivot form of an uncontrolled 2x2
                    p.unit = kUnitOrder[i];
                    p.unit_form = (cls[k] & cls_unit_real(i)) ? 1 : (cls[k] & cls_unit_imag(i)) ? 2 : 0;
                    break;
                }
        }
        if (p.kind == SC_OP_SINGLE || p.kind == SC_OP_MCU) {
            p.is_diag = (cls[k] & kClsDiag) != 0;
            p.is_x = (cls[k] & kClsX) != 0;
            const uint64_t tq = 1ULL << p.q0;
            if (p.is_diag) {
                p.diag = tq | p.controls;
            } else {
                p.targets = tq;
                p.diag = p.controls;
            }
        }
    }
    return SC_OK;
}

// ------------------------------------------------------------------ dense block fusion
void plan_blocks(int n, const std::vector<PlanOp> &ops, int max_k, std::vector<FusedBlock> &blocks) {
    (void)n;
    blocks.clear();
    std::vector<int> remaining(ops.size());
    for (size_t i = 0; i < ops.size(); i++) remaining[i] = (int)i;
    while (!remaining.empty()) {
        uint64_t Q = 0;
        FusedBlock b;
        Blocker blk;
        std::vector<int> next;
        bool restarted = false;
        for (size_t ri = 0; ri < remaining.size(); ri++) {
            int i = remaining[ri];
            const PlanOp &op = ops[i];
            uint64_t qs = op.targets | op.diag;
            bool blocked = blk.blocked(op);
            if (popc(qs) > max_k) {
                if (b.ops.empty() && !blocked) {
                    FusedBlock s;
                    s.standalone = true;
                    s.ops.push_back(i);
                    blocks.push_back(s);
                    for (size_t rj = ri + 1; rj < remaining.size(); rj++) next.push_back(remaining[rj]);
                    restarted = true;
                    break;
                }
                blk.mark(op);
                next.push_back(i);
                continue;
            }
            if (blocked || popc(Q | qs) > max_k) {
                blk.mark(op);
                next.push_back(i);
                continue;
            }
            Q |= qs;
            b.ops.push_back(i);
        }
        if (!b.ops.empty()) {
            for (int q = 0; q < 64; q++)
                if ((Q >> q) & 1ULL) b.qubits.push_back(q);
            blocks.push_back(b);
        }
        (void)restarted;
        remaining.swap(next);
    }
}

namespace {

// Apply one reference op to a small vector whose local bit j is qubit qubits[j].
void apply_local(const PlanOp &op, const double *mats, const std::vector<int> &qubits, std::vector<double> &v) {
    const int k = (int)qubits.size();
    const uint64_t D = 1ULL << k;
    auto lp = [&](int q) -> uint64_t {
        for (int j = 0; j < k; j++)
            if (qubits[j] == q) return 1ULL << j;
        return 0;
    };
    uint64_t cmask = 0;
    for (int q = 0; q < 64; q++)
        if ((op.controls >> q) & 1ULL) cmask |= lp(q);
    const double *m = op.mat >= 0 ? mats + 2 * (int64_t)op.mat : nullptr;
    auto c = [&](uint64_t i) { return &v[2 * i]; };
    switch (op.kind) {
    case SC_OP_SINGLE:
    case SC_OP_MCU:
    case SC_OP_CNOT: {
        uint64_t t = op.kind == SC_OP_CNOT ? lp(op.q1) : lp(op.q0);
        for (uint64_t i = 0; i < D; i++) {
            if ((i & t) || (i & cmask) != cmask) continue;
            double *a0 = c(i), *a1 = c(i | t);
            double x0r = a0[0], x0i = a0[1], x1r = a1[0], x1i = a1[1];
            if (op.kind == SC_OP_CNOT) {
                a0[0] = x1r, a0[1] = x1i, a1[0] = x0r, a1[1] = x0i;
            } else {
                a0[0] = m[0] * x0r - m[1] * x0i + m[2] * x1r - m[3] * x1i;
                a0[1] = m[0] * x0i + m[1] * x0r + m[2] * x1i + m[3] * x1r;
                a1[0] = m[4] * x0r - m[5] * x0i + m[6] * x1r - m[7] * x1i;
                a1[1] = m[4] * x0i + m[5] * x0r + m[6] * x1i + m[7] * x1r;
            }
        }
        break;
    }
    case SC_OP_TWO: {
        uint64_t fm = lp(op.q0), sm = lp(op.q1);
        for (uint64_t b = 0; b < D; b++) {
            if ((b & fm) || (b & sm)) continue;
            uint64_t idx[4] = {b, b | sm, b | fm, b | fm | sm};
            double xr[4], xi[4];
            for (int j = 0; j < 4; j++) xr[j] = c(idx[j])[0], xi[j] = c(idx[j])[1];
            for (int r = 0; r < 4; r++) {
                double re = 0, im = 0;
                for (int j = 0; j < 4; j++) {
                    double mr = m[8 * r + 2 * j], mi = m[8 * r + 2 * j + 1];
                    re += mr * xr[j] - mi * xi[j];
                    im += mr * xi[j] + mi * xr[j];
                }
                c(idx[r])[0] = re, c(idx[r])[1] = im;
            }
        }
        break;
    }
    case SC_OP_DIAG2: {
        uint64_t fm = lp(op.q0), sm = lp(op.q1);
        for (uint64_t i = 0; i < D; i++) {
            int l = ((i & fm) ? 2 : 0) | ((i & sm) ? 1 : 0);
            double dr = m[2 * l], di = m[2 * l + 1];
            double xr = c(i)[0], xi = c(i)[1];
            c(i)[0] = dr * xr - di * xi;
            c(i)[1] = dr * xi + di * xr;
        }
        break;
    }
    case SC_OP_SWAP: {
        uint64_t fm = lp(op.q0), sm = lp(op.q1);
        for (uint64_t b = 0; b < D; b++) {
            if ((b & fm) || (b & sm)) continue;
            std::swap(c(b | fm)[0], c(b | sm)[0]);
            std::swap(c(b | fm)[1], c(b | sm)[1]);
        }
        break;
    }
    }
}

}  // namespace

void block_matrix(const FusedBlock &b, const std::vector<PlanOp> &ops, const double *mats, std::vector<double> &out) {
    const int k = (int)b.qubits.size();
    const uint64_t D = 1ULL << k;
    out.assign(2 * D * D, 0.0);
    std::vector<double> col(2 * D);
    for (uint64_t cidx = 0; cidx < D; cidx++) {
        std::fill(col.begin(), col.end(), 0.0);
        col[2 * cidx] = 1.0;
        for (int oi : b.ops) apply_local(ops[oi], mats, b.qubits, col);
        for (uint64_t r = 0; r < D; r++) {
            out[2 * (r * D + cidx)] = col[2 * r];
            out[2 * (r * D + cidx) + 1] = col[2 * r + 1];
        }
    }
}

// ------------------------------------------------------------------ sweeps
namespace {

struct StageBuild {
    uint64_t R = 0;  // global qubits that are register bits (targets)
    std::vector<int> ops;
};

// Mirror of the device run_op (sc_sweep.hip): applies one register op to the 16 complex values (32 doubles)
// of thread `tid` in the chunk whose outer offset is cb.  m is the op's matrix / factor table.  SW_DIAGL may
// also carry the planner's form with both diagonal sources in register bits (k1 >= 0); plans hand such ops to
// the device only inside a SW_DIAG16 table.
void host_apply(const SweepOp &op, const double *m, double *a, uint32_t tid, uint64_t cb) {
    if ((cb & op.octrl) != op.octrl) return;
    auto cm = [](const double *g, const double *x, double *y) {  // y = g * x
        y[0] = g[0] * x[0] - g[1] * x[1];
        y[1] = g[0] * x[1] + g[1] * x[0];
    };
    const bool lane = (tid & op.tctrl) == op.tctrl;
    double out[2 * kRegs];
    std::memcpy(out, a, sizeof(out));
    switch (op.kind) {
    case SW_U1:
    case SW_U1R:
    case SW_U1C:
    case SW_X:
    case SW_XL:
    case SW_CX: {
        if (!lane) break;
        const bool x = op.kind == SW_X || op.kind == SW_XL || op.kind == SW_CX;
        const uint32_t rc = op.kind == SW_CX ? 1u << op.k1 : op.rctrl;
        for (int r0 = 0; r0 < kRegs; r0++) {
            if (r0 & (1 << op.k0)) continue;
            if (((uint32_t)r0 & rc) != rc) continue;
            const int r1 = r0 | (1 << op.k0);
            const double *xv = a + 2 * r0, *yv = a + 2 * r1;
            if (x) {
                out[2 * r0] = yv[0], out[2 * r0 + 1] = yv[1], out[2 * r1] = xv[0], out[2 * r1 + 1] = xv[1];
            } else {
                double t0[2], t1[2], t2[2], t3[2];
                cm(m + 0, xv, t0);
                cm(m + 2, yv, t1);
                cm(m + 4, xv, t2);
                cm(m + 6, yv, t3);
                out[2 * r0] = t0[0] + t1[0], out[2 * r0 + 1] = t0[1] + t1[1];
                out[2 * r1] = t2[0] + t3[0], out[2 * r1 + 1] = t2[1] + t3[1];
            }
        }
        break;
    }
    case SW_U1U:
        for (int r0 = 0; r0 < kRegs; r0++) {
            if (r0 & (1 << op.k0)) continue;
            const int r1 = r0 | (1 << op.k0);
            for (int row = 0; row < 2; row++) {
                const int c = (op.ta >> row) & 1;
                const double *pv = a + 2 * (c ? r1 : r0), *ov = a + 2 * (c ? r0 : r1);
                double t[2];
                cm(m + 2 * row, ov, t);
                const int dst = row ? r1 : r0;
                out[2 * dst] = pv[0] + t[0], out[2 * dst + 1] = pv[1] + t[1];
            }
        }
        break;
    case SW_SCALE:
        for (int r = 0; r < kRegs; r++) cm(m, a + 2 * r, out + 2 * r);
        break;
    case SW_U2:
    case SW_U2R:
        for (int b = 0; b < kRegs; b++) {
            if (b & ((1 << op.k0) | (1 << op.k1))) continue;
            const int idx[4] = {b, b | (1 << op.k1), b | (1 << op.k0), b | (1 << op.k0) | (1 << op.k1)};
            for (int r = 0; r < 4; r++) {
                double re = 0, im = 0;
                for (int c = 0; c < 4; c++) {
                    double t[2];
                    cm(m + 2 * (4 * r + c), a + 2 * idx[c], t);
                    re += t[0], im += t[1];
                }
                out[2 * idx[r]] = re, out[2 * idx[r] + 1] = im;
            }
        }
        break;
    case SW_SWAP:
        for (int r = 0; r < kRegs; r++) {
            const int b0 = (r >> op.k0) & 1, b1 = (r >> op.k1) & 1;
            const int s = (r & ~((1 << op.k0) | (1 << op.k1))) | (b1 << op.k0) | (b0 << op.k1);
            out[2 * r] = a[2 * s], out[2 * r + 1] = a[2 * s + 1];
        }
        break;
    case SW_D4:
    case SW_D4R:
        for (int r = 0; r < kRegs; r++) {
            double re = 0, im = 0;
            for (int c = 0; c < kRegs; c++) {
                double t[2];
                cm(m + 2 * (r * kRegs + c), a + 2 * c, t);
                re += t[0], im += t[1];
            }
            out[2 * r] = re, out[2 * r + 1] = im;
        }
        break;
    case SW_DIAG16:
        for (int r = 0; r < kRegs; r++) cm(m + 2 * r, a + 2 * r, out + 2 * r);
        break;
    case SW_DIAGL: {
        if (!lane) break;
        const int bl = ((tid & diag_tb(op)) != 0u || (op.ob >= 0 && ((cb >> op.ob) & 1ULL))) ? 1 : 0;
        const int al = ((tid & diag_ta(op)) != 0u || (op.oa >= 0 && ((cb >> op.oa) & 1ULL))) ? 1 : 0;
        for (int r = 0; r < kRegs; r++) {
            if (((uint32_t)r & op.rctrl) != op.rctrl) continue;
            const int A = op.k0 >= 0 ? (r >> op.k0) & 1 : al;
            const int B = op.k1 >= 0 ? (r >> op.k1) & 1 : bl;
            cm(m + 2 * (2 * A + B), a + 2 * r, out + 2 * r);
        }
        break;
    }
    default: break;
    }
    std::memcpy(a, out, sizeof(out));
}

// Qubits an op needs as register bits of its stage (local = the pass's local qubits): its targets and, for
// a non-diagonal controlled op whose qubits fit in the register bits, its local controls too -- it then
// needs no per-lane condition and can fuse with its neighbours.  Chunk-fixed controls are uniform per
// workgroup and stay outside; diagonal factors and controls of diagonal ops may come from thread-index bits
// at little cost.
uint64_t reg_need(const PlanOp &op, uint64_t local) {
    if (op.kind == kPlanGrad) return op.targets | op.diag;  // probes read their generator's qubits per register
    if (op.is_diag || op.controls == 0) return op.targets;
    const uint64_t both = op.targets | (op.controls & local);
    return popc(both) <= kRegBits ? both : op.targets;
}

// FP64 operations per amplitude of a register op (complex multiply-add = 4); exchanges do no arithmetic.
double op_fp64(const SweepOp &op) {
    switch (op.kind) {
    case SW_U1: return 8.0;
    case SW_U1R: return 4.0;
    case SW_U1C: return 8.0;
    case SW_U2: return 16.0;
    case SW_U2R: return 8.0;
    case SW_D4: return 64.0;
    case SW_D4R: return 32.0;
    case SW_DIAG16: return 4.0;
    case SW_DIAGL: return 4.0;
    case SW_U1U: return op.tb == 0 ? 4.0 : 2.0;
    case SW_SCALE: return 2.0;
    default: return 0.0;
    }
}

// Planning cost of a register op in VALU operations per amplitude.  In the compiled sweep kernel every op
// also pays a fixed issue cost that fusing ops saves (scalar loads of its record and matrix and about a
// dozen uniform dispatch branches), and exchanges with register controls or thread-index conditions are
// per-lane selects.  In a runtime-compiled pass kernel the op is a constant: exchanges under register
// controls are register renames and the issue cost is little more than the matrix's scalar loads.
double op_cost(const SweepOp &op, const SweepCostModel &cm) {
    double c = op_fp64(op);
    switch (op.kind) {
    case SW_U1C: c += op.tctrl ? 1.0 : 0.0; break;
    case SW_X: c = (cm.const_ops || !op.rctrl) ? 0.0 : 2.0; break;
    case SW_XL: c = 2.0; break;
    case SW_CX:
    case SW_SWAP: c = cm.const_ops ? 0.0 : 1.0; break;
    case SW_DIAGL: c += op.rctrl ? 2.0 : 0.0; break;
    default: break;
    }
    return c + cm.op_overhead;
}

// A fused segment: a diagonal over the register index (any number of bits) or a dense block on its bits.
double fused_cost(int bits, bool real, bool diag, const SweepCostModel &cm) {
    if (diag) return 4.0 + cm.op_overhead;
    double c = bits <= 1 ? 8.0 : bits == 2 ? 16.0 : 64.0;
    return (real ? 0.5 * c : c) + cm.op_overhead;
}

// Qubits an op needs among a pass's local qubits: its targets (and a probe's generator qubits).
uint64_t local_need(const PlanOp &op) { return op.kind == kPlanGrad ? (op.targets | op.diag) : op.targets; }

// FP64 operations per amplitude an op is expected to cost after fusion: diagonals share tables (a
// quarter of one each), exchanges are free, real matrices cost half.
double plan_op_fp64(const PlanOp &op) {
    if (op.kind == kPlanGrad) return 4.0;
    if (op.is_x || op.kind == SC_OP_CNOT || op.kind == SC_OP_SWAP) return 0.0;
    if (op.is_diag || op.kind == SC_OP_DIAG2) return 1.0;
    if (op.kind == SC_OP_TWO) return op.is_real ? 8.0 : 16.0;
    if (op.unit >= 0 && op.kind == SC_OP_SINGLE && !op.is_x && !op.is_diag) return op.unit_form == 0 ? 4.0 : 2.0;
    return op.is_real ? 4.0 : 8.0;
}

// Register bits a register-internal op reads or writes.
uint32_t reg_support(const SweepOp &op) {
    switch (op.kind) {
    case SW_U1:
    case SW_U1R:
    case SW_U1C:
    case SW_X: return (1u << op.k0) | op.rctrl;
    case SW_U1U: return 1u << op.k0;
    case SW_CX:
    case SW_U2:
    case SW_U2R:
    case SW_SWAP: return (1u << op.k0) | (1u << op.k1);
    case SW_DIAGL: return (op.k0 >= 0 ? 1u << op.k0 : 0u) | (op.k1 >= 0 ? 1u << op.k1 : 0u) | op.rctrl;
    default: return (1u << kRegBits) - 1u;
    }
}

bool is_diag_kind(const SweepOp &op) { return op.kind == SW_DIAGL || op.kind == SW_DIAG16; }

// Register bits whose values an op changes (it acts non-diagonally on them): its targets.
uint32_t nondiag_bits(const SweepOp &op) {
    switch (op.kind) {
    case SW_U1:
    case SW_U1R:
    case SW_U1C:
    case SW_X:
    case SW_XL:
    case SW_U1U:
    case SW_CX: return 1u << op.k0;
    case SW_U2:
    case SW_U2R:
    case SW_SWAP: return (1u << op.k0) | (1u << op.k1);
    case SW_DIAGL:
    case SW_DIAG16: return 0u;
    // a probe of <lambda| G |psi> sees the same value before and after a diagonal applied to both states when
    // G is diagonal (Z, ZZ) or acts on other bits (X, Y on k0)
    case SW_GRAD: return (op.ta == SC_GEN_X || op.ta == SC_GEN_Y) ? 1u << op.k0 : 0u;
    default: return (1u << kRegBits) - 1u;
    }
}

// Ops that only read and write register bits (no thread-index or chunk-fixed dependence): they can be folded
// into a fused block or diagonal table.
bool register_internal(const SweepOp &op) {
    if (op.tctrl != 0u || op.octrl != 0ULL) return false;
    if (op.kind == SW_DIAGL) return diag_ta(op) == 0u && diag_tb(op) == 0u && op.oa < 0 && op.ob < 0 && op.k0 >= 0;
    return op.kind == SW_U1 || op.kind == SW_U1R || op.kind == SW_U1C || op.kind == SW_X || op.kind == SW_CX ||
           op.kind == SW_U2 || op.kind == SW_U2R || op.kind == SW_SWAP || op.kind == SW_U1U;
}

}  // namespace

double sweep_op_fp64(const SweepOp &op) { return op_fp64(op); }

// Least-squares fit (relative error) of 97 runtime-compiled passes timed by rocprofv3 --kernel-trace on
// MI355X: the 30-qubit random layered circuit (depth 20) and the 33-qubit Grover circuit (8 iterations), each
// planned with 12 chunk bits / 4 low qubits, 13 / 4 and 13 / 3 (every pass but the first of each run).  Relative
// RMS error 7% per pass; per run the predicted sums are within 6% of the measured ones.  On the Grover circuit
// they rank the geometries as measured (13 / 3 fastest, 22% below 12 / 4); on the layered circuit, where 12 / 4
// and 13 / 4 measure within 1% of each other, they put 13 / 3 last, as measured.
// A pass runs at the slower of the HBM floor of its chunk size (a 13-bit chunk keeps one 512-thread
// workgroup per CU, which hides less latency than two 256-thread ones) and its arithmetic and LDS work.
double predicted_pass_ns_per_amp(const SweepPlan &plan, const SweepPass &pass) {
    constexpr double kFloor12 = 6.014e-3, kFloor13 = 6.601e-3;  // ns per amplitude
    constexpr double kBase = 2.550e-3, kPerFp64 = 3.560e-5, kPerRoundTrip = 4.597e-4;
    double f = 0.0;
    int in_reg = 0;
    for (int s = 0; s < pass.nstages; s++) {
        for (int o = pass.stages[s].op_begin; o < pass.stages[s].op_end; o++) f += op_fp64(plan.ops[o]);
        // an in-register transition: 32 lane exchanges per register bit and lane (2 per amplitude)
        const int pm = pass.stages[s].perm;
        if (pm) in_reg++, f += 2.0 * ((pm & 7) != 0) + 2.0 * (((pm >> 3) & 7) != 0);
    }
    // LDS round trips: one before each stage and the store phase's, less those the direct HBM paths and the
    // in-register transitions skip
    const int rt = pass.nstages + 1 - (pass.direct & 1) - ((pass.direct >> 1) & 1) - in_reg;
    const double floor = plan.chunk_bits == kMaxChunkBits ? kFloor13 : kFloor12;
    return std::max(floor, kBase + kPerFp64 * f + kPerRoundTrip * rt);
}

double predicted_plan_ns_per_amp(const SweepPlan &plan) {
    double t = 0.0;
    for (const SweepPass &p : plan.passes) t += predicted_pass_ns_per_amp(plan, p);
    return t;
}

int plan_sweeps_geometry(int n, const std::vector<PlanOp> &ops, int low_qubits, int chunk_bits, bool remap,
                         bool compiled, SweepPlan &plan, int max_fuse_bits, const SweepCostModel &cm,
                         GeometryPick *pick) {
    // candidates in order of preference on equal predictions: the measured geometries
    static const int kCandidates[3][2] = {{kChunkBits, 4}, {kMaxChunkBits, 4}, {kMaxChunkBits, 3}};
    std::vector<std::pair<int, int>> cand;
    if (!compiled) {
        cand.push_back({kChunkBits, low_qubits < 0 ? kDefaultLowQubits : low_qubits});
    } else if (chunk_bits > 0 && low_qubits >= 0) {
        cand.push_back({chunk_bits, low_qubits});
    } else {
        for (const auto &c : kCandidates) {
            int cb = chunk_bits > 0 ? chunk_bits : c[0];
            int lq = low_qubits >= 0 ? low_qubits : c[1];
            if (cb == kMaxChunkBits && n < kMaxChunkBits) continue;
            if (chunk_bits > 0 && cb != c[0]) continue;  // fixed chunk: its own measured low counts
            if (std::find(cand.begin(), cand.end(), std::make_pair(cb, lq)) == cand.end()) cand.push_back({cb, lq});
        }
        if (cand.empty()) cand.push_back({chunk_bits > 0 ? chunk_bits : kChunkBits,
                                          low_qubits >= 0 ? low_qubits : kDefaultLowQubits});
    }
    // A later candidate replaces the pick only when predicted at least kMargin faster: closer predictions are
    // within the model's error (6% per run), and the earlier geometry (two workgroups per CU first) is kept.
    constexpr double kMargin = 0.03;
    double best = 0.0;
    int best_i = -1;
    for (size_t i = 0; i < cand.size(); i++) {
        SweepPlan p;
        int rc = plan_sweeps(n, ops, cand[i].second, remap, p, max_fuse_bits, cm, cand[i].first);
        if (rc) return rc;
        const double t = predicted_plan_ns_per_amp(p);
        if (best_i < 0 || t < best * (1.0 - kMargin)) {
            best = t;
            best_i = (int)i;
            plan = std::move(p);
        }
    }
    if (pick) {
        pick->chunk_bits = cand[best_i].first;
        pick->low_qubits = cand[best_i].second;
        pick->ns_per_amp = best;
        pick->candidates = (int)cand.size();
    }
    return SC_OK;
}

int plan_sweeps(int n, const std::vector<PlanOp> &ops, int low_qubits, bool remap, SweepPlan &plan,
                int max_fuse_bits, const SweepCostModel &cm, int chunk_bits) {
    plan = SweepPlan();
    if (chunk_bits != kChunkBits && chunk_bits != kMaxChunkBits)
        return fail(SC_E_ARG, "sweep chunks hold 12 or 13 qubits");
    const int CB = chunk_bits, TB = CB - kRegBits;  // local qubits per chunk, thread-index bits
    plan.chunk_bits = CB;
    if (n < CB) return fail(SC_E_ARG, "sweep plans need at least " + std::to_string(CB) + " qubits");
    const uint64_t full = (n == 64) ? ~0ULL : ((1ULL << n) - 1ULL);
    const int lb = std::max(0, std::min(low_qubits, CB));
    const uint64_t lowmask = (1ULL << lb) - 1ULL;

    std::map<std::tuple<int, int, int>, int> recipe_index;  // (kind, mat, aux) -> recipe
    auto recipe_for = [&](int kind, int mat, int count, int aux = 0) -> int {
        auto key = std::make_tuple(kind, mat, aux);
        auto it = recipe_index.find(key);
        if (it != recipe_index.end()) return it->second;
        PoolRecipe r;
        r.kind = kind;
        r.mat = mat;
        r.count = count;
        r.aux = aux;
        plan.pool.push_back(r);
        int idx = (int)plan.pool.size() - 1;
        recipe_index[key] = idx;
        return idx;
    };
    std::vector<uint64_t> recipe_off;  // filled at the end

    std::vector<int> remaining(ops.size());
    for (size_t i = 0; i < ops.size(); i++) remaining[i] = (int)i;
    // Qubit placement: logical qubit q lives at physical index bit phys[q]; logi is the inverse.  Ops are in
    // logical qubits, kernels address physical bits.  A pass may permute the qubits of its chunk (the store
    // phase writes each local qubit to another of the chunk's positions, in place) so that the qubits the
    // next pass needs occupy the low, coalesced positions.  The placement is the identity again when the
    // plan ends (fix-up passes without ops restore it if needed).
    int phys[64], logi[64];
    for (int q = 0; q < 64; q++) phys[q] = logi[q] = q;
    auto identity = [&]() {
        for (int q = 0; q < n; q++)
            if (phys[q] != q) return false;
        return true;
    };
    // SweepOps hold recipe indices in .mat until offsets are known
    while (!remaining.empty() || !identity()) {
        const bool fixup = remaining.empty();
        // ---- choose the pass's op set and local (logical) qubits; the low physical positions are forced
        uint64_t L = 0;
        for (int p = 0; p < lb; p++) L |= 1ULL << logi[p];
        const uint64_t fixed = L;
        // A pass stops taking on arithmetic at the cost model's cap: the HBM sweep a pass costs hides only
        // so much FP64 work, and ops left over on qubits that stay local (the low positions) fill the
        // arithmetic headroom of later passes instead.
        // Ops a pass with local qubits LL takes on (program order; an op left behind blocks its qubits).
        auto take = [&](uint64_t LL, bool grow, uint64_t *L_out, std::vector<int> *s_out,
                        std::vector<int> *n_out) -> int {
            Blocker b;
            double wk = 0.0;
            int cnt = 0;
            for (int i : remaining) {
                const PlanOp &op = ops[i];
                const double w = plan_op_fp64(op);
                const uint64_t need = local_need(op);
                const bool fits = grow ? popc(LL | need) <= CB : (need & ~LL) == 0;
                if (b.blocked(op) || !fits || popc(op.targets) > kRegBits ||
                    (cm.pass_cap > 0.0 && w > 0.0 && wk + w > cm.pass_cap)) {
                    b.mark(op);
                    if (n_out) n_out->push_back(i);
                    continue;
                }
                wk += w;
                cnt++;
                if (grow) LL |= need;
                if (s_out) s_out->push_back(i);
            }
            if (L_out) *L_out = LL;
            return cnt;
        };
        // first guess: the targets of the earliest takeable ops, filled with the lowest positions
        take(L, true, &L, nullptr, nullptr);
        if (!fixup) {
            for (int p = 0; p < n && popc(L) < CB; p++) L |= 1ULL << logi[p];
            // hill climbing: exchange one free local qubit for an outer target qubit while that lets the
            // pass take on more ops (on layered circuits this halves the number of passes)
            uint64_t cand = 0;
            for (int i : remaining) cand |= local_need(ops[i]);
            int best = take(L, false, nullptr, nullptr, nullptr);
            for (int iter = 0; iter < 4 * CB; iter++) {
                bool improved = false;
                for (uint64_t fa = L & ~fixed; fa && !improved; fa &= fa - 1) {
                    const uint64_t a = fa & (~fa + 1ULL);
                    for (uint64_t fb = cand & ~L; fb; fb &= fb - 1) {
                        const uint64_t L2 = (L & ~a) | (fb & (~fb + 1ULL));
                        const int c = take(L2, false, nullptr, nullptr, nullptr);
                        if (c > best) {
                            best = c;
                            L = L2;
                            improved = true;
                            break;
                        }
                    }
                }
                if (!improved) break;
            }
        }
        std::vector<int> sched, next;
        take(L, false, nullptr, &sched, &next);
        if (!fixup && sched.empty()) return fail(SC_E_STATE, "internal error: a sweep pass took on no op");
        if (fixup) {
            // bring logical 0..lb-1 home, then misplaced qubits together with the occupant of their home
            for (int q = 0; q < lb; q++)
                if (popc(L | (1ULL << q)) <= CB) L |= 1ULL << q;
            for (int q = 0; q < n; q++) {
                if (phys[q] == q) continue;
                uint64_t add = (1ULL << q) | (1ULL << logi[q]);
                if (popc(L | add) <= CB) L |= add;
            }
        }
        // ---- local qubits: fill to 12 with the qubits at the lowest unused physical positions
        for (int p = 0; p < n && popc(L) < CB; p++) L |= 1ULL << logi[p];
        // ---- split into stages
        std::vector<StageBuild> stages;
        std::vector<int> rest = sched;
        // Ops a stage with register qubits RR takes on (program order, blocking as for passes); with grow the
        // register set grows with the ops' needs up to 4 qubits.
        auto stage_take = [&](uint64_t RR, bool grow, StageBuild *sb, std::vector<int> *rest_out) -> int {
            Blocker sblk;
            int cnt = 0;
            for (int i : rest) {
                const PlanOp &op = ops[i];
                const uint64_t need = reg_need(op, L);
                const bool fits = grow ? popc(RR | need) <= kRegBits : (need & ~RR) == 0;
                if (sblk.blocked(op) || !fits) {
                    sblk.mark(op);
                    if (rest_out) rest_out->push_back(i);
                    continue;
                }
                if (grow) RR |= need;
                cnt++;
                if (sb) sb->ops.push_back(i);
            }
            if (sb) sb->R = RR;
            return cnt;
        };
        while (!rest.empty() && (int)stages.size() < kMaxStages) {
            StageBuild sb;
            stage_take(0, true, &sb, nullptr);
            // hill climbing on the register qubits, as for the pass's local qubits: every exchange through
            // LDS between stages costs a read and a write of the whole chunk
            uint64_t R = sb.R, cand = 0;
            for (int i : rest) cand |= reg_need(ops[i], L);
            int best = stage_take(R, false, nullptr, nullptr);
            for (int iter = 0; iter < 16; iter++) {
                bool improved = false;
                for (uint64_t fq = cand & ~R; fq && !improved; fq &= fq - 1) {
                    const uint64_t q = fq & (~fq + 1ULL);
                    if (popc(R) < kRegBits) {
                        const int c = stage_take(R | q, false, nullptr, nullptr);
                        if (c > best) {
                            best = c;
                            R |= q;
                            improved = true;
                        }
                        continue;
                    }
                    for (uint64_t fr = R; fr; fr &= fr - 1) {
                        const uint64_t R2 = (R & ~(fr & (~fr + 1ULL))) | q;
                        const int c = stage_take(R2, false, nullptr, nullptr);
                        if (c > best) {
                            best = c;
                            R = R2;
                            improved = true;
                            break;
                        }
                    }
                }
                if (!improved) break;
            }
            StageBuild chosen;
            std::vector<int> rest2;
            stage_take(R, false, &chosen, &rest2);
            if (chosen.ops.empty()) return fail(SC_E_STATE, "internal error: a sweep stage took on no op");
            chosen.R = R;
            stages.push_back(chosen);
            rest.swap(rest2);
        }
        if (!rest.empty()) {  // truncated: return to the pool in program order
            std::vector<int> merged;
            std::merge(rest.begin(), rest.end(), next.begin(), next.end(), std::back_inserter(merged));
            next.swap(merged);
        }
        // ---- local qubits: fill to 12 with the qubits at the lowest unused physical positions
        for (int p = 0; p < n && popc(L) < CB; p++) L |= 1ULL << logi[p];
        if (popc(L) != CB) return fail(SC_E_STATE, "internal error: a sweep pass needs 12 local qubits");
        if (plan.passes.size() > 4 * ops.size() + 64) return fail(SC_E_STATE, "internal error: sweep planning stalled");
        int lq[kMaxChunkBits];  // local position j -> physical position (ascending)
        int lp[64];          // logical qubit -> local position, -1 when outer
        uint64_t Lphys = 0;
        for (int q = 0; q < 64; q++) lp[q] = -1;
        {
            int j = 0;
            for (int p = 0; p < n; p++)
                if ((L >> logi[p]) & 1ULL) {
                    lq[j] = p;
                    lp[logi[p]] = j;
                    Lphys |= 1ULL << p;
                    j++;
                }
        }
        SweepPass pass;
        std::memset(&pass, 0, sizeof(pass));
        pass.outer_mask = full & ~Lphys;
        pass.nstages = (int)stages.size();

        // ---- placement after the pass: out_q[k] is the logical qubit written to chunk position k.
        int out_q[kMaxChunkBits];
        {
            bool used[kMaxChunkBits] = {false};
            uint64_t placed = 0;
            std::vector<int> front;  // qubits wanted at the low positions, in priority order
            if (!remap) {
                for (int k = 0; k < CB; k++) {
                    out_q[k] = logi[lq[k]];
                    used[k] = true;
                    placed |= 1ULL << out_q[k];
                }
            } else if (fixup || next.empty()) {
                // restore: every local qubit whose home is one of the chunk's positions goes home
                for (int k = 0; k < CB; k++)
                    if ((L >> lq[k]) & 1ULL) {
                        out_q[k] = lq[k];
                        used[k] = true;
                        placed |= 1ULL << lq[k];
                    }
            } else {
                // lookahead: the targets of the next pass, in order of first use
                Blocker b2;
                uint64_t T = 0;
                for (int i : next) {
                    const PlanOp &op = ops[i];
                    if (b2.blocked(op) || popc(T | op.targets) > CB) {
                        b2.mark(op);
                        continue;
                    }
                    for (int q : {op.q0, op.q1})
                        if (q >= 0 && ((op.targets >> q) & 1ULL) && !((T >> q) & 1ULL) && ((L >> q) & 1ULL))
                            front.push_back(q);
                    T |= op.targets;
                }
                for (int k = 0; k < lb && k < (int)front.size(); k++) {
                    out_q[k] = front[k];
                    used[k] = true;
                    placed |= 1ULL << front[k];
                }
                // everything else stays where it is when its position is free
                for (int k = 0; k < CB; k++) {
                    int q = logi[lq[k]];
                    if (!used[k] && !((placed >> q) & 1ULL)) {
                        out_q[k] = q;
                        used[k] = true;
                        placed |= 1ULL << q;
                    }
                }
            }
            // leftovers fill the remaining positions in order
            int k = 0;
            for (int j = 0; j < CB; j++) {
                int q = logi[lq[j]];
                if ((placed >> q) & 1ULL) continue;
                while (used[k]) k++;
                out_q[k] = q;
                used[k] = true;
                placed |= 1ULL << q;
            }
        }
        int pi[kMaxChunkBits];  // chunk position k reads local position pi[k]
        for (int k = 0; k < CB; k++) pi[k] = lp[out_q[k]];

        // per stage: register positions / thread positions
        struct StageLayout {
            int rpos[kRegBits];
            int tpos[kMaxThreadBits];
            bool wave_local;
        };
        std::vector<StageLayout> layouts(stages.size());
        std::vector<uint32_t> stage_rp(stages.size(), 0u);
        std::vector<int> stage_perm(stages.size(), 0);  // SweepStage::perm
        int pass_wave_local = 0;
        for (size_t s = 0; s < stages.size(); s++) {
            uint32_t rp = 0;
            for (int q = 0; q < n; q++)
                if ((stages[s].R >> q) & 1ULL) rp |= 1u << lp[q];
            // spare register bits go to the local qubits the stage's ops use as controls / diagonal sources
            // (most uses first): such ops then resolve per register index instead of per lane, and can fuse
            if (popc(rp) < kRegBits) {
                int uses[kMaxChunkBits] = {0};
                for (int oi : stages[s].ops)
                    for (int q = 0; q < n; q++)
                        if (((ops[oi].diag >> q) & 1ULL) && lp[q] >= 0) uses[lp[q]]++;
                while (popc(rp) < kRegBits) {
                    int best = -1;
                    for (int p = 0; p < CB; p++)
                        if (!((rp >> p) & 1u) && uses[p] > 0 && (best < 0 || uses[p] > uses[best])) best = p;
                    if (best < 0) break;
                    rp |= 1u << best;
                }
            }
            for (int p = CB - 1; p >= 0 && popc(rp) < kRegBits; p--)
                if (!((rp >> p) & 1u)) rp |= 1u << p;
            stage_rp[s] = rp;
        }
        // ---- wave-index bits: thread bits 6 and 7 select the wavefront.  When they sit on the same two
        // local positions before and after an exchange, every wavefront reads back only what it wrote and
        // the exchange needs no workgroup barrier.  A dynamic programme over the stages picks the ordered
        // pair (a, b) per stage (both outside the stage's register positions) with the fewest changes,
        // starting from the load phase's pair (6, 7) and ending at the store phase's (pi[6], pi[7]).
        // ---- direct HBM access at the pass's ends (runtime-compiled passes): when the chunk's four lowest
        // local positions are physical positions 0-3 and a first / last stage keeps them off its register
        // bits (and its wave-index bits), its layout puts them on lane bits 0-3, and HBM can be read into or
        // written from that stage's registers with 256-byte runs per 16 lanes: no LDS round trip there.
        bool direct_load = false, direct_store = false;
        {
            const int S = (int)stages.size();
            bool low_ok = cm.direct_io && S > 0 && CB >= 4;
            for (int k = 0; k < 4 && low_ok; k++) low_ok = lq[k] == k;
            bool pi_identity = true;
            for (int k = 0; k < CB; k++) pi_identity = pi_identity && pi[k] == k;
            direct_load = low_ok && (stage_rp[0] & 0xFu) == 0;
            direct_store = low_ok && pi_identity && (stage_rp[S - 1] & 0xFu) == 0;
        }
        {
            const int S = (int)stages.size();
            const int kPairs = CB * CB;
            const int INF = 1 << 28;
            std::vector<std::vector<int>> cost(S, std::vector<int>(kPairs, INF)), from(S, std::vector<int>(kPairs, -1));
            auto ok = [&](int s, int w) {
                const int a = w / CB, b = w % CB;
                if (a == b || ((stage_rp[s] >> a) & 1u) || ((stage_rp[s] >> b) & 1u)) return false;
                // a stage that touches HBM directly keeps the low positions on lane bits
                if (((s == 0 && direct_load) || (s == S - 1 && direct_store)) && (a < 4 || b < 4)) return false;
                return true;
            };
            const int w_load = 6 * CB + 7, w_store = pi[6] * CB + pi[7];
            for (int w = 0; w < kPairs; w++)
                if (S > 0 && ok(0, w)) cost[0][w] = direct_load ? 0 : (w != w_load);
            for (int s = 1; s < S; s++)
                for (int w = 0; w < kPairs; w++) {
                    if (!ok(s, w)) continue;
                    for (int v = 0; v < kPairs; v++) {
                        if (cost[s - 1][v] >= INF) continue;
                        const int c = cost[s - 1][v] + (v != w);
                        if (c < cost[s][w] || (c == cost[s][w] && v == w)) {
                            cost[s][w] = c;
                            from[s][w] = v;
                        }
                    }
                }
            std::vector<int> wsel(S, -1);
            if (S > 0) {
                int best = -1;
                auto end_cost = [&](int w) { return cost[S - 1][w] + (direct_store ? 0 : (w != w_store)); };
                for (int w = 0; w < kPairs; w++) {
                    if (cost[S - 1][w] >= INF) continue;
                    if (best < 0 || end_cost(w) < end_cost(best)) best = w;
                }
                for (int s = S - 1; s >= 0; s--) {
                    wsel[s] = best;
                    best = from[s][best];
                }
            }
            for (int s = 0; s < S; s++) {
                const uint32_t rp = stage_rp[s];
                const int a = wsel[s] / CB, b = wsel[s] % CB;
                int rj = 0, tj = 0;
                for (int p = 0; p < CB; p++) {
                    if ((rp >> p) & 1u) {
                        layouts[s].rpos[rj++] = p;
                    } else if (p != a && p != b) {
                        layouts[s].tpos[tj++] = p;
                        if (tj == 6) tj = 8;  // 512 threads: thread bit 8 is a third wave-index bit
                    }
                }
                layouts[s].tpos[6] = a;
                layouts[s].tpos[7] = b;
                // with 512 threads the exchange stays in the wavefront only if bit 8 keeps its position too:
                // such chunks always synchronise the workgroup
                layouts[s].wave_local = TB == 8 && wsel[s] == (s == 0 ? w_load : wsel[s - 1]);
            }
            // ---- in-register transitions (runtime-compiled passes, 256 threads).  A stage whose register
            // qubits differ from the previous stage's in at most two, the incoming ones on the previous
            // layout's lane bits 4 / 5 and every other lane and wave bit keeping its qubit, is entered by
            // exchanging register bits with those lane bits inside each wavefront (v_permlane32_swap /
            // v_permlane16_swap, one VALU instruction per dword pair): no LDS round trip and no barrier.  A
            // layout not fixed by such a transition moves the next stage's incoming qubits onto its lane bits
            // 4 / 5 when that makes the next transition one.
            if (cm.permlane && TB == 8 && S > 1) {
                std::vector<char> locked(S, 0);
                for (int s = 0; s + 1 < S; s++) {
                    if (wsel[s] != wsel[s + 1]) continue;
                    StageLayout &A = layouts[s];
                    uint32_t rpA = 0;
                    for (int j = 0; j < kRegBits; j++) rpA |= 1u << A.rpos[j];
                    const uint32_t in = stage_rp[s + 1] & ~rpA;
                    if (popc(in) > 2) continue;
                    // a stage that touches HBM directly keeps positions 0-3 on lane bits 0-3, and a transition
                    // keeps the lane bits 0-3 it finds
                    const bool needs_low = (s == 0 && direct_load) || (s + 1 == S - 1 && direct_store);
                    auto low_ok = [&](const StageLayout &L) {
                        for (int i = 0; i < 4; i++)
                            if (L.tpos[i] != i) return false;
                        return true;
                    };
                    if ((in & ~((1u << A.tpos[4]) | (1u << A.tpos[5]))) != 0) {
                        if (locked[s]) continue;
                        // re-order A's lane bits 0-5: the incoming qubits on 5 (and 4), the rest ascending
                        StageLayout C = A;
                        std::vector<int> rest, inc;
                        for (int i = 0; i < 6; i++) ((in >> A.tpos[i]) & 1u ? inc : rest).push_back(A.tpos[i]);
                        std::sort(rest.begin(), rest.end());
                        std::sort(inc.begin(), inc.end());
                        for (int i = 0; i < 4; i++) C.tpos[i] = rest[i];
                        if (inc.size() == 2) {
                            C.tpos[4] = inc[0];
                            C.tpos[5] = inc[1];
                        } else {
                            C.tpos[4] = rest[4];
                            C.tpos[5] = inc[0];
                        }
                        if (needs_low && !low_ok(C)) continue;
                        A = C;
                    } else if (needs_low && !low_ok(A)) {
                        continue;
                    }
                    // B: A with each outgoing register qubit exchanged with an incoming lane-bit qubit
                    StageLayout B = A;
                    int perm = 64;
                    std::vector<int> outj;
                    for (int j = 0; j < kRegBits; j++)
                        if (!((stage_rp[s + 1] >> A.rpos[j]) & 1u)) outj.push_back(j);
                    size_t k = 0;
                    for (int l : {5, 4}) {
                        if (!((in >> A.tpos[l]) & 1u)) continue;
                        const int j = outj[k++];
                        B.rpos[j] = A.tpos[l];
                        B.tpos[l] = A.rpos[j];
                        perm |= (j + 1) << (l == 5 ? 0 : 3);
                    }
                    B.wave_local = true;
                    layouts[s + 1] = B;
                    stage_perm[s + 1] = perm;
                    locked[s + 1] = 1;
                }
            }
            pass_wave_local = 0;
            if (TB == 8) {
                if (S > 0 ? wsel[S - 1] == w_store : w_load == w_store) pass_wave_local |= 1;
                // a chunk's last LDS read and the next chunk's first LDS write use the same wave-index
                // positions (or one of them does not exist): no wavefront overwrites slots another still reads
                const int first_w = !direct_load ? w_load : (S > 1 || !direct_store) ? wsel[0] : -1;
                const int last_r = !direct_store ? w_store : (S > 1 || !direct_load) ? wsel[S - 1] : -1;
                if (first_w < 0 || last_r < 0 || first_w == last_r) pass_wave_local |= 2;
            }
            pass.wave_local = pass_wave_local;
        }
        // ---- swizzle: slot(l) = l ^ (XOR of hv[p] over the set bits p of l), hv[p] below bit min(p, kSwzBits)
        // (a unitriangular map: one slot per local index), chosen so every stage's lanes 0..3 land on distinct
        // slots mod 16 and its lanes 0..2 on distinct slots mod 8: a 16-byte LDS access then serves 8 (or 16)
        // lanes per cycle without bank conflicts.  Low positions take part too: a stage whose lane bits 0-2
        // include position 3 (slot bit 3, 0 mod 8) needs hv[3] to keep its 8-lane groups apart.
        uint32_t hv[kMaxChunkBits] = {0};
        {
            uint64_t rng = 0x243F6A8885A308D3ULL;
            auto next_rand = [&]() {
                rng ^= rng << 13;
                rng ^= rng >> 7;
                rng ^= rng << 17;
                return rng;
            };
            // lane bits 0..3 of every stage, then of the store phase
            std::vector<std::array<int, 4>> lanes;
            for (size_t s = 0; s <= stages.size(); s++) {
                std::array<int, 4> a;
                for (int i = 0; i < 4; i++) a[i] = s < stages.size() ? layouts[s].tpos[i] : pi[i];
                lanes.push_back(a);
            }
            // columns independent over GF(2): no nonempty subset XORs to zero (reachable sums as a 16-bit set)
            auto independent = [](const uint32_t *c, int k) {
                uint32_t reach = 1u;  // the empty sum
                for (int i = 0; i < k; i++) {
                    uint32_t moved = 0;
                    for (uint32_t v = 0; v < 16; v++)
                        if ((reach >> v) & 1u) moved |= 1u << (v ^ c[i]);
                    if (moved & reach) return false;
                    reach |= moved;
                }
                return true;
            };
            int best_bad = 1 << 30;
            uint32_t best[kMaxChunkBits] = {0};
            for (int attempt = 0; attempt < 100000 && best_bad > 0; attempt++) {
                uint32_t cand[kMaxChunkBits] = {0};
                for (int p = 1; p < CB; p++)
                    cand[p] = attempt == 0 ? 0u : (uint32_t)(next_rand() & ((1u << std::min(p, kSwzBits)) - 1u));
                int bad = 0;
                for (const auto &a : lanes) {
                    uint32_t c16[4], c8[3];
                    for (int i = 0; i < 4; i++) {
                        c16[i] = ((a[i] < kSwzBits ? (1u << a[i]) : 0u) ^ cand[a[i]]) & 15u;
                        if (i < 3) c8[i] = c16[i] & 7u;
                    }
                    if (!independent(c16, 4)) bad += 2;
                    if (!independent(c8, 3)) bad += 1;
                    if (bad >= best_bad) break;
                }
                if (bad < best_bad) {
                    best_bad = bad;
                    std::memcpy(best, cand, sizeof(best));
                }
            }
            std::memcpy(hv, best, sizeof(hv));
        }
        auto slot_of = [&](uint32_t l) -> uint32_t {
            uint32_t h = 0;
            for (int p = 1; p < CB; p++)
                if ((l >> p) & 1u) h ^= hv[p];
            return l ^ h;
        };
        // load / store phase layout: thread bit i -> local i, register bit j -> local 8 + j
        for (int i = 0; i < TB; i++) {
            pass.glob_thread[i] = 1ULL << lq[i];
            pass.io_thread_slot[i] = slot_of(1u << i);
        }
        for (int r = 0; r < kRegs; r++) {
            uint64_t g = 0;
            uint32_t l = 0;
            for (int j = 0; j < kRegBits; j++)
                if ((r >> j) & 1) {
                    g |= 1ULL << lq[TB + j];
                    l |= 1u << (TB + j);
                }
            pass.glob_reg[r] = g;
            pass.io_reg_slot[r] = slot_of(l);
        }
        // direct loads / stores: global offsets of the first / last stage's thread and register bits
        pass.direct = (direct_load ? 1 : 0) | (direct_store ? 2 : 0);
        auto stage_offsets = [&](const StageLayout &lay, uint64_t *th, uint64_t *rg) {
            for (int i = 0; i < TB; i++) th[i] = 1ULL << lq[lay.tpos[i]];
            for (int r = 0; r < kRegs; r++) {
                uint64_t g = 0;
                for (int j = 0; j < kRegBits; j++)
                    if ((r >> j) & 1) g |= 1ULL << lq[lay.rpos[j]];
                rg[r] = g;
            }
        };
        if (direct_load) stage_offsets(layouts[0], pass.ld_thread, pass.ld_reg);
        if (direct_store) stage_offsets(layouts[stages.size() - 1], pass.sd_thread, pass.sd_reg);
        // store phase: chunk position k takes the value of local position pi[k]
        for (int i = 0; i < TB; i++) pass.st_thread_slot[i] = slot_of(1u << pi[i]);
        for (int r = 0; r < kRegs; r++) {
            uint32_t l = 0;
            for (int j = 0; j < kRegBits; j++)
                if ((r >> j) & 1) l |= 1u << pi[TB + j];
            pass.st_reg_slot[r] = slot_of(l);
        }
        // ---- stage ops
        std::vector<std::pair<int, int>> pass_scale;  // pivots of the pass's unit-pivot ops: (mat, column)
        for (size_t s = 0; s < stages.size(); s++) {
            SweepStage &stg = pass.stages[s];
            const StageLayout &lay = layouts[s];
            stg.wave_local = lay.wave_local ? 1 : 0;
            stg.perm = stage_perm[s];
            for (int i = 0; i < TB; i++) stg.thread_slot[i] = slot_of(1u << lay.tpos[i]);
            for (int r = 0; r < kRegs; r++) {
                uint32_t l = 0;
                for (int j = 0; j < kRegBits; j++)
                    if ((r >> j) & 1) l |= 1u << lay.rpos[j];
                stg.lds_off[r] = slot_of(l);
            }
            auto reg_of = [&](int q) -> int {
                int p = lp[q];
                for (int j = 0; j < kRegBits; j++)
                    if (lay.rpos[j] == p) return j;
                return -1;
            };
            // where a qubit's index bit comes from in this stage: a register bit, a thread-index bit or an
            // outer (chunk-fixed) qubit at its physical position
            struct Src {
                int reg = -1;     // register bit
                uint32_t t = 0;   // thread-index mask
                int opos = -1;    // outer qubit's global position
            };
            auto src_of = [&](int q) -> Src {
                Src b;
                if (lp[q] >= 0) {
                    b.reg = reg_of(q);
                    if (b.reg < 0)
                        for (int i = 0; i < TB; i++)
                            if (lay.tpos[i] == lp[q]) b.t = 1u << i;
                } else {
                    b.opos = phys[q];
                }
                return b;
            };
            auto add_controls = [&](SweepOp &so, uint64_t controls) {
                for (int q = 0; q < n; q++) {
                    if (!((controls >> q) & 1ULL)) continue;
                    const Src b = src_of(q);
                    if (b.reg >= 0) so.rctrl |= (uint16_t)(1u << b.reg);
                    so.tctrl |= (uint16_t)b.t;
                    if (b.opos >= 0) so.octrl |= 1ULL << b.opos;
                }
            };
            // an exchange whose only control is one register bit: a CNOT inside the registers
            auto set_exchange = [&](SweepOp &so, int target) {
                so.k0 = (int8_t)reg_of(target);
                if (so.tctrl) {
                    so.kind = SW_XL;
                } else if (so.octrl == 0 && popc(so.rctrl) == 1) {
                    so.kind = SW_CX;
                    so.k1 = (int8_t)__builtin_ctz(so.rctrl);
                    so.rctrl = 0;
                } else {
                    so.kind = SW_X;
                }
            };
            auto new_op = []() {
                SweepOp so;
                std::memset(&so, 0, sizeof(so));
                so.mat = -1;
                so.k0 = so.k1 = -1;
                so.oa = so.ob = -1;
                return so;
            };
            std::vector<SweepOp> sops;
            std::vector<bool> sreal;  // op's matrix is real
            for (int oi : stages[s].ops) {
                const PlanOp &op = ops[oi];
                if (op.skip) continue;  // part of the product-state input
                SweepOp so = new_op();
                switch (op.kind) {
                case kPlanGrad:
                    so.kind = SW_GRAD;
                    so.k0 = (int8_t)reg_of(op.q0);
                    so.k1 = (int8_t)(op.q1 >= 0 ? reg_of(op.q1) : -1);
                    so.ta = (uint8_t)op.gen;
                    so.slot = (uint64_t)op.mat;
                    break;
                case SC_OP_SINGLE:
                case SC_OP_MCU:
                    add_controls(so, op.controls);
                    if (op.is_diag) {
                        // entries (g00, g00, g11, g11): factor m[2A + 0] with A the target's bit
                        const Src a = src_of(op.q0);
                        so.kind = SW_DIAGL;
                        so.k0 = (int8_t)a.reg;
                        so.ta = (uint8_t)a.t;
                        so.thi = (uint8_t)((a.t >> 8) & 1u);
                        so.oa = (int8_t)a.opos;
                        so.mat = recipe_for(RC_DIAG1, op.mat, 4);
                    } else if (op.is_x) {
                        set_exchange(so, op.q0);
                    } else if (cm.unit_1q && op.unit >= 0 && op.kind == SC_OP_SINGLE) {
                        so.kind = SW_U1U;
                        so.k0 = (int8_t)reg_of(op.q0);
                        so.ta = (uint8_t)op.unit;
                        so.tb = (uint8_t)op.unit_form;
                        so.mat = recipe_for(RC_UNIT, op.mat, 2, op.unit);
                        pass_scale.push_back({op.mat, op.unit & 1});
                    } else {
                        so.kind = (so.rctrl | so.tctrl) ? SW_U1C : (op.is_real ? SW_U1R : SW_U1);
                        so.k0 = (int8_t)reg_of(op.q0);
                        so.mat = recipe_for(RC_COPY, op.mat, 4);
                    }
                    break;
                case SC_OP_CNOT:
                    add_controls(so, 1ULL << op.q0);
                    set_exchange(so, op.q1);
                    break;
                case SC_OP_TWO: {
                    // kernels take the more significant register bit first; the other order transposes the
                    // matrix's qubit order
                    const int r0 = reg_of(op.q0), r1 = reg_of(op.q1);
                    so.kind = op.is_real ? SW_U2R : SW_U2;
                    so.k0 = (int8_t)std::max(r0, r1);
                    so.k1 = (int8_t)std::min(r0, r1);
                    so.mat = recipe_for(r0 > r1 ? RC_COPY : RC_SWAP4, op.mat, 16);
                    break;
                }
                case SC_OP_DIAG2: {
                    // factor m[2A + B]: when only the second qubit is a register bit the table is transposed so
                    // that A is the register source
                    Src a = src_of(op.q0), b = src_of(op.q1);
                    int rk = RC_COPY;
                    if (a.reg < 0 && b.reg >= 0) {
                        std::swap(a, b);
                        rk = RC_SWAPDIAG;
                    }
                    so.kind = SW_DIAGL;
                    so.k0 = (int8_t)a.reg;
                    so.ta = (uint8_t)a.t;
                    so.oa = (int8_t)a.opos;
                    so.k1 = (int8_t)b.reg;
                    so.tb = (uint8_t)b.t;
                    so.ob = (int8_t)b.opos;
                    so.thi = (uint8_t)(((a.t >> 8) & 1u) | (((b.t >> 8) & 1u) << 1));
                    so.mat = recipe_for(rk, op.mat, 4);
                    break;
                }
                case SC_OP_SWAP: {
                    const int r0 = reg_of(op.q0), r1 = reg_of(op.q1);
                    so.kind = SW_SWAP;
                    so.k0 = (int8_t)std::max(r0, r1);
                    so.k1 = (int8_t)std::min(r0, r1);
                    break;
                }
                }
                sops.push_back(so);
                sreal.push_back(so.kind == SW_U1U ? op.unit_form == 1
                                                  : (op.is_real || op.kind == SC_OP_CNOT || op.kind == SC_OP_SWAP ||
                                                     op.is_x));
            }
            // ---- gather diagonals: a register diagonal moves back to the previous one when every op in
            // between leaves its register bits' values alone (acts non-diagonally only on other bits) ...
            auto reg_diag = [&](size_t k) { return register_internal(sops[k]) && is_diag_kind(sops[k]); };
            for (size_t i = 1; i < sops.size(); i++) {
                if (!reg_diag(i)) continue;
                const uint32_t dep = reg_support(sops[i]);
                size_t j = i;
                while (j > 0 && !reg_diag(j - 1) && (nondiag_bits(sops[j - 1]) & dep) == 0u) j--;
                if (j == 0 || j == i || !reg_diag(j - 1)) continue;
                const SweepOp d = sops[i];
                const bool dr = sreal[i];
                for (size_t t = i; t > j; t--) {
                    sops[t] = sops[t - 1];
                    sreal[t] = sreal[t - 1];
                }
                sops[j] = d;
                sreal[j] = dr;
            }
            // ... and forward to the next one, so runs such as Ry Rz Ry Rz .. become Ry Ry .. Rz Rz .., whose
            // diagonals fuse into one table.  Processing from the back lets whole groups form.
            for (size_t i = sops.size(); i-- > 0;) {
                if (!register_internal(sops[i]) || !is_diag_kind(sops[i])) continue;
                const uint32_t dep = reg_support(sops[i]);
                size_t e = i + 1;
                while (e < sops.size() && !(register_internal(sops[e]) && is_diag_kind(sops[e])) &&
                       (nondiag_bits(sops[e]) & dep) == 0u)
                    e++;
                if (e == sops.size() || e == i + 1 || !(register_internal(sops[e]) && is_diag_kind(sops[e])))
                    continue;
                const SweepOp d = sops[i];
                const bool dr = sreal[i];
                for (size_t t = i; t + 1 < e; t++) {
                    sops[t] = sops[t + 1];
                    sreal[t] = sreal[t + 1];
                }
                sops[e - 1] = d;
                sreal[e - 1] = dr;
            }
            // ---- fusion: split the stage's op list into consecutive segments, each either left as is or
            // (when every op in it stays inside the register bits) replaced by one op: a diagonal table over
            // the register index when every op is diagonal (4 VALU ops per amplitude), else a dense block on
            // the segment's register bits -- 2x2 (8), 4x4 (16) or 16x16 (64), half that when real -- choosing
            // the segmentation with the lowest cost (dynamic programming over segment ends).  A diagonal on
            // two register bits has no kernel form of its own and always becomes a table.
            const size_t L = sops.size();
            std::vector<double> best(L + 1, 0.0);
            std::vector<int> from(L + 1, -1);  // segment start, or -1 for "op left as is"
            for (size_t e = 1; e <= L; e++) {
                const SweepOp &last = sops[e - 1];
                const bool table = last.kind == SW_DIAGL && last.k1 >= 0;
                best[e] = best[e - 1] + (table ? fused_cost(0, false, true, cm) : op_cost(last, cm));
                from[e] = table ? (int)(e - 1) : -1;
                uint32_t sup = 0;
                bool real = true, diag = true;
                for (size_t s0 = e; s0-- > 0;) {
                    if (!register_internal(sops[s0])) break;
                    sup |= reg_support(sops[s0]);
                    real = real && sreal[s0];
                    diag = diag && is_diag_kind(sops[s0]);
                    if (e - s0 < 2 || max_fuse_bits <= 0 || (!diag && popc(sup) > max_fuse_bits)) continue;
                    double c = best[s0] + fused_cost(popc(sup), real, diag, cm);
                    if (c < best[e] - 1e-9) {
                        best[e] = c;
                        from[e] = (int)s0;
                    }
                }
            }
            std::vector<std::pair<size_t, size_t>> segs;  // [begin, end) of fused segments
            std::vector<bool> seg_start(L + 1, false);
            for (size_t e = L; e > 0;) {
                if (from[e] < 0) {
                    e--;
                } else {
                    segs.push_back({(size_t)from[e], e});
                    e = (size_t)from[e];
                }
            }
            std::sort(segs.begin(), segs.end());
            std::vector<SweepOp> fused;
            size_t si = 0;
            for (size_t i = 0; i < L;) {
                if (si < segs.size() && segs[si].first == i) {
                    size_t j = segs[si].second;
                    uint32_t sup = 0;
                    bool real = true, diag = true;
                    for (size_t t = i; t < j; t++) {
                        sup |= reg_support(sops[t]);
                        real = real && sreal[t];
                        diag = diag && is_diag_kind(sops[t]);
                    }
                    PoolRecipe r;
                    r.mat = -1;
                    SweepOp d = new_op();
                    if (diag) {
                        r.kind = RC_DIAG16;
                        r.count = kRegs;
                        d.kind = SW_DIAG16;
                    } else {
                        r.kind = RC_FUSED;
                        std::vector<int> bits;
                        for (int b = kRegBits - 1; b >= 0; b--)
                            if ((sup >> b) & 1u) bits.push_back(b);
                        if (bits.size() == 1) {
                            d.kind = real ? SW_U1R : SW_U1;
                            d.k0 = bits[0];
                        } else if (bits.size() == 2) {
                            d.kind = real ? SW_U2R : SW_U2;
                            d.k0 = bits[0];
                            d.k1 = bits[1];
                        } else {
                            d.kind = real ? SW_D4R : SW_D4;
                            bits = {3, 2, 1, 0};
                            plan.fused_blocks++;
                        }
                        r.fused_bits = bits;
                        r.count = 1 << (2 * bits.size());
                    }
                    for (size_t t = i; t < j; t++) {
                        r.fused_ops.push_back(sops[t]);
                        r.fused_srcs.push_back(sops[t].mat);
                    }
                    plan.pool.push_back(r);
                    d.mat = (int)plan.pool.size() - 1;
                    fused.push_back(d);
                    i = j;
                    si++;
                } else {
                    fused.push_back(sops[i]);
                    i++;
                }
            }
            stg.op_begin = (int32_t)plan.ops.size();
            for (auto &so : fused) plan.ops.push_back(so);
            stg.op_end = (int32_t)plan.ops.size();
        }
        // ---- the unit-pivot ops' common factor: every amplitude of the chunk is multiplied by it once, in
        // the table of the first uncontrolled matrix op of the pass (a copy of its recipe, scaled), else by a
        // scaling op at the end of the last stage.  The factor is real and commutes with every op.
        bool pass_probes = false;
        for (int s = 0; s < pass.nstages; s++)
            for (int i = pass.stages[s].op_begin; i < pass.stages[s].op_end; i++)
                pass_probes = pass_probes || plan.ops[i].kind == SW_GRAD;
        if (!pass_scale.empty() && pass_probes) {
            // the states a probe sees lack the pivots of the unit-pivot ops before it (fused ones included)
            std::vector<std::pair<int, int>> done;
            auto unit_pivot = [&](const SweepOp &so) {
                if (so.kind == SW_U1U && so.mat >= 0)
                    done.push_back({plan.pool[so.mat].mat, plan.pool[so.mat].aux & 1});
            };
            for (int s = 0; s < pass.nstages; s++)
                for (int i = pass.stages[s].op_begin; i < pass.stages[s].op_end; i++) {
                    const SweepOp &so = plan.ops[i];
                    if (so.kind == SW_GRAD) {
                        plan.probe_pivots.push_back({so.slot, done});
                    } else if (so.mat >= 0 && plan.pool[so.mat].kind == RC_FUSED) {
                        for (const SweepOp &f : plan.pool[so.mat].fused_ops) unit_pivot(f);
                    } else {
                        unit_pivot(so);
                    }
                }
        }
        if (!pass_scale.empty()) {
            int target = -1;
            for (int s = 0; s < pass.nstages && target < 0 && !pass_probes; s++)
                for (int i = pass.stages[s].op_begin; i < pass.stages[s].op_end; i++) {
                    const SweepOp &so = plan.ops[i];
                    const bool full_matrix = so.kind == SW_U1 || so.kind == SW_U1R || so.kind == SW_U2 ||
                                             so.kind == SW_U2R || so.kind == SW_D4 || so.kind == SW_D4R ||
                                             so.kind == SW_DIAG16 || so.kind == SW_DIAGL;
                    if (full_matrix && so.mat >= 0 && so.rctrl == 0 && so.tctrl == 0 && so.octrl == 0) {
                        target = i;
                        break;
                    }
                }
            PoolRecipe r;
            if (target >= 0) {
                r = plan.pool[plan.ops[target].mat];
            } else {
                r.kind = RC_SCALE;
                r.mat = -1;
                r.count = 1;
                SweepOp so;
                std::memset(&so, 0, sizeof(so));
                so.kind = SW_SCALE;
                so.k0 = so.k1 = -1;
                so.oa = so.ob = -1;
                plan.ops.push_back(so);
                pass.stages[pass.nstages - 1].op_end = (int32_t)plan.ops.size();
                target = (int)plan.ops.size() - 1;
            }
            r.scale_srcs.insert(r.scale_srcs.end(), pass_scale.begin(), pass_scale.end());
            plan.pool.push_back(r);
            plan.ops[target].mat = (int)plan.pool.size() - 1;
        }
        plan.passes.push_back(pass);
        for (int k = 0; k < CB; k++) {
            phys[out_q[k]] = lq[k];
            logi[lq[k]] = out_q[k];
        }
        remaining.swap(next);
    }
    // ---- pool offsets: replace recipe indices by complex offsets
    recipe_off.resize(plan.pool.size());
    uint64_t off = 0;
    for (size_t r = 0; r < plan.pool.size(); r++) {
        recipe_off[r] = off;
        off += (uint64_t)plan.pool[r].count;
    }
    plan.pool_complex = off;
    for (auto &so : plan.ops) {
        // kernel contract: register bit positions in range, two-bit kinds ordered, one register diagonal source
        const bool two = so.kind == SW_U2 || so.kind == SW_U2R || so.kind == SW_SWAP;
        const bool one = so.kind == SW_U1 || so.kind == SW_U1R || so.kind == SW_U1C || so.kind == SW_X ||
                         so.kind == SW_XL || so.kind == SW_U1U;
        bool ok = so.kind >= SW_U1 && so.kind <= SW_SCALE;
        if (so.kind == SW_U1U) ok = ok && so.ta < 4 && so.tb <= 2 && so.mat >= 0;
        if (so.kind == SW_SCALE) ok = ok && so.mat >= 0;
        if (so.kind == SW_GRAD)
            ok = so.k0 >= 0 && so.k0 < kRegBits && so.k1 < kRegBits && so.k1 != so.k0 && so.mat < 0 &&
                 so.ta >= SC_GEN_X && so.ta <= SC_GEN_ZZ && ((so.ta == SC_GEN_ZZ) == (so.k1 >= 0));
        if (one) ok = ok && so.k0 >= 0 && so.k0 < kRegBits;
        if (two) ok = ok && so.k1 >= 0 && so.k0 > so.k1 && so.k0 < kRegBits;
        if (so.kind == SW_CX) ok = ok && so.k0 >= 0 && so.k0 < kRegBits && so.k1 >= 0 && so.k1 < kRegBits &&
                                    so.k0 != so.k1;
        if (so.kind == SW_DIAGL) ok = ok && so.k1 < 0 && so.k0 < kRegBits;
        if (!ok) return fail(SC_E_STATE, "internal error: sweep plan produced an invalid register op");
        plan.fp64_ops += op_fp64(so);
    }
    for (auto &so : plan.ops)
        if (so.mat >= 0) so.mat = (int32_t)recipe_off[so.mat];
    // one record past the last op: the kernel reads the record after the current op ahead of time
    SweepOp sentinel;
    std::memset(&sentinel, 0, sizeof(sentinel));
    sentinel.mat = -1;
    plan.ops.push_back(sentinel);
    // fused recipes keep recipe indices in fused_srcs; record their offsets in the SweepOps' .mat for build_pool
    for (auto &r : plan.pool)
        for (size_t t = 0; t < r.fused_ops.size(); t++) r.fused_ops[t].mat = r.fused_srcs[t];
    return SC_OK;
}

namespace {

// Mirror of the device run_op (sc_sweep.hip) on one thread's 16 registers; kernels only see ops whose
// diagonals have at most one register source.
void host_run_op(const SweepOp &op, const double *pool, double *a, uint32_t tid, uint64_t chunk_base) {
    const double *m = op.mat >= 0 ? pool + 2 * (int64_t)op.mat : nullptr;
    host_apply(op, m, a, tid, chunk_base);
}

// <lambda| G |psi> of a probe over one thread's register amplitudes (conj(lambda) * (G psi)).
void host_grad(const SweepOp &op, const double *psi, const double *lam, double *acc) {
    for (int r = 0; r < kRegs; r++) {
        const int b0 = (r >> op.k0) & 1;
        double pr, pi;
        switch (op.ta) {
        case SC_GEN_X: pr = psi[2 * (r ^ (1 << op.k0))], pi = psi[2 * (r ^ (1 << op.k0)) + 1]; break;
        case SC_GEN_Y: {  // Y psi: (-i psi1, i psi0)
            const double xr = psi[2 * (r ^ (1 << op.k0))], xi = psi[2 * (r ^ (1 << op.k0)) + 1];
            pr = b0 ? -xi : xi;
            pi = b0 ? xr : -xr;
            break;
        }
        case SC_GEN_Z: pr = b0 ? -psi[2 * r] : psi[2 * r], pi = b0 ? -psi[2 * r + 1] : psi[2 * r + 1]; break;
        default: {
            const bool neg = (b0 ^ ((r >> op.k1) & 1)) != 0;
            pr = neg ? -psi[2 * r] : psi[2 * r];
            pi = neg ? -psi[2 * r + 1] : psi[2 * r + 1];
        }
        }
        const double lr = lam[2 * r], li = lam[2 * r + 1];
        acc[0] += lr * pr + li * pi;
        acc[1] += lr * pi - li * pr;
    }
}

}  // namespace

void execute_plan_host(int n, const SweepPlan &plan, const double *pool, double *state, double *lam, double *grads) {
    const int CB = plan.chunk_bits, TB = CB - kRegBits;
    const uint32_t threads = 1u << TB;
    const int log_chunks = n - CB;
    const int ns = lam ? 2 : 1;
    double *states[2] = {state, lam};
    std::vector<double> lds[2] = {std::vector<double>(2ULL << CB), std::vector<double>(2ULL << CB)};
    double a[2][2 * kRegs];
    for (const SweepPass &pass : plan.passes) {
        for (uint64_t chunk = 0; chunk < (1ULL << log_chunks); chunk++) {
            uint64_t chunk_base = 0, om = pass.outer_mask, c = chunk;
            while (om) {
                uint64_t low = om & (~om + 1ULL);
                if (c & 1ULL) chunk_base |= low;
                c >>= 1;
                om &= om - 1ULL;
            }
            const bool dl = (pass.direct & 1) != 0, ds = (pass.direct & 2) != 0;
            const int S = pass.nstages;
            // a direct first stage reads from HBM and a direct last stage writes to HBM: each chunk reads and
            // writes only its own addresses, so the in-place update below is safe
            auto direct_offset = [&](const uint64_t *th, uint32_t tid) {
                uint64_t g = 0;
                for (int i = 0; i < TB; i++)
                    if ((tid >> i) & 1u) g += th[i];
                return g;
            };
            for (int k = 0; k < ns && !dl; k++)
                for (uint32_t tid = 0; tid < threads; tid++) {
                    uint64_t g = 0;
                    uint32_t s = 0;
                    for (int i = 0; i < TB; i++)
                        if ((tid >> i) & 1u) g += pass.glob_thread[i], s ^= pass.io_thread_slot[i];
                    for (int r = 0; r < kRegs; r++) {
                        uint64_t gi = chunk_base + g + pass.glob_reg[r];
                        uint32_t sl = s ^ pass.io_reg_slot[r];
                        lds[k][2 * sl] = states[k][2 * gi];
                        lds[k][2 * sl + 1] = states[k][2 * gi + 1];
                    }
                }
            // a stage's reads of every thread precede its writes (the kernel's barrier): buffer the writes of
            // a direct last stage until the stage is done
            std::vector<std::pair<uint64_t, double>> pending;
            for (int st = 0; st < S; st++) {
                const SweepStage &stg = pass.stages[st];
                for (uint32_t tid = 0; tid < threads; tid++) {
                    uint32_t ts = 0;
                    for (int i = 0; i < TB; i++)
                        if ((tid >> i) & 1u) ts ^= stg.thread_slot[i];
                    for (int k = 0; k < ns; k++)
                        for (int r = 0; r < kRegs; r++) {
                            if (st == 0 && dl) {
                                const uint64_t gi = chunk_base + direct_offset(pass.ld_thread, tid) + pass.ld_reg[r];
                                a[k][2 * r] = states[k][2 * gi], a[k][2 * r + 1] = states[k][2 * gi + 1];
                                continue;
                            }
                            uint32_t sl = ts ^ stg.lds_off[r];
                            a[k][2 * r] = lds[k][2 * sl], a[k][2 * r + 1] = lds[k][2 * sl + 1];
                        }
                    for (int o = stg.op_begin; o < stg.op_end; o++) {
                        const SweepOp &op = plan.ops[o];
                        if (op.kind == SW_GRAD) {
                            if (lam && grads) host_grad(op, a[0], a[1], grads + 2 * op.slot);
                            continue;
                        }
                        for (int k = 0; k < ns; k++) host_run_op(op, pool, a[k], tid, chunk_base);
                    }
                    for (int k = 0; k < ns; k++)
                        for (int r = 0; r < kRegs; r++) {
                            if (st == S - 1 && ds) {
                                const uint64_t gi = chunk_base + direct_offset(pass.sd_thread, tid) + pass.sd_reg[r];
                                pending.push_back({2 * gi, a[k][2 * r]});
                                pending.push_back({2 * gi + 1, a[k][2 * r + 1]});
                                continue;
                            }
                            uint32_t sl = ts ^ stg.lds_off[r];
                            lds[k][2 * sl] = a[k][2 * r], lds[k][2 * sl + 1] = a[k][2 * r + 1];
                        }
                }
            }
            for (const auto &pw : pending) states[0][pw.first] = pw.second;
            for (int k = 0; k < ns && !ds; k++)
                for (uint32_t tid = 0; tid < threads; tid++) {
                    uint64_t g = 0;
                    uint32_t s = 0;
                    for (int i = 0; i < TB; i++)
                        if ((tid >> i) & 1u) g += pass.glob_thread[i], s ^= pass.st_thread_slot[i];
                    for (int r = 0; r < kRegs; r++) {
                        uint64_t gi = chunk_base + g + pass.glob_reg[r];
                        uint32_t sl = s ^ pass.st_reg_slot[r];
                        states[k][2 * gi] = lds[k][2 * sl];
                        states[k][2 * gi + 1] = lds[k][2 * sl + 1];
                    }
                }
        }
    }
}

PlanCache::Entry *PlanCache::find(const std::string &key) {
    for (auto &e : entries)
        if (e->key == key) return e.get();
    return nullptr;
}

PlanCache::Entry *PlanCache::insert(std::unique_ptr<Entry> e) {
    if (entries.size() >= kMaxEntries) {
        size_t lru = 0;
        for (size_t i = 1; i < entries.size(); i++)
            if (entries[i]->used < entries[lru]->used) lru = i;
        entries.erase(entries.begin() + (long)lru);
    }
    entries.push_back(std::move(e));
    return entries.back().get();
}

std::string plan_cache_key(int n, const std::vector<PlanOp> &ops, int low_qubits, bool remap, int max_fuse_bits,
                           int jit, double pass_cap, int chunk_bits) {
    std::string k;
    k.reserve(32 + ops.size() * 48);
    auto put = [&](const void *p, size_t bytes) { k.append((const char *)p, bytes); };
    const int32_t head[6] = {n, low_qubits, remap ? 1 : 0, max_fuse_bits, jit, chunk_bits};
    put(head, sizeof head);
    put(&pass_cap, sizeof pass_cap);
    for (const PlanOp &op : ops) {
        const int32_t f[4] = {op.kind, op.q0, op.q1, op.mat};
        const uint64_t m[3] = {op.controls, op.targets, op.diag};
        const uint8_t c = (uint8_t)((op.is_x ? 1 : 0) | (op.is_diag ? 2 : 0) | (op.is_real ? 4 : 0) | (op.gen << 3));
        const uint8_t u = (uint8_t)((op.unit + 1) | (op.unit_form << 3) | (op.skip ? 0x40 : 0));
        put(f, sizeof f);
        put(m, sizeof m);
        put(&c, 1);
        put(&u, 1);
    }
    return k;
}

void probe_factors(const SweepPlan &plan, const double *mats, std::vector<std::pair<uint64_t, double>> &out) {
    out.clear();
    for (const auto &pp : plan.probe_pivots) {
        double f = 1.0;
        for (const auto &src : pp.second) f *= mats[2 * ((int64_t)src.first + src.second)];
        out.push_back({pp.first, f * f});
    }
}

namespace {

// The first pass's input layout of a product plan: global positions of its register bits, thread-index bits
// and chunk-index bits (the chunk index deposits into the outer positions in ascending order).
struct ProductLayout {
    int tb = 0, lc = 0, tl = 0, th = 0, h = 0, hh = 0;
    int rpos[kRegBits] = {0};
    int tpos[kMaxThreadBits] = {0};
    std::vector<int> opos;
};

ProductLayout product_layout(int n, const SweepPlan &plan) {
    ProductLayout L;
    const SweepPass &ps = plan.passes[0];
    const bool direct = ps.nstages > 0 && (ps.direct & 1) != 0;  // the kernel fills the first stage's layout
    const uint64_t *lt = direct ? ps.ld_thread : ps.glob_thread, *lr = direct ? ps.ld_reg : ps.glob_reg;
    L.tb = plan.chunk_bits - kRegBits;
    for (int i = 0; i < L.tb; i++) L.tpos[i] = __builtin_ctzll(lt[i]);
    for (int j = 0; j < kRegBits; j++) L.rpos[j] = __builtin_ctzll(lr[1 << j]);
    for (int q = 0; q < n; q++)
        if ((ps.outer_mask >> q) & 1ULL) L.opos.push_back(q);
    L.lc = (int)L.opos.size();
    L.tl = plan.prod_tl;
    L.th = L.tb - L.tl;
    L.h = plan.prod_h;
    L.hh = L.lc - L.h;
    return L;
}

inline void cmul_host(const double *m, const double *x, double *out) {
    const double re = m[0] * x[0] - m[1] * x[1], im = m[0] * x[1] + m[1] * x[0];
    out[0] = re;
    out[1] = im;
}

}  // namespace

void absorb_leading_ops(int n, std::vector<PlanOp> &ops, std::vector<std::pair<int, int>> &absorbed) {
    std::vector<char> open((size_t)n, 1);
    for (PlanOp &op : ops) {
        if (op.kind == SC_OP_SINGLE && op.controls == 0 && op.q0 >= 0 && op.q0 < n && open[op.q0] && op.mat >= 0) {
            op.skip = true;
            absorbed.push_back({op.q0, op.mat});
            continue;
        }
        uint64_t touched = op.targets | op.diag | op.controls;
        if (op.q0 >= 0) touched |= 1ULL << op.q0;
        if (op.q1 >= 0) touched |= 1ULL << op.q1;
        for (int q = 0; q < n; q++)
            if ((touched >> q) & 1ULL) open[q] = 0;
    }
}

void finalize_product(int n, SweepPlan &plan, const std::vector<std::pair<int, int>> &absorbed) {
    plan.product = true;
    plan.product_ops = absorbed;
    const int tb = plan.chunk_bits - kRegBits, lc = n - plan.chunk_bits;
    plan.prod_tl = tb / 2;
    plan.prod_h = lc / 2;
    plan.prod_off = plan.pool_complex;
    plan.pool_complex += (uint64_t)kRegs + (1ULL << plan.prod_tl) + (1ULL << (tb - plan.prod_tl)) +
                         (1ULL << plan.prod_h) + (1ULL << (lc - plan.prod_h));
}

void product_input_host(int n, const SweepPlan &plan, const double *pool, double *state) {
    const ProductLayout L = product_layout(n, plan);
    const double *R = pool + 2 * plan.prod_off, *TL = R + 2 * kRegs, *TH = TL + 2 * (1 << L.tl);
    const double *CL = TH + 2 * (1 << L.th), *CH = CL + 2 * (1ULL << L.h);
    for (uint64_t i = 0; i < (1ULL << n); i++) {
        uint64_t ci = 0, t = 0, r = 0;
        for (int k = 0; k < L.lc; k++) ci |= ((i >> L.opos[k]) & 1ULL) << k;
        for (int k = 0; k < L.tb; k++) t |= ((i >> L.tpos[k]) & 1ULL) << k;
        for (int k = 0; k < kRegBits; k++) r |= ((i >> L.rpos[k]) & 1ULL) << k;
        double pc[2], pt[2], pct[2];
        cmul_host(CL + 2 * (ci & ((1ULL << L.h) - 1ULL)), CH + 2 * (ci >> L.h), pc);
        cmul_host(TL + 2 * (t & ((1ULL << L.tl) - 1ULL)), TH + 2 * (t >> L.tl), pt);
        cmul_host(pc, pt, pct);
        cmul_host(pct, R + 2 * r, state + 2 * i);
    }
}

void build_pool(const SweepPlan &plan, const double *mats, double *out, long long basis) {
    // first pass: plain recipes (kinds 0, 1) at their offsets
    std::vector<uint64_t> offs(plan.pool.size());
    uint64_t off = 0;
    for (size_t r = 0; r < plan.pool.size(); r++) {
        offs[r] = off;
        off += (uint64_t)plan.pool[r].count;
    }
    for (size_t r = 0; r < plan.pool.size(); r++) {
        const PoolRecipe &rc = plan.pool[r];
        double *dst = out + 2 * offs[r];
        const double *g = rc.mat >= 0 ? mats + 2 * (int64_t)rc.mat : nullptr;
        auto put = [&](int k, int src) { dst[2 * k] = g[2 * src], dst[2 * k + 1] = g[2 * src + 1]; };
        switch (rc.kind) {
        case RC_COPY: std::memcpy(dst, g, sizeof(double) * 2 * rc.count); break;
        case RC_DIAG1:  // (g00, g00, g11, g11) of a diagonal 2x2
            put(0, 0), put(1, 0), put(2, 3), put(3, 3);
            break;
        case RC_SWAP4:  // 4x4 with the qubit order exchanged: M'[2b+a][2d+c] = M[2a+b][2c+d]
            for (int row = 0; row < 4; row++)
                for (int col = 0; col < 4; col++) {
                    const int rs = ((row & 1) << 1) | (row >> 1), cs = ((col & 1) << 1) | (col >> 1);
                    put(4 * row + col, 4 * rs + cs);
                }
            break;
        case RC_SWAPDIAG:  // 4-entry diagonal with the qubit order exchanged
            put(0, 0), put(1, 2), put(2, 1), put(3, 3);
            break;
        case RC_UNIT:  // off-pivot entry / real pivot, per row
            for (int row = 0; row < 2; row++) {
                const int c = (rc.aux >> row) & 1;
                const double pv = g[2 * (2 * row + c)];
                dst[2 * row] = g[2 * (2 * row + 1 - c)] / pv;
                dst[2 * row + 1] = g[2 * (2 * row + 1 - c) + 1] / pv;
            }
            break;
        case RC_SCALE:
            dst[0] = 1.0, dst[1] = 0.0;
            break;
        default: break;
        }
    }
    // second pass: diagonal tables and fused blocks from the plain entries.
    for (size_t r = 0; r < plan.pool.size(); r++) {
        const PoolRecipe &rc = plan.pool[r];
        double *dst = out + 2 * offs[r];
        if (rc.kind == RC_DIAG16) {
            // every fused op is diagonal: applied to the all-ones vector they leave the table
            double col[2 * kRegs];
            for (int reg = 0; reg < kRegs; reg++) col[2 * reg] = 1.0, col[2 * reg + 1] = 0.0;
            for (size_t t = 0; t < rc.fused_ops.size(); t++) {
                const SweepOp &so = rc.fused_ops[t];
                host_apply(so, so.mat >= 0 ? out + 2 * offs[so.mat] : nullptr, col, 0u, 0ULL);
            }
            std::memcpy(dst, col, sizeof(col));
            continue;
        }
        // Block index l has bit (m-1-j) <-> register bit fused_bits[j]; the block's ops touch no other
        // register bit, so columns with those bits 0 suffice.
        if (rc.kind != RC_FUSED) continue;
        const int m = (int)rc.fused_bits.size(), d = 1 << m;
        auto reg_of = [&](int l) {
            int reg = 0;
            for (int j = 0; j < m; j++)
                if ((l >> (m - 1 - j)) & 1) reg |= 1 << rc.fused_bits[j];
            return reg;
        };
        for (int c = 0; c < d; c++) {
            double col[2 * kRegs] = {0};
            col[2 * reg_of(c)] = 1.0;
            for (size_t t = 0; t < rc.fused_ops.size(); t++) {
                const SweepOp &so = rc.fused_ops[t];
                host_apply(so, so.mat >= 0 ? out + 2 * offs[so.mat] : nullptr, col, 0u, 0ULL);
            }
            for (int row = 0; row < d; row++) {
                dst[2 * (row * d + c)] = col[2 * reg_of(row)];
                dst[2 * (row * d + c) + 1] = col[2 * reg_of(row) + 1];
            }
        }
    }
    // third pass: pass factors of unit-pivot ops (real pivots m[0][c])
    for (size_t r = 0; r < plan.pool.size(); r++) {
        const PoolRecipe &rc = plan.pool[r];
        if (rc.scale_srcs.empty()) continue;
        double f = 1.0;
        for (const auto &src : rc.scale_srcs) f *= mats[2 * ((int64_t)src.first + src.second)];
        double *dst = out + 2 * offs[r];
        for (int k = 0; k < 2 * rc.count; k++) dst[k] *= f;
    }
    if (!plan.product || plan.passes.empty()) return;
    // product input: v_q = the absorbed ops on qubit q applied to the basis state's bit, then the tables
    int n = plan.chunk_bits;
    for (int q = 0; q < 64; q++)
        if ((plan.passes[0].outer_mask >> q) & 1ULL) n++;
    std::vector<double> v((size_t)4 * n, 0.0);
    for (int q = 0; q < n; q++) v[4 * q + 2 * (int)(((uint64_t)basis >> q) & 1ULL)] = 1.0;
    for (const auto &a : plan.product_ops) {
        const double *g = mats + 2 * (int64_t)a.second;
        double *x = &v[4 * a.first];
        double y0[2], y1[2], t[2];
        cmul_host(g, x, y0), cmul_host(g + 2, x + 2, t), y0[0] += t[0], y0[1] += t[1];
        cmul_host(g + 4, x, y1), cmul_host(g + 6, x + 2, t), y1[0] += t[0], y1[1] += t[1];
        x[0] = y0[0], x[1] = y0[1], x[2] = y1[0], x[3] = y1[1];
    }
    const ProductLayout L = product_layout(n, plan);
    double *R = out + 2 * plan.prod_off, *TL = R + 2 * kRegs, *TH = TL + 2 * (1 << L.tl);
    double *CL = TH + 2 * (1 << L.th), *CH = CL + 2 * (1ULL << L.h);
    // entry e of a table over positions pos[0..k): prod_j v_{pos[j]}[bit j of e], multiplied in order of j
    auto table = [&](const int *pos, int k, double *dst) {
        for (uint64_t e = 0; e < (1ULL << k); e++) {
            double acc[2] = {1.0, 0.0};
            for (int j = 0; j < k; j++) cmul_host(acc, &v[4 * pos[j] + 2 * (int)((e >> j) & 1ULL)], acc);
            dst[2 * e] = acc[0], dst[2 * e + 1] = acc[1];
        }
    };
    table(L.rpos, kRegBits, R);
    table(L.tpos, L.tl, TL);
    table(L.tpos + L.tl, L.th, TH);
    table(L.opos.data(), L.h, CL);
    table(L.opos.data() + L.h, L.hh, CH);
}

}  // namespace sc
