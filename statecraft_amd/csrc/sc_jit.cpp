// sc_jit.cpp -- sweep passes compiled at run time (hipRTC).
//
// The compiled sweep kernel (sc_sweep.hip) interprets a pass: it reads every op record, branches on its
// kind and bit positions, and holds the code of every op variant.  At 30 qubits and above that dispatch
// costs about as much as an op's arithmetic.  Here each pass of a plan becomes its own kernel whose op
// sequence, bit positions, LDS slots and HBM offsets are compile-time constants: the ops inline to
// straight-line FP64 code, exchanges (X, CNOT, SWAP) become register renames, and the matrices stay in
// the pass pool, so one kernel serves every parameter value of a circuit structure.
//
// Kernels are cached per device by their source text for the life of the process; a plan whose passes
// are all cached launches without compiling.  The op library (sc_ops.hpp) is compiled into the library
// as text (kOpsHeaderText) and handed to hipRTC as a header.
#include <hip/hiprtc.h>

#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "sc_jit.hpp"
#include "sc_plan.hpp"

namespace sc {

extern const char kOpsHeaderText[];

namespace {

std::string hex64(uint64_t v) {
    char b[32];
    std::snprintf(b, sizeof b, "0x%llxULL", (unsigned long long)v);
    return b;
}
std::string hex32(uint32_t v) {
    char b[24];
    std::snprintf(b, sizeof b, "0x%xu", v);
    return b;
}

// Per-lane value sum_i (bit i of `var` ? scale * tab[i] : 0), combined with `op` ("+" or "^"), as an
// expression.
template <typename T>
std::string lane_expr(const char *var, const T *tab, int bits, const char *op, bool wide, uint32_t scale = 1) {
    std::string e;
    for (int i = 0; i < bits; i++) {
        if (tab[i] == 0) continue;
        if (!e.empty()) e += std::string(" ") + op + " ";
        e += std::string("((") + var + " & " + std::to_string(1u << i) + "u) ? " +
             (wide ? hex64((uint64_t)tab[i] * scale) : hex32((uint32_t)tab[i] * scale)) + " : " +
             (wide ? "0ULL" : "0u") + ")";
    }
    if (e.empty()) e = wide ? "0ULL" : "0u";
    return e;
}

// Expression depositing the bits of `var` (a chunk index) at the set positions of `mask`, one shift per
// run of consecutive positions.
std::string deposit_expr(const char *var, uint64_t mask) {
    std::string e;
    int j = 0;  // next chunk-index bit
    for (int pos = 0; pos < 64;) {
        if (!((mask >> pos) & 1ULL)) {
            pos++;
            continue;
        }
        int len = 0;
        while (pos + len < 64 && ((mask >> (pos + len)) & 1ULL)) len++;
        const uint64_t lm = len >= 64 ? ~0ULL : ((1ULL << len) - 1ULL);
        if (!e.empty()) e += " | ";
        e += "(((" + std::string(var) + " >> " + std::to_string(j) + ") & " + hex64(lm) + ") << " +
             std::to_string(pos) + ")";
        j += len;
        pos += len;
    }
    if (e.empty()) e = "0ULL";
    return e;
}

// The call that applies op o from register set S to D: the per-kind entry point of sc_ops.hpp that run_op
// would dispatch to, with every field a literal.
std::string op_call(const SweepOp &o, const char *S, const char *D) {
    std::ostringstream s;
    const std::string sd = std::string(S) + ", " + D;
    const std::string m = "emats + " + std::to_string(o.mat);
    const int k0 = o.k0, k1 = o.k1;
    switch (o.kind) {
    case SW_U1: s << "op_u1m<" << k0 << ">(" << sd << ", " << m << ");"; break;
    case SW_U1R: s << "op_u1r<" << k0 << ">(" << sd << ", " << m << ");"; break;
    case SW_U1C:
        s << "op_u1c<" << k0 << ">(" << sd << ", " << m << ", " << (unsigned)o.rctrl << "u, " << (unsigned)o.tctrl
          << "u, tid);";
        break;
    case SW_X: s << "op_x<" << k0 << ", false>(" << sd << ", " << (unsigned)o.rctrl << "u, true);"; break;
    case SW_XL:
        s << "op_xl<" << k0 << ">(" << sd << ", " << (unsigned)o.rctrl << "u, " << (unsigned)o.tctrl << "u, tid);";
        break;
    case SW_U2: s << "op_u2<" << k0 << ", " << k1 << ", false>(" << sd << ", " << m << ");"; break;
    case SW_U2R: s << "op_u2<" << k0 << ", " << k1 << ", true>(" << sd << ", " << m << ");"; break;
    case SW_SWAP: s << "op_swap<" << k0 << ", " << k1 << ">(" << sd << ");"; break;
    case SW_D4: s << "op_d4<false>(" << sd << ", " << m << ");"; break;
    case SW_D4R: s << "op_d4<true>(" << sd << ", " << m << ");"; break;
    case SW_DIAG16: s << "op_diag16(" << sd << ", " << m << ");"; break;
    case SW_DIAGL:
        s << "op_diagl_sel<" << (k0 < 0 ? kRegBits : k0) << ">(" << sd << ", " << m << ", " << (unsigned)o.rctrl
          << "u, " << (unsigned)o.tctrl << "u, " << (unsigned)(o.ta | ((o.thi & 1u) << 8)) << "u, "
          << (unsigned)(o.tb | (((o.thi >> 1) & 1u) << 8)) << "u, " << (int)o.oa
          << ", " << (int)o.ob << ", tid, cb);";
        break;
    case SW_CX: s << "op_cx<" << k0 << ", " << k1 << ">(" << sd << ");"; break;
    case SW_U1U:
        s << "op_u1u<" << k0 << ", " << (o.ta & 1) << ", " << ((o.ta >> 1) & 1) << ", " << (int)o.tb << ">(" << sd
          << ", " << m << ");";
        break;
    case SW_SCALE: s << "op_scale(" << sd << ", " << m << ");"; break;
    default: s << "copy_all(" << sd << ");"; break;
    }
    if (o.octrl == 0) return s.str();
    // outer (chunk-fixed) controls: identity on chunks where they are not all set
    return "if ((cb & " + hex64(o.octrl) + ") == " + hex64(o.octrl) + ") { " + s.str() + " } else { copy_all(" + sd +
           "); }";
}

// Source of one pass kernel; "@K@" stands for the kernel name.  The structure is that of k_sweep
// (sc_sweep.hip) with every table read replaced by its value.  With `prefetch` the HBM loads of a
// workgroup's next chunk are issued before the stages of the current one, into a third register set, so
// that they arrive while the chunk is computed; otherwise they are issued after the stages, as in k_sweep.
// A pass with direct loads (SweepPass::direct bit 0) reads HBM straight into its first stage's registers,
// one with direct stores (bit 1) writes its last stage's registers straight to HBM; the LDS round trips of
// the load and store phases are then skipped.
std::string pass_source(const SweepPlan &plan, const SweepPass &ps, bool prefetch, int sched = 0,
                        const std::vector<uint64_t> *zterms = nullptr) {
    std::ostringstream s;
    const int TB = plan.chunk_bits - kRegBits;  // thread-index bits: 8 (256 threads) or 9 (512)
    const int S = ps.nstages;
    const bool DL = S > 0 && (ps.direct & 1) != 0, DS = S > 0 && (ps.direct & 2) != 0;
    const int NT = zterms ? (int)zterms->size() : 0;
    s << "extern \"C\" __global__ void __launch_bounds__(" << (1 << TB)
      << ") __attribute__((amdgpu_waves_per_eu(2, 2)))\n"
         // amp and out may be the same buffer: every chunk is read once and written once, by one workgroup,
         // so no access of one pointer can observe an access of the other
         "@K@(sc::v2d *__restrict__ amp, sc::v2d *__restrict__ out, int n, int log_chunks, uint64_t total_chunks,\n"
         "    const sc::v2d *__restrict__ mats_arg, uint64_t mat_stride, int elem_begin, long long basis"
      << (NT ? ",\n    double *__restrict__ tpart, int tparts" : "") << ") {\n"
         "  using namespace sc;\n";
    // Z-string terms (zterms, one mask each): the pass also sums sign_t(i) |a_i|^2 over every amplitude it stores,
    // per batch element, into tpart (see below)
    if (NT) s << "  __shared__ double tsum[" << (1 << TB) / 64 << "][" << NT << "];\n";
    // stage st > 0 with SweepStage::perm is entered inside the registers, with no LDS round trip
    auto in_reg = [&](int st) { return st > 0 && st < S && ps.stages[st].perm != 0; };
    bool uses_lds = !DL || !DS;
    for (int st = 1; st < S; st++) uses_lds = uses_lds || !in_reg(st);
    if (uses_lds) s << "  __shared__ v2d lds[" << (1 << plan.chunk_bits) << "];\n";
    s << "  CV2 *mats = (CV2 *)mats_arg;\n"
         "  const uint32_t tid = threadIdx.x;\n"
         // with one matrix pool per element (mat_stride != 0) a workgroup takes a contiguous range of chunks,
         // so that consecutive chunks share their element's matrices in the scalar cache; otherwise chunks
         // are dealt round-robin
         "  const bool blocked = mat_stride != 0;\n"
         "  const uint64_t span = blocked ? (total_chunks + gridDim.x - 1) / gridDim.x : 0;\n"
         "  uint64_t c = blocked ? blockIdx.x * span : blockIdx.x;\n"
         "  const uint64_t c_end = blocked ? (c + span < total_chunks ? c + span : total_chunks) : total_chunks;\n"
         "  if (c >= c_end) return;\n";
    // wave priority: one of the two workgroups a CU holds issues first when both have VALU work, so the two
    // drift apart and one's LDS exchanges overlap the other's arithmetic (1: the grid's second half, 2: odd
    // workgroups)
    if (sched == 1) s << "  if (blockIdx.x >= gridDim.x / 2u) __builtin_amdgcn_s_setprio(1);\n";
    if (sched == 2) s << "  if (blockIdx.x & 1u) __builtin_amdgcn_s_setprio(1);\n";
    s << "  const uint64_t chunk_mask = (1ULL << log_chunks) - 1ULL;\n"
         "  auto cbase = [](uint64_t x) -> uint64_t { return "
      << deposit_expr("x", ps.outer_mask) << "; };\n";
    // global offsets of the load side (first stage's layout when direct, else the load phase's) and of the
    // store side (last stage's layout when direct, else the store phase's)
    const uint64_t *l_thread = DL ? ps.ld_thread : ps.glob_thread, *l_reg = DL ? ps.ld_reg : ps.glob_reg;
    const uint64_t *s_thread = DS ? ps.sd_thread : ps.glob_thread, *s_reg = DS ? ps.sd_reg : ps.glob_reg;
    s << "  const uint64_t gload = " << lane_expr("tid", l_thread, TB, "+", true) << ";\n";
    s << "  const uint64_t gstore = " << lane_expr("tid", s_thread, TB, "+", true) << ";\n";
    // term t's sign at index cb + gstore + s_reg[r] is (-1)^(parity(z_t & gstore) + parity(z_t & cb) +
    // parity(z_t & s_reg[r])): the lane's part once here, the chunk's part per chunk (uniform), the register
    // part through a Walsh-Hadamard transform of the 16 probabilities over the register bits
    if (NT) {
        s << "  uint32_t tlane = 0u;\n  double tacc[" << NT << "];\n";
        for (int t = 0; t < NT; t++)
            s << "  tlane |= (uint32_t)(__builtin_popcountll(" << hex64((*zterms)[t]) << " & gstore) & 1) << " << t
              << ";\n  tacc[" << t << "] = 0.0;\n";
    }
    // LDS addresses are byte offsets (slot * 16), so that each access costs one XOR
    if (uses_lds) {
        s << "  char *const ldsb = (char *)lds;\n";
        s << "  const uint32_t sthread = " << lane_expr("tid", ps.io_thread_slot, TB, "^", false, 16) << ";\n";
        s << "  const uint32_t sout = " << lane_expr("tid", ps.st_thread_slot, TB, "^", false, 16) << ";\n";
    }
    auto regs_load = [&](const char *dst, const char *srcp) {
        for (int r = 0; r < kRegs; r++) s << "    " << dst << "[" << r << "] = " << srcp << "[" << l_reg[r] << "];\n";
    };
    // basis >= 0: the input of every element is the basis state |basis>, generated instead of loaded
    auto regs_basis = [&](const char *dst, const char *chunk_base) {
        s << "    const uint64_t i0 = " << chunk_base << " + gload;\n";
        for (int r = 0; r < kRegs; r++)
            s << "    " << dst << "[" << r << "] = (i0 + " << l_reg[r]
              << "ULL == (uint64_t)basis) ? make_v2d(1.0, 0.0) : make_v2d(0.0, 0.0);\n";
    };
    // product plans: the first pass generates prod_q v_q[i_q] from the pool's tables (SweepPlan::product):
    // the chunk's factor CL[low chunk bits] CH[high chunk bits], the thread's TL[low tid bits] TH[high tid
    // bits], then one product per register with R[r]
    const bool product = plan.product && &ps == &plan.passes[0];
    const int ptl = plan.prod_tl, pth = TB - plan.prod_tl, ph = plan.prod_h;
    const uint64_t oTL = kRegs, oTH = oTL + (1ULL << ptl), oCL = oTH + (1ULL << pth), oCH = oCL + (1ULL << ph);
    // The thread factor only changes with the batch element: it is kept in registers (pel: its element),
    // so that a chunk's generation waits on no vector load.
    auto regs_product = [&](const char *dst, const char *cvar) {
        s << "    {\n"
             "    const uint64_t pe = (uint64_t)elem_begin + (" << cvar << " >> log_chunks);\n"
             "    CV2 *pm = mats + pe * mat_stride + " << plan.prod_off << "ULL;\n"
             "    if (pe != pel) {\n"
             "      pel = pe;\n"
             "      pt = cmul(pm[" << oTL << "u + (tid & " << ((1u << ptl) - 1u) << "u)], pm[" << oTH << "u + (tid >> "
          << ptl << ")]);\n"
             "    }\n"
             "    const uint64_t pci = " << cvar << " & chunk_mask;\n"
             "    const v2d pc = cmul(pm[" << oCL << "ULL + (pci & " << hex64((1ULL << ph) - 1ULL) << ")], pm[" << oCH
          << "ULL + (pci >> " << ph << ")]);\n"
             "    const v2d pct = cmul(pc, pt);\n";
        for (int r = 0; r < kRegs; r++) s << "    " << dst << "[" << r << "] = cmul(pct, pm[" << r << "]);\n";
        s << "    }\n";
    };
    auto regs_store = [&](const char *dstp, const char *src) {
        for (int r = 0; r < kRegs; r++) s << "    " << dstp << "[" << s_reg[r] << "] = " << src << "[" << r << "];\n";
    };
    auto lds_write = [&](const char *base, const uint32_t *off, const char *src) {
        for (int r = 0; r < kRegs; r++)
            s << "    *(v2d *)(ldsb + (" << base << " ^ " << off[r] * 16u << "u)) = " << src << "[" << r << "];\n";
    };
    auto lds_read = [&](const char *dst, const char *base, const uint32_t *off) {
        for (int r = 0; r < kRegs; r++)
            s << "    " << dst << "[" << r << "] = *(const v2d *)(ldsb + (" << base << " ^ " << off[r] * 16u << "u));\n";
    };
    auto copy_regs = [&](const char *dst, const char *src) {
        for (int r = 0; r < kRegs; r++) s << "    " << dst << "[" << r << "] = " << src << "[" << r << "];\n";
    };
    // Exchanges in which every wavefront reads back only the LDS slots it wrote need no workgroup barrier.
    auto sync = [&](bool wave_local) { s << (wave_local ? "    wave_sync();\n" : "    __syncthreads();\n"); };
    // the chunk's input: one register set; the prefetching variant loads the next chunk into a third one
    const char *in0 = prefetch ? "q" : "p";
    s << (prefetch ? "  Regs a, p, q;\n" : "  Regs a, p;\n");
    if (product) s << "  uint64_t pel = ~0ULL;\n  v2d pt = make_v2d(0.0, 0.0);\n";
    s << "  uint64_t cb = cbase(c & chunk_mask);\n"
         "  if (basis >= 0) {\n";
    if (product) regs_product(in0, "c");
    else regs_basis(in0, "cb");
    s << "  } else {\n"
         "    const v2d *src = amp + (((uint64_t)elem_begin + (c >> log_chunks)) << n) + cb + opaque(gload);\n";
    regs_load(in0, "src");
    s << "  }\n";
    s << "  for (;;) {\n"
         "    const uint64_t elem = (uint64_t)elem_begin + (c >> log_chunks);\n"
         "    CV2 *emats = mats + elem * mat_stride;\n"
         "    const uint64_t next = blocked ? c + 1 : c + gridDim.x;\n"
         "    const uint64_t ncb = next < c_end ? cbase(next & chunk_mask) : 0ULL;\n";
    auto next_loads = [&](const char *dst) {
        s << "    if (next < c_end && basis >= 0) {\n";
        if (product) regs_product(dst, "next");
        else regs_basis(dst, "ncb");
        s << "    } else if (next < c_end) {\n"
             "    const v2d *src = amp + (((uint64_t)elem_begin + (next >> log_chunks)) << n) + ncb + opaque(gload);\n";
        regs_load(dst, "src");
        s << "    }\n";
    };
    // The first LDS write of a chunk follows the previous chunk's last LDS read: a barrier unless both use the
    // same wave-index positions (SweepPass::wave_local bit 1).
    const bool war_local = (ps.wave_local & 2) != 0;
    bool first_write_done = false;
    auto before_first_write = [&]() {
        if (!first_write_done) sync(war_local);
        first_write_done = true;
    };
    const char *cur = nullptr;  // register set holding the current values
    if (!DL) {
        before_first_write();
        s << "    {\n";
        lds_write("sthread", ps.io_reg_slot, in0);
        s << "    }\n";
        if (prefetch) next_loads("q");
        sync(S > 0 ? ps.stages[0].wave_local != 0 : (ps.wave_local & 1) != 0);
    } else {
        if (prefetch) {
            s << "    {\n";
            copy_regs("p", "q");
            s << "    }\n";
            next_loads("q");
        }
        cur = "p";
    }
    for (int st = 0; st < S; st++) {
        const SweepStage &stg = ps.stages[st];
        s << "    {\n"
             "    const uint32_t t = opaque(tid);\n";
        const bool from_lds = !(st == 0 && DL) && !in_reg(st);
        const bool to_lds = !(st == S - 1 && DS) && !in_reg(st + 1);
        if (from_lds || to_lds)
            s << "    const uint32_t ts = " << lane_expr("t", stg.thread_slot, TB, "^", false, 16) << ";\n";
        // wave priority raised over part of the stage (3: arithmetic; 4: LDS reads and arithmetic; 5: reads,
        // arithmetic and writes; 6: arithmetic and writes)
        const bool hi_read = sched == 4 || sched == 5, hi_write = sched == 5 || sched == 6;
        const bool hi_arith = sched >= 3;
        bool in_a;
        if (hi_read) s << "    __builtin_amdgcn_s_setprio(1);\n";
        if (from_lds) {
            lds_read("a", "ts", stg.lds_off);
            in_a = true;
        } else {
            in_a = std::string(cur) == "a";
        }
        if (hi_arith && !hi_read) s << "    __builtin_amdgcn_s_setprio(1);\n";
        if (in_reg(st)) {
            const int j5 = (stg.perm & 7) - 1, j4 = ((stg.perm >> 3) & 7) - 1;
            if (j5 >= 0) s << "    lane_exchange<" << j5 << ", false>(" << cur << ");\n";
            if (j4 >= 0) s << "    lane_exchange<" << j4 << ", true>(" << cur << ");\n";
        }
        for (int o = stg.op_begin; o < stg.op_end; o++) {
            s << "    " << op_call(plan.ops[o], in_a ? "a" : "p", in_a ? "p" : "a") << "\n";
            in_a = !in_a;
        }
        if (hi_arith && !(hi_write && to_lds)) s << "    __builtin_amdgcn_s_setprio(0);\n";
        cur = in_a ? "a" : "p";
        if (to_lds) {
            before_first_write();
            s << "    const uint32_t tw = opaque(ts);\n";
            lds_write("tw", stg.lds_off, cur);
            if (hi_write) s << "    __builtin_amdgcn_s_setprio(0);\n";
            sync(st + 1 < S ? ps.stages[st + 1].wave_local != 0 : (ps.wave_local & 1) != 0);
        }
        s << "    }\n";
    }
    if (!DS) {
        s << "    {\n"
             "    const uint32_t sl = opaque(sout);\n";
        lds_read("a", "sl", ps.st_reg_slot);
        s << "    }\n";
        cur = "a";
    }
    const std::string fin = cur ? cur : "a";
    if (NT) {
        s << "    {\n    double w[16];\n";
        for (int r = 0; r < kRegs; r++)
            s << "    w[" << r << "] = fma(" << fin << "[" << r << "].x, " << fin << "[" << r << "].x, " << fin << "[" << r
              << "].y * " << fin << "[" << r << "].y);\n";
        for (int b = 0; b < kRegBits; b++)
            for (int r = 0; r < kRegs; r++)
                if (!((r >> b) & 1))
                    s << "    { const double u = w[" << r << "], v = w[" << (r | (1 << b)) << "]; w[" << r
                      << "] = u + v; w[" << (r | (1 << b)) << "] = u - v; }\n";
        s << "    uint32_t tsig = tlane;\n";
        for (int t = 0; t < NT; t++) {
            const uint64_t z = (*zterms)[t];
            int m = 0;
            for (int j = 0; j < kRegBits; j++)
                if (s_reg[1 << j] & z) m |= 1 << j;
            s << "    tsig ^= (uint32_t)(__builtin_popcountll(" << hex64(z) << " & cb) & 1) << " << t << ";\n"
              << "    tacc[" << t << "] += ((tsig >> " << t << ") & 1u) ? -w[" << m << "] : w[" << m << "];\n";
        }
        s << "    }\n";
    }
    // without prefetch the next chunk's loads go into p: after the store when p holds the results
    if (!prefetch && fin != "p") next_loads("p");
    s << "    {\n"
         "    v2d *dst = out + (elem << n) + cb + opaque(gstore);\n";
    regs_store("dst", fin.c_str());
    s << "    }\n";
    if (!prefetch && fin == "p") next_loads("p");
    // Term sums leave a workgroup when its contiguous chunk range moves to another element (or ends): reduced
    // over each wavefront, then over the wavefronts in a fixed order, into tpart[(element * NT + t) * tparts +
    // slot], slot = this workgroup's place among those whose ranges cover the element (blocked assignment only).
    if (NT) {
        s << "    {\n    const uint64_t ce = c >> log_chunks;\n"
             "    if (next >= c_end || (next >> log_chunks) != ce) {\n";
        for (int t = 0; t < NT; t++)
            s << "      { double v = tacc[" << t << "]; for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);\n"
                 "        if ((tid & 63u) == 0u) tsum[tid >> 6][" << t << "] = v; tacc[" << t << "] = 0.0; }\n";
        s << "      __syncthreads();\n"
             "      if (tid < " << NT << "u) {\n"
             "        double v = tsum[0][tid];\n";
        for (int wv = 1; wv < (1 << TB) / 64; wv++) s << "        v += tsum[" << wv << "][tid];\n";
        s << "        const uint64_t wf = (ce << log_chunks) / span;\n"
             "        tpart[(ce * " << NT << "ULL + tid) * (uint64_t)tparts + (blockIdx.x - wf)] = v;\n"
             "      }\n"
             "      __syncthreads();\n"
             "    }\n    }\n";
    }
    s << "    if (next >= c_end) break;\n"
         "    c = next;\n"
         "    cb = ncb;\n";
    s << "  }\n"
         "}\n";
    return s.str();
}

// Source of one adjoint pass: the pass applied to two states, psi and lambda, at once (every op on both, the
// matrices shared), with each gradient probe's <lambda| G |psi> reduced per wavefront and accumulated per
// workgroup in LDS.  Chunks move through one 64 KiB LDS buffer, psi then lambda.  Workgroup w serves batch
// element e0 + w % nb and that element's chunks sub, sub + gsub, ... (sub = w / nb); at the end it writes
// its sum for probe k to part[(slot_k * B + e) * gsub + sub].
std::string adjoint_pass_source(const SweepPlan &plan, const SweepPass &ps, bool prio = false) {
    std::ostringstream s;
    std::vector<uint64_t> probes;
    for (int st = 0; st < ps.nstages; st++)
        for (int o = ps.stages[st].op_begin; o < ps.stages[st].op_end; o++)
            if (plan.ops[o].kind == SW_GRAD) probes.push_back(plan.ops[o].slot);
    const int K = (int)probes.size();
    s << "extern \"C\" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))\n"
         "@K@(sc::v2d *__restrict__ psi, sc::v2d *__restrict__ lam, int n, int log_chunks,\n"
         "    const sc::v2d *__restrict__ mats_arg, uint64_t mat_stride, int e0, int nb, int gsub,\n"
         "    double *__restrict__ part, int B) {\n"
         "  using namespace sc;\n"
         "  __shared__ v2d lds[kChunk];\n"
      << "  __shared__ double gacc[4][" << std::max(K, 1) << "][2];\n"
      << "  char *const ldsb = (char *)lds;\n"
         "  CV2 *mats = (CV2 *)mats_arg;\n"
         "  const uint32_t tid = threadIdx.x;\n"
         "  const int e = e0 + (int)(blockIdx.x % (uint32_t)nb);\n"
         "  const uint64_t sub = blockIdx.x / (uint32_t)nb;\n"
         "  const uint64_t nchunks = 1ULL << log_chunks;\n"
         "  CV2 *emats = mats + (uint64_t)e * mat_stride;\n"
         "  v2d *P = psi + ((uint64_t)e << n);\n"
         "  v2d *L = lam + ((uint64_t)e << n);\n"
      << "  for (int i = tid; i < " << 8 * std::max(K, 1) << "; i += kThreads) (&gacc[0][0][0])[i] = 0.0;\n"
      << "  auto cbase = [](uint64_t x) -> uint64_t { return " << deposit_expr("x", ps.outer_mask) << "; };\n";
    s << "  const uint64_t gthread = " << lane_expr("tid", ps.glob_thread, kThreadBits, "+", true) << ";\n";
    s << "  const uint32_t sthread = " << lane_expr("tid", ps.io_thread_slot, kThreadBits, "^", false, 16) << ";\n";
    s << "  const uint32_t sout = " << lane_expr("tid", ps.st_thread_slot, kThreadBits, "^", false, 16) << ";\n";
    auto emit_regs = [&](const char *fmt_l, const uint32_t *off, const char *base, const char *arr, bool write) {
        (void)fmt_l;
        for (int r = 0; r < kRegs; r++) {
            if (write)
                s << "    *(v2d *)(ldsb + (" << base << " ^ " << off[r] * 16u << "u)) = " << arr << "[" << r << "];\n";
            else
                s << "    " << arr << "[" << r << "] = *(const v2d *)(ldsb + (" << base << " ^ " << off[r] * 16u
                  << "u));\n";
        }
    };
    s << "  Regs a, p, b, q;\n"
         "  __syncthreads();\n"
         "  for (uint64_t cl = sub; cl < nchunks; cl += (uint64_t)gsub) {\n"
         "    const uint64_t cb = cbase(cl);\n";
    // load phase of one state into stage 0's layout (or straight to the store layout when there are no stages)
    auto load_state = [&](const char *base_ptr, const char *arr) {
        s << "    {\n    const v2d *src = " << base_ptr << " + cb + opaque(gthread);\n";
        for (int r = 0; r < kRegs; r++) s << "    " << arr << "[" << r << "] = src[" << ps.glob_reg[r] << "];\n";
        emit_regs("", ps.io_reg_slot, "sthread", arr, true);
        s << "    }\n    __syncthreads();\n";
        if (ps.nstages > 0) {
            const SweepStage &stg = ps.stages[0];
            s << "    {\n    const uint32_t t = opaque(tid);\n    const uint32_t ts = "
              << lane_expr("t", stg.thread_slot, kThreadBits, "^", false, 16) << ";\n";
            emit_regs("", stg.lds_off, "ts", arr, false);
            s << "    }\n    __syncthreads();\n";
        }
    };
    load_state("P", "a");
    load_state("L", "b");
    bool in_a = true;  // current registers: a/b (else p/q)
    int k = 0;
    for (int st = 0; st < ps.nstages; st++) {
        const SweepStage &stg = ps.stages[st];
        if (st > 0) {
            // exchange both states from stage st-1's layout to stage st's
            const SweepStage &prev = ps.stages[st - 1];
            for (int which = 0; which < 2; which++) {
                const char *arr = which == 0 ? (in_a ? "a" : "p") : (in_a ? "b" : "q");
                s << "    {\n    const uint32_t t = opaque(tid);\n    const uint32_t tw = "
                  << lane_expr("t", prev.thread_slot, kThreadBits, "^", false, 16) << ";\n";
                emit_regs("", prev.lds_off, "tw", arr, true);
                s << "    __syncthreads();\n    const uint32_t t2 = opaque(tid);\n    const uint32_t ts = "
                  << lane_expr("t2", stg.thread_slot, kThreadBits, "^", false, 16) << ";\n";
                emit_regs("", stg.lds_off, "ts", arr, false);
                s << "    __syncthreads();\n    }\n";
            }
        }
        // wave priority raised over the stage's arithmetic, as in the forward passes
        if (prio) s << "    __builtin_amdgcn_s_setprio(1);\n";
        for (int o = stg.op_begin; o < stg.op_end; o++) {
            const SweepOp &op = plan.ops[o];
            const char *P0 = in_a ? "a" : "p", *P1 = in_a ? "p" : "a";
            const char *L0 = in_a ? "b" : "q", *L1 = in_a ? "q" : "b";
            if (op.kind == SW_GRAD) {
                // only Im <lambda|G|psi> enters a gradient: the real part stays zero
                s << "    {\n    double re = 0.0, im = 0.0;\n    grad_acc<" << (int)op.ta << ", " << (int)op.k0 << ", "
                  << (op.k1 >= 0 ? (int)op.k1 : 0) << ">(" << P0 << ", " << L0 << ", re, im);\n"
                  << "    for (int m = 32; m >= 1; m >>= 1) im += __shfl_xor(im, m);\n"
                  << "    if ((tid & 63u) == 0u) gacc[tid >> 6][" << k << "][1] += im;\n    }\n";
                k++;
                continue;
            }
            s << "    " << op_call(op, P0, P1) << "\n";
            s << "    " << op_call(op, L0, L1) << "\n";
            in_a = !in_a;
        }
        if (prio) s << "    __builtin_amdgcn_s_setprio(0);\n";
    }
    // store phase: each state through LDS into the store layout, then to HBM
    const uint32_t *last_off = ps.nstages > 0 ? ps.stages[ps.nstages - 1].lds_off : ps.io_reg_slot;
    for (int which = 0; which < 2; which++) {
        const char *arr = which == 0 ? (in_a ? "a" : "p") : (in_a ? "b" : "q");
        s << "    {\n";
        if (ps.nstages > 0) {
            const SweepStage &lst = ps.stages[ps.nstages - 1];
            s << "    const uint32_t t = opaque(tid);\n    const uint32_t tw = "
              << lane_expr("t", lst.thread_slot, kThreadBits, "^", false, 16) << ";\n";
            emit_regs("", last_off, "tw", arr, true);
        } else {
            emit_regs("", ps.io_reg_slot, "sthread", arr, true);  // the registers still hold the load layout
        }
        s << "    __syncthreads();\n";
        s << "    const uint32_t sl = opaque(sout);\n";
        emit_regs("", ps.st_reg_slot, "sl", arr, false);
        s << "    v2d *dst = " << (which == 0 ? "P" : "L") << " + cb + opaque(gthread);\n";
        for (int r = 0; r < kRegs; r++) s << "    dst[" << ps.glob_reg[r] << "] = " << arr << "[" << r << "];\n";
        s << "    __syncthreads();\n    }\n";
    }
    s << "  }\n";
    // per-workgroup probe sums, in a fixed order over the wavefronts
    s << "  __syncthreads();\n";
    for (int j = 0; j < K; j++) {
        s << "  if (tid < 2u) part[((uint64_t)" << probes[j] << "ULL * (uint64_t)B + (uint64_t)e) * (uint64_t)gsub * 2ULL"
          << " + sub * 2ULL + tid] = ((gacc[0][" << j << "][tid] + gacc[1][" << j << "][tid]) + gacc[2][" << j
          << "][tid]) + gacc[3][" << j << "][tid];\n";
    }
    s << "}\n";
    return s.str();
}

std::string replace_name(const std::string &body, const std::string &name) {
    std::string out = body;
    const size_t at = out.find("@K@");
    out.replace(at, 3, name);
    return out;
}

const char *const kRtcOptions[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                                   "-Xclang", "-packed-fp32-ops"};

// ---- on-disk cache of compiled modules.  A module's code object is kept under the cache directory, named by a
// 64-bit FNV-1a hash of everything that determines it (the hipRTC version, the options, the op library text and
// the module source); the file starts with a magic word, a second hash of the same material and the code size,
// so a truncated, foreign or stale file is recompiled instead of loaded.  Directory: $SC_KERNEL_CACHE_DIR (empty:
// no cache), else $XDG_CACHE_HOME/statecraft_amd, else $HOME/.cache/statecraft_amd.  Files are written under a
// temporary name and renamed, so concurrent processes never see a partial file.
uint64_t fnv1a(const char *p, size_t n, uint64_t h) {
    for (size_t i = 0; i < n; i++) {
        h ^= (uint8_t)p[i];
        h *= 0x100000001b3ULL;
    }
    return h;
}

std::string kernel_cache_dir() {
    if (const char *e = std::getenv("SC_KERNEL_CACHE_DIR")) return e;
    if (const char *x = std::getenv("XDG_CACHE_HOME"); x && *x) return std::string(x) + "/statecraft_amd";
    if (const char *h = std::getenv("HOME"); h && *h) return std::string(h) + "/.cache/statecraft_amd";
    return "";
}

constexpr char kCacheMagic[8] = {'S', 'C', 'K', 'C', 'O', 'D', 'E', '1'};

struct CacheKey {
    uint64_t name = 0, check = 0;
};

CacheKey cache_key(const std::string &src) {
    int major = 0, minor = 0;
    (void)hiprtcVersion(&major, &minor);
    std::string material = std::to_string(major) + "." + std::to_string(minor) + "\n";
    for (const char *o : kRtcOptions) material += std::string(o) + "\n";
    CacheKey k;
    for (uint64_t seed : {0xcbf29ce484222325ULL, 0x84222325cbf29ce4ULL}) {
        uint64_t h = fnv1a(material.data(), material.size(), seed);
        h = fnv1a(kOpsHeaderText, std::strlen(kOpsHeaderText), h);
        h = fnv1a(src.data(), src.size(), h);
        (k.name ? k.check : k.name) = h;
    }
    return k;
}

std::string cache_path(const std::string &dir, const CacheKey &k) {
    char b[40];
    std::snprintf(b, sizeof b, "/sc_%016llx.co", (unsigned long long)k.name);
    return dir + b;
}

bool cache_load(const std::string &dir, const CacheKey &k, std::vector<char> &code) {
    FILE *f = std::fopen(cache_path(dir, k).c_str(), "rb");
    if (!f) return false;
    char magic[8];
    uint64_t check = 0, size = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kCacheMagic, 8) == 0 &&
              std::fread(&check, 8, 1, f) == 1 && check == k.check && std::fread(&size, 8, 1, f) == 1 &&
              size > 0 && size < (1ULL << 30);
    if (ok) {
        code.resize(size);
        ok = std::fread(code.data(), 1, size, f) == size && std::fgetc(f) == EOF;
    }
    std::fclose(f);
    if (!ok) code.clear();
    return ok;
}

void cache_store(const std::string &dir, const CacheKey &k, const std::vector<char> &code) {
    // create the directory and its parent (0700); failures only mean no cache
    const size_t slash = dir.find_last_of('/');
    if (slash != std::string::npos && slash > 0) (void)mkdir(dir.substr(0, slash).c_str(), 0700);
    (void)mkdir(dir.c_str(), 0700);
    const std::string path = cache_path(dir, k);
    char suffix[64];
    std::snprintf(suffix, sizeof suffix, ".tmp.%ld.%zx", (long)getpid(),
                  std::hash<std::thread::id>()(std::this_thread::get_id()));
    const std::string tmp = path + suffix;
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return;
    const uint64_t size = code.size();
    bool ok = std::fwrite(kCacheMagic, 1, 8, f) == 8 && std::fwrite(&k.check, 8, 1, f) == 1 &&
              std::fwrite(&size, 8, 1, f) == 1 && std::fwrite(code.data(), 1, size, f) == size;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) (void)std::remove(tmp.c_str());
}

// Compiles one module (or takes it from the on-disk cache); code receives the code object.
int compile_module(const std::string &src, std::vector<char> &code) {
    const std::string dir = kernel_cache_dir();
    CacheKey key;
    if (!dir.empty()) {
        key = cache_key(src);
        if (cache_load(dir, key, code)) return SC_OK;
    }
    hiprtcProgram prog = nullptr;
    const char *headers[] = {kOpsHeaderText};
    const char *names[] = {"sc_ops.hpp"};
    hiprtcResult r = hiprtcCreateProgram(&prog, src.c_str(), "sc_passes.hip", 1, headers, names);
    if (r != HIPRTC_SUCCESS) return fail(SC_E_HIP, std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r));
    r = hiprtcCompileProgram(prog, (int)(sizeof kRtcOptions / sizeof kRtcOptions[0]), kRtcOptions);
    if (r != HIPRTC_SUCCESS) {
        size_t n = 0;
        std::string log;
        if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
            log.resize(n);
            (void)hiprtcGetProgramLog(prog, &log[0]);
        }
        (void)hiprtcDestroyProgram(&prog);
        return fail(SC_E_HIP, std::string("sweep pass compilation failed: ") + hiprtcGetErrorString(r) + "\n" +
                                  log.substr(0, 4000));
    }
    size_t n = 0;
    r = hiprtcGetCodeSize(prog, &n);
    if (r == HIPRTC_SUCCESS) {
        code.resize(n);
        r = hiprtcGetCode(prog, code.data());
    }
    (void)hiprtcDestroyProgram(&prog);
    if (r != HIPRTC_SUCCESS) return fail(SC_E_HIP, std::string("hiprtcGetCode: ") + hiprtcGetErrorString(r));
    if (!dir.empty()) cache_store(dir, key, code);
    return SC_OK;
}

// Compiles the given kernel bodies, split over up to `threads` modules compiled in parallel.  Kernel i of
// the result is named <prefix><i>; modules[j] holds kernels [first[j], first[j+1]).
int compile_bodies(const std::vector<std::string> &bodies, int threads, std::vector<std::vector<char>> &codes,
                   std::vector<size_t> &first, const std::string &prefix = "sc_pass_") {
    const size_t nb = bodies.size();
    const size_t groups = std::max<size_t>(1, std::min<size_t>((size_t)std::max(threads, 1), nb));
    first.assign(groups + 1, 0);
    for (size_t g = 0; g <= groups; g++) first[g] = nb * g / groups;
    std::vector<std::string> srcs(groups);
    for (size_t g = 0; g < groups; g++) {
        std::string m = "#include \"sc_ops.hpp\"\n";
        for (size_t i = first[g]; i < first[g + 1]; i++) m += replace_name(bodies[i], prefix + std::to_string(i));
        srcs[g] = std::move(m);
    }
    codes.assign(groups, {});
    std::vector<int> rcs(groups, SC_OK);
    std::vector<std::string> errs(groups);
    auto work = [&](size_t g) {
        rcs[g] = compile_module(srcs[g], codes[g]);
        if (rcs[g]) errs[g] = sc_last_error();
    };
    if (groups == 1) {
        work(0);
    } else {
        std::vector<std::thread> pool;
        for (size_t g = 0; g < groups; g++) pool.emplace_back(work, g);
        for (auto &t : pool) t.join();
    }
    for (size_t g = 0; g < groups; g++)
        if (rcs[g]) return fail(rcs[g], errs[g]);
    return SC_OK;
}

struct JitCache {
    std::mutex mu;
    std::unordered_map<std::string, hipFunction_t> fns;  // key: device id + kernel body
    std::vector<hipModule_t> modules;                    // loaded for the life of the process
};
JitCache &cache() {
    static JitCache *c = new JitCache();  // never destroyed: kernels may run until the process ends
    return *c;
}

int compile_threads() {
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(8u, hc ? hc : 1u));
}

}  // namespace

std::string jit_pass_source(const SweepPlan &plan, size_t pass, int variant) {
    return pass_source(plan, plan.passes[pass], variant == 2);
}

namespace {

// Kernels for the given bodies ("@K@" names the kernel), compiled on first use and cached per device.  The
// kernels compiled by one call are named <prefix><i> (i: position among the bodies this call compiles).
int jit_kernels(Context *ctx, const std::vector<std::string> &bodies, std::vector<hipFunction_t> &out,
                const std::string &prefix = "sc_pass_") {
    JitCache &jc = cache();
    std::lock_guard<std::mutex> lock(jc.mu);
    const std::string dev = std::to_string(ctx->device) + "\n";
    std::vector<std::string> keys(bodies.size());
    std::vector<std::string> missing;
    std::unordered_map<std::string, size_t> missing_at;
    for (size_t p = 0; p < bodies.size(); p++) {
        keys[p] = dev + bodies[p];
        if (!jc.fns.count(keys[p]) && !missing_at.count(keys[p])) {
            missing_at[keys[p]] = missing.size();
            missing.push_back(bodies[p]);
        }
    }
    if (!missing.empty()) {
        if (jc.fns.size() + missing.size() > kJitMaxKernels)
            return fail(SC_E_UNSUPPORTED, "runtime-compiled sweep kernel cache is full (" +
                                              std::to_string(kJitMaxKernels) + " kernels)");
        std::vector<std::vector<char>> codes;
        std::vector<size_t> first;
        int rc = compile_bodies(missing, compile_threads(), codes, first, prefix);
        if (rc) return rc;
        for (size_t g = 0; g + 1 < first.size(); g++) {
            hipModule_t mod = nullptr;
            SC_HIP_CHECK(hipModuleLoadData(&mod, codes[g].data()));
            jc.modules.push_back(mod);
            for (size_t i = first[g]; i < first[g + 1]; i++) {
                hipFunction_t f = nullptr;
                SC_HIP_CHECK(hipModuleGetFunction(&f, mod, (prefix + std::to_string(i)).c_str()));
                jc.fns[dev + missing[i]] = f;
            }
        }
    }
    out.resize(bodies.size());
    for (size_t p = 0; p < bodies.size(); p++) out[p] = jc.fns.at(keys[p]);
    return SC_OK;
}

}  // namespace

int jit_last_pass_terms_kernel(Context *ctx, const SweepPlan &plan, const std::vector<uint64_t> &zterms,
                               hipFunction_t &out) {
    std::vector<std::string> bodies{pass_source(plan, plan.passes.back(), ctx->jit == 2, ctx->pass_sched, &zterms)};
    std::vector<hipFunction_t> fns;
    int rc = jit_kernels(ctx, bodies, fns, "sc_terms_pass_");
    if (rc) return rc;
    out = fns[0];
    return SC_OK;
}

int jit_pass_kernels(Context *ctx, const SweepPlan &plan, std::vector<hipFunction_t> &out) {
    std::vector<std::string> bodies;
    for (size_t p = 0; p < plan.passes.size(); p++)
        bodies.push_back(pass_source(plan, plan.passes[p], ctx->jit == 2, ctx->pass_sched));
    return jit_kernels(ctx, bodies, out);
}

std::vector<uint64_t> jit_pass_probes(const SweepPlan &plan, size_t pass) {
    std::vector<uint64_t> slots;
    const SweepPass &ps = plan.passes[pass];
    for (int st = 0; st < ps.nstages; st++)
        for (int o = ps.stages[st].op_begin; o < ps.stages[st].op_end; o++)
            if (plan.ops[o].kind == SW_GRAD) slots.push_back(plan.ops[o].slot);
    return slots;
}

int jit_adjoint_kernels(Context *ctx, const SweepPlan &plan, std::vector<hipFunction_t> &out) {
    std::vector<std::string> bodies;
    for (size_t p = 0; p < plan.passes.size(); p++)
        bodies.push_back(adjoint_pass_source(plan, plan.passes[p], ctx->pass_sched >= 3));
    return jit_kernels(ctx, bodies, out);
}

int jit_compile_host(const SweepPlan &plan, int variant, int threads, double *seconds, uint64_t *code_bytes,
                     const char *dump_dir) {
    std::vector<std::string> bodies;
    for (size_t p = 0; p < plan.passes.size(); p++)
        bodies.push_back(variant == 3 ? adjoint_pass_source(plan, plan.passes[p])
                                      : pass_source(plan, plan.passes[p], variant == 2));
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::vector<char>> codes;
    std::vector<size_t> first;
    int rc = compile_bodies(bodies, threads, codes, first);
    if (rc) return rc;
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    uint64_t total = 0;
    for (auto &c : codes) total += c.size();
    if (code_bytes) *code_bytes = total;
    if (dump_dir && *dump_dir) {
        for (size_t g = 0; g < codes.size(); g++) {
            const std::string base = std::string(dump_dir) + "/sc_passes_" + std::to_string(g);
            std::string src = "#include \"sc_ops.hpp\"\n";
            for (size_t i = first[g]; i < first[g + 1]; i++) src += replace_name(bodies[i], "sc_pass_" + std::to_string(i));
            FILE *f = std::fopen((base + ".hip").c_str(), "w");
            if (!f) return fail(SC_E_ARG, "cannot write " + base + ".hip");
            std::fwrite(src.data(), 1, src.size(), f);
            std::fclose(f);
            f = std::fopen((base + ".co").c_str(), "wb");
            if (!f) return fail(SC_E_ARG, "cannot write " + base + ".co");
            std::fwrite(codes[g].data(), 1, codes[g].size(), f);
            std::fclose(f);
        }
    }
    return SC_OK;
}

uint64_t sweep_jit_grid(Context *ctx, int chunk_bits, uint64_t total_chunks) {
    // 12-qubit chunks: as many 256-thread workgroups per CU as k_sweep's occupancy allows (two 64 KiB LDS
    // chunks); 13-qubit chunks: one 512-thread workgroup per CU (128 KiB of LDS)
    const uint64_t resident = chunk_bits == kChunkBits ? (uint64_t)sweep_grid_size(ctx) : (uint64_t)ctx->cu_count;
    return std::min<uint64_t>(total_chunks, resident);
}

int sweep_terms_parts(Context *ctx, int n, int chunk_bits, int elem_count) {
    const uint64_t cpe = 1ULL << (n - chunk_bits), total = (uint64_t)elem_count * cpe;
    const uint64_t grid = sweep_jit_grid(ctx, chunk_bits, total);
    const uint64_t span = (total + grid - 1) / grid;
    // an element's chunks [e cpe, (e+1) cpe) meet at most ceil(cpe / span) + 1 workgroup ranges
    return (int)((cpe + span - 1) / span + 1);
}

namespace {

// rows[e * nt + t] = the slots of (e, t) in workgroup order (only the slots the covering workgroups wrote)
__global__ void k_terms_rows(const double *__restrict__ tpart, uint64_t elems, int nt, int tparts, uint64_t span,
                             int log_chunks, double *__restrict__ rows) {
    const uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (r >= elems * (uint64_t)nt) return;
    const uint64_t e = r / (uint64_t)nt;
    const uint64_t w0 = (e << log_chunks) / span, w1 = (((e + 1) << log_chunks) - 1) / span;
    double v = 0.0;
    for (uint64_t k = 0; k <= w1 - w0; k++) v += tpart[r * (uint64_t)tparts + k];
    rows[r] = v;
}

}  // namespace

int launch_sweep_jit_terms(Context *ctx, State *st, hipFunction_t fn, const double *mats_dev, size_t mat_stride,
                           int chunk_bits, int n_terms, double *tpart, int tparts, double *rows) {
    if (mat_stride == 0) return fail(SC_E_STATE, "internal error: term sums need one matrix pool per element");
    int log_chunks = st->n - chunk_bits;
    uint64_t total = (uint64_t)st->batch << log_chunks;
    const int threads = 1 << (chunk_bits - kRegBits);
    const uint64_t grid = sweep_jit_grid(ctx, chunk_bits, total);
    const uint64_t span = (total + grid - 1) / grid;
    if ((uint64_t)tparts < ((1ULL << log_chunks) + span - 1) / span + 1)
        return fail(SC_E_STATE, "internal error: too few term partial slots");
    v2d *amp = reinterpret_cast<v2d *>(st->amp);
    v2d *out = amp;
    const v2d *mats = reinterpret_cast<const v2d *>(mats_dev);
    int n = st->n, elem_begin = 0;
    uint64_t stride = (uint64_t)mat_stride;
    long long basis = -1;
    void *args[] = {&amp, &out, &n, &log_chunks, &total, &mats, &stride, &elem_begin, &basis, &tpart, &tparts};
    {
        ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_SWEEP, 32.0 * (double)((uint64_t)st->batch << st->n));
        SC_HIP_CHECK(hipModuleLaunchKernel(fn, (unsigned)grid, 1, 1, threads, 1, 1, 0, ctx->stream, args, nullptr));
    }
    const uint64_t nrows = (uint64_t)st->batch * (uint64_t)n_terms;
    hipLaunchKernelGGL(k_terms_rows, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, ctx->stream, tpart,
                       (uint64_t)st->batch, n_terms, tparts, span, log_chunks, rows);
    SC_HIP_CHECK(hipGetLastError());
    return SC_OK;
}

int launch_sweep_jit(Context *ctx, State *st, hipFunction_t fn, const double *mats_dev, int elem_begin,
                     int elem_count, size_t mat_stride, int chunk_bits, const double2 *src, double2 *dst,
                     long long basis) {
    int log_chunks = st->n - chunk_bits;
    uint64_t total = (uint64_t)elem_count << log_chunks;
    const int threads = 1 << (chunk_bits - kRegBits);
    const uint64_t grid = sweep_jit_grid(ctx, chunk_bits, total);
    // a pass reads every chunk from src and writes it to dst (the same addresses; in place when equal)
    v2d *amp = reinterpret_cast<v2d *>(const_cast<double2 *>(src ? src : st->amp));
    v2d *out = reinterpret_cast<v2d *>(dst ? dst : st->amp);
    const v2d *mats = reinterpret_cast<const v2d *>(mats_dev);
    int n = st->n;
    uint64_t stride = (uint64_t)mat_stride;
    void *args[] = {&amp, &out, &n, &log_chunks, &total, &mats, &stride, &elem_begin, &basis};
    // algorithmic bytes: a read and a write of every amplitude, or the write alone when the pass generates
    // its basis-state input instead of loading it
    const double per_amp = basis >= 0 ? 16.0 : 32.0;
    ProfScope ps(&ctx->prof, ctx->stream, SC_KFAM_SWEEP, per_amp * (double)((uint64_t)elem_count << st->n));
    SC_HIP_CHECK(hipModuleLaunchKernel(fn, (unsigned)grid, 1, 1, threads, 1, 1, 0, ctx->stream, args, nullptr));
    return SC_OK;
}

}  // namespace sc
