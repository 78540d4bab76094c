// sc_jit.hpp -- sweep passes compiled at run time (sc_jit.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "sc_internal.hpp"

namespace sc {

struct SweepPlan;

// Most pass kernels kept loaded per process (each holds its pass's straight-line code, ~20-200 KB).
constexpr size_t kJitMaxKernels = 4096;

// Kernel source of one pass of a plan (the kernel is named "@K@"); variant = the SC_OPT_JIT value.
std::string jit_pass_source(const SweepPlan &plan, size_t pass, int variant);
// One kernel per pass of the plan, compiled on first use and cached per device by source text.
int jit_pass_kernels(Context *ctx, const SweepPlan &plan, std::vector<hipFunction_t> &out);
// The plan's last pass with Z-string term sums (masks zterms, at most 32): besides storing the states it writes
// tpart[(element * terms + t) * tparts + slot] = sum over the element's amplitudes in this workgroup's chunk
// range of (-1)^parity(z_t & i) |a_i|^2.  Launched by launch_sweep_jit_terms with one matrix pool per element
// (contiguous chunk ranges per workgroup); sweep_terms_reduce adds the slots.
int jit_last_pass_terms_kernel(Context *ctx, const SweepPlan &plan, const std::vector<uint64_t> &zterms,
                               hipFunction_t &out);
// Compiles every pass of a plan without a device (timing, and the code objects for inspection when
// dump_dir is set).
// variant 3: the plan is an adjoint plan (gradient probes) and compiles as adjoint passes.
int jit_compile_host(const SweepPlan &plan, int variant, int threads, double *seconds, uint64_t *code_bytes,
                     const char *dump_dir);
// The reverse sweep of an adjoint gradient over ops[0, n_ops) for one batch element, planned as sweep passes
// (sc_adjoint.hip); inv receives the inverse matrices the plan's pool is built from, slot each op's
// gradient slot (-1 for none).
int adjoint_plan_host(int n, const sc_op *ops, uint64_t n_ops, const double *mats, uint64_t n_mat_complex,
                      int n_params, int low_qubits, SweepPlan &plan, std::vector<double> &inv, std::vector<int> &slot);
// Adjoint passes (plans with gradient probes): one kernel per pass over psi and lambda; probe slots of a pass
// in the order its kernel writes them.
int jit_adjoint_kernels(Context *ctx, const SweepPlan &plan, std::vector<hipFunction_t> &out);
std::vector<uint64_t> jit_pass_probes(const SweepPlan &plan, size_t pass);
// src / dst: where the pass reads and writes the states (nullptr: st->amp; equal pointers: in place).
// basis >= 0: the pass's input is |basis> in every element (generated, src unread).
int launch_sweep_jit(Context *ctx, State *st, hipFunction_t fn, const double *mats_dev, int elem_begin,
                     int elem_count, size_t mat_stride, int chunk_bits, const double2 *src = nullptr,
                     double2 *dst = nullptr, long long basis = -1);
// Workgroups of a sweep launch that are resident at once (occupancy x CUs, capped by the context option).
int sweep_grid_size(Context *ctx);
// Grid of a runtime-compiled sweep launch over total chunks, and the chunks per workgroup of its blocked
// (one matrix pool per element) assignment.
uint64_t sweep_jit_grid(Context *ctx, int chunk_bits, uint64_t total_chunks);
// Partial-sum slots per (element, term) of a terms pass over elem_count elements of n qubits.
int sweep_terms_parts(Context *ctx, int n, int chunk_bits, int elem_count);
// launch_sweep_jit for a kernel of jit_last_pass_terms_kernel (in place, per-element pools, all elements in one
// launch), then the slot sums into rows[element * n_terms + t].
int launch_sweep_jit_terms(Context *ctx, State *st, hipFunction_t fn, const double *mats_dev, size_t mat_stride,
                           int chunk_bits, int n_terms, double *tpart, int tparts, double *rows);

}  // namespace sc
