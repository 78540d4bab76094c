/*
 * sv_oracle.c -- CPU restatement of StateCraft's `statevector` engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP engine is
 * compared against; nothing in the product links or calls it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * It follows the reference Java text operation by operation, in plain scalar
 * C with 64-bit indices, compiled with -ffp-contract=off so that every product
 * is rounded separately like Java's `double` arithmetic.  Paths are relative to
 * the StateCraft tree:
 *   E = engines/src/main/java/com/omaarr90/statecraft/engines/statevector/StatevectorEngine.java
 *   O = engines/src/main/java/com/omaarr90/statecraft/engines/statevector/StatevectorOps.java
 *   A = core/src/main/java/com/omaarr90/statecraft/core/math/ComplexArrays.java
 *
 * Every entry point is pinned by tests/test_oracle_*.py against golden vectors
 * (reference test expectations, closed forms, brute-force Kronecker matrices,
 * published SplitMix64 outputs); see the header of each test.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/statecraft_hip.h"

#define ORC_OK 0
#define ORC_E_ARG 2
#define ORC_E_UNSUPPORTED 3
#define ORC_E_STATE 4

/* ------------------------------------------------------------------------- */
/* java.util.SplittableRandom (JDK, not in the reference tree; call sites      */
/* E:81-85, E:238-240, E:245, E:252, E:489).  Published algorithm: SplitMix64, */
/* Steele, Lea & Flood, "Fast splittable pseudorandom number generators",      */
/* OOPSLA 2014: seed += GOLDEN_GAMMA; output = mix64(seed) (Stafford variant   */
/* 13); nextDouble = (nextLong >>> 11) * 2^-53; nextDouble(bound) multiplies   */
/* and clamps to nextDown(bound).                                             */
/* ------------------------------------------------------------------------- */
#define SR_GOLDEN_GAMMA 0x9e3779b97f4a7c15ULL

static uint64_t sr_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

uint64_t orc_sr_next_long(uint64_t *seed) {
    *seed += SR_GOLDEN_GAMMA;
    return sr_mix64(*seed);
}

double orc_sr_next_double(uint64_t *seed) {
    return (double)(orc_sr_next_long(seed) >> 11) * 0x1.0p-53;
}

double orc_sr_next_double_bounded(uint64_t *seed, double bound) {
    double r = orc_sr_next_double(seed);
    r = r * bound;
    if (r >= bound) r = nextafter(bound, -INFINITY);
    return r;
}

/* Fill `out` with `count` nextDouble() values of SplittableRandom(seed). */
void orc_sr_doubles(uint64_t seed, double *out, int64_t count) {
    uint64_t s = seed;
    for (int64_t i = 0; i < count; i++) out[i] = orc_sr_next_double(&s);
}

/* ------------------------------------------------------------------------- */
/* Gate kernels (scalar forms of StatevectorOps).                             */
/* ------------------------------------------------------------------------- */

/* O:62-82 (scalar tail of applySingleGate) and O:830-849 (applySingleGatePair). */
void orc_apply_single(double *state, int n, int target, const double *g) {
    const double g00r = g[0], g00i = g[1], g01r = g[2], g01i = g[3];
    const double g10r = g[4], g10i = g[5], g11r = g[6], g11i = g[7];
    const uint64_t dimension = 1ULL << n;
    const uint64_t stride = 1ULL << target;
    const uint64_t period = stride << 1;
    for (uint64_t base = 0; base < dimension; base += period) {
        for (uint64_t offset = 0; offset < stride; offset++) {
            uint64_t off0 = (base + offset) << 1;
            uint64_t off1 = (base + offset + stride) << 1;
            double a0r = state[off0], a0i = state[off0 + 1];
            double a1r = state[off1], a1i = state[off1 + 1];
            double new0r = g00r * a0r - g00i * a0i + g01r * a1r - g01i * a1i;
            double new0i = g00r * a0i + g00i * a0r + g01r * a1i + g01i * a1r;
            double new1r = g10r * a0r - g10i * a0i + g11r * a1r - g11i * a1i;
            double new1i = g10r * a0i + g10i * a0r + g11r * a1i + g11i * a1r;
            state[off0] = new0r;
            state[off0 + 1] = new0i;
            state[off1] = new1r;
            state[off1 + 1] = new1i;
        }
    }
}

/* O:128-145 (serial applyCnot). */
void orc_apply_cnot(double *state, int n, int control, int target) {
    const uint64_t controlMask = 1ULL << control, targetMask = 1ULL << target;
    const uint64_t dimension = 1ULL << n;
    for (uint64_t index = 0; index < dimension; index++) {
        if ((index & controlMask) == controlMask && (index & targetMask) == 0) {
            uint64_t a = index << 1, b = (index | targetMask) << 1;
            double tr = state[a], ti = state[a + 1];
            state[a] = state[b];
            state[a + 1] = state[b + 1];
            state[b] = tr;
            state[b + 1] = ti;
        }
    }
}

/* O:851-929 (applyTwoQubitUnitaryGroups): local index l = 2*bit(first) + bit(second). m = 16 complex row-major. */
void orc_apply_two(double *state, int n, int first, int second, const double *m) {
    const uint64_t fm = 1ULL << first, sm = 1ULL << second;
    const uint64_t dimension = 1ULL << n;
    for (uint64_t base = 0; base < dimension; base++) {
        if ((base & fm) || (base & sm)) continue;
        uint64_t off[4] = {base << 1, (base | sm) << 1, (base | fm) << 1, (base | fm | sm) << 1};
        double ar[4], ai[4];
        for (int c = 0; c < 4; c++) {
            ar[c] = state[off[c]];
            ai[c] = state[off[c] + 1];
        }
        for (int r = 0; r < 4; r++) {
            const double *row = m + 8 * r;
            /* left-to-right sums exactly as written at O:906-928 */
            double re = row[0] * ar[0] - row[1] * ai[0] + row[2] * ar[1] - row[3] * ai[1] + row[4] * ar[2] -
                        row[5] * ai[2] + row[6] * ar[3] - row[7] * ai[3];
            double im = row[0] * ai[0] + row[1] * ar[0] + row[2] * ai[1] + row[3] * ar[1] + row[4] * ai[2] +
                        row[5] * ar[2] + row[6] * ai[3] + row[7] * ar[3];
            state[off[r]] = re;
            state[off[r] + 1] = im;
        }
    }
}

/* O:931-972 (applyTwoQubitDiagonalGroups). d = 4 complex. */
void orc_apply_diag2(double *state, int n, int first, int second, const double *d) {
    const uint64_t fm = 1ULL << first, sm = 1ULL << second;
    const uint64_t dimension = 1ULL << n;
    for (uint64_t idx = 0; idx < dimension; idx++) {
        int l = ((idx & fm) ? 2 : 0) | ((idx & sm) ? 1 : 0);
        uint64_t o = idx << 1;
        double ar = state[o], ai = state[o + 1];
        double dr = d[2 * l], di = d[2 * l + 1];
        state[o] = dr * ar - di * ai;
        state[o + 1] = dr * ai + di * ar;
    }
}

/* O:543-629 (applySwap): exchange |..1..0..> and |..0..1..>. */
void orc_apply_swap(double *state, int n, int first, int second) {
    const uint64_t fm = 1ULL << first, sm = 1ULL << second;
    const uint64_t dimension = 1ULL << n;
    for (uint64_t base = 0; base < dimension; base++) {
        if ((base & fm) || (base & sm)) continue;
        uint64_t a = (base | fm) << 1, b = (base | sm) << 1;
        double tr = state[a], ti = state[a + 1];
        state[a] = state[b];
        state[a + 1] = state[b + 1];
        state[b] = tr;
        state[b + 1] = ti;
    }
}

/* O:631-703 and O:786-828 (applyMultiControlledSingleGate): 2x2 on pairs whose index has every control bit. */
void orc_apply_mcu(double *state, int n, int target, uint64_t controlMask, const double *g) {
    const double g00r = g[0], g00i = g[1], g01r = g[2], g01i = g[3];
    const double g10r = g[4], g10i = g[5], g11r = g[6], g11i = g[7];
    const uint64_t dimension = 1ULL << n;
    const uint64_t stride = 1ULL << target;
    for (uint64_t idx0 = 0; idx0 < dimension; idx0++) {
        if (idx0 & stride) continue;
        if ((idx0 & controlMask) != controlMask) continue;
        uint64_t off0 = idx0 << 1, off1 = (idx0 | stride) << 1;
        double a0r = state[off0], a0i = state[off0 + 1];
        double a1r = state[off1], a1i = state[off1 + 1];
        state[off0] = g00r * a0r - g00i * a0i + g01r * a1r - g01i * a1i;
        state[off0 + 1] = g00r * a0i + g00i * a0r + g01r * a1i + g01i * a1r;
        state[off1] = g10r * a0r - g10i * a0i + g11r * a1r - g11i * a1i;
        state[off1 + 1] = g10r * a0i + g10i * a0r + g11r * a1i + g11i * a1r;
    }
}

/* The op loop of E:90-151, over the same sc_op encoding the HIP engine takes.
 * Returns ORC_E_UNSUPPORTED for a unitary after a measurement (E:92-95). */
int orc_run_ops(double *state, int n, const sc_op *ops, int64_t n_ops, const double *mats) {
    int measurementSeen = 0;
    for (int64_t i = 0; i < n_ops; i++) {
        const sc_op *op = &ops[i];
        if (op->kind == SC_OP_MEASURE) {
            measurementSeen = 1;
            continue;
        }
        if (measurementSeen) return ORC_E_UNSUPPORTED;
        const double *m = op->mat >= 0 ? mats + 2 * (int64_t)op->mat : NULL;
        switch (op->kind) {
        case SC_OP_SINGLE: orc_apply_single(state, n, op->q0, m); break;
        case SC_OP_CNOT: orc_apply_cnot(state, n, op->q0, op->q1); break;
        case SC_OP_TWO: orc_apply_two(state, n, op->q0, op->q1, m); break;
        case SC_OP_DIAG2: orc_apply_diag2(state, n, op->q0, op->q1, m); break;
        case SC_OP_SWAP: orc_apply_swap(state, n, op->q0, op->q1); break;
        case SC_OP_MCU: orc_apply_mcu(state, n, op->q0, op->controls, m); break;
        default: return ORC_E_ARG;
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* Measurement (E:229-338).                                                    */
/* ------------------------------------------------------------------------- */

/* E:274-296 computeMeasurementProbabilities: accumulate in basis order. */
void orc_probabilities(const double *state, int n, const int32_t *measured, int m, double *out) {
    const uint64_t outcomes = 1ULL << m;
    memset(out, 0, outcomes * sizeof(double));
    const uint64_t dimension = 1ULL << n;
    int identity = (m == n);
    for (int bit = 0; identity && bit < m; bit++) identity = measured[bit] == bit;
    if (identity) { /* every qubit in order: outcome == basis, probabilities[basis] = 0.0 + p */
        for (uint64_t basis = 0; basis < dimension; basis++) {
            double re = state[basis << 1], im = state[(basis << 1) + 1];
            double p = (re * re) + (im * im);
            if (p != 0.0) out[basis] += p;
        }
        return;
    }
    for (uint64_t basis = 0; basis < dimension; basis++) {
        double re = state[basis << 1], im = state[(basis << 1) + 1];
        double p = (re * re) + (im * im);
        if (p == 0.0) continue;
        uint64_t outcome = 0;
        for (int bit = 0; bit < m; bit++)
            if ((basis >> measured[bit]) & 1ULL) outcome |= (1ULL << bit);
        out[outcome] += p;
    }
}

/* E:298-312 normalize: sequential sum; rescale by 1/sum only when |sum-1| > 1e-12. */
int orc_normalize(double *p, uint64_t count) {
    double sum = 0.0;
    for (uint64_t i = 0; i < count; i++) sum += p[i];
    if (sum == 0.0) return ORC_E_STATE;
    if (fabs(sum - 1.0) > 1e-12) {
        double scale = 1.0 / sum;
        for (uint64_t i = 0; i < count; i++) p[i] *= scale;
    }
    return ORC_OK;
}

/* E:314-323 cumulative: running sum, last entry forced to 1.0. */
void orc_cumulative(const double *p, double *cum, uint64_t count) {
    double running = 0.0;
    for (uint64_t i = 0; i < count; i++) {
        running += p[i];
        cum[i] = running;
    }
    cum[count - 1] = 1.0;
}

/* java.util.Arrays.binarySearch(double[], double) (JDK, published behaviour):
 * returns an index of key, or -(insertion point)-1; ties on equal values are
 * broken by the raw bit patterns (so -0.0 < 0.0). */
static int64_t java_binary_search(const double *a, int64_t len, double key) {
    int64_t low = 0, high = len - 1;
    while (low <= high) {
        int64_t mid = (int64_t)(((uint64_t)low + (uint64_t)high) >> 1);
        double midVal = a[mid];
        if (midVal < key) {
            low = mid + 1;
        } else if (midVal > key) {
            high = mid - 1;
        } else {
            int64_t midBits, keyBits;
            memcpy(&midBits, &midVal, 8);
            memcpy(&keyBits, &key, 8);
            if (midBits == keyBits) return mid;
            if (midBits < keyBits) low = mid + 1;
            else high = mid - 1;
        }
    }
    return -(low + 1);
}

/* E:325-338 sampleOutcome. */
int64_t orc_sample_outcome(const double *cum, int64_t len, double value) {
    if (value >= 1.0) return len - 1;
    int64_t index = java_binary_search(cum, len, value);
    if (index >= 0) {
        while (index > 0 && cum[index - 1] >= value) index--;
        return index;
    }
    return -index - 1;
}

/* E:229-257 performMeasurement with the uniforms given explicitly (shot order). */
int orc_sample(const double *state, int n, const int32_t *measured, int m, const double *uniforms, int64_t shots,
               int64_t *outcomes) {
    const uint64_t count = 1ULL << m;
    double *p = (double *)malloc(count * sizeof(double));
    double *cum = (double *)malloc(count * sizeof(double));
    if (!p || !cum) {
        free(p);
        free(cum);
        return ORC_E_STATE;
    }
    orc_probabilities(state, n, measured, m, p);
    int rc = orc_normalize(p, count);
    if (rc == ORC_OK) {
        orc_cumulative(p, cum, count);
        for (int64_t s = 0; s < shots; s++) outcomes[s] = orc_sample_outcome(cum, (int64_t)count, uniforms[s]);
    }
    free(p);
    free(cum);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Noise (E:382-600, A:74-177).                                               */
/* ------------------------------------------------------------------------- */

/* A:74-143 norm2Sq: LASSQ scaled sum of squares, NaN/Inf short-circuit. */
double orc_norm2sq(const double *a, uint64_t count) {
    double scale = 0.0, ssq = 1.0;
    for (uint64_t i = 0; i < 2 * count; i += 2) {
        double re = a[i], im = a[i + 1];
        if (isnan(re) || isnan(im)) return NAN;
        if (isinf(re) || isinf(im)) return INFINITY;
        double vals[2] = {fabs(re), fabs(im)};
        for (int k = 0; k < 2; k++) {
            double av = vals[k];
            if (av != 0.0) {
                if (scale < av) {
                    double r = scale / av;
                    ssq = 1.0 + ssq * r * r;
                    scale = av;
                } else if (scale != 0.0) {
                    double r = av / scale;
                    ssq += r * r;
                } else {
                    scale = av;
                }
            }
        }
    }
    return scale == 0.0 ? 0.0 : (scale * scale) * ssq;
}

/* E:590-600 renormalizeState (+ A:159-177 scal with a real factor). Returns ORC_E_STATE on an invalid norm. */
int orc_renormalize(double *state, int n, double *norm_sq_out) {
    const uint64_t dimension = 1ULL << n;
    double normSq = orc_norm2sq(state, dimension);
    if (norm_sq_out) *norm_sq_out = normSq;
    if (!isfinite(normSq) || normSq == 0.0) return ORC_E_STATE;
    if (fabs(normSq - 1.0) > 1e-12) {
        double scale = 1.0 / sqrt(normSq);
        for (uint64_t i = 0; i < dimension; i++) {
            double xr = state[2 * i], xi = state[2 * i + 1];
            state[2 * i] = scale * xr - 0.0 * xi;
            state[2 * i + 1] = scale * xi + 0.0 * xr;
        }
    }
    return ORC_OK;
}

/* E:511-530 computeBranchWeight for one single-qubit Kraus operator k (4 complex). */
double orc_kraus_weight(const double *state, int n, int target, const double *k) {
    const uint64_t dimension = 1ULL << n, tm = 1ULL << target;
    double weight = 0.0;
    for (uint64_t basis = 0; basis < dimension; basis++) {
        if (basis & tm) continue;
        double in[4] = {state[basis << 1], state[(basis << 1) + 1], state[(basis | tm) << 1],
                        state[((basis | tm) << 1) + 1]};
        double out[4];
        for (int row = 0; row < 2; row++) { /* E:550-567 applyKrausMatrix */
            double sr = 0.0, si = 0.0;
            for (int col = 0; col < 2; col++) {
                double cr = k[2 * (2 * row + col)], ci = k[2 * (2 * row + col) + 1];
                sr += cr * in[2 * col] - ci * in[2 * col + 1];
                si += cr * in[2 * col + 1] + ci * in[2 * col];
            }
            out[2 * row] = sr;
            out[2 * row + 1] = si;
        }
        double normSq = 0.0; /* E:569-577 blockNormSquared */
        for (int j = 0; j < 4; j += 2) normSq += out[j] * out[j] + out[j + 1] * out[j + 1];
        weight += normSq;
    }
    return weight;
}

/* E:579-588 sampleWeightedIndex. */
int orc_sample_weighted_index(const double *weights, int count, double value) {
    double cumulative = 0.0;
    for (int i = 0; i < count; i++) {
        cumulative += weights[i];
        if (value < cumulative) return i;
    }
    return count - 1;
}

/* One single-qubit channel on one target: E:463-491 sampleKrausOperator then E:426-439 applySingleQubitKraus.
 * kraus: k operators x 4 complex.  Draws one rng.nextDouble(total) from *seed.  chosen (may be NULL) receives
 * the operator index. */
int orc_apply_channel_1q(double *state, int n, int target, const double *kraus, int k, uint64_t *seed,
                         int *chosen) {
    double weights[16];
    double total = 0.0;
    if (k > 16) return ORC_E_ARG;
    for (int i = 0; i < k; i++) {
        double w = orc_kraus_weight(state, n, target, kraus + 8 * i);
        if (!isfinite(w) || w < 0.0) return ORC_E_STATE;
        weights[i] = w;
        total += w;
    }
    if (!isfinite(total) || total <= 1e-15) return ORC_E_STATE; /* MIN_KRAUS_WEIGHT E:36 */
    int idx = orc_sample_weighted_index(weights, k, orc_sr_next_double_bounded(seed, total));
    if (chosen) *chosen = idx;
    orc_apply_single(state, n, target, kraus + 8 * idx);
    return orc_renormalize(state, n, NULL);
}

/* A noisy run with an explicit schedule: after op i, for j in [sched_begin[i], sched_begin[i+1]) apply channel
 * sched_channel[j] (an index into `channels`, each `channel_k[c]` operators starting at channel_off[c] in
 * `kraus`) on qubit sched_target[j].  This is the schedule NoiseModel.scheduledChannelsAfter
 * (core/.../noise/NoiseModel.java:211-241) yields, flattened by the host. */
int orc_run_noisy(double *state, int n, const sc_op *ops, int64_t n_ops, const double *mats,
                  const int64_t *sched_begin, const int32_t *sched_channel, const int32_t *sched_target,
                  const double *kraus, const int32_t *channel_off, const int32_t *channel_k, uint64_t seed) {
    uint64_t s = seed;
    int measurementSeen = 0;
    for (int64_t i = 0; i < n_ops; i++) {
        if (ops[i].kind == SC_OP_MEASURE) {
            measurementSeen = 1;
            continue;
        }
        if (measurementSeen) return ORC_E_UNSUPPORTED;
        int rc = orc_run_ops(state, n, &ops[i], 1, mats);
        if (rc) return rc;
        for (int64_t j = sched_begin[i]; j < sched_begin[i + 1]; j++) {
            int c = sched_channel[j];
            rc = orc_apply_channel_1q(state, n, sched_target[j], kraus + 2 * (int64_t)channel_off[c], channel_k[c],
                                      &s, NULL);
            if (rc) return rc;
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* Observables (no reference counterpart; closed-form Pauli algebra).          */
/* ------------------------------------------------------------------------- */

/* <psi|P|psi> for P = i^{|x&z|} X^x Z^z, i.e. Y wherever both masks have the bit.  Returns the real part. */
double orc_expect_pauli_term(const double *state, int n, uint64_t xmask, uint64_t zmask) {
    const uint64_t dimension = 1ULL << n;
    int ny = __builtin_popcountll(xmask & zmask) & 3;
    double acc_r = 0.0, acc_i = 0.0;
    for (uint64_t i = 0; i < dimension; i++) {
        uint64_t j = i ^ xmask;
        double sign = (__builtin_popcountll(i & zmask) & 1) ? -1.0 : 1.0;
        /* conj(psi_j) * psi_i * sign */
        double pr = state[2 * j] * state[2 * i] + state[2 * j + 1] * state[2 * i + 1];
        double pi = state[2 * j] * state[2 * i + 1] - state[2 * j + 1] * state[2 * i];
        acc_r += sign * pr;
        acc_i += sign * pi;
    }
    /* multiply by i^ny */
    switch (ny) {
    case 0: return acc_r;
    case 1: return -acc_i;
    case 2: return -acc_r;
    default: return acc_i;
    }
}

double orc_expect_pauli(const double *state, int n, const uint64_t *xmask, const uint64_t *zmask,
                        const double *coeff, int n_terms) {
    double e = 0.0;
    for (int t = 0; t < n_terms; t++) e += coeff[t] * orc_expect_pauli_term(state, n, xmask[t], zmask[t]);
    return e;
}

/* Probability checksum weight w(i) = (mix64((i + 1) * GOLDEN_GAMMA) >>> 11) * 2^-53, i.e. the first
 * nextDouble() of SplittableRandom(i * GOLDEN_GAMMA); the HIP engine uses the same definition. */
double orc_checksum_weight(uint64_t i) {
    uint64_t s = i * SR_GOLDEN_GAMMA;
    return (double)(orc_sr_next_long(&s) >> 11) * 0x1.0p-53;
}

void orc_probability_checksum(const double *state, int n, double out[2]) {
    const uint64_t dimension = 1ULL << n;
    double s0 = 0.0, s1 = 0.0;
    for (uint64_t i = 0; i < dimension; i++) {
        double p = state[2 * i] * state[2 * i] + state[2 * i + 1] * state[2 * i + 1];
        s0 += p;
        s1 += p * orc_checksum_weight(i);
    }
    out[0] = s0;
    out[1] = s1;
}
