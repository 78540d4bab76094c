// sc_seqsum.hpp -- exact sequential running sums computed segment-parallel (sc_seqsum.hip).
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "sc_internal.hpp"

namespace sc {
namespace seq {

constexpr int kSegBits = 12;
constexpr uint64_t kSeg = 1ULL << kSegBits;  // values per segment
constexpr int32_t kNoBinade = -100000;
// Segments whose running sum lies outside [2^kMinBinade, 2^(kMaxBinade+1)) are summed serially.
constexpr int kMinBinade = -900;
constexpr int kMaxBinade = 900;

enum : uint8_t { kZero = 0, kFast = 1, kSerial = 2 };

// Where the values come from.  State source: value j of sequence o is |amp[pdep(o, mmask) | pdep(j, cmask)]|^2
// (contiguous: |amp[j]|^2); array source: arr[o * len + j].  Either is multiplied by scale when scale != 1.
struct Source {
    const double *amp;  // interleaved complex amplitudes, or null
    const double *arr;  // values, or null
    uint64_t mmask, cmask;
    uint64_t len;   // values per sequence
    uint64_t nseq;  // sequences
    uint64_t segs_per_seq;
    int contiguous;
    double scale;
};

// The exact running sums, per segment: the sum before its first value, and how the segment is summed
// (kFast: integers on the grid of `binade`; kSerial: the sequential loop; kZero: all values zero).
struct Plan {
    std::vector<double> start;
    std::vector<int32_t> binade;
    std::vector<uint8_t> mode;
    std::vector<double> total;  // per sequence: the sequential sum of all its values
    uint64_t fast = 0, serial = 0, zero = 0, gather_rounds = 0;
};

// The three per-segment queries the chain needs (device kernels, or their host equivalents for tests).
struct Provider {
    virtual ~Provider() = default;
    virtual int approx(std::vector<double> &out) = 0;
    virtual int int_sums(const std::vector<int32_t> &cand, std::vector<double> &isum, std::vector<uint32_t> &ties) = 0;
    virtual int gather(const std::vector<uint64_t> &segs, std::vector<double> &vals) = 0;
};

struct SampleStats {
    uint64_t fast = 0, serial = 0, zero = 0, gather_rounds = 0;
    bool rescaled = false;
};

int plan_sequences(Provider &pv, const Source &src, Plan &plan);

// Exact marginal probabilities (StatevectorEngine.computeMeasurementProbabilities :274-296) of batch element
// state `amp` into out_dev (2^m doubles).
int exact_marginals(Context *ctx, const double2 *amp, int n, const int32_t *measured, int m, double *out_dev);

// normalize + cumulative + sampleOutcome (:298-338) for every uniform, exactly: over the whole state in index
// order when marginals_dev is null (every qubit measured in order), else over the 2^m marginals_dev values.
// u_host and u_dev hold the same uniforms (the host groups the shots by segment).
int exact_sample(Context *ctx, const double2 *amp, int n, const int32_t *measured, int m, const double *marginals_dev,
                 const double *u_host, const double *u_dev, int64_t shots, int64_t *out_dev, double *gap_dev,
                 SampleStats *stats);

}  // namespace seq
}  // namespace sc
