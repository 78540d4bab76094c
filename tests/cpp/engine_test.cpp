// engine_test.cpp -- the reference's StatevectorEngineTest / StatevectorNoiseTest /
// StatevectorAlgorithmSuiteTest cases, restated against HipStatevectorEngine through the C++ mirror.
// Expected states come from `apply_scalar`, a scalar restatement of QuantumCircuit.apply
// (core/src/main/java/com/omaarr90/statecraft/quantum/QuantumCircuit.java:144-360), the reference's own
// test oracle.  Exit status = number of failed checks.
#include <cmath>
#include <cstdio>
#include <functional>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "../../statecraft_amd/engine/statecraft.hpp"

using namespace statecraft;

static int failures = 0;
static constexpr double EPS = 1e-12;

#define CHECK(cond, what)                                                      \
    do {                                                                       \
        if (!(cond)) {                                                         \
            std::printf("FAIL %s: %s (line %d)\n", current, what, __LINE__);   \
            failures++;                                                        \
        }                                                                      \
    } while (0)

static const char *current = "";

// QuantumCircuit.apply (scalar ComplexNumber evaluator).
static std::vector<Complex> apply_scalar(const QuantumCircuit &c, std::vector<Complex> s) {
    for (const Operation &op : c.operations()) {
        const size_t dim = s.size();
        switch (op.kind) {
        case Operation::Kind::SingleGate:
        case Operation::Kind::MultiControl: {
            size_t t = size_t(1) << op.q0, cm = 0;
            for (int q : op.qubits) cm |= size_t(1) << q;
            if (op.kind == Operation::Kind::SingleGate) cm = 0;
            for (size_t b = 0; b < dim; b++) {
                if ((b & t) || (b & cm) != cm) continue;
                Complex a0 = s[b], a1 = s[b | t];
                s[b] = op.matrix[0] * a0 + op.matrix[1] * a1;
                s[b | t] = op.matrix[2] * a0 + op.matrix[3] * a1;
            }
            break;
        }
        case Operation::Kind::Cnot: {
            size_t cmask = size_t(1) << op.q0, tm = size_t(1) << op.q1;
            for (size_t b = 0; b < dim; b++)
                if ((b & cmask) && !(b & tm)) std::swap(s[b], s[b | tm]);
            break;
        }
        case Operation::Kind::TwoQubitGate:
        case Operation::Kind::TwoQubitDiagonal: {
            size_t fm = size_t(1) << op.q0, sm = size_t(1) << op.q1;
            for (size_t b = 0; b < dim; b++) {
                if ((b & fm) || (b & sm)) continue;
                size_t idx[4] = {b, b | sm, b | fm, b | fm | sm};
                Complex a[4] = {s[idx[0]], s[idx[1]], s[idx[2]], s[idx[3]]};
                for (int r = 0; r < 4; r++) {
                    if (op.kind == Operation::Kind::TwoQubitDiagonal) {
                        s[idx[r]] = op.matrix[r] * a[r];
                    } else {
                        Complex acc = 0;
                        for (int k = 0; k < 4; k++) acc += op.matrix[4 * r + k] * a[k];
                        s[idx[r]] = acc;
                    }
                }
            }
            break;
        }
        case Operation::Kind::Swap: {
            size_t fm = size_t(1) << op.q0, sm = size_t(1) << op.q1;
            for (size_t b = 0; b < dim; b++)
                if (!(b & fm) && !(b & sm)) std::swap(s[b | fm], s[b | sm]);
            break;
        }
        case Operation::Kind::Measure: throw UnsupportedOperationException("measure");
        }
    }
    return s;
}

static std::vector<Complex> zero(int n) {
    std::vector<Complex> s(size_t(1) << n, 0.0);
    s[0] = 1.0;
    return s;
}

static bool matches(const std::vector<Complex> &want, const StateVector &got) {
    for (size_t i = 0; i < want.size(); i++)
        if (std::fabs(want[i].real() - got.amplitude(i).real()) > EPS ||
            std::fabs(want[i].imag() - got.amplitude(i).imag()) > EPS)
            return false;
    return true;
}

static StateVector to_state(int n, const std::vector<Complex> &v) {
    std::vector<double> d;
    for (auto &c : v) d.push_back(c.real()), d.push_back(c.imag());
    return StateVector(n, d);
}

static std::vector<Complex> random_state(int n, std::mt19937_64 &g) {
    std::normal_distribution<double> N;
    std::vector<Complex> v(size_t(1) << n);
    double ns = 0;
    for (auto &c : v) c = Complex(N(g), N(g)), ns += std::norm(c);
    for (auto &c : v) c /= std::sqrt(ns);
    return v;
}

static std::vector<Complex> random_unitary4(std::mt19937_64 &g) {
    // product of random 1q rotations and CNOTs on 2 qubits, as the reference test builds its matrices
    QuantumCircuit c(2);
    std::uniform_real_distribution<double> U(0, 6.283185307179586);
    for (int i = 0; i < 6; i++) {
        c = c.append(rotationY(U(g)), i % 2).append(rotationZ(U(g)), (i + 1) % 2);
        if (i % 2) c = c.appendCnot(0, 1);
    }
    std::vector<Complex> m(16);
    for (int col = 0; col < 4; col++) {
        std::vector<Complex> e(4, 0.0);
        // column index l = 2*bit(first=0) + bit(second=1) -> basis index bit0 = first, bit1 = second
        size_t basis = ((col >> 1) & 1) | ((col & 1) << 1);
        e[basis] = 1.0;
        auto out = apply_scalar(c, e);
        for (int row = 0; row < 4; row++) {
            size_t b = ((row >> 1) & 1) | ((row & 1) << 1);
            m[4 * row + col] = out[b];
        }
    }
    return m;
}

static void test(const char *name, const std::function<void()> &f) {
    current = name;
    int before = failures;
    try {
        f();
    } catch (const std::exception &e) {
        std::printf("FAIL %s: exception %s\n", name, e.what());
        failures++;
    }
    if (failures == before) std::printf("PASS %s\n", name);
}

int main() {
    HipStatevectorEngine engine;
    std::mt19937_64 g(0x51D5);

    test("idIsStable", [&] { CHECK(engine.id() == "statevector-hip", "id"); });
    test("simulateFromZeroStateMatchesCircuitApply", [&] {
        QuantumCircuit c = QuantumCircuit(2).append(Hadamard(), 0).appendCnot(0, 1);
        auto r = engine.simulate(SimulationRequest::zeroState(c));
        CHECK(matches(apply_scalar(c, zero(2)), *r.finalState), "bell");
    });
    test("simulateFromCustomStateMatchesCircuitApply", [&] {
        QuantumCircuit c = QuantumCircuit(1).append(Hadamard(), 0);
        auto r = engine.simulate(SimulationRequest::withInitialState(c, StateVector::fromArray(1, {0, 0, 1, 0})));
        CHECK(matches(apply_scalar(c, {0.0, 1.0}), *r.finalState), "custom");
    });
    test("simulateFromBasisStateMatchesCircuitApply", [&] {
        QuantumCircuit c = QuantumCircuit(2).append(Hadamard(), 1).append(PauliZ(), 0);
        auto r = engine.simulate(SimulationRequest::zeroState(c).withBasisState({0}));
        CHECK(matches(apply_scalar(c, {0.0, 1.0, 0.0, 0.0}), *r.finalState), "basis");
    });
    test("simulateIdentityCircuitPreservesInitialStateAmplitudes", [&] {
        auto init = random_state(3, g);
        auto r = engine.simulate(SimulationRequest::withInitialState(QuantumCircuit(3), to_state(3, init)));
        CHECK(matches(init, *r.finalState), "identity");
    });
    test("simulateMultiQubitCircuitMatchesReference", [&] {
        QuantumCircuit c = QuantumCircuit(5).append(Hadamard(), 4).append(Hadamard(), 1).append(PauliX(), 0)
                               .appendCnot(4, 2).append(Hadamard(), 2);
        CHECK(matches(apply_scalar(c, zero(5)), *engine.simulate(SimulationRequest::zeroState(c)).finalState), "multi");
    });
    test("simulateTwoQubitUnitaryMatchesReference", [&] {
        for (int t = 0; t < 6; t++) {
            QuantumCircuit c = QuantumCircuit(2).appendTwoQubitUnitary(random_unitary4(g), 0, 1);
            CHECK(matches(apply_scalar(c, zero(2)), *engine.simulate(SimulationRequest::zeroState(c)).finalState), "2q");
        }
    });
    test("simulateTwoQubitUnitaryOnNonAdjacentQubitsMatchesReference", [&] {
        QuantumCircuit c = QuantumCircuit(3).append(Hadamard(), 1).appendTwoQubitUnitary(random_unitary4(g), 2, 0)
                               .append(PauliZ(), 1);
        CHECK(matches(apply_scalar(c, zero(3)), *engine.simulate(SimulationRequest::zeroState(c)).finalState), "2q-na");
    });
    test("simulateControlledPhaseMatchesReference", [&] {
        for (int dir = 0; dir < 2; dir++) {
            QuantumCircuit c = QuantumCircuit(2).append(Hadamard(), 0).append(Hadamard(), 1)
                                   .appendControlledPhase(M_PI / 3.0, dir, 1 - dir);
            CHECK(matches(apply_scalar(c, zero(2)), *engine.simulate(SimulationRequest::zeroState(c)).finalState), "cp");
        }
    });
    test("simulateSwapMatchesReference", [&] {
        QuantumCircuit c = QuantumCircuit(3).append(Hadamard(), 0).append(PauliX(), 2).appendSwap(0, 2)
                               .append(Hadamard(), 1);
        CHECK(matches(apply_scalar(c, zero(3)), *engine.simulate(SimulationRequest::zeroState(c)).finalState), "swap");
    });
    test("simulateMultiControlWithThreeControlsMatchesReference", [&] {
        QuantumCircuit c = QuantumCircuit(4).append(Hadamard(), 0).append(Hadamard(), 1).append(Hadamard(), 2)
                               .appendMultiControl(PauliX(), 3, {0, 1, 2});
        CHECK(matches(apply_scalar(c, zero(4)), *engine.simulate(SimulationRequest::zeroState(c)).finalState), "mc");
    });
    test("simulateGhzCircuitMatchesReference", [&] {
        QuantumCircuit c = QuantumCircuit(3).append(Hadamard(), 0).appendCnot(0, 1).appendMultiControl(PauliX(), 2, {0, 1});
        CHECK(matches(apply_scalar(c, zero(3)), *engine.simulate(SimulationRequest::zeroState(c)).finalState), "ghz");
    });
    test("simulateQftThreeQubitsMatchesReference", [&] {
        QuantumCircuit c(3);
        for (int t = 0; t < 3; t++) {
            c = c.append(Hadamard(), t);
            for (int k = t + 1; k < 3; k++) c = c.appendControlledPhase(M_PI / (1 << (k - t)), k, t);
        }
        c = c.appendSwap(0, 2);
        auto init = random_state(3, g);
        auto r = engine.simulate(SimulationRequest::withInitialState(c, to_state(3, init)));
        CHECK(matches(apply_scalar(c, init), *r.finalState), "qft");
    });
    test("mixedOperationsOnLargerRegisterMatchesReference", [&] {
        // exercises the sweep planner (>= 12 qubits) with every op kind
        QuantumCircuit c(14);
        std::uniform_int_distribution<int> Q(0, 13);
        std::uniform_real_distribution<double> U(0, 6.28);
        for (int i = 0; i < 120; i++) {
            int a = Q(g), b = (a + 1 + Q(g) % 13) % 14;
            switch (i % 6) {
            case 0: c = c.append(rotationY(U(g)), a); break;
            case 1: c = c.appendCnot(a, b); break;
            case 2: c = c.appendTwoQubitUnitary(random_unitary4(g), a, b); break;
            case 3: c = c.appendControlledPhase(U(g), a, b); break;
            case 4: c = c.appendSwap(a, b); break;
            default: {
                int c2 = (b + 1) % 14 == a ? (b + 2) % 14 : (b + 1) % 14;
                c = c.appendMultiControl(rotationX(U(g)), a, {b, c2});
                break;
            }
            }
        }
        auto init = random_state(14, g);
        auto r = engine.simulate(SimulationRequest::withInitialState(c, to_state(14, init)));
        CHECK(matches(apply_scalar(c, init), *r.finalState), "mixed14");
    });
    test("bellStateHistogramShowsBalancedCounts", [&] {
        QuantumCircuit c = QuantumCircuit(2).append(Hadamard(), 0).appendCnot(0, 1);
        auto r = engine.simulate(SimulationRequest::zeroState(c).withMeasurement(
            MeasurementInstruction::countsAll(10000).withSeed(1234), false));
        auto &h = r.measurement->counts;
        CHECK(h.size() == 2 && h.count(0) && h.count(3), "outcomes");
        CHECK(std::abs(h[0] - 5000) < 400 && std::abs(h[3] - 5000) < 400, "balance");
    });
    test("seededMeasurementIsDeterministic", [&] {
        QuantumCircuit c = QuantumCircuit(3).append(Hadamard(), 0).append(Hadamard(), 2);
        auto req = SimulationRequest::zeroState(c).withMeasurement(MeasurementInstruction::countsAll(512).withSeed(7));
        CHECK(engine.simulate(req).measurement->counts == engine.simulate(req).measurement->counts, "det");
    });
    test("bernsteinVaziraniFindsSecret", [&] {
        QuantumCircuit c = QuantumCircuit(5).append(PauliX(), 4).append(Hadamard(), 4);
        for (int q = 0; q < 4; q++) c = c.append(Hadamard(), q);
        for (int q = 0; q < 4; q++)
            if ((0b1011 >> q) & 1) c = c.appendCnot(q, 4);
        for (int q = 0; q < 4; q++) c = c.append(Hadamard(), q);
        auto r = engine.simulate(SimulationRequest::zeroState(c).withMeasurement(
            MeasurementInstruction::counts(4096, {0, 1, 2, 3}).withSeed(5), false));
        CHECK(r.measurement->counts.size() == 1 && r.measurement->counts.at(0b1011) == 4096, "bv");
    });
    test("phaseFlipOnPlusGivesMinus", [&] {
        QuantumCircuit c = QuantumCircuit(1).append(Hadamard(), 0);
        NoiseModel m;
        m.afterAllGates(ErrorChannel::phaseFlip(1.0, 0));
        auto r = engine.simulate(SimulationRequest::zeroState(c).withNoiseModel(m).withNoiseSeed(3));
        double s = 1.0 / std::sqrt(2.0);
        CHECK(matches({s, -s}, *r.finalState), "minus");
    });
    test("thermalRelaxationRelaxesToGround", [&] {
        QuantumCircuit c = QuantumCircuit(1).append(PauliX(), 0);
        NoiseModel m;
        m.afterAllGates(ErrorChannel::thermalRelaxation(1e-6, 1e-6, 1.0, 0));
        auto r = engine.simulate(SimulationRequest::zeroState(c).withNoiseModel(m).withNoiseSeed(11));
        CHECK(matches({1.0, 0.0}, *r.finalState), "ground");
    });
    test("seededNoiseIsDeterministic", [&] {
        QuantumCircuit c = QuantumCircuit(3).append(Hadamard(), 0).appendCnot(0, 1).appendCnot(1, 2);
        NoiseModel m;
        m.afterAllGates(ErrorChannel::depolarizing(0.3, 1)).onQubits(ErrorChannel::amplitudeDamping(0.2, 0), {2});
        auto req = SimulationRequest::zeroState(c).withNoiseModel(m).withNoiseSeed(99);
        auto a = engine.simulate(req), b = engine.simulate(req);
        CHECK(a.finalState->data() == b.finalState->data(), "noise determinism");
    });
    test("errorsMapToReferenceExceptions", [&] {
        bool threw = false;
        try {
            QuantumCircuit(2).append(Hadamard(), 2);
        } catch (const IllegalArgumentException &) {
            threw = true;
        }
        CHECK(threw, "IAE");
        threw = false;
        try {
            QuantumCircuit c = QuantumCircuit(2).measure({0}).append(Hadamard(), 1);
            engine.simulate(SimulationRequest::zeroState(c));
        } catch (const UnsupportedOperationException &) {
            threw = true;
        }
        CHECK(threw, "UOE after measure");
        threw = false;
        try {
            engine.simulate(SimulationRequest::withInitialState(QuantumCircuit(2), StateVector::fromArray(2, std::vector<double>(8, 0.0)))
                                .withMeasurement(MeasurementInstruction::countsAll(4)));
        } catch (const IllegalStateException &) {
            threw = true;
        }
        CHECK(threw, "ISE zero probability");
        CHECK(SimulatorEngines::require("statevector-hip")->id() == "statevector-hip", "registry");
    });
    // ---- StatevectorEngineTest measurement and limit cases (StatevectorEngineTest.java:282-380) ----
    test("partialMeasurementTargetsSpecifiedQubits", [&] {
        QuantumCircuit c = QuantumCircuit(3).append(Hadamard(), 0).append(Hadamard(), 1).appendCnot(0, 2);
        auto r = engine.simulate(SimulationRequest::zeroState(c).withMeasurement(
            MeasurementInstruction::counts(5000, {2}).withSeed(0xDEAD)));
        auto &m = *r.measurement;
        CHECK(m.type == MeasurementResult::Type::Histogram, "histogram");
        CHECK(m.measuredQubits == std::vector<int>{2}, "measured qubits");
        CHECK(m.counts.size() == 2, "two outcomes");
        const int zeros = m.counts.count(0) ? m.counts.at(0) : 0, ones = m.counts.count(1) ? m.counts.at(1) : 0;
        CHECK(zeros + ones == 5000 && zeros > 0 && ones > 0, "counts");
    });
    test("measurementOperationsDefineDefaultTargets", [&] {
        QuantumCircuit c = QuantumCircuit(3).append(Hadamard(), 0).append(PauliX(), 2).measure({0, 2});
        auto r = engine.simulate(SimulationRequest::zeroState(c).withMeasurement(
            MeasurementInstruction::countsAll(1000).withSeed(42)));
        auto &m = *r.measurement;
        CHECK((m.measuredQubits == std::vector<int>{0, 2}), "targets from MeasureOperation");
        int total = 0;
        for (auto &kv : m.counts) {
            total += kv.second;
            CHECK((kv.first & 2) == 2, "qubit 2 is |1>");
        }
        CHECK(total == 1000, "shots");
    });
    test("shotSamplingSupportsRawOutcomes", [&] {
        QuantumCircuit c = QuantumCircuit(2).append(Hadamard(), 0).appendCnot(0, 1);
        auto r = engine.simulate(SimulationRequest::zeroState(c).withMeasurement(
            MeasurementInstruction::samplesAll(16).withSeed(123)));
        auto &m = *r.measurement;
        CHECK(m.type == MeasurementResult::Type::Samples && m.shots == 16 && m.samples.size() == 16, "samples");
        for (int64_t o : m.samples) CHECK(o == 0 || o == 3, "bell outcome");
    });
    test("unitaryAfterMeasurementThrows", [&] {
        bool threw = false;
        try {
            QuantumCircuit c = QuantumCircuit(1).append(Hadamard(), 0).measure({0}).append(PauliX(), 0);
            engine.simulate(SimulationRequest::zeroState(c).withMeasurement(MeasurementInstruction::countsAll(32)));
        } catch (const UnsupportedOperationException &) {
            threw = true;
        }
        CHECK(threw, "UOE");
    });
    test("largeCircuitNearTargetLimitRemainsNumericallySane", [&] {
        const int n = 22;
        QuantumCircuit c = QuantumCircuit(n).append(PauliX(), 0).append(PauliX(), 5).append(PauliX(), 21)
                               .appendCnot(21, 1).appendCnot(0, 2).appendSwap(5, 3);
        auto st = *engine.simulate(SimulationRequest::zeroState(c)).finalState;
        const uint64_t expected = (1u << 0) | (1u << 1) | (1u << 2) | (1u << 3) | (1u << 21);
        CHECK(std::fabs(st.amplitude(expected).real() - 1.0) <= EPS && std::fabs(st.amplitude(expected).imag()) <= EPS,
              "basis amplitude");
        double ns = 0.0;
        for (uint64_t i = 0; i < (1ULL << n); i++) ns += std::norm(st.amplitude(i));
        CHECK(std::fabs(ns - 1.0) <= EPS, "normalised");
    });
    test("twentyQubitRyQftWithDefaultOptionsMatchesReference", [&] {
        // the 20-qubit Ry layer + QFT workload (thetas 2 pi u from SplittableRandom(42)) through the engine with
        // its default options (runtime-compiled passes, geometry chosen per circuit), 4096 shots, seed 42
        const int n = 20;
        SplittableRandom rng(42);
        QuantumCircuit c(n);
        for (int q = 0; q < n; q++) c = c.append(rotationY(2.0 * M_PI * rng.nextDouble()), q);
        for (int t = n - 1; t >= 0; t--) {
            c = c.append(Hadamard(), t);
            for (int k = t - 1; k >= 0; k--) c = c.appendControlledPhase(M_PI / double(1 << (t - k)), k, t);
        }
        for (int q = 0; q < n / 2; q++) c = c.appendSwap(q, n - 1 - q);
        std::vector<Complex> init(size_t(1) << n, Complex(0.0, 0.0));
        init[0] = 1.0;
        auto req = SimulationRequest::zeroState(c).withMeasurement(MeasurementInstruction::countsAll(4096).withSeed(42));
        auto r = engine.simulate(req);
        CHECK(matches(apply_scalar(c, init), *r.finalState), "20-qubit state");
        const auto &h = r.measurement->counts;
        int total = 0;
        for (const auto &kv : h) total += kv.second;
        CHECK(r.measurement->shots == 4096 && total == 4096, "4096 shots");
        auto again = engine.simulate(req);
        CHECK(again.measurement->counts == h, "same seed, same counts");
    });
    test("rejectsQubitCountAboveDenseLimit", [&] {
        bool threw = false;
        try {
            engine.simulate(SimulationRequest::zeroState(QuantumCircuit(35)));
        } catch (const UnsupportedOperationException &e) {
            threw = std::string(e.what()).find("supports up to 34 qubits") != std::string::npos;
        }
        CHECK(threw, "UOE above 34 qubits");
    });
    test("wideMeasurementIsBitstringHistogramMsbLeft", [&] {
        // 32 measured qubits (beyond Histogram's Integer keys): the BitstringHistogram of TensorNetworkEngine
        // (MSB-left strings); a basis state gives one deterministic outcome
        const int n = 32;
        QuantumCircuit c(n);
        for (int q : {0, 3, 17, 30, 31}) c = c.append(PauliX(), q);
        auto r = engine.simulate(SimulationRequest::zeroState(c).withMeasurement(
            MeasurementInstruction::countsAll(64).withSeed(9), false));
        auto &m = *r.measurement;
        std::string want(n, '0');
        for (int q : {0, 3, 17, 30, 31}) want[n - 1 - q] = '1';
        CHECK(m.type == MeasurementResult::Type::BitstringHistogram, "bitstring histogram");
        CHECK(m.bitstringCounts.size() == 1 && m.bitstringCounts.count(want) && m.bitstringCounts.at(want) == 64,
              "deterministic MSB-left outcome");
        CHECK(!r.finalState, "no final state requested");
    });
    // ---- StatevectorNoiseTest cases (StatevectorNoiseTest.java:84-207) ----
    test("amplitudeDampingDistributionMatchesProbability", [&] {
        NoiseModel m;
        m.afterAllGates(ErrorChannel::amplitudeDamping(0.25, 0));
        QuantumCircuit c = QuantumCircuit(1).append(PauliX(), 0);
        int zero = 0;
        for (int t = 0; t < 1000; t++) {
            auto st = *engine.simulate(SimulationRequest::zeroState(c).withNoiseModel(m).withNoiseSeed(0xBAD0 + t)).finalState;
            if (std::norm(st.amplitude(0)) > 0.5) zero++;
        }
        CHECK(std::fabs(zero / 1000.0 - 0.25) <= 0.05, "gamma +- 0.05");
    });
    test("idleDecoherenceUsesMappedQubitInsteadOfChannelTemplate", [&] {
        QuantumCircuit c = QuantumCircuit(2).append(Hadamard(), 0);
        NoiseModel m;
        m.onQubits(ErrorChannel::amplitudeDamping(1.0, 0), {0, 1});
        auto r = engine.simulate(SimulationRequest::withInitialState(c, StateVector::fromArray(2, {0, 0, 0, 0, 0, 0, 1, 0}))
                                     .withNoiseModel(m));
        const double s = 1.0 / std::sqrt(2.0);
        CHECK(matches({s, -s, 0.0, 0.0}, *r.finalState), "idle channel on qubit 1");
    });
    test("bellDemoAmplitudeDampingSeedThreeKeepsStateNormalized", [&] {
        NoiseModel m;
        m.afterAllGates(ErrorChannel::amplitudeDamping(0.9, 0)).afterAllGates(ErrorChannel::amplitudeDamping(0.9, 1));
        QuantumCircuit c = QuantumCircuit(2).append(Hadamard(), 0).appendCnot(0, 1);
        auto st = *engine.simulate(SimulationRequest::zeroState(c).withNoiseModel(m).withNoiseSeed(3)).finalState;
        double ns = 0.0;
        for (uint64_t i = 0; i < 4; i++) ns += std::norm(st.amplitude(i));
        CHECK(std::fabs(ns - 1.0) <= EPS, "normalised");
    });
    test("amplitudeAndPhaseDampingOnGroundStateSkipZeroBranch", [&] {
        for (int kind = 0; kind < 2; kind++) {
            NoiseModel m;
            m.afterAllGates(kind == 0 ? ErrorChannel::amplitudeDamping(0.9, 1) : ErrorChannel::phaseDamping(0.9, 1));
            QuantumCircuit c = QuantumCircuit(2).append(Hadamard(), 0);
            auto r = engine.simulate(SimulationRequest::zeroState(c).withNoiseModel(m).withNoiseSeed(3));
            const double s = 1.0 / std::sqrt(2.0);
            CHECK(matches({s, s, 0.0, 0.0}, *r.finalState), kind == 0 ? "amplitude damping" : "phase damping");
        }
    });
    test("scheduledNoisePathsMatchGoldenState", [&] {
        NoiseModel m;
        m.afterGate(Operation::Kind::SingleGate, ErrorChannel::phaseFlip(1.0, 0))
            .onQubits(ErrorChannel::amplitudeDamping(1.0, 0), {1})
            .afterAllGates(ErrorChannel::phaseFlip(1.0, 0));
        QuantumCircuit c = QuantumCircuit(2).append(Hadamard(), 0);
        auto r = engine.simulate(SimulationRequest::zeroState(c).withBasisState({0, 1}).withNoiseModel(m));
        const double s = 1.0 / std::sqrt(2.0);
        CHECK(matches({s, -s, 0.0, 0.0}, *r.finalState), "golden");
    });
    test("noisyMeasurementPathMatchesGoldenHistogram", [&] {
        NoiseModel m;
        m.afterAllGates(ErrorChannel::amplitudeDamping(1.0, 0));
        QuantumCircuit c = QuantumCircuit(1).append(PauliX(), 0);
        auto r = engine.simulate(SimulationRequest::zeroState(c).withNoiseModel(m).withMeasurement(
            MeasurementInstruction::countsAll(16).withSeed(5), false));
        CHECK(r.measurement->shots == 16 && r.measurement->counts == (std::map<int64_t, int>{{0, 16}}), "{0: 16}");
    });
    std::printf("%d failure(s)\n", failures);
    return failures;
}
