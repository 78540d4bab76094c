// statecraft.hpp -- C++ mirror of StateCraft's engine API over the gfx950 C-ABI.
//
// The reference is Java (no JDK in this image); this header restates, in C++, the types a
// StatevectorEngine caller touches, with the same names, argument meaning and exception classes:
//   SimulatorEngine            core/src/main/java/com/omaarr90/statecraft/core/engine/SimulatorEngine.java:3-9
//   SimulatorEngines           core/.../core/engine/SimulatorEngines.java:23-51
//   SimulationRequest          core/.../core/engine/SimulationRequest.java:17-130
//   MeasurementInstruction     core/.../core/engine/MeasurementInstruction.java:12-85
//   MeasurementResult          core/.../core/engine/MeasurementResult.java:10-173
//   SimulationResult           core/.../core/engine/SimulationResult.java:10-37
//   QuantumCircuit / Operation core/src/main/java/com/omaarr90/statecraft/quantum/QuantumCircuit.java:9-574
//   SingleQubitGate & gates    core/.../quantum/{SingleQubitGate,Hadamard,PauliX,PauliY,PauliZ,SGate,SdgGate,
//                              MatrixSingleQubitGate}.java
//   StateVector                core/.../quantum/StateVector.java:9-82
//   NoiseModel / ErrorChannel  core/src/main/java/com/omaarr90/statecraft/core/noise/*.java
// HipStatevectorEngine (id "statevector-hip") is the drop-in for StatevectorEngine
// (engines/.../statevector/StatevectorEngine.java:32-167): it lowers the circuit to the sc_op buffer and
// runs it in HBM through libstatecraft_hip.so.
#pragma once

#include <complex>
#include <cstdint>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/statecraft_hip.h"

namespace statecraft {

using Complex = std::complex<double>;

// ---- exception classes of the reference --------------------------------------------------------
struct NullPointerException : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};
struct IllegalArgumentException : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};
struct UnsupportedOperationException : std::logic_error {
    using std::logic_error::logic_error;
};
struct IllegalStateException : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Maps an sc_* status to the reference's exception class (throws on failure).
void check_status(int rc);

// ---- java.util.SplittableRandom (SplitMix64) ------------------------------------------------------
class SplittableRandom {
  public:
    explicit SplittableRandom(int64_t seed) : seed_((uint64_t)seed) {}
    int64_t nextLong();
    double nextDouble();
    double nextDouble(double bound);

  private:
    uint64_t seed_;
};

// ---- gates --------------------------------------------------------------------------------------------
class SingleQubitGate {
  public:
    SingleQubitGate(std::string name, Complex m00, Complex m01, Complex m10, Complex m11)
        : name_(std::move(name)), m_{m00, m01, m10, m11} {}
    const std::string &name() const { return name_; }
    Complex element(int row, int col) const;
    const Complex *data() const { return m_; }

  private:
    std::string name_;
    Complex m_[4];
};

SingleQubitGate Hadamard();
SingleQubitGate PauliX();
SingleQubitGate PauliY();
SingleQubitGate PauliZ();
SingleQubitGate SGate();
SingleQubitGate SdgGate();
SingleQubitGate MatrixSingleQubitGate(const std::string &name, Complex m00, Complex m01, Complex m10, Complex m11);
// OpenQasmCircuitParser.rotationX / rotationY / rotationZ / phaseGate (OpenQasmCircuitParser.java:646-684)
SingleQubitGate rotationX(double angle);
SingleQubitGate rotationY(double angle);
SingleQubitGate rotationZ(double angle);
SingleQubitGate phaseGate(double angle);

// ---- circuit -----------------------------------------------------------------------------------------
struct Operation {
    enum class Kind { SingleGate, Cnot, TwoQubitGate, TwoQubitDiagonal, Swap, MultiControl, Measure };
    Kind kind;
    int q0 = -1;  // qubit / control / first / target (MultiControl)
    int q1 = -1;  // target (Cnot) / second
    std::vector<int> qubits;      // controls (MultiControl, sorted) or measured qubits (Measure, sorted)
    std::vector<Complex> matrix;  // 4 (gate / diagonal) or 16 (two-qubit unitary)
    std::vector<int> affectedQubits() const;
};

class QuantumCircuit {
  public:
    explicit QuantumCircuit(int qubitCount);
    int qubitCount() const { return n_; }
    const std::vector<Operation> &operations() const { return ops_; }

    QuantumCircuit append(const SingleQubitGate &gate, int targetQubit) const;
    QuantumCircuit appendCnot(int controlQubit, int targetQubit) const;  // append(CnotGate, c, t)
    QuantumCircuit appendTwoQubitUnitary(const std::vector<Complex> &matrix, int firstQubit, int secondQubit) const;
    QuantumCircuit appendDiagonalTwoQubit(const std::vector<Complex> &diagonal, int firstQubit, int secondQubit) const;
    QuantumCircuit appendControlledPhase(double angle, int controlQubit, int targetQubit) const;
    QuantumCircuit appendControlled(const SingleQubitGate &gate, int controlQubit, int targetQubit) const;
    QuantumCircuit appendControlledX(int controlQubit, int targetQubit) const;
    QuantumCircuit appendControlledY(int controlQubit, int targetQubit) const;
    QuantumCircuit appendControlledZ(int controlQubit, int targetQubit) const;
    QuantumCircuit appendToffoli(int firstControlQubit, int secondControlQubit, int targetQubit) const;
    QuantumCircuit appendSwap(int firstQubit, int secondQubit) const;
    QuantumCircuit appendMultiControl(const SingleQubitGate &gate, int targetQubit, std::vector<int> controls) const;
    QuantumCircuit measure(std::vector<int> qubits) const;

    // Flat C-ABI encoding (sc_op records + interleaved complex matrix pool).
    void lower(std::vector<sc_op> &ops, std::vector<double> &mats) const;

  private:
    QuantumCircuit withOp(Operation op) const;
    void validate(const Operation &op) const;
    int n_;
    std::vector<Operation> ops_;
};

// ---- state snapshot ----------------------------------------------------------------------------------
class StateVector {
  public:
    static constexpr int MAX_QUBITS = 34;
    StateVector(int qubitCount, std::vector<double> data);  // interleaved (re, im)
    static StateVector fromArray(int qubitCount, const std::vector<double> &data) { return StateVector(qubitCount, data); }
    int qubitCount() const { return n_; }
    const std::vector<double> &data() const { return data_; }
    Complex amplitude(uint64_t index) const { return {data_[2 * index], data_[2 * index + 1]}; }

  private:
    int n_;
    std::vector<double> data_;
};

// ---- noise -------------------------------------------------------------------------------------------
class ErrorChannel {
  public:
    virtual ~ErrorChannel() = default;
    virtual std::vector<int> affectedQubits() const = 0;
    // Single-qubit Kraus operators, 4 complex each (row-major); empty for composites.
    virtual const std::vector<std::vector<Complex>> &kraus() const = 0;
    virtual const std::vector<std::shared_ptr<ErrorChannel>> *components() const { return nullptr; }

    static std::shared_ptr<ErrorChannel> depolarizing(double probability, int qubit);
    static std::shared_ptr<ErrorChannel> amplitudeDamping(double gamma, int qubit);
    static std::shared_ptr<ErrorChannel> phaseFlip(double probability, int qubit);
    static std::shared_ptr<ErrorChannel> phaseDamping(double lambda, int qubit);
    static std::shared_ptr<ErrorChannel> thermalRelaxation(double t1, double t2, double gateTime, int qubit);
    static std::shared_ptr<ErrorChannel> compose(std::vector<std::shared_ptr<ErrorChannel>> channels);
};

class NoiseModel {
  public:
    struct ScheduledChannel {
        std::shared_ptr<ErrorChannel> channel;
        std::vector<int> targetQubits;
    };
    NoiseModel &afterGate(Operation::Kind kind, std::shared_ptr<ErrorChannel> channel);
    NoiseModel &onQubits(std::shared_ptr<ErrorChannel> channel, std::vector<int> qubits);
    NoiseModel &afterAllGates(std::shared_ptr<ErrorChannel> channel);
    bool hasNoise() const;
    std::vector<ScheduledChannel> scheduledChannelsAfter(const Operation &op) const;

  private:
    std::map<Operation::Kind, std::vector<std::shared_ptr<ErrorChannel>>> gate_;
    std::vector<std::pair<int, std::vector<std::shared_ptr<ErrorChannel>>>> qubit_;  // insertion order
    std::vector<std::shared_ptr<ErrorChannel>> global_;
};

// ---- requests / results ----------------------------------------------------------------------------
struct MeasurementInstruction {
    enum class Mode { COUNTS, SAMPLES };
    int shots;
    std::optional<std::vector<int>> measuredQubits;
    std::optional<int64_t> seed;
    Mode mode;
    MeasurementInstruction(int shots, std::optional<std::vector<int>> measuredQubits, std::optional<int64_t> seed,
                           Mode mode);
    static MeasurementInstruction countsAll(int shots);
    static MeasurementInstruction counts(int shots, std::vector<int> qubits);
    static MeasurementInstruction samplesAll(int shots);
    static MeasurementInstruction samples(int shots, std::vector<int> qubits);
    MeasurementInstruction withSeed(int64_t newSeed) const;
};

struct MeasurementResult {
    enum class Type { Histogram, Samples, BitstringHistogram, BitstringSamples };
    Type type;
    std::vector<int> measuredQubits;
    int shots = 0;
    std::map<int64_t, int> counts;              // Histogram
    std::vector<int64_t> samples;               // Samples
    std::map<std::string, int> bitstringCounts; // BitstringHistogram (MSB-left)
    std::vector<std::string> bitstringSamples;  // BitstringSamples
};

class SimulationRequest {
  public:
    SimulationRequest(QuantumCircuit circuit, std::optional<StateVector> initialState,
                      std::optional<std::vector<int>> basisStateQubits,
                      std::optional<MeasurementInstruction> measurement, std::optional<NoiseModel> noiseModel,
                      std::optional<int64_t> noiseSeed, bool returnFinalState);
    static SimulationRequest zeroState(const QuantumCircuit &circuit);
    static SimulationRequest withInitialState(const QuantumCircuit &circuit, const StateVector &state);
    SimulationRequest withMeasurement(const MeasurementInstruction &instruction) const;
    SimulationRequest withMeasurement(const MeasurementInstruction &instruction, bool includeFinalState) const;
    SimulationRequest withoutFinalState() const;
    SimulationRequest withNoiseModel(const NoiseModel &model) const;
    SimulationRequest withNoiseSeed(int64_t seed) const;
    SimulationRequest withBasisState(std::vector<int> qubits) const;

    const QuantumCircuit &circuit() const { return circuit_; }
    const std::optional<StateVector> &initialState() const { return initial_; }
    const std::optional<std::vector<int>> &basisStateQubits() const { return basis_; }
    const std::optional<MeasurementInstruction> &measurement() const { return measurement_; }
    const std::optional<NoiseModel> &noiseModel() const { return noise_; }
    const std::optional<int64_t> &noiseSeed() const { return noiseSeed_; }
    bool returnFinalState() const { return returnFinalState_; }

  private:
    QuantumCircuit circuit_;
    std::optional<StateVector> initial_;
    std::optional<std::vector<int>> basis_;
    std::optional<MeasurementInstruction> measurement_;
    std::optional<NoiseModel> noise_;
    std::optional<int64_t> noiseSeed_;
    bool returnFinalState_;
};

struct SimulationResult {
    std::optional<StateVector> finalState;
    std::optional<MeasurementResult> measurement;
};

// ---- engines -----------------------------------------------------------------------------------------
class SimulatorEngine {
  public:
    virtual ~SimulatorEngine() = default;
    virtual std::string id() const = 0;
    virtual SimulationResult simulate(const SimulationRequest &request) = 0;
};

// StatevectorExecutionConfig analogue: device and execution mode of the HIP engine.
struct HipStatevectorConfig {
    int device = 0;
    int execMode = SC_EXEC_SWEEP;
    int maxFusedQubits = 4;
    static HipStatevectorConfig automatic() { return HipStatevectorConfig(); }
};

class HipStatevectorEngine : public SimulatorEngine {
  public:
    static constexpr const char *ID = "statevector-hip";
    HipStatevectorEngine();
    explicit HipStatevectorEngine(const HipStatevectorConfig &config);
    ~HipStatevectorEngine() override;
    HipStatevectorEngine(const HipStatevectorEngine &) = delete;
    HipStatevectorEngine &operator=(const HipStatevectorEngine &) = delete;
    std::string id() const override { return ID; }
    SimulationResult simulate(const SimulationRequest &request) override;

  private:
    sc_ctx *ctx_ = nullptr;
};

// SimulatorEngines.loadAll / findById / require for engines linked into this library.
struct SimulatorEngines {
    static std::vector<std::string> ids();
    static std::unique_ptr<SimulatorEngine> findById(const std::string &id);
    static std::unique_ptr<SimulatorEngine> require(const std::string &id);
};

}  // namespace statecraft
