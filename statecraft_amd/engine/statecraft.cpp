// statecraft.cpp -- C++ mirror of StateCraft's engine API (see statecraft.hpp for the reference map).
#include "statecraft.hpp"

#include <algorithm>
#include <cmath>
#include <exception>
#include <random>
#include <set>

namespace statecraft {

namespace {
[[noreturn]] void iae(const std::string &m) { throw IllegalArgumentException(m); }
}  // namespace

void check_status(int rc) {
    if (rc == SC_OK) return;
    std::string msg = sc_last_error();
    switch (rc) {
    case SC_E_NULL: throw NullPointerException(msg);
    case SC_E_ARG: throw IllegalArgumentException(msg);
    case SC_E_UNSUPPORTED:
    case SC_E_OOM: throw UnsupportedOperationException(msg);
    default: throw IllegalStateException(msg);
    }
}

// ---- SplittableRandom --------------------------------------------------------------------------
int64_t SplittableRandom::nextLong() {
    uint64_t z = (seed_ += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return (int64_t)(z ^ (z >> 31));
}
double SplittableRandom::nextDouble() { return (double)(((uint64_t)nextLong()) >> 11) * 0x1.0p-53; }
double SplittableRandom::nextDouble(double bound) {
    double r = nextDouble() * bound;
    if (r >= bound) r = std::nextafter(bound, -INFINITY);
    return r;
}

// ---- gates -------------------------------------------------------------------------------------
Complex SingleQubitGate::element(int row, int col) const {
    if (row < 0 || row > 1) iae("row must be 0 or 1: " + std::to_string(row));
    if (col < 0 || col > 1) iae("col must be 0 or 1: " + std::to_string(col));
    return m_[2 * row + col];
}

SingleQubitGate Hadamard() {
    const double s = 1.0 / std::sqrt(2.0);
    return SingleQubitGate("H", {s, 0}, {s, 0}, {s, 0}, {-s, 0});
}
SingleQubitGate PauliX() { return SingleQubitGate("X", {0, 0}, {1, 0}, {1, 0}, {0, 0}); }
SingleQubitGate PauliY() { return SingleQubitGate("Y", {0, 0}, {0, -1}, {0, 1}, {0, 0}); }
SingleQubitGate PauliZ() { return SingleQubitGate("Z", {1, 0}, {0, 0}, {0, 0}, {-1, 0}); }
SingleQubitGate SGate() { return SingleQubitGate("S", {1, 0}, {0, 0}, {0, 0}, {0, 1}); }
SingleQubitGate SdgGate() { return SingleQubitGate("Sdg", {1, 0}, {0, 0}, {0, 0}, {0, -1}); }
SingleQubitGate MatrixSingleQubitGate(const std::string &name, Complex m00, Complex m01, Complex m10, Complex m11) {
    return SingleQubitGate(name, m00, m01, m10, m11);
}
static Complex cis(double a) { return {std::cos(a), std::sin(a)}; }
SingleQubitGate rotationX(double angle) {
    double h = angle / 2.0, c = std::cos(h), s = std::sin(h);
    return SingleQubitGate("rx", {c, 0}, {0, -s}, {0, -s}, {c, 0});
}
SingleQubitGate rotationY(double angle) {
    double h = angle / 2.0, c = std::cos(h), s = std::sin(h);
    return SingleQubitGate("ry", {c, 0}, {-s, 0}, {s, 0}, {c, 0});
}
SingleQubitGate rotationZ(double angle) {
    double h = angle / 2.0;
    return SingleQubitGate("rz", cis(-h), {0, 0}, {0, 0}, cis(h));
}
SingleQubitGate phaseGate(double angle) { return SingleQubitGate("p", {1, 0}, {0, 0}, {0, 0}, cis(angle)); }

// ---- circuit -----------------------------------------------------------------------------------
std::vector<int> Operation::affectedQubits() const {
    std::vector<int> q;
    switch (kind) {
    case Kind::SingleGate: q = {q0}; break;
    case Kind::Cnot:
    case Kind::TwoQubitGate:
    case Kind::TwoQubitDiagonal:
    case Kind::Swap: q = {q0, q1}; break;
    case Kind::MultiControl:
        q = qubits;
        q.push_back(q0);
        break;
    case Kind::Measure: q = qubits; break;
    }
    return q;
}

QuantumCircuit::QuantumCircuit(int qubitCount) : n_(qubitCount) {
    if (qubitCount <= 0) iae("qubitCount must be positive");
}

void QuantumCircuit::validate(const Operation &op) const {
    auto range = [&](int q, const char *what) {
        if (q >= n_) iae(std::string(what) + " out of range: " + std::to_string(q));
    };
    switch (op.kind) {
    case Operation::Kind::SingleGate:
        if (op.q0 < 0) iae("qubit index must be non-negative");
        range(op.q0, "target qubit");
        break;
    case Operation::Kind::Cnot:
        if (op.q0 < 0) iae("control qubit index must be non-negative");
        if (op.q1 < 0) iae("target qubit index must be non-negative");
        if (op.q0 == op.q1) iae("control and target qubits must differ");
        range(op.q0, "control qubit");
        range(op.q1, "target qubit");
        break;
    case Operation::Kind::TwoQubitGate:
    case Operation::Kind::TwoQubitDiagonal:
    case Operation::Kind::Swap:
        if (op.q0 < 0 || op.q1 < 0) iae("qubit indices must be non-negative");
        if (op.q0 == op.q1) iae(op.kind == Operation::Kind::Swap ? "cannot swap qubit with itself" : "qubits must be distinct");
        range(op.q0, "first qubit");
        range(op.q1, "second qubit");
        break;
    case Operation::Kind::MultiControl:
        if (op.q0 < 0) iae("target qubit index must be non-negative");
        if (op.qubits.empty()) iae("at least one control qubit is required");
        for (size_t i = 0; i < op.qubits.size(); i++) {
            if (op.qubits[i] < 0) iae("control qubit index must be non-negative");
            if (op.qubits[i] == op.q0) iae("control and target qubits must differ");
            if (i > 0 && op.qubits[i] == op.qubits[i - 1]) iae("duplicate control qubit: " + std::to_string(op.qubits[i]));
        }
        range(op.q0, "target qubit");
        for (int c : op.qubits) range(c, "control qubit");
        break;
    case Operation::Kind::Measure:
        if (op.qubits.empty()) iae("at least one qubit must be measured");
        for (size_t i = 0; i < op.qubits.size(); i++) {
            if (op.qubits[i] < 0) iae("measured qubit index must be non-negative");
            if (i > 0 && op.qubits[i] == op.qubits[i - 1]) iae("duplicate measured qubit: " + std::to_string(op.qubits[i]));
        }
        for (int q : op.qubits) range(q, "measured qubit");
        break;
    }
}

QuantumCircuit QuantumCircuit::withOp(Operation op) const {
    validate(op);
    QuantumCircuit c = *this;
    c.ops_.push_back(std::move(op));
    return c;
}

QuantumCircuit QuantumCircuit::append(const SingleQubitGate &gate, int target) const {
    Operation op{Operation::Kind::SingleGate};
    op.q0 = target;
    op.matrix.assign(gate.data(), gate.data() + 4);
    return withOp(op);
}
QuantumCircuit QuantumCircuit::appendCnot(int control, int target) const {
    Operation op{Operation::Kind::Cnot};
    op.q0 = control;
    op.q1 = target;
    return withOp(op);
}
QuantumCircuit QuantumCircuit::appendTwoQubitUnitary(const std::vector<Complex> &m, int first, int second) const {
    if (m.size() != 16) iae("two-qubit matrix must contain 16 elements");
    Operation op{Operation::Kind::TwoQubitGate};
    op.q0 = first;
    op.q1 = second;
    op.matrix = m;
    return withOp(op);
}
QuantumCircuit QuantumCircuit::appendDiagonalTwoQubit(const std::vector<Complex> &d, int first, int second) const {
    if (d.size() != 4) iae("two-qubit diagonal must contain 4 elements");
    Operation op{Operation::Kind::TwoQubitDiagonal};
    op.q0 = first;
    op.q1 = second;
    op.matrix = d;
    return withOp(op);
}
QuantumCircuit QuantumCircuit::appendControlledPhase(double angle, int control, int target) const {
    return appendDiagonalTwoQubit({1.0, 1.0, 1.0, Complex(std::cos(angle), std::sin(angle))}, control, target);
}
QuantumCircuit QuantumCircuit::appendControlled(const SingleQubitGate &g, int control, int target) const {
    return appendMultiControl(g, target, {control});
}
QuantumCircuit QuantumCircuit::appendControlledX(int c, int t) const { return appendControlled(PauliX(), c, t); }
QuantumCircuit QuantumCircuit::appendControlledY(int c, int t) const { return appendControlled(PauliY(), c, t); }
QuantumCircuit QuantumCircuit::appendControlledZ(int c, int t) const { return appendControlled(PauliZ(), c, t); }
QuantumCircuit QuantumCircuit::appendToffoli(int c1, int c2, int t) const { return appendMultiControl(PauliX(), t, {c1, c2}); }
QuantumCircuit QuantumCircuit::appendSwap(int first, int second) const {
    Operation op{Operation::Kind::Swap};
    op.q0 = first;
    op.q1 = second;
    return withOp(op);
}
QuantumCircuit QuantumCircuit::appendMultiControl(const SingleQubitGate &g, int target, std::vector<int> controls) const {
    if (controls.empty()) iae("at least one control qubit is required");
    std::sort(controls.begin(), controls.end());
    Operation op{Operation::Kind::MultiControl};
    op.q0 = target;
    op.qubits = std::move(controls);
    op.matrix.assign(g.data(), g.data() + 4);
    return withOp(op);
}
QuantumCircuit QuantumCircuit::measure(std::vector<int> qubits) const {
    std::sort(qubits.begin(), qubits.end());
    Operation op{Operation::Kind::Measure};
    op.qubits = std::move(qubits);
    return withOp(op);
}

void QuantumCircuit::lower(std::vector<sc_op> &ops, std::vector<double> &mats) const {
    ops.clear();
    mats.clear();
    for (const Operation &o : ops_) {
        sc_op r{};
        r.q0 = o.q0;
        r.q1 = o.q1;
        r.mat = -1;
        r.param = -1;
        auto put = [&]() {
            r.mat = (int32_t)(mats.size() / 2);
            for (const Complex &c : o.matrix) {
                mats.push_back(c.real());
                mats.push_back(c.imag());
            }
        };
        switch (o.kind) {
        case Operation::Kind::SingleGate:
            r.kind = SC_OP_SINGLE;
            put();
            break;
        case Operation::Kind::Cnot:
            r.kind = SC_OP_CNOT;
            r.controls = 1ULL << o.q0;
            break;
        case Operation::Kind::TwoQubitGate:
            r.kind = SC_OP_TWO;
            put();
            break;
        case Operation::Kind::TwoQubitDiagonal:
            r.kind = SC_OP_DIAG2;
            put();
            break;
        case Operation::Kind::Swap: r.kind = SC_OP_SWAP; break;
        case Operation::Kind::MultiControl:
            r.kind = SC_OP_MCU;
            for (int c : o.qubits) r.controls |= 1ULL << c;
            put();
            break;
        case Operation::Kind::Measure:
            r.kind = SC_OP_MEASURE;
            for (int q : o.qubits) r.controls |= 1ULL << q;
            break;
        }
        ops.push_back(r);
    }
}

// ---- StateVector --------------------------------------------------------------------------------
StateVector::StateVector(int qubitCount, std::vector<double> data) : n_(qubitCount), data_(std::move(data)) {
    if (qubitCount <= 0) iae("qubitCount must be positive");
    if (qubitCount > MAX_QUBITS) iae("qubitCount exceeds " + std::to_string(MAX_QUBITS));
    if (data_.size() != (size_t(2) << qubitCount))
        iae("Expected state vector of length " + std::to_string(size_t(2) << qubitCount) + ", got " +
            std::to_string(data_.size()));
}

// ---- noise --------------------------------------------------------------------------------------
namespace {

struct KrausChannel : ErrorChannel {
    std::vector<int> q;
    std::vector<std::vector<Complex>> k;
    std::vector<int> affectedQubits() const override { return q; }
    const std::vector<std::vector<Complex>> &kraus() const override { return k; }
};

struct Composite : ErrorChannel {
    std::vector<std::shared_ptr<ErrorChannel>> parts;
    std::vector<int> q;
    std::vector<std::vector<Complex>> none;
    std::vector<int> affectedQubits() const override { return q; }
    const std::vector<std::vector<Complex>> &kraus() const override {
        throw UnsupportedOperationException("CompositeChannel does not provide a single Kraus decomposition. "
                                            "Use getChannels() for sequential application.");
    }
    const std::vector<std::shared_ptr<ErrorChannel>> *components() const override { return &parts; }
};

double prob(const char *name, double v) {
    if (!(v >= 0.0 && v <= 1.0)) iae(std::string(name) + " must be in [0,1], got " + std::to_string(v));
    return v;
}

std::shared_ptr<ErrorChannel> make(int qubit, std::vector<std::vector<Complex>> k) {
    auto c = std::make_shared<KrausChannel>();
    c->q = {qubit};
    c->k = std::move(k);
    return c;
}

double bounded(const char *name, double v) {
    if (!std::isfinite(v)) throw IllegalStateException(std::string(name) + " must be finite");
    if (v < 0.0) {
        if (v >= -1e-12) return 0.0;
        throw IllegalStateException(std::string(name) + " must be in [0,1]");
    }
    if (v > 1.0) {
        if (v <= 1.0 + 1e-12) return 1.0;
        throw IllegalStateException(std::string(name) + " must be in [0,1]");
    }
    return v;
}

}  // namespace

// DepolarizingChannel.java:49-76
std::shared_ptr<ErrorChannel> ErrorChannel::depolarizing(double p, int qubit) {
    prob("probability", p);
    double a = std::sqrt(1.0 - p), s = std::sqrt(p / 3.0);
    return make(qubit, {{a, 0.0, 0.0, a}, {0.0, s, s, 0.0}, {0.0, Complex(0, -s), Complex(0, s), 0.0}, {s, 0.0, 0.0, -s}});
}
// AmplitudeDampingChannel.java:47-62
std::shared_ptr<ErrorChannel> ErrorChannel::amplitudeDamping(double g, int qubit) {
    prob("gamma", g);
    return make(qubit, {{1.0, 0.0, 0.0, std::sqrt(1.0 - g)}, {0.0, std::sqrt(g), 0.0, 0.0}});
}
// PhaseFlipChannel.java:50-64
std::shared_ptr<ErrorChannel> ErrorChannel::phaseFlip(double p, int qubit) {
    prob("probability", p);
    double a = std::sqrt(1.0 - p), s = std::sqrt(p);
    return make(qubit, {{a, 0.0, 0.0, a}, {s, 0.0, 0.0, -s}});
}
// PhaseDampingChannel.java:51-64
std::shared_ptr<ErrorChannel> ErrorChannel::phaseDamping(double l, int qubit) {
    prob("lambda", l);
    return make(qubit, {{1.0, 0.0, 0.0, std::sqrt(1.0 - l)}, {0.0, 0.0, 0.0, std::sqrt(l)}});
}
// ThermalRelaxationChannel.java:61-132
std::shared_ptr<ErrorChannel> ErrorChannel::thermalRelaxation(double t1, double t2, double gt, int qubit) {
    if (!(t1 > 0.0) || !std::isfinite(t1)) iae("T1 must be positive and finite");
    if (!(t2 > 0.0) || !std::isfinite(t2)) iae("T2 must be positive and finite");
    if (t2 > 2.0 * t1) iae("T2 must satisfy T2 <= 2*T1 (physical constraint)");
    if (gt < 0.0 || !std::isfinite(gt)) iae("gate time must be non-negative and finite");
    double surv = 1.0, deph = 1.0;
    if (gt != 0.0) {
        surv = bounded("relaxation survival", std::exp(-gt / t1));
        double it2 = 1.0 / t2, hit1 = 1.0 / (2.0 * t1), rate = it2 - hit1;
        if (rate < 0.0) {
            if (rate >= -1e-12 * std::max(1.0, std::fabs(it2) + std::fabs(hit1))) rate = 0.0;
            else throw IllegalStateException("pure dephasing rate must be non-negative");
        }
        deph = bounded("pure dephasing factor", std::exp(-gt * rate));
    }
    double decay = bounded("decay probability", 1.0 - surv), ss = std::sqrt(surv);
    double plus = bounded("thermal plus factor", 0.5 * (1.0 + deph));
    double minus = bounded("thermal minus factor", 0.5 * (1.0 - deph));
    double k0 = std::sqrt(plus), k1 = std::sqrt(minus);
    return make(qubit, {{k0, 0.0, 0.0, ss * k0}, {k1, 0.0, 0.0, -(ss * k1)}, {0.0, std::sqrt(decay), 0.0, 0.0}});
}
// CompositeChannel.java:19-79
std::shared_ptr<ErrorChannel> ErrorChannel::compose(std::vector<std::shared_ptr<ErrorChannel>> channels) {
    if (channels.empty()) iae("composite channel requires at least one channel");
    auto c = std::make_shared<Composite>();
    std::set<int> qs;
    for (auto &ch : channels)
        for (int q : ch->affectedQubits()) qs.insert(q);
    c->q.assign(qs.begin(), qs.end());
    c->parts = std::move(channels);
    return c;
}

NoiseModel &NoiseModel::afterGate(Operation::Kind kind, std::shared_ptr<ErrorChannel> ch) {
    if (!ch) throw NullPointerException("channel");
    gate_[kind].push_back(std::move(ch));
    return *this;
}
NoiseModel &NoiseModel::onQubits(std::shared_ptr<ErrorChannel> ch, std::vector<int> qubits) {
    if (!ch) throw NullPointerException("channel");
    if (qubits.empty()) iae("must specify at least one qubit");
    for (int q : qubits) {
        auto it = std::find_if(qubit_.begin(), qubit_.end(), [&](auto &e) { return e.first == q; });
        if (it == qubit_.end()) qubit_.push_back({q, {ch}});
        else it->second.push_back(ch);
    }
    return *this;
}
NoiseModel &NoiseModel::afterAllGates(std::shared_ptr<ErrorChannel> ch) {
    if (!ch) throw NullPointerException("channel");
    global_.push_back(std::move(ch));
    return *this;
}
bool NoiseModel::hasNoise() const { return !gate_.empty() || !qubit_.empty() || !global_.empty(); }

// NoiseModel.scheduledChannelsAfter (NoiseModel.java:211-241); idle channels iterate in java.util.HashMap order.
std::vector<NoiseModel::ScheduledChannel> NoiseModel::scheduledChannelsAfter(const Operation &op) const {
    std::vector<ScheduledChannel> out;
    auto g = gate_.find(op.kind);
    if (g != gate_.end())
        for (auto &ch : g->second) out.push_back({ch, ch->affectedQubits()});
    std::vector<int> aff = op.affectedQubits();
    size_t cap = 16;
    while (qubit_.size() > cap * 3 / 4) cap *= 2;
    std::vector<size_t> order(qubit_.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    auto bucket = [&](int k) {
        uint32_t h = (uint32_t)k;
        return (h ^ (h >> 16)) & (uint32_t)(cap - 1);
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return bucket(qubit_[a].first) < bucket(qubit_[b].first); });
    for (size_t i : order) {
        int q = qubit_[i].first;
        if (std::find(aff.begin(), aff.end(), q) != aff.end()) continue;
        for (auto &ch : qubit_[i].second) out.push_back({ch, {q}});
    }
    for (auto &ch : global_) out.push_back({ch, ch->affectedQubits()});
    return out;
}

// ---- requests ----------------------------------------------------------------------------------
MeasurementInstruction::MeasurementInstruction(int s, std::optional<std::vector<int>> mq, std::optional<int64_t> sd,
                                               Mode m)
    : shots(s), measuredQubits(std::move(mq)), seed(sd), mode(m) {
    if (shots <= 0) iae("shots must be positive");
    if (measuredQubits) {
        auto &q = *measuredQubits;
        if (q.empty()) iae("measured qubit list must not be empty");
        for (int x : q)
            if (x < 0) iae("qubit index must be non-negative: " + std::to_string(x));
        std::sort(q.begin(), q.end());
        for (size_t i = 1; i < q.size(); i++)
            if (q[i] == q[i - 1]) iae("duplicate qubit index: " + std::to_string(q[i]));
    }
}
MeasurementInstruction MeasurementInstruction::countsAll(int s) { return {s, std::nullopt, std::nullopt, Mode::COUNTS}; }
MeasurementInstruction MeasurementInstruction::counts(int s, std::vector<int> q) {
    if (q.empty()) iae("at least one qubit must be provided");
    return {s, std::move(q), std::nullopt, Mode::COUNTS};
}
MeasurementInstruction MeasurementInstruction::samplesAll(int s) { return {s, std::nullopt, std::nullopt, Mode::SAMPLES}; }
MeasurementInstruction MeasurementInstruction::samples(int s, std::vector<int> q) {
    if (q.empty()) iae("at least one qubit must be provided");
    return {s, std::move(q), std::nullopt, Mode::SAMPLES};
}
MeasurementInstruction MeasurementInstruction::withSeed(int64_t s) const { return {shots, measuredQubits, s, mode}; }

SimulationRequest::SimulationRequest(QuantumCircuit circuit, std::optional<StateVector> initialState,
                                     std::optional<std::vector<int>> basis, std::optional<MeasurementInstruction> m,
                                     std::optional<NoiseModel> noise, std::optional<int64_t> noiseSeed, bool ret)
    : circuit_(std::move(circuit)), initial_(std::move(initialState)), basis_(std::move(basis)), measurement_(std::move(m)),
      noise_(std::move(noise)), noiseSeed_(noiseSeed), returnFinalState_(ret) {
    const int n = circuit_.qubitCount();
    if (initial_ && initial_->qubitCount() != n)
        iae("Initial state qubit count " + std::to_string(initial_->qubitCount()) +
            " does not match circuit qubit count " + std::to_string(n));
    if (basis_) {
        auto &q = *basis_;
        for (int x : q)
            if (x < 0 || x >= n) iae("basis-state qubit out of range: " + std::to_string(x));
        std::sort(q.begin(), q.end());
        for (size_t i = 1; i < q.size(); i++)
            if (q[i] == q[i - 1]) iae("duplicate basis-state qubit: " + std::to_string(q[i]));
    }
    if (initial_ && basis_) iae("initialState and basisStateQubits are mutually exclusive");
    if (measurement_ && measurement_->measuredQubits)
        for (int q : *measurement_->measuredQubits)
            if (q < 0 || q >= n) iae("measurement qubit out of range: " + std::to_string(q));
    if (!returnFinalState_ && !measurement_)
        iae("simulation must request the final state or provide a measurement instruction");
}
SimulationRequest SimulationRequest::zeroState(const QuantumCircuit &c) {
    return {c, std::nullopt, std::nullopt, std::nullopt, std::nullopt, std::nullopt, true};
}
SimulationRequest SimulationRequest::withInitialState(const QuantumCircuit &c, const StateVector &s) {
    return {c, s, std::nullopt, std::nullopt, std::nullopt, std::nullopt, true};
}
SimulationRequest SimulationRequest::withMeasurement(const MeasurementInstruction &m) const {
    return {circuit_, initial_, basis_, m, noise_, noiseSeed_, returnFinalState_};
}
SimulationRequest SimulationRequest::withMeasurement(const MeasurementInstruction &m, bool inc) const {
    return {circuit_, initial_, basis_, m, noise_, noiseSeed_, inc};
}
SimulationRequest SimulationRequest::withoutFinalState() const {
    if (!measurement_) throw IllegalStateException("cannot drop final state without measurement instruction");
    return {circuit_, initial_, basis_, measurement_, noise_, noiseSeed_, false};
}
SimulationRequest SimulationRequest::withNoiseModel(const NoiseModel &m) const {
    return {circuit_, initial_, basis_, measurement_, m, noiseSeed_, returnFinalState_};
}
SimulationRequest SimulationRequest::withNoiseSeed(int64_t s) const {
    return {circuit_, initial_, basis_, measurement_, noise_, s, returnFinalState_};
}
SimulationRequest SimulationRequest::withBasisState(std::vector<int> q) const {
    return {circuit_, std::nullopt, std::move(q), measurement_, noise_, noiseSeed_, returnFinalState_};
}

// ---- engine ------------------------------------------------------------------------------------
HipStatevectorEngine::HipStatevectorEngine() : HipStatevectorEngine(HipStatevectorConfig::automatic()) {}

HipStatevectorEngine::HipStatevectorEngine(const HipStatevectorConfig &cfg) {
    check_status(sc_ctx_create(cfg.device, &ctx_));
    check_status(sc_ctx_set_option(ctx_, SC_OPT_EXEC_MODE, cfg.execMode));
    check_status(sc_ctx_set_option(ctx_, SC_OPT_MAX_FUSED, cfg.maxFusedQubits));
}

HipStatevectorEngine::~HipStatevectorEngine() { sc_ctx_destroy(ctx_); }

namespace {
struct StateHandle {
    sc_state *s = nullptr;
    ~StateHandle() { sc_state_destroy(s); }
};
uint64_t random_seed() {
    std::random_device rd;
    return ((uint64_t)rd() << 32) ^ rd();
}
}  // namespace

// StatevectorEngine.simulate (StatevectorEngine.java:63-167)
SimulationResult HipStatevectorEngine::simulate(const SimulationRequest &req) {
    const QuantumCircuit &c = req.circuit();
    const int n = c.qubitCount();
    StateHandle st;
    check_status(sc_state_create(ctx_, n, 1, &st.s));
    if (req.initialState()) {
        check_status(sc_state_upload(st.s, req.initialState()->data().data(), 1ULL << n));
    } else if (req.basisStateQubits()) {
        uint64_t idx = 0;
        for (int q : *req.basisStateQubits()) idx |= 1ULL << q;
        check_status(sc_state_set_basis(st.s, idx));
    }
    std::vector<sc_op> ops;
    std::vector<double> mats;
    c.lower(ops, mats);
    if (mats.empty()) mats.assign(2, 0.0);
    const auto &noise = req.noiseModel();
    if (noise && noise->hasNoise()) {
        // Every scheduled channel and target becomes an sc_run_noisy event whose nextDouble() is drawn here in
        // circuit order; the branch is chosen on the device from the live weights.  A validation error ends
        // the event list where the reference raises it, after the events before it have run.
        SplittableRandom rng(req.noiseSeed() ? *req.noiseSeed() : (int64_t)random_seed());
        std::vector<sc_noise_event> events;
        std::vector<double> kraus;
        std::exception_ptr pending;
        size_t end = ops.size();
        struct Collect {
            std::vector<sc_noise_event> &events;
            std::vector<double> &kraus;
            int n;
            uint64_t after;
            // applyErrorChannel (StatevectorEngine.java:382-407), composite recursion included
            void operator()(const ErrorChannel &ch, const std::vector<int> *mapped) {
                if (auto comps = ch.components()) {
                    const std::vector<int> *m = (mapped && mapped->size() == 1) ? mapped : nullptr;
                    for (auto &part : *comps) (*this)(*part, m);
                    return;
                }
                std::vector<int> targets = mapped ? *mapped : ch.affectedQubits();
                if (targets.empty()) iae("noise channel must target at least one qubit");
                std::set<int> seen;
                for (int t : targets) {
                    if (t < 0 || t >= n) iae("noise channel qubit out of range: " + std::to_string(t));
                    if (!seen.insert(t).second) iae("noise channel contains duplicate qubit: " + std::to_string(t));
                }
                const uint64_t first = kraus.size() / 2;
                for (auto &op : ch.kraus())
                    for (auto &v : op) kraus.push_back(v.real()), kraus.push_back(v.imag());
                for (int t : targets) events.push_back({after, t, (int32_t)ch.kraus().size(), first});
            }
        };
        bool measured = false;
        for (size_t i = 0; i < ops.size() && !pending; i++) {
            if (ops[i].kind == SC_OP_MEASURE) {
                measured = true;
                continue;
            }
            if (measured) {
                pending = std::make_exception_ptr(
                    UnsupportedOperationException("Unitary operations cannot follow measurement operations in the circuit"));
                end = i;
                break;
            }
            try {
                Collect collect{events, kraus, n, (uint64_t)i};
                for (auto &sch : noise->scheduledChannelsAfter(c.operations()[i])) collect(*sch.channel, &sch.targetQubits);
            } catch (const IllegalArgumentException &) {
                pending = std::current_exception();
                end = i + 1;
            }
        }
        std::vector<double> u(events.size());
        for (auto &x : u) x = rng.nextDouble();
        if (kraus.empty()) kraus.assign(8, 0.0);
        if (u.empty()) u.assign(1, 0.0);
        sc_noise_failure failure;
        const int rc = sc_run_noisy(st.s, ops.data(), end, mats.data(), mats.size() / 2, events.data(), events.size(),
                                    kraus.data(), kraus.size() / 2, u.data(), &failure);
        if (rc == SC_E_STATE && failure.event >= 0) {
            if (failure.kind == 1) throw IllegalStateException("invalid Kraus branch weight");
            if (failure.kind == 2) throw IllegalStateException("invalid Kraus branch weights");
            throw IllegalStateException("state collapsed to invalid norm after Kraus application");
        }
        check_status(rc);
        if (pending) std::rethrow_exception(pending);
    } else {
        check_status(sc_apply_ops(st.s, ops.data(), ops.size(), mats.data(), mats.size() / 2));
    }
    SimulationResult res;
    if (req.returnFinalState()) {
        std::vector<double> data(size_t(2) << n);
        check_status(sc_state_download(st.s, data.data(), 1ULL << n));
        res.finalState = StateVector(n, std::move(data));
    }
    if (req.measurement()) {
        const MeasurementInstruction &mi = *req.measurement();
        std::vector<int> measured;
        if (mi.measuredQubits) {
            measured = *mi.measuredQubits;
        } else {
            std::set<int> mq;
            for (auto &op : c.operations())
                if (op.kind == Operation::Kind::Measure) mq.insert(op.qubits.begin(), op.qubits.end());
            if (!mq.empty()) measured.assign(mq.begin(), mq.end());
            else
                for (int q = 0; q < n; q++) measured.push_back(q);
        }
        if (measured.empty()) throw IllegalStateException("No qubits available for measurement");
        SplittableRandom rng(mi.seed ? *mi.seed : (int64_t)random_seed());
        std::vector<double> u(mi.shots);
        for (auto &x : u) x = rng.nextDouble();
        std::vector<int64_t> outs(mi.shots);
        std::vector<int32_t> m32(measured.begin(), measured.end());
        check_status(sc_sample(st.s, 0, m32.data(), (int)m32.size(), u.data(), mi.shots, outs.data(), nullptr));
        MeasurementResult mr;
        mr.measuredQubits = measured;
        mr.shots = mi.shots;
        const int m = (int)measured.size();
        auto bits = [&](int64_t v) {
            std::string s(m, '0');
            for (int k = 0; k < m; k++)
                if ((v >> k) & 1) s[m - 1 - k] = '1';
            return s;
        };
        if (mi.mode == MeasurementInstruction::Mode::COUNTS) {
            if (m <= 31) {
                mr.type = MeasurementResult::Type::Histogram;
                for (int64_t o : outs) mr.counts[o]++;
            } else {
                mr.type = MeasurementResult::Type::BitstringHistogram;
                for (int64_t o : outs) mr.bitstringCounts[bits(o)]++;
            }
        } else {
            if (m <= 31) {
                mr.type = MeasurementResult::Type::Samples;
                mr.samples = outs;
            } else {
                mr.type = MeasurementResult::Type::BitstringSamples;
                for (int64_t o : outs) mr.bitstringSamples.push_back(bits(o));
            }
        }
        res.measurement = std::move(mr);
    }
    return res;
}

std::vector<std::string> SimulatorEngines::ids() { return {HipStatevectorEngine::ID}; }
std::unique_ptr<SimulatorEngine> SimulatorEngines::findById(const std::string &id) {
    if (id == HipStatevectorEngine::ID) return std::make_unique<HipStatevectorEngine>();
    return nullptr;
}
std::unique_ptr<SimulatorEngine> SimulatorEngines::require(const std::string &id) {
    auto e = findById(id);
    if (!e) iae("No simulator engine registered with id '" + id + "'");
    return e;
}

}  // namespace statecraft
