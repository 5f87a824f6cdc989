// Host C++ layer of sparqsim_amd: the reference's register table (System),
// the device-resident SparseState and the BaseOperator hierarchy
// (SparQ/include/basic_components.h:221-830), driving the MI355X engine
// through the C ABI in include/sparq_hip.h.
//
// Registers are packed into the key most-significant-first in register-id
// order, each field inside one 64-bit word, so that unsigned lexicographic
// order of the packed words equals System::operator<
// (SparQ/src/basic_components.cpp:461-474).
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../../include/sparq_hip.h"

namespace sparqsim {

using complex_t = std::complex<double>;
using u22_t = std::array<complex_t, 4>;   // row-major [u00,u01,u10,u11]

enum StateStorageType { General = 0, UnsignedInteger = 1, SignedInteger = 2, Boolean = 3, Rational = 4 };

constexpr double pi = 3.141592653589793238462643383279502884;
constexpr double sqrt2inv = 0.7071067811865475244008443621048490;
constexpr double epsilon = 1e-14;

inline uint64_t width_mask(size_t w) { return w >= 64 ? ~uint64_t(0) : (uint64_t(1) << w) - 1; }

// Throws the exception matching an engine status code.
void check(int status);

// ------------------------------------------------------------ register table
using StateInfoType = std::tuple<std::string, StateStorageType, size_t, bool>;

// One register value of a host-side entry, with the reference's typed views
// (SparQ/include/basic_components.h:90-131): Rational = value / 2^w, SignedInteger
// = two's complement of the w-bit value, Boolean = value != 0.
struct StateStorage {
    uint64_t value = 0;
    template <typename Ty> Ty as(size_t size) const {
        const uint64_t v = value & width_mask(size);
        if constexpr (std::is_floating_point_v<Ty>) {
            return size >= 64 ? Ty(v * 1.0 / 2 / 9223372036854775808.0) : Ty(1.0 * v / double(uint64_t(1) << size));
        } else if constexpr (std::is_same_v<Ty, bool>) {
            return v != 0;
        } else if constexpr (std::is_signed_v<Ty>) {
            if (size == 0 || size >= 64) return Ty(int64_t(v));
            return Ty((v >> (size - 1)) & 1 ? int64_t(v | ~width_mask(size)) : int64_t(v));
        } else {
            return Ty(v);
        }
    }
    void flip(size_t digit) { value ^= uint64_t(1) << digit; }
    bool operator==(const StateStorage& o) const { return value == o.value; }
    bool operator!=(const StateStorage& o) const { return value != o.value; }
    bool operator<(const StateStorage& o) const { return value < o.value; }
    bool operator>(const StateStorage& o) const { return value > o.value; }
};

// Process-global register table with the reference's id-reuse rules
// (SparQ/src/basic_components.cpp:341-459).  An instance is one host-side
// entry (amplitude + a value per register id), the element of the
// std::vector<System> view that host-side operators receive.
struct System {
    complex_t amplitude = 1.0;
    mutable std::vector<StateStorage> registers;
    System() : registers(name_register_map.size()) {}
    StateStorage& get(size_t id) {
        if (id >= name_register_map.size()) throw std::runtime_error("Register not found.");
        if (registers.size() < name_register_map.size()) registers.resize(name_register_map.size());
        return registers[id];
    }
    const StateStorage& get(size_t id) const {
        if (id >= name_register_map.size()) throw std::runtime_error("Register not found.");
        if (registers.size() < name_register_map.size()) registers.resize(name_register_map.size());
        return registers[id];
    }
    StateStorage& last_register() { return get(get_last_activated_register()); }
    const StateStorage& last_register() const { return get(get_last_activated_register()); }
    // Canonical order: active registers in ascending id (basic_components.cpp:461-490).
    bool operator<(const System& rhs) const;
    bool operator==(const System& rhs) const;
    bool operator!=(const System& rhs) const { return !(*this == rhs); }
    std::string to_string() const;

    static std::vector<StateInfoType> name_register_map;
    static std::unordered_map<std::string, size_t> name_to_index;
    static std::vector<size_t> temporal_registers;
    static std::vector<size_t> reusable_registers;
    static size_t max_qubit_count, max_register_count, max_system_size;
    static uint64_t layout_version;   // bumps on every table change

    static void clear();
    static size_t get_qubit_count();
    static size_t get_activated_register_size();
    static size_t get_last_activated_register();
    static void update_max_size(size_t n);
    static size_t get(std::string_view name);
    static StateInfoType get_register_info(std::string_view name);
    static const std::string& name_of(size_t id);
    static size_t size_of(std::string_view name);
    static size_t size_of(size_t id);
    static StateStorageType type_of(std::string_view name);
    static StateStorageType type_of(size_t id);
    static bool status_of(std::string_view name);
    static bool status_of(size_t id);
    static size_t add_register(std::string_view name, StateStorageType type, size_t size);
    static void remove_register(size_t id);
    static void remove_register(std::string_view name);
    static void set_size(size_t id, size_t size);
    static void set_type(size_t id, StateStorageType t);
};

// Bit-field assignment of the active registers.
struct Layout {
    uint32_t words = 1;
    std::vector<sparq_field_t> field;   // by register id; width 0 = inactive
    static Layout from_registry();
    bool operator==(const Layout& o) const;
    sparq_field_t at(size_t id) const;
};

// ------------------------------------------------------------ state
class SparseState {
public:
    SparseState();                                 // |0...0>, amplitude 1
    explicit SparseState(size_t) = delete;
    SparseState(const SparseState& other);
    SparseState& operator=(const SparseState& other);
    SparseState(SparseState&& other) noexcept;
    SparseState& operator=(SparseState&& other) noexcept;
    ~SparseState();

    size_t size();
    bool empty() { return size() == 0; }
    sparq_state_t handle() { return h_; }

    // Bring the device keys to the current register-table layout (new
    // registers read as 0, removed ones are dropped).  Flushes pending work.
    const Layout& sync();
    const Layout& layout() const { return layout_; }
    // Replace the layout after an explicit field move (Split/Combine).
    void relayout_with(const Layout& target, const std::vector<std::pair<sparq_field_t, sparq_field_t>>& moves);

    // Record that register `id` is zero in every entry (checked by the caller);
    // the next layout change then drops its field without clearing it.
    void mark_zero(size_t id) { zeroed_.push_back(id); }

    // Queue a per-entry op; consecutive ops run fused in one kernel.
    void push_pointwise(const sparq_pw_op_t& op, const std::vector<sparq_cond_t>& conds);
    void flush();

    // Host copies: values[i][k] for the active register ids (ascending), amplitudes.
    void export_entries(std::vector<size_t>& ids, std::vector<uint64_t>& values, std::vector<complex_t>& amps);
    void import_entries(const std::vector<uint64_t>& values, const std::vector<complex_t>& amps);
    // Host view as the reference's std::vector<System> (download / upload, entry order kept).
    void to_host(std::vector<System>& entries);
    void from_host(const std::vector<System>& entries);
    // Dense post-selection on the device (GetOutput without the renormalisation):
    // out[index] for entries whose zero_ids registers are 0, index = the index_ids
    // registers concatenated (first least significant); returns the selected weight.
    // A nonzero `length` keeps only indices below it (out has `length` values).
    double post_select_dense(const std::vector<size_t>& zero_ids, const std::vector<size_t>& index_ids,
                             std::vector<complex_t>& out, uint64_t length = 0);
    std::string to_string(int precision = 6);

private:
    sparq_state_t h_ = nullptr;
    Layout layout_;
    uint64_t layout_version_ = ~0ull;   // System::layout_version the layout was checked against
    std::vector<sparq_pw_op_t> pend_ops_;
    std::vector<sparq_cond_t> pend_conds_;
    std::vector<size_t> zeroed_;   // registers known to be zero in every entry (see mark_zero)
};

// ------------------------------------------------------------ conditions
struct Controllable {
    std::vector<size_t> condition_variable_nonzeros;
    std::vector<size_t> condition_variable_all_ones;
    std::vector<std::pair<size_t, size_t>> condition_variable_by_bit;
    // bit == 0 (no counterpart in the reference's Controllable; see SPARQ_COND_BIT_CLEAR)
    std::vector<std::pair<size_t, size_t>> condition_variable_by_clear_bit;
    std::vector<std::pair<size_t, uint64_t>> condition_variable_by_value;
    bool has_condition() const {
        return !condition_variable_nonzeros.empty() || !condition_variable_all_ones.empty() ||
               !condition_variable_by_bit.empty() || !condition_variable_by_value.empty() ||
               !condition_variable_by_clear_bit.empty();
    }
    void clear_control_nonzeros() { condition_variable_nonzeros.clear(); }
    void clear_control_all_ones() { condition_variable_all_ones.clear(); }
    void clear_control_by_bit() { condition_variable_by_bit.clear(); }
    void clear_control_by_clear_bit() { condition_variable_by_clear_bit.clear(); }
    void clear_control_by_value() { condition_variable_by_value.clear(); }
    // Resolve to engine predicates for the given layout.
    std::vector<sparq_cond_t> conditions(const Layout& L) const;
    template <class T> void copy_control_conditions_to(T& t) const {
        t.condition_variable_nonzeros = condition_variable_nonzeros;
        t.condition_variable_all_ones = condition_variable_all_ones;
        t.condition_variable_by_bit = condition_variable_by_bit;
        t.condition_variable_by_clear_bit = condition_variable_by_clear_bit;
        t.condition_variable_by_value = condition_variable_by_value;
    }
};

// ------------------------------------------------------------ operators
// An operator acts on the device state (operator()(SparseState&)) or on the
// reference's host view (operator()(std::vector<System>&),
// basic_components.h:748-830); a class overrides either.  The other form is
// derived: a host-side operator applied to a SparseState downloads the entries,
// runs, and uploads them; an engine operator applied to a std::vector<System>
// uploads it into a temporary state, runs on the device, and downloads.  An
// operator overriding neither form (or no dag) throws.
struct BaseOperator {
    virtual ~BaseOperator() = default;
    virtual void operator()(SparseState& state) const;
    virtual void operator()(std::vector<System>& state) const;
    virtual void dag(SparseState& state) const;
    virtual void dag(std::vector<System>& state) const;
    virtual Controllable* controls() { return nullptr; }
};
struct SelfAdjointOperator : BaseOperator {
    using BaseOperator::operator();
    void dag(SparseState& state) const override { (*this)(state); }
    void dag(std::vector<System>& state) const override { (*this)(state); }
};
// Re-exposes both forms in a class that overrides one of them.
#define SPQ_BOTH_VIEWS using BaseOperator::operator(); using BaseOperator::dag;

// A fixed list of operators applied in order (dag: their adjoints in reverse
// order), so that a circuit built once is run by one native call.  No
// counterpart in the reference, whose C++ callers loop over operators
// themselves.
struct OperatorSequence : BaseOperator {
    SPQ_BOTH_VIEWS
    std::vector<std::shared_ptr<BaseOperator>> ops;
    explicit OperatorSequence(std::vector<std::shared_ptr<BaseOperator>> o);
    void operator()(SparseState& state) const override;
    void dag(SparseState& state) const override;
};

// Plugin entry points catch exceptions (they must not cross the C boundary) and
// park the message here; the Python proxy takes it and raises.
void record_plugin_error(const char* message);
std::string take_plugin_error();

double& clearzero_default_eps();

// Per-entry operator described by an engine opcode over register ids.
constexpr size_t kNoRegister = ~size_t(0);   // an absent operand in PointwiseOperator::regs
struct PointwiseOperator : BaseOperator, Controllable {
    std::string name;
    uint32_t opcode = 0, dag_opcode = 0;    // dag_opcode 0: self-adjoint
    std::vector<size_t> regs;               // operand register ids (f0..f3)
    uint64_t c[3] = {0, 0, 0};
    uint64_t dag_c[3] = {0, 0, 0};
    double d[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double dag_d[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::shared_ptr<void> table_owner;      // keeps a QRAM table alive
    const void* table = nullptr;
    uint32_t table_word_bytes = 0;
    std::vector<std::pair<size_t, uint64_t>> extra_zero_regs;   // implicit "== value" predicates
    std::vector<std::pair<size_t, uint64_t>> extra_bit_regs;    // implicit "bit set" predicates
    bool has_dag = true;
    Controllable* controls() override { return this; }
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override { emit(state, false); }
    void dag(SparseState& state) const override;
    void emit(SparseState& state, bool dagger) const;
};

// ------------------------------------------------------------ captured sequences
// A fixed operator sequence recorded once on one state and replayed as a HIP
// graph with no host synchronisation (no counterpart in the reference, whose
// callers loop over operators and wait after every branching one).  The body
// runs twice: once to plan each branching operator from a snapshot of the
// state, once to record the kernels.  It must leave the register layout as it
// found it and use only per-entry operators, Hadamard_Int / Hadamard_Int_Full,
// register add/remove that keeps fields in place, PauliNoise and
// AdvancePauliNoise.  Replays assume what the plan saw (the key bits that vary
// outside each Hadamard's target, its segments, no duplicate keys); a replay
// that breaks it is reported by the next access to the state.  The state must
// outlive the sequence, and its arrays must not be reallocated between
// launches (an eager branching operator does).
class CapturedSequence {
public:
    CapturedSequence(SparseState& state, const std::function<void(SparseState&)>& body);
    ~CapturedSequence();
    CapturedSequence(const CapturedSequence&) = delete;
    CapturedSequence& operator=(const CapturedSequence&) = delete;
    void launch(size_t times = 1);   // queue replays; no synchronisation
    void sync();                     // wait for them and check their assumptions
    uint64_t capacity() const;       // entries the sequence may hold
    uint64_t nodes() const;          // kernels (graph nodes) per replay
    SparseState& state() { return *st_; }

private:
    SparseState* st_;
    sparq_graph_t g_ = nullptr;
};

// Seeded Pauli X/Z insertions on one register, per (iteration, segment), held
// in a device table so that a noise channel whose events change every
// iteration replays inside one captured sequence.  x and z are [iteration][row]
// bit masks; the table's iteration counter starts at 0.
struct PauliNoiseTable {
    std::shared_ptr<struct sparq_table_s> table;
    size_t rows = 0, iterations = 0;
    PauliNoiseTable(size_t rows, size_t iterations, const std::vector<uint64_t>& x, const std::vector<uint64_t>& z);
    void set_iteration(uint64_t k);
};
// For the entries whose `segment` register holds row s < rows: `target` ^= the
// X mask of (current iteration, s), then a Z (phase pi, as Z_Bool) per set bit
// of the Z mask on the flipped value, lowest bit first: the same amplitudes as
// X_Bool(target, i) then Z_Bool(target, i) for the drawn i in ascending order,
// each conditioned on segment == s.
struct PauliNoise : PointwiseOperator {
    std::shared_ptr<struct sparq_table_s> keep;
    PauliNoise(std::string_view segment, std::string_view target, const PauliNoiseTable& t)
        : PauliNoise(System::get(segment), System::get(target), t) {}
    PauliNoise(size_t segment, size_t target, const PauliNoiseTable& t);
};
// The table's iteration counter += 1, ordered on the state's stream.
struct AdvancePauliNoise : BaseOperator {
    SPQ_BOTH_VIEWS
    std::shared_ptr<struct sparq_table_s> table;
    explicit AdvancePauliNoise(const PauliNoiseTable& t) : table(t.table) {}
    void operator()(SparseState& state) const override;
};

struct QRAMCircuit_qutrit {
    size_t address_size, data_size;
    std::vector<uint64_t> memory;
    std::shared_ptr<struct sparq_table_s> table;   // device copy
    uint32_t word_bytes = 8;
    uint64_t max_word = 0;
    QRAMCircuit_qutrit(size_t addr, size_t data, std::vector<uint64_t> mem);
};

// ---- per-entry operator classes (names and constructor arguments follow the
// reference classes in quantum_arithmetic.h, basic_gates.h, qram.h,
// parallel_phase_operations.h and dark_magic.h)
#define SPQ_PW_CLASS_3(NAME)                                                                        \
    struct NAME : PointwiseOperator {                                                               \
        NAME(std::string_view a, std::string_view b, std::string_view c)                            \
            : NAME(System::get(a), System::get(b), System::get(c)) {}                               \
        NAME(size_t a, size_t b, size_t c);                                                         \
    };
#define SPQ_PW_CLASS_2(NAME)                                                                        \
    struct NAME : PointwiseOperator {                                                               \
        NAME(std::string_view a, std::string_view b) : NAME(System::get(a), System::get(b)) {}      \
        NAME(size_t a, size_t b);                                                                   \
    };
#define SPQ_PW_CLASS_4(NAME)                                                                        \
    struct NAME : PointwiseOperator {                                                               \
        NAME(std::string_view a, std::string_view b, std::string_view c, std::string_view d)        \
            : NAME(System::get(a), System::get(b), System::get(c), System::get(d)) {}               \
        NAME(size_t a, size_t b, size_t c, size_t d);                                               \
    };
#define SPQ_PW_CLASS_RCR(NAME)                                                                      \
    struct NAME : PointwiseOperator {                                                               \
        NAME(std::string_view a, uint64_t k, std::string_view b) : NAME(System::get(a), k, System::get(b)) {} \
        NAME(size_t a, uint64_t k, size_t b);                                                       \
    };
#define SPQ_PW_CLASS_RC(NAME)                                                                       \
    struct NAME : PointwiseOperator {                                                               \
        NAME(std::string_view a, uint64_t k) : NAME(System::get(a), k) {}                           \
        NAME(size_t a, uint64_t k);                                                                 \
    };

SPQ_PW_CLASS_3(Add_UInt_UInt)
SPQ_PW_CLASS_3(Sub_UInt_UInt)
SPQ_PW_CLASS_3(Mul_UInt_UInt)
SPQ_PW_CLASS_3(Div_UInt_UInt)
SPQ_PW_CLASS_3(And_UInt_UInt)
SPQ_PW_CLASS_3(Or_UInt_UInt)
SPQ_PW_CLASS_3(Xor_UInt_UInt)
SPQ_PW_CLASS_3(Less_SInt_SInt)
SPQ_PW_CLASS_3(Less_UInt_UInt)
SPQ_PW_CLASS_3(GetMid_UInt_UInt)
SPQ_PW_CLASS_3(Div_Sqrt_Arccos_UInt_UInt)
SPQ_PW_CLASS_3(Sqrt_Div_Arccos_Int_UInt)
SPQ_PW_CLASS_3(GetRotateAngle_Int_Int)
SPQ_PW_CLASS_2(Add_UInt_UInt_InPlace)
SPQ_PW_CLASS_2(Neg_UInt)
SPQ_PW_CLASS_2(Abs_SInt)
SPQ_PW_CLASS_2(Sqrt_UInt)
SPQ_PW_CLASS_2(IsZero_UInt)
SPQ_PW_CLASS_2(Negative_SInt)
SPQ_PW_CLASS_2(Add_AnyInt_AnyInt_InPlace)
SPQ_PW_CLASS_2(Assign)
SPQ_PW_CLASS_2(Swap_General_General)
SPQ_PW_CLASS_4(Select_Bool_UInt_UInt)
SPQ_PW_CLASS_4(Carry_UInt_UInt)
SPQ_PW_CLASS_4(Overflow_SInt_SInt)
SPQ_PW_CLASS_4(MulOverflow_UInt_UInt)
SPQ_PW_CLASS_4(Compare_UInt_UInt)
SPQ_PW_CLASS_RCR(Mult_UInt_ConstUInt)
SPQ_PW_CLASS_RCR(Add_Mult_UInt_ConstUInt_InPlace)
SPQ_PW_CLASS_RCR(Add_UInt_ConstUInt)
SPQ_PW_CLASS_RC(Add_ConstUInt_InPlace)
SPQ_PW_CLASS_RC(ShiftLeft_InPlace)
SPQ_PW_CLASS_RC(ShiftRight_InPlace)
SPQ_PW_CLASS_RC(X_Bool)
SPQ_PW_CLASS_RC(Y_Bool)
SPQ_PW_CLASS_RC(Init_Unsafe)

struct FlipBools : PointwiseOperator {
    explicit FlipBools(std::string_view r) : FlipBools(System::get(r)) {}
    explicit FlipBools(size_t r);
};
struct Mod_Mult_UInt_ConstUInt_InPlace : PointwiseOperator {
    Mod_Mult_UInt_ConstUInt_InPlace(std::string_view r, uint64_t a, uint64_t x, uint64_t N)
        : Mod_Mult_UInt_ConstUInt_InPlace(System::get(r), a, x, N) {}
    Mod_Mult_UInt_ConstUInt_InPlace(size_t r, uint64_t a, uint64_t x, uint64_t N);
};
struct Swap_Bool_Bool : PointwiseOperator {
    Swap_Bool_Bool(std::string_view a, size_t d1, std::string_view b, size_t d2)
        : Swap_Bool_Bool(System::get(a), d1, System::get(b), d2) {}
    Swap_Bool_Bool(size_t a, size_t d1, size_t b, size_t d2);
};
struct Phase_Bool : PointwiseOperator {
    Phase_Bool(std::string_view r, size_t digit, double lambda) : Phase_Bool(System::get(r), digit, lambda) {}
    Phase_Bool(size_t r, size_t digit, double lambda);
};
// Z/S/T: Phase_Bool with lambda = pi, pi/2, pi/4 on bit `digit` (basic_gates.h:390-450)
struct Z_Bool : Phase_Bool {
    explicit Z_Bool(std::string_view r, size_t digit = 0) : Phase_Bool(r, digit, pi) {}
    explicit Z_Bool(size_t r, size_t digit = 0) : Phase_Bool(r, digit, pi) {}
};
struct S_Bool : Phase_Bool {
    explicit S_Bool(std::string_view r, size_t digit = 0) : Phase_Bool(r, digit, pi / 2) {}
    explicit S_Bool(size_t r, size_t digit = 0) : Phase_Bool(r, digit, pi / 2) {}
};
struct T_Bool : Phase_Bool {
    explicit T_Bool(std::string_view r, size_t digit = 0) : Phase_Bool(r, digit, pi / 4) {}
    explicit T_Bool(size_t r, size_t digit = 0) : Phase_Bool(r, digit, pi / 4) {}
};
struct RZ_Bool : PointwiseOperator {
    RZ_Bool(std::string_view r, size_t digit, double angle) : RZ_Bool(System::get(r), digit, angle) {}
    RZ_Bool(size_t r, size_t digit, double angle);
};
struct QRAMLoad : PointwiseOperator {
    QRAMLoad(const QRAMCircuit_qutrit* q, std::string_view a, std::string_view d) : QRAMLoad(q, System::get(a), System::get(d)) {}
    QRAMLoad(const QRAMCircuit_qutrit* q, size_t a, size_t d);
};
using QRAMLoadFast = QRAMLoad;
struct ZeroConditionalPhaseFlip : PointwiseOperator {
    explicit ZeroConditionalPhaseFlip(const std::vector<size_t>& ids);
    explicit ZeroConditionalPhaseFlip(const std::vector<std::string>& names);
};
struct Reflection_Bool : PointwiseOperator {
    bool inverse;
    Reflection_Bool(const std::vector<size_t>& ids, bool inverse = false);
    Reflection_Bool(const std::vector<std::string>& names, bool inverse = false);
    Reflection_Bool(std::string_view name, bool inverse = false) : Reflection_Bool(std::vector<size_t>{System::get(name)}, inverse) {}
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
    void dag(SparseState& state) const override { (*this)(state); }
};
struct GlobalPhase : PointwiseOperator {
    explicit GlobalPhase(complex_t c);
};
// main += 1 modulo 2^w, flipping `overflow` on wrap-around
// (SparQ_Algorithm/include/BlockEncoding/block_encoding_tridiagonal.h:30-60)
struct PlusOneAndOverflow : PointwiseOperator {
    PlusOneAndOverflow(std::string_view main, std::string_view overflow) : PlusOneAndOverflow(System::get(main), System::get(overflow)) {}
    PlusOneAndOverflow(size_t main, size_t overflow);
};

// Branching / structural operators
struct Hadamard_Int : SelfAdjointOperator, Controllable {
    size_t id, n_digits;
    Hadamard_Int(std::string_view reg, size_t n) : Hadamard_Int(System::get(reg), n) {}
    Hadamard_Int(size_t reg, size_t n) : id(reg), n_digits(n) {}
    Controllable* controls() override { return this; }
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};
struct Hadamard_Int_Full : SelfAdjointOperator, Controllable {
    size_t id;
    explicit Hadamard_Int_Full(std::string_view reg) : Hadamard_Int_Full(System::get(reg)) {}
    explicit Hadamard_Int_Full(size_t reg) : id(reg) {}
    Controllable* controls() override { return this; }
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};
struct Hadamard_Bool : SelfAdjointOperator, Controllable {
    size_t out_id;
    explicit Hadamard_Bool(std::string_view reg) : Hadamard_Bool(System::get(reg)) {}
    explicit Hadamard_Bool(size_t reg);
    Controllable* controls() override { return this; }
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};
struct CondRot_Rational_Bool : BaseOperator, Controllable {
    size_t register_in, register_out;
    CondRot_Rational_Bool(std::string_view in, std::string_view out) : CondRot_Rational_Bool(System::get(in), System::get(out)) {}
    CondRot_Rational_Bool(size_t in, size_t out);
    Controllable* controls() override { return this; }
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override { apply(state, 1.0); }
    void dag(SparseState& state) const override { apply(state, -1.0); }
    void apply(SparseState& state, double sign) const;
};
struct CondRot_Fixed_Bool : CondRot_Rational_Bool {
    using CondRot_Rational_Bool::CondRot_Rational_Bool;
};
// CondRot_General_Bool_Fast with the 2x2 given per value of the input register
// (tabulated on the host for inputs of up to 24 bits).
struct CondRot_General_Bool : BaseOperator {
    size_t in_id, out_id;
    std::function<u22_t(uint64_t)> func;
    CondRot_General_Bool(size_t in, size_t out, std::function<u22_t(uint64_t)> f);
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};
struct Rot_Bool : BaseOperator, Controllable {
    size_t id, digit;
    u22_t mat;
    Rot_Bool(std::string_view reg, size_t digit, u22_t m) : Rot_Bool(System::get(reg), digit, m) {}
    Rot_Bool(size_t reg, size_t digit, u22_t m);
    Controllable* controls() override { return this; }
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
    void dag(SparseState& state) const override;
};
// Named 2x2 gates (basic_gates.cpp:422-520): fixed matrices on Rot_Bool.
struct RX_Bool : Rot_Bool { RX_Bool(std::string_view r, size_t digit, double theta); RX_Bool(size_t r, size_t digit, double theta); };
struct RY_Bool : Rot_Bool { RY_Bool(std::string_view r, size_t digit, double theta); RY_Bool(size_t r, size_t digit, double theta); };
struct SX_Bool : Rot_Bool { explicit SX_Bool(std::string_view r, size_t digit = 0); explicit SX_Bool(size_t r, size_t digit = 0); };
struct U2_Bool : Rot_Bool { U2_Bool(std::string_view r, size_t digit, double phi, double lambda); U2_Bool(size_t r, size_t digit, double phi, double lambda); };
struct U3_Bool : Rot_Bool { U3_Bool(std::string_view r, size_t digit, double theta, double phi, double lambda); U3_Bool(size_t r, size_t digit, double theta, double phi, double lambda); };
u22_t make_func(uint64_t value, size_t n_digit);
u22_t make_func_inv(uint64_t value, size_t n_digit);

// Dense unitary on a whole register (SparQ/include/rot.h:16-120), row-major 2^w x 2^w.
struct Rot_GeneralUnitary : BaseOperator, Controllable {
    size_t id;
    std::vector<complex_t> mat;
    Rot_GeneralUnitary(std::string_view reg, std::vector<complex_t> m) : Rot_GeneralUnitary(System::get(reg), std::move(m)) {}
    Rot_GeneralUnitary(size_t reg, std::vector<complex_t> m);
    Controllable* controls() override { return this; }
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override { apply(state, false); }
    void dag(SparseState& state) const override { apply(state, true); }
    void apply(SparseState& state, bool dagger) const;
};
// Unitary whose first column is the normalised target (Householder reflection,
// SparQ/src/rot.cpp:469-503).
std::vector<complex_t> stateprep_unitary(const std::vector<complex_t>& target);
struct Rot_GeneralStatePrep : Rot_GeneralUnitary {
    Rot_GeneralStatePrep(std::string_view reg, const std::vector<complex_t>& vec) : Rot_GeneralStatePrep(System::get(reg), vec) {}
    Rot_GeneralStatePrep(size_t reg, const std::vector<complex_t>& vec);
};

struct ClearZero : SelfAdjointOperator {
    double eps;
    ClearZero() : eps(clearzero_default_eps()) {}
    explicit ClearZero(double e) : eps(e) {}
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};
struct Normalize : SelfAdjointOperator {
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};
struct SortUnconditional : SelfAdjointOperator {
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};
// Grouping sorts (SparQ/src/sort_state.cpp:10-171), stable on the device; the
// reference's std::sort leaves the order of entries equal under the ordering
// unspecified.
struct SortExceptKey : SelfAdjointOperator {       // others in id order, then key
    size_t id;
    explicit SortExceptKey(std::string_view key) : SortExceptKey(System::get(key)) {}
    explicit SortExceptKey(size_t key) : id(key) {}
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};
struct SortByKey : SelfAdjointOperator {           // key first (the rest in canonical order)
    size_t register_key;
    explicit SortByKey(std::string_view key) : SortByKey(System::get(key)) {}
    explicit SortByKey(size_t key) : register_key(key) {}
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};
struct SortExceptBit : SelfAdjointOperator {       // others, then key with bit `digit` masked, then the bit
    size_t id, digit;
    SortExceptBit(std::string_view key, size_t d) : SortExceptBit(System::get(key), d) {}
    SortExceptBit(size_t key, size_t d);
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};
struct SortExceptKeyHadamard : SelfAdjointOperator {   // key's listed qubits moved last
    size_t id;
    std::vector<size_t> qubit_ids;                 // ascending, distinct
    SortExceptKeyHadamard(std::string_view key, const std::vector<size_t>& q) : SortExceptKeyHadamard(System::get(key), q) {}
    SortExceptKeyHadamard(size_t key, const std::vector<size_t>& q);
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};
struct SortByAmplitude : SelfAdjointOperator {     // |a|^2 descending
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};
struct SortByKey2 : SelfAdjointOperator {          // key1, key2, then the others
    size_t id1, id2;
    SortByKey2(std::string_view k1, std::string_view k2) : SortByKey2(System::get(k1), System::get(k2)) {}
    SortByKey2(size_t k1, size_t k2) : id1(k1), id2(k2) {}
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};

// Hadamard on the listed bits of a register (SparQ/include/hadamard.h:255-330,
// SparQ/src/hadamard.cpp:470-712): H on every listed qubit, i.e. the
// Walsh-Hadamard transform over those bits of each group of entries equal
// elsewhere (the reference's sort-based path, hadamard.cpp:573-597).  Its
// hash-based path pairs entries as for a single qubit; the two agree for one
// position.  ClearZero once at the end, as the reference does.
struct Hadamard_Partial : SelfAdjointOperator, Controllable {
    size_t id;
    std::vector<size_t> qubit_positions;           // ascending, distinct
    Hadamard_Partial(std::string_view reg, const std::vector<size_t>& pos) : Hadamard_Partial(System::get(reg), pos) {}
    Hadamard_Partial(size_t reg, const std::vector<size_t>& pos);
    Controllable* controls() override { return this; }
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};

struct CheckDuplicateKey : SelfAdjointOperator {
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};
struct TestRemovable : SelfAdjointOperator {
    size_t register_id;
    explicit TestRemovable(std::string_view n) : register_id(System::get(n)) {}
    explicit TestRemovable(size_t id) : register_id(id) {}
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};

struct AddRegister {
    std::string register_name;
    StateStorageType type;
    size_t size;
    AddRegister(std::string_view n, StateStorageType t, size_t s);
    size_t operator()(SparseState& state) const;
};
struct AddRegisterWithHadamard {
    std::string register_name;
    StateStorageType type;
    size_t size;
    AddRegisterWithHadamard(std::string_view n, StateStorageType t, size_t s) : register_name(n), type(t), size(s) {}
    size_t operator()(SparseState& state) const;
};
struct RemoveRegister {
    size_t register_id;
    explicit RemoveRegister(std::string_view n) : register_id(System::get(n)) {}
    explicit RemoveRegister(size_t id) : register_id(id) {}
    void operator()(SparseState& state) const;
};
// MoveBackRegister (SparQ/src/system_operations.cpp:223-249): the register
// takes the last id of the table and the register holding that id takes its
// place; the values move with them in `state` only (like the reference, other
// states are not updated: it must not run while other states are in use).
struct MoveBackRegister {
    size_t register_id;
    explicit MoveBackRegister(std::string_view reg) : register_id(System::get(reg)) {}
    explicit MoveBackRegister(size_t reg) : register_id(reg) {}
    void operator()(SparseState& state) const;
};
// Push / Pop (SparQ/src/system_operations.cpp:338-369): Push adds a register
// `garbage` of reg's type and size, swaps it with reg (reg becomes 0) and stacks
// its id; Pop swaps reg with the last stacked register and removes that one
// (it must then be zero).  Push's dag is Pop(reg); Pop has no dag.
struct Push : BaseOperator {
    size_t reg_id;
    std::string garbage_name;
    Push(std::string_view reg, std::string_view garbage) : Push(System::get(reg), garbage) {}
    Push(size_t reg, std::string_view garbage) : reg_id(reg), garbage_name(garbage) {}
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
    void dag(SparseState& state) const override;
};
struct Pop : BaseOperator {
    size_t reg_id;
    explicit Pop(std::string_view reg) : Pop(System::get(reg)) {}
    explicit Pop(size_t reg) : reg_id(reg) {}
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};

struct SplitRegister {
    std::string first_name, second_name;
    size_t second_size;
    SplitRegister(std::string_view f, std::string_view s, size_t sz) : first_name(f), second_name(s), second_size(sz) {}
    SplitRegister(size_t f, std::string_view s, size_t sz) : first_name(System::name_of(f)), second_name(s), second_size(sz) {}
    size_t operator()(SparseState& state) const;
};
struct CombineRegister {
    std::string first_name, second_name;
    CombineRegister(std::string_view f, std::string_view s) : first_name(f), second_name(s) {}
    CombineRegister(size_t f, size_t s) : first_name(System::name_of(f)), second_name(System::name_of(s)) {}
    size_t operator()(SparseState& state) const;
};

// ------------------------------------------------------------ debug checks
// CheckNormalization / CheckNan (SparQ/src/debugger.cpp:13-70): one device
// reduction; they throw on a bad state.  Like the reference's release build
// (QRAM_Release, PySparQ/CMakeLists.txt:8) they are no-ops unless
// debug_checks() is set.
bool& debug_checks();
struct CheckNormalization : SelfAdjointOperator {
    double threshold = 1e-5;
    CheckNormalization() = default;
    explicit CheckNormalization(double t) : threshold(t) {}
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};
struct CheckNan : SelfAdjointOperator {
    SPQ_BOTH_VIEWS
    void operator()(SparseState& state) const override;
};

// ------------------------------------------------------------ post-selection
// PartialTraceSelect / PartialTraceSelectRange (SparQ/include/partial_trace.h:83-145,
// SparQ/src/partial_trace.cpp:148-225): keep the entries whose registers hold
// the selected values (or lie in [first, second]), renormalise, and return the
// factor 1/sqrt(kept probability) (+inf if nothing is kept, 0 for an empty state
// or an empty register list).
struct PartialTraceSelect {
    std::vector<size_t> partial_trace_registers;
    std::vector<uint64_t> select_values;
    explicit PartialTraceSelect(const std::map<std::string, uint64_t>& sel);
    explicit PartialTraceSelect(const std::map<size_t, uint64_t>& sel);
    PartialTraceSelect(const std::vector<size_t>& regs, const std::vector<uint64_t>& values);
    PartialTraceSelect(const std::vector<std::string>& regs, const std::vector<uint64_t>& values);
    double operator()(SparseState& state) const;
    double operator()(std::vector<System>& state) const;
};
struct PartialTraceSelectRange {
    size_t partial_trace_register;
    std::pair<size_t, size_t> select_range;
    PartialTraceSelectRange(std::string_view reg, std::pair<size_t, size_t> range)
        : partial_trace_register(System::get(reg)), select_range(range) {}
    PartialTraceSelectRange(size_t reg, std::pair<size_t, size_t> range) : partial_trace_register(reg), select_range(range) {}
    double operator()(SparseState& state) const;
    double operator()(std::vector<System>& state) const;
};

// ------------------------------------------------------------ measurement
// Seedable global random engine (the role of random_engine.h in the reference,
// PySparQ/core.cpp:404-424): splitmix64, uniform01 = top 53 bits / 2^53.  The
// reference's engine lives in the absent submodule, so streams differ from it.
namespace random_engine {
void set_seed(uint64_t seed);
uint64_t get_seed();
double uniform01();
}  // namespace random_engine

// MeasureZ / Reset / Probability (SparQ/include/measurement.h:60-140,
// SparQ/src/measurement.cpp).  MeasureZ sorts the state into canonical order
// and walks the entries in that order with r = uniform01(), so the outcome for
// a given seed does not depend on the engine's internal entry order (the
// reference walks its own entry order); it then keeps the entries with the
// sampled register values and renormalises.  Returns (outcome, probability).
struct MeasureZ {
    std::vector<size_t> registers;
    explicit MeasureZ(const std::vector<std::string>& names);
    explicit MeasureZ(const std::vector<size_t>& ids);
    explicit MeasureZ(std::string_view name);
    explicit MeasureZ(size_t id);
    std::pair<std::vector<uint64_t>, double> operator()(SparseState& state) const;
};
struct Reset {
    std::vector<size_t> registers;
    std::vector<uint64_t> target_values;
    explicit Reset(const std::vector<std::string>& names);
    Reset(const std::vector<std::string>& names, const std::vector<uint64_t>& targets);
    explicit Reset(const std::vector<size_t>& ids);
    Reset(const std::vector<size_t>& ids, const std::vector<uint64_t>& targets);
    explicit Reset(std::string_view name, uint64_t target = 0);
    explicit Reset(size_t id, uint64_t target = 0);
    std::vector<uint64_t> operator()(SparseState& state) const;
};
struct Probability {
    std::vector<size_t> registers;
    std::vector<uint64_t> values;
    explicit Probability(const std::map<std::string, uint64_t>& assignments);
    explicit Probability(const std::map<size_t, uint64_t>& assignments);
    Probability(const std::vector<std::string>& names, const std::vector<uint64_t>& values);
    Probability(const std::vector<size_t>& ids, const std::vector<uint64_t>& values);
    Probability(std::string_view name, uint64_t value);
    Probability(size_t id, uint64_t value);
    double operator()(SparseState& state) const;
    static std::map<uint64_t, double> distribution(SparseState& state, size_t id);
    static std::map<uint64_t, double> distribution(SparseState& state, std::string_view name);
};

// ------------------------------------------------------------ CKS quantum walk primitives
// (SparQ_Algorithm/include/hamiltonian_simulation.h, SparQ_Algorithm/src/hamiltonian_simulation.cpp)

// Compact QRAM storage of a sparse matrix with nnz_col slots per row: slot s of
// row i holds its quantised element at memory[i*nnz_col + s] and its column at
// memory[get_sparsity_offset() + i*nnz_col + s]; columns ascend within a row so
// the binary search of SparseMatrixOracle2 finds them.  get_data() is that
// memory, zero-padded to a power of two (the QRAM address space).  The layout is
// the one TridiagonalSpec::combined_mem reproduces
// (SparQ_Algorithm/include/hamiltonian_simulation_tridiagonal.h:169-190); the
// constructor from real values quantises x to floor(x*2^ds) (positive_only) or
// the ds-bit two's complement of trunc(x*2^(ds-1)) (signed), clipped like
// SparseMatrix.from_dense (PySparQ/pysparq/algorithms/cks_solver.py:288-349).
struct SparseMatrix {
    std::vector<uint64_t> elements;
    std::vector<size_t> sparsity;
    size_t data_size = 0, nnz_col = 0, n_row = 0;
    bool positive_only = true;
    SparseMatrix() = default;
    SparseMatrix(const std::vector<double>& values, const std::vector<size_t>& sparsity, size_t data_size, size_t nnz_col,
                 size_t n_row, bool positive_only);
    SparseMatrix(const std::vector<uint64_t>& encoded, const std::vector<size_t>& sparsity, size_t data_size, size_t nnz_col,
                 size_t n_row, bool positive_only = true);
    std::vector<uint64_t> get_data() const;
    size_t get_sparsity_offset() const { return elements.size(); }
};

// row_offset ^= offset + row_sz*row (hamiltonian_simulation.cpp:222-237)
struct GetRowAddr : PointwiseOperator {
    GetRowAddr(std::string_view offset, std::string_view row, size_t row_sz, std::string_view row_offset)
        : GetRowAddr(System::get(offset), System::get(row), row_sz, System::get(row_offset)) {}
    GetRowAddr(size_t offset, size_t row, size_t row_sz, size_t row_offset);
};
// data_offset ^= offset + row_sz*row + col_sparse (hamiltonian_simulation.cpp:239-255)
struct GetDataAddr : PointwiseOperator {
    GetDataAddr(std::string_view offset, std::string_view row, std::string_view col_sparse, size_t row_sz,
                std::string_view data_offset)
        : GetDataAddr(System::get(offset), System::get(row), System::get(col_sparse), row_sz, System::get(data_offset)) {}
    GetDataAddr(size_t offset, size_t row, size_t col_sparse, size_t row_sz, size_t data_offset);
};
// result ^= address of target in memory[offset, offset+total_length), 0 when absent
// (hamiltonian_simulation.cpp:182-220; max_step = floor(log2(total_length)) + 1)
struct QuantumBinarySearch_Fast : PointwiseOperator {
    size_t total_length, max_step;
    QuantumBinarySearch_Fast(const QRAMCircuit_qutrit* q, std::string_view offset, size_t total_length, std::string_view target,
                             std::string_view result)
        : QuantumBinarySearch_Fast(q, System::get(offset), total_length, System::get(target), System::get(result)) {}
    QuantumBinarySearch_Fast(const QRAMCircuit_qutrit* q, size_t offset, size_t total_length, size_t target, size_t result);
};
// out ^= rational(acos(sqrt(|a|/Amax)) / 2pi) of the quantised element in `data`
// (hamiltonian_simulation.cpp:257-290); row and col are not read
struct GetQWRotateAngle_Int_Int_Int : PointwiseOperator {
    GetQWRotateAngle_Int_Int_Int(std::string_view data, std::string_view row, std::string_view col, std::string_view out,
                                 const SparseMatrix* mat)
        : GetQWRotateAngle_Int_Int_Int(System::get(data), System::get(row), System::get(col), System::get(out), mat) {}
    GetQWRotateAngle_Int_Int_Int(size_t data, size_t row, size_t col, size_t out, const SparseMatrix* mat);
};

// The tridiagonal walk's direct oracles (hamiltonian_simulation_tridiagonal.h:104-396,
// SparQ_Algorithm/src/hamiltonian_tridiagonal.cpp:16-105).
struct TridiagonalSpec {
    size_t n_row = 0, data_size = 0;
    uint64_t alpha_enc = 0, beta_enc = 0;
    size_t sparse_offset = 0;
    uint64_t addr_mask = 0;
    size_t k_width = 0;
    static constexpr size_t slot_count = 4;
    size_t window_base(size_t j) const;
    size_t col_of(size_t j, size_t s) const { return window_base(j) + s; }
    uint64_t value_of(size_t j, size_t s) const;
    uint64_t combined_mem(uint64_t idx) const;
};
std::pair<SparseMatrix, TridiagonalSpec> make_tridiagonal_matrix_spec(double alpha, double beta, size_t n, size_t ds);
// out ^= M(offset + 4*row + slot)
struct TridiagOracle1 : PointwiseOperator {
    TridiagOracle1(std::string_view offset, std::string_view row, std::string_view slot, std::string_view out,
                   const TridiagonalSpec& spec)
        : TridiagOracle1(System::get(offset), System::get(row), System::get(slot), System::get(out), spec) {}
    TridiagOracle1(size_t offset, size_t row, size_t slot, size_t out, const TridiagonalSpec& spec);
};
// forward: column -> slot; dag: slot -> column (garbage branches tagged in sr)
struct TridiagOracle2 : PointwiseOperator {
    TridiagOracle2(std::string_view offset, std::string_view row, std::string_view k, std::string_view sr,
                   const TridiagonalSpec& spec)
        : TridiagOracle2(System::get(offset), System::get(row), System::get(k), System::get(sr), spec) {}
    TridiagOracle2(size_t offset, size_t row, size_t k, size_t sr, const TridiagonalSpec& spec);
};
// amplitude negated where bit 0 of the register is 1
struct TridiagParityPhase : PointwiseOperator {
    explicit TridiagParityPhase(std::string_view reg) : TridiagParityPhase(System::get(reg)) {}
    explicit TridiagParityPhase(size_t reg);
};

// split_systems / combine_systems (SparQ/src/system_operations.cpp:53-106):
// the entries failing the conditions move to the returned state; combine
// appends `from`'s entries to `to`.
SparseState split_systems(SparseState& state, const std::vector<size_t>& nonzeros, const std::vector<size_t>& all_ones,
                          const std::vector<std::pair<size_t, size_t>>& by_bit,
                          const std::vector<std::pair<size_t, uint64_t>>& by_value);
void combine_systems(SparseState& to, SparseState& from);
// merge_system / remove_system (SparQ/src/basic_components.cpp:517-526): a takes
// the amplitude sum of two equal entries and b's becomes 0; the merge's prune
// test is |a|^2 < epsilon.
void merge_system(System& a, System& b);
bool remove_system(const System& s);

// add_systems (system_operations.h:157-181)
void add_systems(SparseState& current, SparseState& new_state, double coef, double eps = epsilon);
// reset_systems (system_operations.cpp:108-117): back to the single entry |0...0>, amplitude 1.
void reset_systems(SparseState& state);

}  // namespace sparqsim
