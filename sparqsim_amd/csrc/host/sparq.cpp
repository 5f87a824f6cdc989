// Host layer implementation: register table, layout, SparseState and the
// operator classes.  Semantics cite the reference files they follow.
#include "sparq.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <sstream>

namespace sparqsim {

void check(int status) {
    if (status == SPARQ_OK) return;
    std::string msg = sparq_last_error();
    if (status == SPARQ_INVALID) throw std::invalid_argument(msg);
    if (status == SPARQ_OOM) throw std::runtime_error("out of device memory: " + msg);
    throw std::runtime_error(msg);
}

double& clearzero_default_eps() {   // system_operations.h:577
    static double eps = epsilon;
    return eps;
}

// ------------------------------------------------------------ register table
std::vector<StateInfoType> System::name_register_map;
std::unordered_map<std::string, size_t> System::name_to_index;
std::vector<size_t> System::temporal_registers;
std::vector<size_t> System::reusable_registers;
size_t System::max_qubit_count = 0, System::max_register_count = 0, System::max_system_size = 0;
uint64_t System::layout_version = 0;

static const StateInfoType& info(size_t id) {
    if (id >= System::name_register_map.size()) throw std::runtime_error("Register not found.");
    return System::name_register_map[id];
}

void System::clear() {   // basic_components.cpp:150-161
    name_register_map.clear();
    name_to_index.clear();
    temporal_registers.clear();
    reusable_registers.clear();
    max_qubit_count = max_register_count = max_system_size = 0;
    ++layout_version;
}
size_t System::get_qubit_count() {
    size_t c = 0;
    for (auto& m : name_register_map) if (std::get<3>(m)) c += std::get<2>(m);
    return c;
}
size_t System::get_activated_register_size() {
    size_t c = 0;
    for (auto& m : name_register_map) if (std::get<3>(m)) ++c;
    return c;
}
size_t System::get_last_activated_register() {   // basic_components.cpp:185-195
    if (name_register_map.empty()) throw std::runtime_error("No register is activated.");
    for (size_t i = name_register_map.size(); i-- > 0;) if (status_of(i)) return i;
    return SIZE_MAX;
}
void System::update_max_size(size_t n) { if (n > max_system_size) max_system_size = n; }
size_t System::get(std::string_view name) {   // basic_components.cpp:206-228
    auto it = name_to_index.find(std::string(name));
    if (it != name_to_index.end() && it->second < name_register_map.size() && status_of(it->second)) return it->second;
    for (size_t i = name_register_map.size(); i-- > 0;) {
        if (!status_of(i)) continue;
        if (std::get<0>(name_register_map[i]) == name) {
            name_to_index[std::string(name)] = i;
            return i;
        }
    }
    return SIZE_MAX;
}
StateInfoType System::get_register_info(std::string_view name) {
    for (auto it = name_register_map.rbegin(); it != name_register_map.rend(); ++it)
        if (std::get<0>(*it) == name) return *it;
    throw std::runtime_error("Register not found.");
}
const std::string& System::name_of(size_t id) { return std::get<0>(info(id)); }
size_t System::size_of(std::string_view name) { return std::get<2>(get_register_info(name)); }
size_t System::size_of(size_t id) { return std::get<2>(info(id)); }
StateStorageType System::type_of(std::string_view name) { return std::get<1>(get_register_info(name)); }
StateStorageType System::type_of(size_t id) { return std::get<1>(info(id)); }
bool System::status_of(std::string_view name) { return status_of(get(name)); }
bool System::status_of(size_t id) { return std::get<3>(info(id)); }
size_t System::add_register(std::string_view name, StateStorageType type, size_t size) {   // :344-381
    size_t id;
    if (!reusable_registers.empty()) {
        id = reusable_registers.back();
        reusable_registers.pop_back();
        name_register_map[id] = StateInfoType{std::string(name), type, size, true};
    } else {
        name_register_map.emplace_back(std::string(name), type, size, true);
        id = name_register_map.size() - 1;
    }
    name_to_index[std::string(name)] = id;
    max_qubit_count = std::max(max_qubit_count, get_qubit_count());
    max_register_count = std::max(max_register_count, get_activated_register_size());
    ++layout_version;
    return id;
}
void System::remove_register(size_t id) {   // :405-415
    if (id >= name_register_map.size()) throw std::runtime_error("Register not found.");
    if (!std::get<3>(name_register_map[id])) throw std::runtime_error("Register is not activated.");
    name_to_index.erase(std::get<0>(name_register_map[id]));
    std::get<3>(name_register_map[id]) = false;
    reusable_registers.push_back(id);
    ++layout_version;
}
void System::remove_register(std::string_view name) { remove_register(get(name)); }
void System::set_size(size_t id, size_t size) {
    if (id >= name_register_map.size()) throw std::runtime_error("Register not found.");
    std::get<2>(name_register_map[id]) = size;
    ++layout_version;
}
void System::set_type(size_t id, StateStorageType t) {
    if (id >= name_register_map.size()) throw std::runtime_error("Register not found.");
    std::get<1>(name_register_map[id]) = t;
    ++layout_version;   // the signed flag of the field follows the type
}

// ------------------------------------------------------------ layout
Layout Layout::from_registry() {
    Layout L;
    L.field.assign(System::name_register_map.size(), sparq_field_t{0, 0, 0, 0});
    int word = 0, room = 64;
    for (size_t id = 0; id < System::name_register_map.size(); ++id) {
        const auto& m = System::name_register_map[id];
        if (!std::get<3>(m)) continue;
        size_t w = std::get<2>(m);
        if (w == 0) continue;
        if (w > 64) throw std::invalid_argument("register wider than 64 bits");
        if ((int)w > room) {
            ++word;
            room = 64;
        }
        room -= (int)w;
        L.field[id] = sparq_field_t{(uint8_t)word, (uint8_t)room, (uint8_t)w,
                                    (uint8_t)(std::get<1>(m) == SignedInteger ? 1 : 0)};
    }
    L.words = word + 1;
    if (L.words > SPARQ_MAX_WORDS) throw std::invalid_argument("active registers exceed 512 key bits");
    return L;
}
bool Layout::operator==(const Layout& o) const {
    if (words != o.words || field.size() != o.field.size()) return false;
    for (size_t i = 0; i < field.size(); ++i)
        if (std::memcmp(&field[i], &o.field[i], sizeof(sparq_field_t)) != 0) return false;
    return true;
}
sparq_field_t Layout::at(size_t id) const {
    if (id >= field.size() || field[id].width == 0) throw std::runtime_error("Register not found.");
    return field[id];
}

// ------------------------------------------------------------ state
SparseState::SparseState() {
    layout_ = Layout::from_registry();
    check(sparq_state_create(layout_.words, 1, &h_));
    check(sparq_state_set_basis(h_, nullptr));
}
SparseState::SparseState(const SparseState& other) {
    SparseState& o = const_cast<SparseState&>(other);
    o.flush();
    layout_ = o.layout_;
    check(sparq_state_clone(o.h_, &h_));
}
SparseState& SparseState::operator=(const SparseState& other) {
    if (this == &other) return *this;
    SparseState tmp(other);
    std::swap(h_, tmp.h_);
    std::swap(layout_, tmp.layout_);
    pend_ops_.clear();
    pend_conds_.clear();
    zeroed_.clear();
    layout_version_ = ~0ull;
    return *this;
}
SparseState::SparseState(SparseState&& other) noexcept
    : h_(other.h_), layout_(std::move(other.layout_)), pend_ops_(std::move(other.pend_ops_)), pend_conds_(std::move(other.pend_conds_)) {
    other.h_ = nullptr;
}
SparseState& SparseState::operator=(SparseState&& other) noexcept {
    std::swap(h_, other.h_);
    std::swap(layout_, other.layout_);
    std::swap(pend_ops_, other.pend_ops_);
    std::swap(pend_conds_, other.pend_conds_);
    std::swap(zeroed_, other.zeroed_);
    std::swap(layout_version_, other.layout_version_);
    return *this;
}
SparseState::~SparseState() {
    if (h_) sparq_state_destroy(h_);
}

size_t SparseState::size() {
    flush();
    uint64_t n = 0;
    check(sparq_state_size(h_, &n));
    return n;
}

void SparseState::flush() {
    if (pend_ops_.empty()) return;
    int rc = sparq_pointwise(h_, pend_ops_.data(), (uint32_t)pend_ops_.size(), pend_conds_.data(), (uint32_t)pend_conds_.size());
    pend_ops_.clear();
    pend_conds_.clear();
    check(rc);
}

const Layout& SparseState::sync() {
    if (layout_version_ == System::layout_version) return layout_;
    Layout T = Layout::from_registry();
    layout_version_ = System::layout_version;
    if (T == layout_) {
        zeroed_.clear();
        return layout_;
    }
    flush();
    std::vector<sparq_field_t> from, to, clear;
    bool in_place = true;   // every surviving field keeps its word and shift
    for (size_t id = 0; id < layout_.field.size(); ++id) {
        sparq_field_t a = layout_.field[id];
        if (a.width == 0) continue;
        if (id < T.field.size() && T.field[id].width == a.width) {
            sparq_field_t b = T.field[id];
            from.push_back(a);
            to.push_back(b);
            if (a.word != b.word || a.shift != b.shift) in_place = false;
        } else if (a.word < T.words && std::find(zeroed_.begin(), zeroed_.end(), id) == zeroed_.end()) {
            clear.push_back(a);   // a dropped field in a surviving word must read as zero
        }
    }
    zeroed_.clear();
    if (in_place) {
        // Fields stay put: clear dropped non-zero fields in place (one fused
        // pointwise pass over their words) and change the word count without
        // moving keys.  New fields land on zero bits (gaps, cleared fields or
        // added words).
        for (size_t i = 0; i < clear.size(); ++i) {
            sparq_pw_op_t op{};
            op.opcode = SPARQ_PW_SET_FIELD;
            op.f[0] = clear[i];
            push_pointwise(op, {});
        }
        flush();
        if (T.words != layout_.words) check(sparq_state_resize_words(h_, T.words));
    } else {
        check(sparq_state_relayout(h_, T.words, from.data(), to.data(), (uint32_t)from.size()));
    }
    layout_ = T;
    return layout_;
}

void SparseState::relayout_with(const Layout& target, const std::vector<std::pair<sparq_field_t, sparq_field_t>>& moves) {
    flush();
    std::vector<sparq_field_t> from, to;
    for (auto& m : moves) {
        from.push_back(m.first);
        to.push_back(m.second);
    }
    check(sparq_state_relayout(h_, target.words, from.data(), to.data(), (uint32_t)from.size()));
    layout_ = target;
    layout_version_ = System::layout_version;
}

static bool same_conds(const sparq_cond_t* a, const sparq_cond_t* b, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (a[i].kind != b[i].kind || std::memcmp(&a[i].field, &b[i].field, sizeof(sparq_field_t)) != 0 || a[i].value != b[i].value)
            return false;
    return true;
}

void SparseState::push_pointwise(const sparq_pw_op_t& op_in, const std::vector<sparq_cond_t>& conds) {
    if (conds.size() > SPARQ_MAX_PROGRAM_CONDS) throw std::invalid_argument("too many conditions on one operator");
    // Consecutive phase ops on the same register with equal phase and conditions merge.
    if (op_in.opcode == SPARQ_PW_PHASE_BITS && !pend_ops_.empty()) {
        sparq_pw_op_t& last = pend_ops_.back();
        if (last.opcode == SPARQ_PW_PHASE_BITS && std::memcmp(&last.f[0], &op_in.f[0], sizeof(sparq_field_t)) == 0 &&
            last.d[0] == op_in.d[0] && last.d[1] == op_in.d[1] && last.ncond == conds.size() &&
            (conds.empty() || same_conds(&pend_conds_[last.cond_first], conds.data(), conds.size())) && (last.c[0] & op_in.c[0]) == 0) {
            last.c[0] |= op_in.c[0];
            return;
        }
    }
    if (pend_ops_.size() >= SPARQ_MAX_PROGRAM_OPS || pend_conds_.size() + conds.size() > SPARQ_MAX_PROGRAM_CONDS) flush();
    sparq_pw_op_t op = op_in;
    op.cond_first = (uint16_t)pend_conds_.size();
    op.ncond = (uint16_t)conds.size();
    pend_conds_.insert(pend_conds_.end(), conds.begin(), conds.end());
    pend_ops_.push_back(op);
}

static uint64_t field_value(const std::vector<uint64_t>& key, sparq_field_t f) {
    return (key[f.word] >> f.shift) & width_mask(f.width);
}

void SparseState::export_entries(std::vector<size_t>& ids, std::vector<uint64_t>& values, std::vector<complex_t>& amps) {
    sync();
    flush();
    uint64_t n = 0;
    check(sparq_state_size(h_, &n));
    std::vector<uint64_t> keys((size_t)n * layout_.words);
    std::vector<double> a((size_t)n * 2);
    check(sparq_state_download(h_, keys.data(), a.data(), n, &n));
    ids.clear();
    for (size_t id = 0; id < layout_.field.size(); ++id) if (layout_.field[id].width) ids.push_back(id);
    values.assign((size_t)n * ids.size(), 0);
    amps.resize(n);
    std::vector<uint64_t> k(layout_.words);
    for (uint64_t i = 0; i < n; ++i) {
        for (uint32_t w = 0; w < layout_.words; ++w) k[w] = keys[i * layout_.words + w];
        for (size_t j = 0; j < ids.size(); ++j) values[i * ids.size() + j] = field_value(k, layout_.field[ids[j]]);
        amps[i] = complex_t(a[2 * i], a[2 * i + 1]);
    }
}

void SparseState::import_entries(const std::vector<uint64_t>& values, const std::vector<complex_t>& amps) {
    sync();
    flush();
    std::vector<size_t> ids;
    for (size_t id = 0; id < layout_.field.size(); ++id) if (layout_.field[id].width) ids.push_back(id);
    size_t n = amps.size();
    if (values.size() != n * ids.size()) throw std::invalid_argument("value array does not match the active registers");
    std::vector<uint64_t> keys(n * layout_.words, 0);
    std::vector<double> a(2 * n);
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = 0; j < ids.size(); ++j) {
            sparq_field_t f = layout_.field[ids[j]];
            uint64_t v = values[i * ids.size() + j];
            if (v & ~width_mask(f.width)) throw std::invalid_argument("register value exceeds its width");
            keys[i * layout_.words + f.word] |= v << f.shift;
        }
        a[2 * i] = amps[i].real();
        a[2 * i + 1] = amps[i].imag();
    }
    check(sparq_state_upload(h_, keys.data(), a.data(), n));
}

void SparseState::to_host(std::vector<System>& entries) {
    std::vector<size_t> ids;
    std::vector<uint64_t> vals;
    std::vector<complex_t> amps;
    export_entries(ids, vals, amps);
    entries.assign(amps.size(), System());
    for (size_t i = 0; i < amps.size(); ++i) {
        entries[i].amplitude = amps[i];
        for (size_t j = 0; j < ids.size(); ++j) entries[i].registers[ids[j]].value = vals[i * ids.size() + j];
    }
}

void SparseState::from_host(const std::vector<System>& entries) {
    sync();
    std::vector<size_t> ids;
    for (size_t id = 0; id < layout_.field.size(); ++id) if (layout_.field[id].width) ids.push_back(id);
    std::vector<uint64_t> vals(entries.size() * ids.size());
    std::vector<complex_t> amps(entries.size());
    for (size_t i = 0; i < entries.size(); ++i) {
        amps[i] = entries[i].amplitude;
        for (size_t j = 0; j < ids.size(); ++j)
            vals[i * ids.size() + j] = ids[j] < entries[i].registers.size() ? entries[i].registers[ids[j]].value : 0;
    }
    import_entries(vals, amps);
}

double SparseState::post_select_dense(const std::vector<size_t>& zero_ids, const std::vector<size_t>& index_ids,
                                      std::vector<complex_t>& out, uint64_t length) {
    const Layout& L = sync();
    flush();
    std::vector<sparq_field_t> z, ix;
    size_t bits = 0;
    for (size_t id : zero_ids) z.push_back(L.at(id));
    for (size_t id : index_ids) {
        ix.push_back(L.at(id));
        bits += ix.back().width;
    }
    if (length == 0 && bits > 40) throw std::invalid_argument("post-selected index wider than 40 bits");
    const uint64_t len = length ? length : uint64_t(1) << bits;
    out.assign(len, complex_t(0, 0));
    double w = 0;
    check(sparq_post_select_dense(h_, z.data(), (uint32_t)z.size(), ix.data(), (uint32_t)ix.size(),
                                  reinterpret_cast<double*>(out.data()), len, &w));
    return w;
}

// ------------------------------------------------------------ host entries
bool System::operator<(const System& rhs) const {
    for (size_t id = 0; id < name_register_map.size(); ++id) {
        if (!status_of(id)) continue;
        const uint64_t a = id < registers.size() ? registers[id].value : 0;
        const uint64_t b = id < rhs.registers.size() ? rhs.registers[id].value : 0;
        if (a != b) return a < b;
    }
    return false;
}

bool System::operator==(const System& rhs) const {
    for (size_t id = 0; id < name_register_map.size(); ++id) {
        if (!status_of(id)) continue;
        const uint64_t a = id < registers.size() ? registers[id].value : 0;
        const uint64_t b = id < rhs.registers.size() ? rhs.registers[id].value : 0;
        if (a != b) return false;
    }
    return true;
}

std::string System::to_string() const {
    std::ostringstream os;
    os << amplitude.real() << (amplitude.imag() < 0 ? " - " : " + ") << std::abs(amplitude.imag()) << "i";
    for (size_t id = 0; id < name_register_map.size(); ++id)
        if (status_of(id)) os << " |" << name_of(id) << "=" << (id < registers.size() ? registers[id].value : 0) << ">";
    return os.str();
}

// ------------------------------------------------------------ the two operator forms
namespace {
// An operator that overrides neither form would bounce between the two derived
// forms; the guard turns that into an error.
thread_local const BaseOperator* t_bridging = nullptr;
struct Bridge {
    const BaseOperator* prev;
    Bridge(const BaseOperator* op, const char* what) : prev(t_bridging) {
        if (t_bridging == op) throw std::runtime_error(what);
        t_bridging = op;
    }
    ~Bridge() { t_bridging = prev; }
};
}  // namespace

namespace {
thread_local std::string t_plugin_error;
}

void record_plugin_error(const char* message) { t_plugin_error = message ? message : "unknown error"; }

std::string take_plugin_error() {
    std::string m;
    m.swap(t_plugin_error);
    return m;
}

void BaseOperator::operator()(SparseState& state) const {
    Bridge b(this, "operator() is not implemented.");
    std::vector<System> host;
    state.to_host(host);
    (*this)(host);
    state.from_host(host);
}

OperatorSequence::OperatorSequence(std::vector<std::shared_ptr<BaseOperator>> o) : ops(std::move(o)) {
    for (const auto& op : ops)
        if (!op) throw std::invalid_argument("OperatorSequence: null operator");
}
void OperatorSequence::operator()(SparseState& state) const {
    for (const auto& op : ops) (*op)(state);
}
void OperatorSequence::dag(SparseState& state) const {
    for (auto it = ops.rbegin(); it != ops.rend(); ++it) (*it)->dag(state);
}

void BaseOperator::operator()(std::vector<System>& state) const {
    Bridge b(this, "operator() is not implemented.");
    SparseState dev;
    dev.from_host(state);
    (*this)(dev);
    dev.to_host(state);
}

void BaseOperator::dag(SparseState& state) const {
    Bridge b(this, "Dagger is not implemented.");
    std::vector<System> host;
    state.to_host(host);
    dag(host);
    state.from_host(host);
}

void BaseOperator::dag(std::vector<System>& state) const {
    Bridge b(this, "Dagger is not implemented.");
    SparseState dev;
    dev.from_host(state);
    dag(dev);
    dev.to_host(state);
}

std::string SparseState::to_string(int precision) {
    std::vector<size_t> ids;
    std::vector<uint64_t> vals;
    std::vector<complex_t> amps;
    export_entries(ids, vals, amps);
    std::ostringstream os;
    os.precision(precision);
    for (size_t i = 0; i < amps.size(); ++i) {
        os << amps[i].real() << (amps[i].imag() < 0 ? " - " : " + ") << std::abs(amps[i].imag()) << "i";
        for (size_t j = 0; j < ids.size(); ++j) os << " |" << System::name_of(ids[j]) << "=" << vals[i * ids.size() + j] << ">";
        os << "\n";
    }
    return os.str();
}

// ------------------------------------------------------------ conditions
std::vector<sparq_cond_t> Controllable::conditions(const Layout& L) const {   // global_macros.h:28-92
    std::vector<sparq_cond_t> out;
    for (size_t id : condition_variable_all_ones) out.push_back(sparq_cond_t{SPARQ_COND_ALL_ONES, L.at(id), 0});
    for (auto& p : condition_variable_by_bit) {
        if (p.second >= 64) throw std::invalid_argument("condition bit out of range");
        out.push_back(sparq_cond_t{SPARQ_COND_BIT, L.at(p.first), p.second});
    }
    for (auto& p : condition_variable_by_clear_bit) {
        if (p.second >= 64) throw std::invalid_argument("condition bit out of range");
        out.push_back(sparq_cond_t{SPARQ_COND_BIT_CLEAR, L.at(p.first), p.second});
    }
    for (size_t id : condition_variable_nonzeros) out.push_back(sparq_cond_t{SPARQ_COND_NONZERO, L.at(id), 0});
    for (auto& p : condition_variable_by_value) out.push_back(sparq_cond_t{SPARQ_COND_VALUE, L.at(p.first), p.second});
    return out;
}

// ------------------------------------------------------------ pointwise operators
void PointwiseOperator::emit(SparseState& state, bool dagger) const {
    const Layout& L = state.sync();
    sparq_pw_op_t op{};
    op.opcode = dagger && dag_opcode ? dag_opcode : opcode;
    for (size_t i = 0; i < regs.size() && i < 4; ++i)
        if (regs[i] != kNoRegister) op.f[i] = L.at(regs[i]);   // absent operand: width-0 field
    const uint64_t* cc = dagger && dag_opcode ? dag_c : c;
    const double* dd = dagger && dag_opcode ? dag_d : d;
    for (int i = 0; i < 3; ++i) op.c[i] = cc[i];
    for (int i = 0; i < 8; ++i) op.d[i] = dd[i];
    op.table = table;
    op.table_word_bytes = table_word_bytes;
    std::vector<sparq_cond_t> conds = conditions(L);
    for (auto& z : extra_zero_regs) conds.push_back(sparq_cond_t{SPARQ_COND_VALUE, L.at(z.first), z.second});
    for (auto& b : extra_bit_regs) conds.push_back(sparq_cond_t{SPARQ_COND_BIT, L.at(b.first), b.second});
    state.push_pointwise(op, conds);
}
void PointwiseOperator::dag(SparseState& state) const {
    if (!has_dag) throw std::runtime_error("Dagger is not implemented.");
    emit(state, true);
}

static void set_op(PointwiseOperator& p, const char* name, uint32_t opcode, std::vector<size_t> regs, uint32_t dag_opcode = 0) {
    p.name = name;
    p.opcode = opcode;
    p.dag_opcode = dag_opcode;
    p.regs = std::move(regs);
}
static void require_ids(const std::vector<size_t>& ids) {
    for (size_t id : ids)
        if (id >= System::name_register_map.size() || !System::status_of(id)) throw std::runtime_error("Register not found.");
}

#define SPQ_DEF3(NAME, OPC)                                   \
    NAME::NAME(size_t a, size_t b, size_t c) {                \
        require_ids({a, b, c});                               \
        set_op(*this, #NAME, OPC, {a, b, c});                 \
    }
#define SPQ_DEF2(NAME, OPC)                                   \
    NAME::NAME(size_t a, size_t b) {                          \
        require_ids({a, b});                                  \
        set_op(*this, #NAME, OPC, {a, b});                    \
    }
#define SPQ_DEF4(NAME, OPC)                                   \
    NAME::NAME(size_t a, size_t b, size_t c, size_t d) {      \
        require_ids({a, b, c, d});                            \
        set_op(*this, #NAME, OPC, {a, b, c, d});              \
    }

// quantum_arithmetic.cpp line ranges are in include/sparq_hip.h next to each opcode.
SPQ_DEF3(Add_UInt_UInt, SPARQ_PW_XOR_ADD)
SPQ_DEF3(Sub_UInt_UInt, SPARQ_PW_XOR_SUB)
SPQ_DEF3(Mul_UInt_UInt, SPARQ_PW_XOR_MUL)
SPQ_DEF3(Div_UInt_UInt, SPARQ_PW_XOR_DIV)
SPQ_DEF3(And_UInt_UInt, SPARQ_PW_XOR_AND)
SPQ_DEF3(Or_UInt_UInt, SPARQ_PW_XOR_OR)
SPQ_DEF3(Xor_UInt_UInt, SPARQ_PW_XOR_XOR)
SPQ_DEF3(Less_SInt_SInt, SPARQ_PW_FLAG_LESS_S)
SPQ_DEF3(Less_UInt_UInt, SPARQ_PW_FLAG_LESS_U)
SPQ_DEF3(GetMid_UInt_UInt, SPARQ_PW_XOR_MID)
SPQ_DEF3(Div_Sqrt_Arccos_UInt_UInt, SPARQ_PW_XOR_DIVSQRT_ACOS)
SPQ_DEF3(Sqrt_Div_Arccos_Int_UInt, SPARQ_PW_XOR_SQRTDIV_ACOS)
SPQ_DEF3(GetRotateAngle_Int_Int, SPARQ_PW_XOR_ROTATE_ANGLE)
SPQ_DEF2(Neg_UInt, SPARQ_PW_XOR_NEG)
SPQ_DEF2(Abs_SInt, SPARQ_PW_XOR_ABS)
SPQ_DEF2(Sqrt_UInt, SPARQ_PW_XOR_ISQRT)
SPQ_DEF2(IsZero_UInt, SPARQ_PW_FLAG_ISZERO)
SPQ_DEF2(Negative_SInt, SPARQ_PW_FLAG_NEGATIVE)
SPQ_DEF4(Select_Bool_UInt_UInt, SPARQ_PW_XOR_SELECT)
SPQ_DEF4(Carry_UInt_UInt, SPARQ_PW_FLAG_CARRY)
SPQ_DEF4(Overflow_SInt_SInt, SPARQ_PW_FLAG_OVERFLOW_S)
SPQ_DEF4(MulOverflow_UInt_UInt, SPARQ_PW_FLAG_MULOVERFLOW)

Add_UInt_UInt_InPlace::Add_UInt_UInt_InPlace(size_t a, size_t b) {   // quantum_arithmetic.cpp:308-349
    require_ids({a, b});
    set_op(*this, "Add_UInt_UInt_InPlace", SPARQ_PW_ADD_FIELD, {a, b}, SPARQ_PW_SUB_FIELD);
}
Add_AnyInt_AnyInt_InPlace::Add_AnyInt_AnyInt_InPlace(size_t a, size_t b) {   // :869-907
    require_ids({a, b});
    set_op(*this, "Add_AnyInt_AnyInt_InPlace", SPARQ_PW_ADD_ANY, {a, b}, SPARQ_PW_SUB_ANY);
}
Assign::Assign(size_t a, size_t b) {   // :909-930 (quantum_arithmetic.h:2777-2783 rejects a == b)
    if (a == b) throw std::invalid_argument("Assign: registers must differ");
    require_ids({a, b});
    set_op(*this, "Assign", SPARQ_PW_XOR_ASSIGN, {a, b});
}
Swap_General_General::Swap_General_General(size_t a, size_t b) {   // :982-999
    require_ids({a, b});
    if (System::size_of(a) != System::size_of(b))
        throw std::invalid_argument("Swap_General_General: registers of different widths");
    set_op(*this, "Swap_General_General", SPARQ_PW_SWAP_FIELDS, {a, b});
}
Compare_UInt_UInt::Compare_UInt_UInt(size_t l, size_t r, size_t lt, size_t eq) {   // :932-959
    if (lt == eq || lt == l || lt == r || eq == l || eq == r) throw std::invalid_argument("Compare_UInt_UInt: bad registers");
    require_ids({l, r, lt, eq});
    set_op(*this, "Compare_UInt_UInt", SPARQ_PW_COMPARE, {l, r, lt, eq});
}
Mult_UInt_ConstUInt::Mult_UInt_ConstUInt(size_t a, uint64_t k, size_t b) {   // :137-155
    require_ids({a, b});
    set_op(*this, "Mult_UInt_ConstUInt", SPARQ_PW_XOR_MUL_CONST, {a, b});
    c[0] = k;
}
Add_Mult_UInt_ConstUInt_InPlace::Add_Mult_UInt_ConstUInt_InPlace(size_t a, uint64_t k, size_t b) {   // :157-195
    require_ids({a, b});
    set_op(*this, "Add_Mult_UInt_ConstUInt_InPlace", SPARQ_PW_ADD_MUL_CONST, {a, b}, SPARQ_PW_ADD_MUL_CONST);
    c[0] = k;
    dag_c[0] = uint64_t(0) - k;   // (r - l*k) mod 2^64 == (r + l*(-k)) mod 2^64
}
Add_UInt_ConstUInt::Add_UInt_ConstUInt(size_t a, uint64_t k, size_t b) {   // :351-369
    require_ids({a, b});
    set_op(*this, "Add_UInt_ConstUInt", SPARQ_PW_XOR_ADD_CONST, {a, b});
    c[0] = k;
}
Add_ConstUInt_InPlace::Add_ConstUInt_InPlace(size_t a, uint64_t k) {   // :371-411
    require_ids({a});
    set_op(*this, "Add_ConstUInt_InPlace", SPARQ_PW_ADD_CONST, {a}, SPARQ_PW_ADD_CONST);
    c[0] = k;
    dag_c[0] = uint64_t(0) - k;
}
ShiftLeft_InPlace::ShiftLeft_InPlace(size_t a, uint64_t k) {   // :63-135
    require_ids({a});
    if (k > System::size_of(a)) throw std::invalid_argument("shift larger than the register");
    set_op(*this, "ShiftLeft_InPlace", SPARQ_PW_ROTL, {a}, SPARQ_PW_ROTR);
    c[0] = dag_c[0] = k;
}
ShiftRight_InPlace::ShiftRight_InPlace(size_t a, uint64_t k) {
    require_ids({a});
    if (k > System::size_of(a)) throw std::invalid_argument("shift larger than the register");
    set_op(*this, "ShiftRight_InPlace", SPARQ_PW_ROTR, {a}, SPARQ_PW_ROTL);
    c[0] = dag_c[0] = k;
}
X_Bool::X_Bool(size_t a, uint64_t digit) {   // basic_gates.cpp:383-400; GateBase digit check basic_gates.h:20-40
    require_ids({a});
    if (digit >= System::size_of(a)) throw std::invalid_argument("digit out of range");
    set_op(*this, "X_Bool", SPARQ_PW_FLIP_BIT, {a});
    c[0] = digit;
}
Y_Bool::Y_Bool(size_t a, uint64_t digit) {   // basic_gates.cpp:402-420
    require_ids({a});
    if (digit >= System::size_of(a)) throw std::invalid_argument("digit out of range");
    set_op(*this, "Y_Bool", SPARQ_PW_Y, {a});
    c[0] = digit;
}
Init_Unsafe::Init_Unsafe(size_t a, uint64_t value) {   // dark_magic.cpp:35-41
    require_ids({a});
    if (value & ~width_mask(System::size_of(a))) throw std::invalid_argument("Init_Unsafe: value exceeds the register width");
    set_op(*this, "Init_Unsafe", SPARQ_PW_SET_FIELD, {a});
    c[0] = value;
}
FlipBools::FlipBools(size_t a) {   // quantum_arithmetic.cpp:11-30
    require_ids({a});
    set_op(*this, "FlipBools", SPARQ_PW_FLIP_ALL, {a});
}
Mod_Mult_UInt_ConstUInt_InPlace::Mod_Mult_UInt_ConstUInt_InPlace(size_t r, uint64_t a, uint64_t x, uint64_t N) {   // :197-285
    require_ids({r});
    if (N == 0 || std::gcd(a, N) > 1) throw std::invalid_argument("a and N must be coprime");
    uint64_t opnum = a % N;
    for (uint64_t i = 0; i < x; ++i) opnum = (opnum * opnum) % N;
    int64_t t = 0, nt = 1, rr = (int64_t)N, nr = (int64_t)opnum;
    while (nr != 0) {
        int64_t q = rr / nr, tt = t - q * nt, tr = rr - q * nr;
        t = nt; rr = nr; nt = tt; nr = tr;
    }
    uint64_t inv = 1;
    if (rr == 1) {
        int64_t tn = t % (int64_t)N;
        if (tn < 0) tn += (int64_t)N;
        inv = (uint64_t)tn;
    }
    set_op(*this, "Mod_Mult_UInt_ConstUInt_InPlace", SPARQ_PW_MODMUL_CONST, {r}, SPARQ_PW_MODMUL_CONST);
    c[0] = opnum; c[1] = N;
    dag_c[0] = inv; dag_c[1] = N;
}
Swap_Bool_Bool::Swap_Bool_Bool(size_t a, size_t d1, size_t b, size_t d2) {   // :32-61
    require_ids({a, b});
    if (d1 >= System::size_of(a) || d2 >= System::size_of(b)) throw std::invalid_argument("digit out of range");
    set_op(*this, "Swap_Bool_Bool", SPARQ_PW_SWAP_BITS, {a, b});
    c[0] = d1; c[1] = d2;
}
Phase_Bool::Phase_Bool(size_t r, size_t digit, double lambda) {   // basic_gates.cpp:11-55
    require_ids({r});
    if (digit >= System::size_of(r)) throw std::invalid_argument("digit out of range");
    set_op(*this, "Phase_Bool", SPARQ_PW_PHASE_BITS, {r}, SPARQ_PW_PHASE_BITS);
    c[0] = dag_c[0] = uint64_t(1) << digit;
    d[0] = std::cos(lambda); d[1] = std::sin(lambda);
    dag_d[0] = std::cos(lambda); dag_d[1] = -std::sin(lambda);
}
RZ_Bool::RZ_Bool(size_t r, size_t digit, double angle) {   // basic_gates.cpp:486-520
    require_ids({r});
    if (digit >= System::size_of(r)) throw std::invalid_argument("digit out of range");
    set_op(*this, "RZ_Bool", SPARQ_PW_RZ, {r}, SPARQ_PW_RZ);
    c[0] = dag_c[0] = digit;
    d[0] = std::cos(angle / 2); d[1] = std::sin(angle / 2); d[2] = std::cos(angle / 2); d[3] = -std::sin(angle / 2);
    double na = -angle;
    dag_d[0] = std::cos(na / 2); dag_d[1] = std::sin(na / 2); dag_d[2] = std::cos(na / 2); dag_d[3] = -std::sin(na / 2);
}

QRAMCircuit_qutrit::QRAMCircuit_qutrit(size_t addr, size_t data, std::vector<uint64_t> mem)
    : address_size(addr), data_size(data), memory(std::move(mem)) {
    if (addr > 40) throw std::invalid_argument("address size too large");
    if (memory.size() < (size_t(1) << addr)) throw std::invalid_argument("memory smaller than 2^address_size");
    for (uint64_t v : memory) max_word = std::max(max_word, v);
    word_bytes = max_word < (1ull << 8) ? 1 : max_word < (1ull << 16) ? 2 : max_word < (1ull << 32) ? 4 : 8;
    std::vector<uint8_t> packed((size_t(1) << addr) * word_bytes);
    for (size_t i = 0; i < (size_t(1) << addr); ++i) {
        uint64_t v = memory[i];
        std::memcpy(packed.data() + i * word_bytes, &v, word_bytes);   // little endian
    }
    sparq_table_t t = nullptr;
    check(sparq_table_create(packed.data(), size_t(1) << addr, word_bytes, &t));
    table = std::shared_ptr<sparq_table_s>(t, [](sparq_table_s* p) { sparq_table_destroy(p); });
}

QRAMLoad::QRAMLoad(const QRAMCircuit_qutrit* q, size_t a, size_t dreg) {   // qram.cpp:15-35, qram.h:49-71
    if (q == nullptr || a == dreg) throw std::invalid_argument("QRAMLoad: bad arguments");
    require_ids({a, dreg});
    if (q->max_word & ~width_mask(System::size_of(dreg)))
        throw std::invalid_argument("QRAMLoad: memory word wider than the data register");
    set_op(*this, "QRAMLoad", SPARQ_PW_QRAM_LOAD, {a, dreg});
    c[0] = q->address_size;
    const void* ptr = nullptr;
    check(sparq_table_device_ptr(q->table.get(), &ptr));
    table = ptr;
    table_word_bytes = q->word_bytes;
    table_owner = q->table;
}

ZeroConditionalPhaseFlip::ZeroConditionalPhaseFlip(const std::vector<size_t>& ids) {   // parallel_phase_operations.cpp:11-31
    require_ids(ids);
    set_op(*this, "ZeroConditionalPhaseFlip", SPARQ_PW_NEG_AMP, {});
    for (size_t id : ids) extra_zero_regs.emplace_back(id, 0);
}
ZeroConditionalPhaseFlip::ZeroConditionalPhaseFlip(const std::vector<std::string>& names) {
    std::vector<size_t> ids;
    for (auto& n : names) ids.push_back(System::get(n));
    *this = ZeroConditionalPhaseFlip(ids);
}
Reflection_Bool::Reflection_Bool(const std::vector<size_t>& ids, bool inv) : inverse(inv) {   // :46-70
    require_ids(ids);
    set_op(*this, "Reflection_Bool", SPARQ_PW_NEG_AMP, {});
    for (size_t id : ids) extra_zero_regs.emplace_back(id, 0);
}
Reflection_Bool::Reflection_Bool(const std::vector<std::string>& names, bool inv) : inverse(inv) {
    std::vector<size_t> ids;
    for (auto& n : names) ids.push_back(System::get(n));
    require_ids(ids);
    set_op(*this, "Reflection_Bool", SPARQ_PW_NEG_AMP, {});
    for (size_t id : ids) extra_zero_regs.emplace_back(id, 0);
}
void Reflection_Bool::operator()(SparseState& state) const {
    // The reference flips when !(inverse ^ all_zero) (parallel_phase_operations.cpp:63-66):
    // inverse = true flips the all-zero entries; inverse = false flips every other
    // entry, done here as "flip all, then flip the all-zero ones back" (two exact
    // sign changes, bit-identical to one conditional flip).
    if (!inverse) {
        const Layout& L = state.sync();
        sparq_pw_op_t op{};
        op.opcode = SPARQ_PW_NEG_AMP;
        state.push_pointwise(op, conditions(L));
    }
    emit(state, false);
}
GlobalPhase::GlobalPhase(complex_t cc) {   // parallel_phase_operations.cpp:72-108
    set_op(*this, "GlobalPhase", SPARQ_PW_MUL_AMP, {}, SPARQ_PW_MUL_AMP);
    d[0] = cc.real(); d[1] = cc.imag();
    complex_t cj = std::conj(cc);
    dag_d[0] = cj.real(); dag_d[1] = cj.imag();
}

PlusOneAndOverflow::PlusOneAndOverflow(size_t main, size_t overflow) {   // block_encoding_tridiagonal.cpp:20-84
    require_ids({main, overflow});
    set_op(*this, "PlusOneAndOverflow", SPARQ_PW_INC_WRAP, {main, overflow}, SPARQ_PW_DEC_WRAP);
}

// ------------------------------------------------------------ branching operators
static std::vector<sparq_cond_t> conds_of(const Controllable* c, const Layout& L) { return c ? c->conditions(L) : std::vector<sparq_cond_t>{}; }

void Hadamard_Int::operator()(SparseState& state) const {   // hadamard.cpp:134-162
    const Layout& L = state.sync();
    state.flush();
    auto conds = conds_of(this, L);
    check(sparq_hadamard_int(state.handle(), L.at(id), (uint32_t)n_digits, conds.data(), (uint32_t)conds.size(), clearzero_default_eps()));
    System::update_max_size(state.size());
}
void Hadamard_Int_Full::operator()(SparseState& state) const {   // hadamard.cpp:181-267
    const Layout& L = state.sync();
    state.flush();
    auto conds = conds_of(this, L);
    check(sparq_hadamard_int_full(state.handle(), L.at(id), conds.data(), (uint32_t)conds.size(), clearzero_default_eps()));
    System::update_max_size(state.size());
}
Hadamard_Bool::Hadamard_Bool(size_t reg) : out_id(reg) {   // hadamard.h:160-175
    if (System::size_of(out_id) != 1) throw std::invalid_argument("Hadamard_Bool: size of output register must be 1");
}
void Hadamard_Bool::operator()(SparseState& state) const {   // hadamard.cpp:366-439
    const Layout& L = state.sync();
    state.flush();
    auto conds = conds_of(this, L);
    sparq_pair_params_t p{};
    p.kind = SPARQ_PAIR_HADAMARD_BOOL;
    check(sparq_pair_op(state.handle(), L.at(out_id), 0, &p, conds.data(), (uint32_t)conds.size(), clearzero_default_eps()));
    System::update_max_size(state.size());
}
CondRot_Rational_Bool::CondRot_Rational_Bool(size_t in, size_t out) : register_in(in), register_out(out) {}
void CondRot_Rational_Bool::apply(SparseState& state, double sign) const {   // condrot.cpp:12-151
    const Layout& L = state.sync();
    state.flush();
    sparq_pair_params_t p{};
    p.kind = SPARQ_PAIR_CONDROT_RATIONAL;
    p.in = L.at(register_in);
    p.mat[0] = sign;
    // Conditions are an extension (the reference class is not controllable): the
    // rotation acts on the entries that satisfy them, the others pass through.
    std::vector<sparq_cond_t> conds = conditions(L);
    check(sparq_pair_op(state.handle(), L.at(register_out), 0, &p, conds.data(), (uint32_t)conds.size(), clearzero_default_eps()));
}

u22_t make_func(uint64_t value, size_t n) {   // condrot.h:19-30
    double th = (n == 64) ? value * 1.0 / 2 / 9223372036854775808.0 : value * 1.0 / double(uint64_t(1) << n);
    th *= (2 * pi);
    return u22_t{complex_t(std::cos(th)), complex_t(-std::sin(th)), complex_t(std::sin(th)), complex_t(std::cos(th))};
}
u22_t make_func_inv(uint64_t value, size_t n) {   // condrot.h:32-41
    double th = (n == 64) ? value * 1.0 / 2 / 9223372036854775808.0 : value * 1.0 / double(uint64_t(1) << n);
    th *= (2 * pi);
    return u22_t{complex_t(std::cos(th)), complex_t(std::sin(th)), complex_t(-std::sin(th)), complex_t(std::cos(th))};
}

CondRot_General_Bool::CondRot_General_Bool(size_t in, size_t out, std::function<u22_t(uint64_t)> f) : in_id(in), out_id(out), func(std::move(f)) {
    if (System::size_of(out_id) != 1) throw std::invalid_argument("CondRot_General_Bool: size of output register must be 1");
}
namespace {
struct DeviceTable {
    double* ptr = nullptr;
    sparq_state_t carrier = nullptr;   // tables live in a helper state's amplitude array
};
}  // namespace
void CondRot_General_Bool::operator()(SparseState& state) const {   // condrot.h:149-332
    const Layout& L = state.sync();
    state.flush();
    size_t w = System::size_of(in_id);
    if (w > 24) throw std::invalid_argument("CondRot_General_Bool: input register wider than 24 bits");
    size_t nv = size_t(1) << w;
    // Tabulate the 2x2 for every input value and upload it as a temporary device state.
    std::vector<double> tab(nv * 8);
    for (size_t v = 0; v < nv; ++v) {
        u22_t m = func(v);
        for (int k = 0; k < 4; ++k) { tab[v * 8 + 2 * k] = m[k].real(); tab[v * 8 + 2 * k + 1] = m[k].imag(); }
    }
    sparq_state_t carrier = nullptr;
    check(sparq_state_create(1, nv * 4, &carrier));
    std::vector<uint64_t> zeros(nv * 4, 0);
    int rc = sparq_state_upload(carrier, zeros.data(), tab.data(), nv * 4);
    if (rc != SPARQ_OK) { sparq_state_destroy(carrier); check(rc); }
    double* dev = nullptr;
    sparq_state_device_arrays(carrier, nullptr, &dev, nullptr);
    sparq_pair_params_t p{};
    p.kind = SPARQ_PAIR_CONDROT_TABLE;
    p.in = L.at(in_id);
    p.table = dev;
    rc = sparq_pair_op(state.handle(), L.at(out_id), 0, &p, nullptr, 0, clearzero_default_eps());
    sparq_state_sync(state.handle());
    sparq_state_destroy(carrier);
    check(rc);
}

Rot_Bool::Rot_Bool(size_t reg, size_t dg, u22_t m) : id(reg), digit(dg), mat(m) {   // basic_gates.h:157-200
    if (digit >= System::size_of(id)) throw std::invalid_argument("digit out of range");
}
static double abs_sqr(complex_t c) { return c.real() * c.real() + c.imag() * c.imag(); }
void Rot_Bool::operator()(SparseState& state) const {   // basic_gates.cpp:74-255
    const Layout& L = state.sync();
    auto conds = conditions(L);
    sparq_pw_op_t op{};
    op.f[0] = L.at(id);
    op.c[0] = digit;
    if (abs_sqr(mat[1]) < epsilon && abs_sqr(mat[2]) < epsilon) {          // diagonal, in place
        op.opcode = SPARQ_PW_ROT_DIAG;
        op.d[0] = mat[0].real(); op.d[1] = mat[0].imag(); op.d[2] = mat[3].real(); op.d[3] = mat[3].imag();
        state.push_pointwise(op, conds);
        return;
    }
    if (abs_sqr(mat[0]) < epsilon && abs_sqr(mat[3]) < epsilon) {          // off-diagonal, flip in place
        op.opcode = SPARQ_PW_ROT_OFFDIAG;
        op.d[0] = mat[2].real(); op.d[1] = mat[2].imag(); op.d[2] = mat[1].real(); op.d[3] = mat[1].imag();
        state.push_pointwise(op, conds);
        return;
    }
    state.flush();
    sparq_pair_params_t p{};
    p.kind = SPARQ_PAIR_ROT_GENERAL;
    for (int k = 0; k < 4; ++k) { p.mat[2 * k] = mat[k].real(); p.mat[2 * k + 1] = mat[k].imag(); }
    check(sparq_pair_op(state.handle(), L.at(id), (uint32_t)digit, &p, conds.data(), (uint32_t)conds.size(), clearzero_default_eps()));
    System::update_max_size(state.size());
}
void Rot_Bool::dag(SparseState& state) const {   // basic_gates.cpp:310-315
    u22_t d{std::conj(mat[0]), std::conj(mat[2]), std::conj(mat[1]), std::conj(mat[3])};
    Rot_Bool inv(id, digit, d);
    copy_control_conditions_to(inv);
    inv(state);
}
// basic_gates.cpp:422-520
static u22_t rx(double a) { return u22_t{complex_t(std::cos(a / 2), 0), complex_t(0, -std::sin(a / 2)), complex_t(0, -std::sin(a / 2)), complex_t(std::cos(a / 2), 0)}; }
static u22_t ry(double a) { return u22_t{complex_t(std::cos(a / 2), 0), complex_t(-std::sin(a / 2), 0), complex_t(std::sin(a / 2), 0), complex_t(std::cos(a / 2), 0)}; }
static u22_t sx() { return u22_t{complex_t(1, 1) / 2.0, complex_t(1, -1) / 2.0, complex_t(1, -1) / 2.0, complex_t(1, 1) / 2.0}; }
static const double kSqrt2 = 1.4142135623730950488016887242096981;
static u22_t u2(double phi, double lambda) {
    return u22_t{complex_t(1, 0) / kSqrt2, complex_t(-std::cos(lambda), -std::sin(lambda)) / kSqrt2,
                 complex_t(std::cos(phi), std::sin(phi)) / kSqrt2, complex_t(std::cos(lambda + phi), std::sin(lambda + phi)) / kSqrt2};
}
static u22_t u3(double t, double phi, double lambda) {
    return u22_t{complex_t(std::cos(t / 2), 0), complex_t(-std::cos(lambda) * std::sin(t / 2), -std::sin(lambda) * std::sin(t / 2)),
                 complex_t(std::cos(phi) * std::sin(t / 2), std::sin(phi) * std::sin(t / 2)),
                 complex_t(std::cos(lambda + phi) * std::cos(t / 2), std::sin(lambda + phi) * std::cos(t / 2))};
}
RX_Bool::RX_Bool(std::string_view r, size_t dg, double a) : Rot_Bool(r, dg, rx(a)) {}
RX_Bool::RX_Bool(size_t r, size_t dg, double a) : Rot_Bool(r, dg, rx(a)) {}
RY_Bool::RY_Bool(std::string_view r, size_t dg, double a) : Rot_Bool(r, dg, ry(a)) {}
RY_Bool::RY_Bool(size_t r, size_t dg, double a) : Rot_Bool(r, dg, ry(a)) {}
SX_Bool::SX_Bool(std::string_view r, size_t dg) : Rot_Bool(r, dg, sx()) {}
SX_Bool::SX_Bool(size_t r, size_t dg) : Rot_Bool(r, dg, sx()) {}
U2_Bool::U2_Bool(std::string_view r, size_t dg, double phi, double lambda) : Rot_Bool(r, dg, u2(phi, lambda)) {}
U2_Bool::U2_Bool(size_t r, size_t dg, double phi, double lambda) : Rot_Bool(r, dg, u2(phi, lambda)) {}
U3_Bool::U3_Bool(std::string_view r, size_t dg, double t, double phi, double lambda) : Rot_Bool(r, dg, u3(t, phi, lambda)) {}
U3_Bool::U3_Bool(size_t r, size_t dg, double t, double phi, double lambda) : Rot_Bool(r, dg, u3(t, phi, lambda)) {}

Rot_GeneralUnitary::Rot_GeneralUnitary(size_t reg, std::vector<complex_t> m) : id(reg), mat(std::move(m)) {   // rot.h:40-60
    require_ids({id});
    size_t full = size_t(1) << System::size_of(id);
    if (mat.size() != full * full) throw std::invalid_argument("Matrix size does not match the register's size.");
}
void Rot_GeneralUnitary::apply(SparseState& state, bool dagger) const {   // rot.cpp:340-467
    const Layout& L = state.sync();
    state.flush();
    size_t full = size_t(1) << System::size_of(id);
    std::vector<double> m(full * full * 2);
    for (size_t i = 0; i < full; ++i)
        for (size_t j = 0; j < full; ++j) {
            complex_t v = dagger ? std::conj(mat[j * full + i]) : mat[i * full + j];
            m[2 * (i * full + j)] = v.real();
            m[2 * (i * full + j) + 1] = v.imag();
        }
    auto conds = conditions(L);
    check(sparq_group_unitary(state.handle(), L.at(id), m.data(), conds.data(), (uint32_t)conds.size(), clearzero_default_eps()));
    System::update_max_size(state.size());
}
std::vector<complex_t> stateprep_unitary(const std::vector<complex_t>& target) {   // rot.cpp:469-503
    const size_t N = target.size();
    if (N == 0) throw std::invalid_argument("Cannot prepare an empty vector.");
    double scale = 0;
    for (const auto& t : target) {
        if (!std::isfinite(t.real()) || !std::isfinite(t.imag())) throw std::invalid_argument("State amplitudes must be finite.");
        scale = std::max(scale, std::max(std::abs(t.real()), std::abs(t.imag())));
    }
    if (scale == 0) throw std::invalid_argument("Cannot prepare the zero vector.");
    std::vector<complex_t> u(N);
    double sq = 0;
    for (size_t i = 0; i < N; ++i) {
        u[i] = target[i] / scale;
        sq += abs_sqr(u[i]);
    }
    const double nrm = std::sqrt(sq);
    for (auto& x : u) x /= nrm;
    const double lead = std::abs(u[0]);
    const complex_t ph = lead == 0 ? complex_t(1) : u[0] / lead;
    u[0] += ph;
    double usq = 0;
    for (const auto& x : u) usq += abs_sqr(x);
    std::vector<complex_t> R(N * N);
    for (size_t r = 0; r < N; ++r)
        for (size_t c = 0; c < N; ++c) {
            R[r * N + c] = complex_t(r == c ? 1 : 0) - (2.0 / usq) * u[r] * std::conj(u[c]);
            if (c == 0) R[r * N + c] *= -ph;
        }
    return R;
}
Rot_GeneralStatePrep::Rot_GeneralStatePrep(size_t reg, const std::vector<complex_t>& vec)   // rot.cpp:511-538
    : Rot_GeneralUnitary(reg, [&] {
          if (vec.size() != (size_t(1) << System::size_of(reg))) throw std::invalid_argument("State size does not match the register's size.");
          return stateprep_unitary(vec);
      }()) {}

// ------------------------------------------------------------ structural operators
void ClearZero::operator()(SparseState& state) const {   // system_operations.cpp:371-382
    state.flush();
    check(sparq_clear_zero(state.handle(), eps));
}
void Normalize::operator()(SparseState& state) const {   // dark_magic.cpp:11-33
    state.flush();
    double n2 = 0;
    check(sparq_normalize(state.handle(), &n2));
}
void SortUnconditional::operator()(SparseState& state) const {   // sort_state.cpp:84-88
    state.sync();
    state.flush();
    check(sparq_sort(state.handle(), nullptr));
}
// ---- grouping sorts (sort_state.cpp:10-171)
static void sort_fields(SparseState& state, const std::vector<sparq_field_t>& major, const std::vector<sparq_field_t>& minor) {
    state.flush();
    check(sparq_sort_by(state.handle(), major.data(), (uint32_t)major.size(), minor.data(), (uint32_t)minor.size()));
}
static sparq_field_t bits_of(sparq_field_t f, uint32_t lo, uint32_t count) {
    return sparq_field_t{f.word, (uint8_t)(f.shift + lo), (uint8_t)count, 0};
}
void SortExceptKey::operator()(SparseState& state) const {
    const Layout& L = state.sync();
    sort_fields(state, {}, {L.at(id)});
}
void SortByKey::operator()(SparseState& state) const {
    const Layout& L = state.sync();
    sort_fields(state, {L.at(register_key)}, {});
}
SortExceptBit::SortExceptBit(size_t key, size_t d) : id(key), digit(d) {
    if (digit >= System::size_of(id)) throw std::invalid_argument("SortExceptBit: digit outside the register");
}
void SortExceptBit::operator()(SparseState& state) const {
    // others in id order, then the register with the bit masked (its bits above,
    // then below, the digit), then the bit itself
    const Layout& L = state.sync();
    const sparq_field_t f = L.at(id);
    std::vector<sparq_field_t> minor;
    if (digit + 1 < f.width) minor.push_back(bits_of(f, (uint32_t)digit + 1, f.width - (uint32_t)digit - 1));
    if (digit > 0) minor.push_back(bits_of(f, 0, (uint32_t)digit));
    minor.push_back(bits_of(f, (uint32_t)digit, 1));
    sort_fields(state, {}, minor);
}
static std::vector<size_t> checked_positions(size_t id, std::vector<size_t> pos, const char* who) {
    std::sort(pos.begin(), pos.end());
    if (std::adjacent_find(pos.begin(), pos.end()) != pos.end()) throw std::invalid_argument(std::string(who) + ": repeated qubit");
    for (size_t q : pos)
        if (q >= System::size_of(id)) throw std::invalid_argument(std::string(who) + ": qubit outside the register");
    return pos;
}
SortExceptKeyHadamard::SortExceptKeyHadamard(size_t key, const std::vector<size_t>& q)
    : id(key), qubit_ids(checked_positions(key, q, "SortExceptKeyHadamard")) {}
void SortExceptKeyHadamard::operator()(SparseState& state) const {
    // the key with the listed qubits masked compares in its id position; the
    // full key decides last, i.e. the listed qubits, most significant first
    const Layout& L = state.sync();
    const sparq_field_t f = L.at(id);
    std::vector<sparq_field_t> minor;
    for (size_t i = qubit_ids.size(); i-- > 0;) minor.push_back(bits_of(f, (uint32_t)qubit_ids[i], 1));
    sort_fields(state, {}, minor);
}
void SortByAmplitude::operator()(SparseState& state) const {
    state.sync();
    state.flush();
    check(sparq_sort_by_amplitude(state.handle()));
}
void SortByKey2::operator()(SparseState& state) const {
    if (id1 == id2) throw std::invalid_argument("SortByKey2: the two keys must differ");
    const Layout& L = state.sync();
    sort_fields(state, {L.at(id1), L.at(id2)}, {});
}

// ---- Hadamard_Partial (hadamard.cpp:470-712)
Hadamard_Partial::Hadamard_Partial(size_t reg, const std::vector<size_t>& pos)
    : id(reg), qubit_positions(checked_positions(reg, pos, "Hadamard_Partial")) {}
void Hadamard_Partial::operator()(SparseState& state) const {
    const Layout& L = state.sync();
    state.flush();
    const auto conds = conds_of(this, L);
    sparq_pair_params_t p{};
    p.kind = SPARQ_PAIR_HADAMARD_BOOL;
    // H on each listed bit in ascending order: the butterfly stages of the
    // transform over those bits; only the last prunes
    for (size_t i = 0; i < qubit_positions.size(); ++i) {
        const double eps = i + 1 == qubit_positions.size() ? clearzero_default_eps() : 0.0;
        check(sparq_pair_op(state.handle(), L.at(id), (uint32_t)qubit_positions[i], &p, conds.data(), (uint32_t)conds.size(), eps));
    }
    System::update_max_size(state.size());
}

void CheckDuplicateKey::operator()(SparseState& state) const {   // debugger.cpp:214-222
    state.sync();
    state.flush();
    int dup = 0;
    check(sparq_check_duplicates(state.handle(), &dup));
    if (dup) throw std::runtime_error("CheckDuplicateKey: duplicate key");
}
void TestRemovable::operator()(SparseState& state) const {   // debugger.cpp:191-212
    const Layout& L = state.sync();
    state.flush();
    int uni = 1;
    uint64_t v = 0;
    check(sparq_field_uniform(state.handle(), L.at(register_id), &uni, &v));
    if (!uni) throw std::runtime_error("Removing failed: register " + System::name_of(register_id) + " is not uniform");
}

// ------------------------------------------------------------ captured sequences
// Same key width and the same field for every active register (the register
// table keeps removed entries, so the id ranges may differ).
static bool same_active_fields(const Layout& a, const Layout& b) {
    if (a.words != b.words) return false;
    const size_t n = std::max(a.field.size(), b.field.size());
    for (size_t i = 0; i < n; ++i) {
        const sparq_field_t fa = i < a.field.size() ? a.field[i] : sparq_field_t{0, 0, 0, 0};
        const sparq_field_t fb = i < b.field.size() ? b.field[i] : sparq_field_t{0, 0, 0, 0};
        if (fa.width == 0 && fb.width == 0) continue;
        if (std::memcmp(&fa, &fb, sizeof(sparq_field_t)) != 0) return false;
    }
    return true;
}

CapturedSequence::CapturedSequence(SparseState& state, const std::function<void(SparseState&)>& body) : st_(&state) {
    const Layout before = state.sync();
    state.flush();
    check(sparq_capture_plan(state.handle()));
    auto run = [&](const char* what) {
        body(state);
        state.flush();
        if (!same_active_fields(state.sync(), before))
            throw std::invalid_argument(std::string("CapturedSequence: the sequence must leave the register layout as it found it (") +
                                        what + ")");
    };
    try {
        run("planning");
        check(sparq_capture_record(state.handle()));
        run("recording");
        check(sparq_capture_end(state.handle(), &g_));
    } catch (...) {
        sparq_capture_abort(state.handle());
        throw;
    }
}
CapturedSequence::~CapturedSequence() {
    if (g_) sparq_graph_destroy(g_);
}
void CapturedSequence::launch(size_t times) {
    st_->flush();
    check(sparq_graph_launch(g_, (uint32_t)times));
}
void CapturedSequence::sync() { (void)st_->size(); }
uint64_t CapturedSequence::capacity() const {
    uint64_t c = 0;
    check(sparq_graph_info(g_, &c, nullptr));
    return c;
}
uint64_t CapturedSequence::nodes() const {
    uint64_t n = 0;
    check(sparq_graph_info(g_, nullptr, &n));
    return n;
}

PauliNoiseTable::PauliNoiseTable(size_t r, size_t it, const std::vector<uint64_t>& x, const std::vector<uint64_t>& z)
    : rows(r), iterations(it) {
    if (r == 0 || it == 0) throw std::invalid_argument("PauliNoiseTable: rows and iterations must be positive");
    if (x.size() != r * it || z.size() != r * it) throw std::invalid_argument("PauliNoiseTable: need rows * iterations masks");
    std::vector<uint64_t> words(1 + 2 * r * it, 0);
    for (size_t i = 0; i < r * it; ++i) {
        words[1 + 2 * i] = x[i];
        words[2 + 2 * i] = z[i];
    }
    sparq_table_t t = nullptr;
    check(sparq_table_create(words.data(), words.size(), 8, &t));
    table = std::shared_ptr<sparq_table_s>(t, [](sparq_table_s* p) { sparq_table_destroy(p); });
}
void PauliNoiseTable::set_iteration(uint64_t k) { check(sparq_table_write(table.get(), 0, &k, 1)); }

PauliNoise::PauliNoise(size_t segment, size_t target, const PauliNoiseTable& t) : keep(t.table) {
    require_ids({segment, target});
    set_op(*this, "PauliNoise", SPARQ_PW_PAULI_TABLE, {segment, target});
    c[0] = t.rows;
    c[1] = t.iterations;
    d[0] = std::cos(pi);   // Z_Bool's phase (Phase_Bool with lambda = pi)
    d[1] = std::sin(pi);
    const void* ptr = nullptr;
    check(sparq_table_device_ptr(t.table.get(), &ptr));
    table = ptr;
    table_word_bytes = 8;
    has_dag = false;
}
void AdvancePauliNoise::operator()(SparseState& state) const {
    state.flush();   // after the noise it follows
    check(sparq_table_advance(state.handle(), table.get()));
}

AddRegister::AddRegister(std::string_view n, StateStorageType t, size_t s) : register_name(n), type(t), size(s) {
    if (size > 64) throw std::invalid_argument("register wider than 64 bits");   // system_operations.cpp:235-245
}
size_t AddRegister::operator()(SparseState& state) const {   // system_operations.cpp:251-262
    state.sync();
    size_t id = System::add_register(register_name, type, size);
    state.sync();
    return id;
}
size_t AddRegisterWithHadamard::operator()(SparseState& state) const {   // system_operations.cpp:264-290
    size_t id = AddRegister(register_name, type, size)(state);
    const Layout& L = state.sync();
    state.flush();
    check(sparq_replicate(state.handle(), L.at(id)));
    return id;
}
void RemoveRegister::operator()(SparseState& state) const {   // system_operations.cpp:303-322
    const Layout& L = state.sync();
    if (register_id >= System::name_register_map.size()) throw std::runtime_error("Register not found.");
    int phase = 0;
    check(sparq_capture_phase(state.handle(), &phase));
    if (phase) {
        // Inside a captured sequence the register must be zero: checked (and
        // cleared) on the device instead of read back.
        sparq_pw_op_t op{};
        op.opcode = SPARQ_PW_ASSERT_ZERO;
        op.f[0] = L.at(register_id);
        state.push_pointwise(op, {});
        state.mark_zero(register_id);
    } else {
        state.flush();
        int uni = 1;
        uint64_t v = 0;
        check(sparq_field_uniform(state.handle(), L.at(register_id), &uni, &v));
        if (!uni) throw std::runtime_error("Removing failed: register " + System::name_of(register_id) + " is not uniform");
        if (v == 0) state.mark_zero(register_id);   // its bits need no clearing when the field is dropped
    }
    System::remove_register(register_id);
    state.sync();
}
void MoveBackRegister::operator()(SparseState& state) const {   // system_operations.cpp:223-249
    auto& map = System::name_register_map;
    if (register_id >= map.size()) throw std::runtime_error("Register not found.");
    const size_t last = map.size() - 1;
    if (register_id == last) return;   // already at the back
    if (!System::status_of(register_id) || !System::status_of(last))
        throw std::invalid_argument("MoveBackRegister: the register and the last table entry must be active");
    const Layout old = state.sync();
    state.flush();
    std::swap(map[register_id], map[last]);
    System::name_to_index[std::get<0>(map[register_id])] = register_id;
    System::name_to_index[std::get<0>(map[last])] = last;
    ++System::layout_version;
    const Layout T = Layout::from_registry();
    std::vector<std::pair<sparq_field_t, sparq_field_t>> moves;
    for (size_t id = 0; id < T.field.size(); ++id) {
        if (T.field[id].width == 0) continue;
        const size_t from = id == register_id ? last : id == last ? register_id : id;
        moves.emplace_back(old.field[from], T.field[id]);
    }
    state.relayout_with(T, moves);
}

void Push::operator()(SparseState& state) const {   // system_operations.cpp:338-354
    const size_t garbage = AddRegister(garbage_name, System::type_of(reg_id), System::size_of(reg_id))(state);
    Swap_General_General(reg_id, garbage)(state);
    System::temporal_registers.push_back(garbage);
}
void Push::dag(SparseState& state) const {
    const Pop pop(reg_id);
    pop(state);
}
void Pop::operator()(SparseState& state) const {   // system_operations.cpp:356-369
    if (System::temporal_registers.empty()) throw std::runtime_error("Pop: no pushed register");
    const size_t garbage = System::temporal_registers.back();
    Swap_General_General(reg_id, garbage)(state);
    const RemoveRegister remove(garbage);
    remove(state);
    System::temporal_registers.pop_back();
}

size_t SplitRegister::operator()(SparseState& state) const {   // system_operations.cpp:134-166
    const Layout old = state.sync();
    state.flush();
    size_t fp = System::get(first_name);
    if (fp == SIZE_MAX) throw std::runtime_error("Register not found.");
    size_t orig = System::size_of(fp);
    if (orig < second_size) throw std::invalid_argument("SplitRegister: second register larger than the first");
    StateStorageType type = System::type_of(fp);
    size_t sp = System::add_register(second_name, type, second_size);
    System::set_size(fp, orig - second_size);
    Layout T = Layout::from_registry();
    std::vector<std::pair<sparq_field_t, sparq_field_t>> moves;
    for (size_t id = 0; id < old.field.size(); ++id) {
        if (old.field[id].width == 0 || id == fp || id == sp) continue;
        if (id < T.field.size() && T.field[id].width == old.field[id].width) moves.emplace_back(old.field[id], T.field[id]);
    }
    sparq_field_t f = old.field[fp];
    if (orig - second_size > 0)
        moves.emplace_back(sparq_field_t{f.word, (uint8_t)(f.shift + second_size), (uint8_t)(orig - second_size), 0}, T.field[fp]);
    if (second_size > 0) moves.emplace_back(sparq_field_t{f.word, f.shift, (uint8_t)second_size, 0}, T.field[sp]);
    state.relayout_with(T, moves);
    return sp;
}
size_t CombineRegister::operator()(SparseState& state) const {   // system_operations.cpp:172-203
    const Layout old = state.sync();
    state.flush();
    size_t fp = System::get(first_name), sp = System::get(second_name);
    if (fp == SIZE_MAX || sp == SIZE_MAX) throw std::runtime_error("Register not found.");
    size_t w1 = System::size_of(fp), w2 = System::size_of(sp);
    if (w1 + w2 > 64) throw std::invalid_argument("CombineRegister: combined register wider than 64 bits");
    System::set_size(fp, w1 + w2);
    System::remove_register(sp);
    Layout T = Layout::from_registry();
    std::vector<std::pair<sparq_field_t, sparq_field_t>> moves;
    for (size_t id = 0; id < old.field.size(); ++id) {
        if (old.field[id].width == 0 || id == fp || id == sp) continue;
        if (id < T.field.size() && T.field[id].width == old.field[id].width) moves.emplace_back(old.field[id], T.field[id]);
    }
    sparq_field_t nf = T.field[fp];
    moves.emplace_back(old.field[fp], sparq_field_t{nf.word, (uint8_t)(nf.shift + w2), (uint8_t)w1, 0});
    moves.emplace_back(old.field[sp], sparq_field_t{nf.word, nf.shift, (uint8_t)w2, 0});
    state.relayout_with(T, moves);
    return fp;
}

// ------------------------------------------------------------ debug checks
bool& debug_checks() {
    static bool on = false;
    return on;
}
void CheckNormalization::operator()(SparseState& state) const {   // debugger.cpp:13-28
    if (!debug_checks()) return;
    state.flush();
    double f = 0;
    check(sparq_norm2(state.handle(), &f));
    if (!(std::fabs(f - 1.0) < threshold) || std::isnan(f))
        throw std::runtime_error("State is not normalized (factor " + std::to_string(f) + ").");
}
void CheckNan::operator()(SparseState& state) const {   // debugger.cpp:57-70
    if (!debug_checks()) return;
    state.flush();
    double f = 0;
    check(sparq_norm2(state.handle(), &f));
    if (std::isnan(f)) throw std::runtime_error("Bad result: NaN amplitude.");
}

// ------------------------------------------------------------ post-selection
PartialTraceSelect::PartialTraceSelect(const std::map<std::string, uint64_t>& sel) {
    for (auto& [name, v] : sel) {
        partial_trace_registers.push_back(System::get(name));
        select_values.push_back(v);
    }
}
PartialTraceSelect::PartialTraceSelect(const std::map<size_t, uint64_t>& sel) {
    for (auto& [id, v] : sel) {
        partial_trace_registers.push_back(id);
        select_values.push_back(v);
    }
}
PartialTraceSelect::PartialTraceSelect(const std::vector<size_t>& regs, const std::vector<uint64_t>& values)
    : partial_trace_registers(regs), select_values(values) {
    if (regs.size() != values.size()) throw std::invalid_argument("PartialTraceSelect: one value per register");
}
PartialTraceSelect::PartialTraceSelect(const std::vector<std::string>& regs, const std::vector<uint64_t>& values)
    : select_values(values) {
    if (regs.size() != values.size()) throw std::invalid_argument("PartialTraceSelect: one value per register");
    for (auto& n : regs) partial_trace_registers.push_back(System::get(n));
}
static double select_on_device(SparseState& state, const std::vector<sparq_cond_t>& conds) {
    state.flush();
    double f = 0;
    check(sparq_select(state.handle(), conds.data(), (uint32_t)conds.size(), &f));
    return f;
}
double PartialTraceSelect::operator()(SparseState& state) const {   // partial_trace.cpp:148-185
    if (partial_trace_registers.empty()) return 0;
    const Layout& L = state.sync();
    std::vector<sparq_cond_t> conds;
    for (size_t i = 0; i < partial_trace_registers.size(); ++i) {
        const size_t id = partial_trace_registers[i];
        if (!System::status_of(id)) throw std::runtime_error("Register not found.");
        // A value wider than the register can match no entry.
        if (select_values[i] & ~width_mask(System::size_of(id))) {
            conds.push_back(sparq_cond_t{SPARQ_COND_VALUE, L.at(id), 0});
            conds.push_back(sparq_cond_t{SPARQ_COND_NONZERO, L.at(id), 0});
        } else {
            conds.push_back(sparq_cond_t{SPARQ_COND_VALUE, L.at(id), select_values[i]});
        }
    }
    return select_on_device(state, conds);
}
double PartialTraceSelect::operator()(std::vector<System>& state) const {
    if (partial_trace_registers.empty() || state.empty()) return 0;
    double sum = 0;
    auto drop = [&](const System& s) {
        for (size_t i = 0; i < partial_trace_registers.size(); ++i)
            if (s.get(partial_trace_registers[i]).value != select_values[i]) return true;
        sum += std::norm(s.amplitude);
        return false;
    };
    state.erase(std::remove_if(state.begin(), state.end(), drop), state.end());
    if (sum == 0) return std::numeric_limits<double>::infinity();
    const double f = 1.0 / std::sqrt(sum);
    for (auto& s : state) s.amplitude *= f;
    return f;
}
double PartialTraceSelectRange::operator()(SparseState& state) const {   // partial_trace.cpp:187-225
    const Layout& L = state.sync();
    if (!System::status_of(partial_trace_register)) throw std::runtime_error("Register not found.");
    const sparq_field_t f = L.at(partial_trace_register);
    std::vector<sparq_cond_t> conds{sparq_cond_t{SPARQ_COND_GE, f, select_range.first},
                                    sparq_cond_t{SPARQ_COND_LE, f, select_range.second}};
    return select_on_device(state, conds);
}
double PartialTraceSelectRange::operator()(std::vector<System>& state) const {
    if (state.empty()) return 0;
    double sum = 0;
    auto drop = [&](const System& s) {
        const uint64_t v = s.get(partial_trace_register).value;
        if (v >= select_range.first && v <= select_range.second) {
            sum += std::norm(s.amplitude);
            return false;
        }
        return true;
    };
    state.erase(std::remove_if(state.begin(), state.end(), drop), state.end());
    if (sum == 0) return std::numeric_limits<double>::infinity();
    const double f = 1.0 / std::sqrt(sum);
    for (auto& s : state) s.amplitude *= f;
    return f;
}

// ------------------------------------------------------------ measurement
namespace random_engine {
static uint64_t g_seed = 0x5EED2026ull, g_state = 0x5EED2026ull;
void set_seed(uint64_t seed) { g_seed = g_state = seed; }
uint64_t get_seed() { return g_seed; }
double uniform01() {
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * (1.0 / 9007199254740992.0);
}
}  // namespace random_engine

namespace {
// measurement.cpp:18-103: ids exist, are active, are distinct; values fit the widths
void validate_ids(const std::vector<size_t>& ids, const char* ctx) {
    for (size_t id : ids)
        if (id >= System::name_register_map.size() || !System::status_of(id))
            throw std::invalid_argument(std::string(ctx) + ": register id " + std::to_string(id) + " is not an active register.");
    std::vector<size_t> sorted(ids);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        throw std::invalid_argument(std::string(ctx) + ": duplicate register id in the same register list.");
}
std::vector<size_t> resolve_names(const std::vector<std::string>& names, const char* ctx) {
    std::vector<size_t> ids;
    for (auto& n : names) {
        size_t id = System::get(n);
        if (id == SIZE_MAX) throw std::invalid_argument(std::string(ctx) + ": register '" + n + "' not found.");
        ids.push_back(id);
    }
    validate_ids(ids, ctx);
    return ids;
}
void validate_fit(const std::vector<size_t>& ids, const std::vector<uint64_t>& v, const char* ctx) {
    if (ids.size() != v.size()) throw std::invalid_argument(std::string(ctx) + ": one value per register.");
    for (size_t i = 0; i < ids.size(); ++i)
        if (v[i] & ~width_mask(System::size_of(ids[i])))
            throw std::invalid_argument(std::string(ctx) + ": value " + std::to_string(v[i]) + " does not fit in register '" +
                                        System::name_of(ids[i]) + "'.");
}
std::vector<sparq_cond_t> value_conds(const Layout& L, const std::vector<size_t>& ids, const std::vector<uint64_t>& v) {
    std::vector<sparq_cond_t> c;
    for (size_t i = 0; i < ids.size(); ++i) c.push_back(sparq_cond_t{SPARQ_COND_VALUE, L.at(ids[i]), v[i]});
    return c;
}
}  // namespace

MeasureZ::MeasureZ(const std::vector<std::string>& names) : registers(resolve_names(names, "MeasureZ")) {}
MeasureZ::MeasureZ(const std::vector<size_t>& ids) : registers(ids) { validate_ids(ids, "MeasureZ"); }
MeasureZ::MeasureZ(std::string_view name) : registers(resolve_names({std::string(name)}, "MeasureZ")) {}
MeasureZ::MeasureZ(size_t id) : registers{id} { validate_ids(registers, "MeasureZ"); }
std::pair<std::vector<uint64_t>, double> MeasureZ::operator()(SparseState& state) const {   // measurement.cpp:127-182
    const size_t n = state.size();
    if (n == 0) throw std::invalid_argument("MeasureZ: cannot measure an empty state.");
    constexpr uint64_t kTile = 4096;
    SortUnconditional{}(state);   // canonical order: the walk below does not depend on the engine's entry order
    const Layout& L = state.sync();
    const uint64_t ntiles = (n + kTile - 1) / kTile;
    std::vector<double> tiles(ntiles);
    check(sparq_tile_norm2(state.handle(), kTile, tiles.data(), ntiles));
    double total = 0;
    for (double t : tiles) total += t;
    if (!std::isfinite(total) || total <= 0.0 || !(std::fabs(total - 1.0) < 1e-5))
        throw std::runtime_error("MeasureZ: input state is not normalized (total probability = " + std::to_string(total) +
                                 "); renormalize (e.g. via Normalize()) before measuring.");
    if (registers.empty()) return {{}, 1.0};
    double r = random_engine::uniform01();
    uint64_t pick = n - 1;   // reference fallback: the last entry
    for (uint64_t t = 0; t < ntiles; ++t) {
        if (!(r < tiles[t])) {
            r -= tiles[t];
            continue;
        }
        const uint64_t first = t * kTile, count = std::min<uint64_t>(kTile, n - first);
        std::vector<double> a(2 * count);
        check(sparq_state_read_range(state.handle(), first, count, nullptr, a.data()));
        pick = first + count - 1;
        for (uint64_t i = 0; i < count; ++i) {
            const double p = a[2 * i] * a[2 * i] + a[2 * i + 1] * a[2 * i + 1];
            if (r < p) {
                pick = first + i;
                break;
            }
            r -= p;
        }
        break;
    }
    std::vector<uint64_t> key(L.words);
    check(sparq_state_read_range(state.handle(), pick, 1, key.data(), nullptr));
    std::vector<uint64_t> outcome;
    for (size_t id : registers) outcome.push_back(field_value(key, L.at(id)));
    const std::vector<sparq_cond_t> conds = value_conds(L, registers, outcome);
    double sum_prob = 0, f = 0;
    check(sparq_cond_norm2(state.handle(), conds.data(), (uint32_t)conds.size(), &sum_prob));
    if (!(sum_prob > 0)) throw std::runtime_error("MeasureZ: sampled outcome has zero support.");
    check(sparq_select(state.handle(), conds.data(), (uint32_t)conds.size(), &f));
    return {outcome, sum_prob};
}

Reset::Reset(const std::vector<std::string>& names) : Reset(names, std::vector<uint64_t>(names.size(), 0)) {}
Reset::Reset(const std::vector<std::string>& names, const std::vector<uint64_t>& targets)
    : registers(resolve_names(names, "Reset")), target_values(targets) {
    validate_fit(registers, target_values, "Reset");
}
Reset::Reset(const std::vector<size_t>& ids) : Reset(ids, std::vector<uint64_t>(ids.size(), 0)) {}
Reset::Reset(const std::vector<size_t>& ids, const std::vector<uint64_t>& targets) : registers(ids), target_values(targets) {
    validate_ids(ids, "Reset");
    validate_fit(registers, target_values, "Reset");
}
Reset::Reset(std::string_view name, uint64_t target) : registers(resolve_names({std::string(name)}, "Reset")), target_values{target} {
    validate_fit(registers, target_values, "Reset");
}
Reset::Reset(size_t id, uint64_t target) : registers{id}, target_values{target} {
    validate_ids(registers, "Reset");
    validate_fit(registers, target_values, "Reset");
}
std::vector<uint64_t> Reset::operator()(SparseState& state) const {   // measurement.cpp:213-228
    if (registers.empty()) return {};
    auto res = MeasureZ(registers)(state);
    const Layout& L = state.sync();
    for (size_t i = 0; i < registers.size(); ++i) {
        sparq_pw_op_t op{};
        op.opcode = SPARQ_PW_SET_FIELD;
        op.f[0] = L.at(registers[i]);
        op.c[0] = target_values[i];
        state.push_pointwise(op, {});
    }
    state.flush();
    return res.first;
}

Probability::Probability(const std::map<std::string, uint64_t>& a) {
    for (auto& [k, v] : a) {
        registers.push_back(resolve_names({k}, "Probability")[0]);
        values.push_back(v);
    }
    validate_fit(registers, values, "Probability");
}
Probability::Probability(const std::map<size_t, uint64_t>& a) {
    for (auto& [k, v] : a) {
        registers.push_back(k);
        values.push_back(v);
    }
    validate_ids(registers, "Probability");
    validate_fit(registers, values, "Probability");
}
Probability::Probability(const std::vector<std::string>& names, const std::vector<uint64_t>& v)
    : registers(resolve_names(names, "Probability")), values(v) {
    validate_fit(registers, values, "Probability");
}
Probability::Probability(const std::vector<size_t>& ids, const std::vector<uint64_t>& v) : registers(ids), values(v) {
    validate_ids(registers, "Probability");
    validate_fit(registers, values, "Probability");
}
Probability::Probability(std::string_view name, uint64_t value)
    : registers(resolve_names({std::string(name)}, "Probability")), values{value} {
    validate_fit(registers, values, "Probability");
}
Probability::Probability(size_t id, uint64_t value) : registers{id}, values{value} {
    validate_ids(registers, "Probability");
    validate_fit(registers, values, "Probability");
}
double Probability::operator()(SparseState& state) const {   // measurement.cpp:281-300
    if (registers.empty()) return 1.0;
    const Layout& L = state.sync();
    state.flush();
    const std::vector<sparq_cond_t> conds = value_conds(L, registers, values);
    double p = 0;
    check(sparq_cond_norm2(state.handle(), conds.data(), (uint32_t)conds.size(), &p));
    return p;
}
std::map<uint64_t, double> Probability::distribution(SparseState& state, size_t id) {   // measurement.cpp:302-318
    validate_ids({id}, "Probability::distribution");
    std::vector<size_t> ids;
    std::vector<uint64_t> vals;
    std::vector<complex_t> amps;
    state.export_entries(ids, vals, amps);
    const size_t col = std::find(ids.begin(), ids.end(), id) - ids.begin();
    std::map<uint64_t, double> dist;
    for (size_t i = 0; i < amps.size(); ++i) dist[vals[i * ids.size() + col]] += std::norm(amps[i]);
    return dist;
}
std::map<uint64_t, double> Probability::distribution(SparseState& state, std::string_view name) {
    return distribution(state, resolve_names({std::string(name)}, "Probability::distribution")[0]);
}

// ------------------------------------------------------------ CKS quantum walk primitives
static uint64_t encode_element(double x, size_t ds, bool positive_only) {
    if (positive_only) {
        const double top = double(width_mask(ds));
        double v = std::floor(x * std::ldexp(1.0, (int)ds));
        v = std::min(std::max(v, 0.0), top);
        return uint64_t(v);
    }
    const double top = double(width_mask(ds - 1));
    double v = std::trunc(x * std::ldexp(1.0, (int)ds - 1));
    v = std::min(std::max(v, -top), top);
    return uint64_t(int64_t(v)) & width_mask(ds);
}
static void check_sparse_shape(size_t ne, size_t ns, size_t ds, size_t nnz, size_t nrow) {
    if (ds < 2 || ds > 63) throw std::invalid_argument("SparseMatrix: data_size must be in [2, 63]");
    if (nnz == 0 || nrow == 0) throw std::invalid_argument("SparseMatrix: empty matrix");
    if (ne != nnz * nrow || ns != nnz * nrow) throw std::invalid_argument("SparseMatrix: need nnz_col entries per row");
}
SparseMatrix::SparseMatrix(const std::vector<double>& values, const std::vector<size_t>& sp, size_t ds, size_t nnz, size_t nrow,
                           bool pos)
    : sparsity(sp), data_size(ds), nnz_col(nnz), n_row(nrow), positive_only(pos) {
    check_sparse_shape(values.size(), sp.size(), ds, nnz, nrow);
    elements.reserve(values.size());
    for (double x : values) elements.push_back(encode_element(x, ds, pos));
}
SparseMatrix::SparseMatrix(const std::vector<uint64_t>& enc, const std::vector<size_t>& sp, size_t ds, size_t nnz, size_t nrow, bool pos)
    : elements(enc), sparsity(sp), data_size(ds), nnz_col(nnz), n_row(nrow), positive_only(pos) {
    check_sparse_shape(enc.size(), sp.size(), ds, nnz, nrow);
}
std::vector<uint64_t> SparseMatrix::get_data() const {
    size_t total = 1;
    while (total < elements.size() + sparsity.size()) total <<= 1;
    std::vector<uint64_t> mem(total, 0);
    std::copy(elements.begin(), elements.end(), mem.begin());
    for (size_t i = 0; i < sparsity.size(); ++i) mem[elements.size() + i] = sparsity[i];
    return mem;
}

GetRowAddr::GetRowAddr(size_t offset, size_t row, size_t row_sz, size_t out) {
    require_ids({offset, row, out});
    set_op(*this, "GetRowAddr", SPARQ_PW_XOR_AFFINE, {offset, row, kNoRegister, out});
    c[0] = row_sz;
}
GetDataAddr::GetDataAddr(size_t offset, size_t row, size_t col, size_t row_sz, size_t out) {
    require_ids({offset, row, col, out});
    set_op(*this, "GetDataAddr", SPARQ_PW_XOR_AFFINE, {offset, row, col, out});
    c[0] = row_sz;
}
QuantumBinarySearch_Fast::QuantumBinarySearch_Fast(const QRAMCircuit_qutrit* q, size_t offset, size_t len, size_t target,
                                                   size_t result)
    : total_length(len), max_step(size_t(std::log2(double(len))) + 1) {
    if (q == nullptr || len == 0) throw std::invalid_argument("QuantumBinarySearch_Fast: bad arguments");
    require_ids({offset, target, result});
    set_op(*this, "QuantumBinarySearch_Fast", SPARQ_PW_XOR_BSEARCH, {offset, target, result});
    c[0] = len;
    c[1] = max_step;
    c[2] = width_mask(q->address_size);
    const void* ptr = nullptr;
    check(sparq_table_device_ptr(q->table.get(), &ptr));
    table = ptr;
    table_word_bytes = q->word_bytes;
    table_owner = q->table;
}
GetQWRotateAngle_Int_Int_Int::GetQWRotateAngle_Int_Int_Int(size_t data, size_t row, size_t col, size_t out, const SparseMatrix* mat) {
    if (mat == nullptr) throw std::invalid_argument("GetQWRotateAngle_Int_Int_Int: null matrix");
    require_ids({data, row, col, out});
    if (mat->data_size < 2 || mat->data_size > 63) throw std::invalid_argument("GetQWRotateAngle_Int_Int_Int: bad data size");
    set_op(*this, "GetQWRotateAngle_Int_Int_Int", SPARQ_PW_XOR_QW_ANGLE, {data, out});
    c[0] = mat->data_size;
    c[1] = mat->positive_only ? 1 : 0;
}

size_t TridiagonalSpec::window_base(size_t j) const {   // hamiltonian_simulation_tridiagonal.h:130-140
    const size_t b = j > 0 ? j - 1 : 0;
    const size_t bmax = n_row - slot_count;
    return b < bmax ? b : bmax;
}
uint64_t TridiagonalSpec::value_of(size_t j, size_t s) const {   // :152-160
    const size_t c = col_of(j, s);
    if (c == j) return alpha_enc;
    if (c + 1 == j || c == j + 1) return beta_enc;
    return 0;
}
uint64_t TridiagonalSpec::combined_mem(uint64_t idx) const {   // :169-178
    idx &= addr_mask;
    const uint64_t half = uint64_t(n_row) * slot_count;
    if (idx < half) return value_of(size_t(idx >> 2), size_t(idx & 3));
    const uint64_t t = idx - half;
    return uint64_t(col_of(size_t(t >> 2), size_t(t & 3)));
}
std::pair<SparseMatrix, TridiagonalSpec> make_tridiagonal_matrix_spec(double alpha, double beta, size_t n, size_t ds) {   // :192-225
    if (n < 4 || (n & (n - 1))) throw std::invalid_argument("tridiagonal walk needs a power-of-two dimension >= 4");
    if (ds < 2 || ds > 63) throw std::invalid_argument("data size must be in [2, 63]");
    const double scale = 1.0 / (std::fabs(alpha) + 2.0 * std::fabs(beta));
    TridiagonalSpec spec;
    spec.n_row = n;
    spec.data_size = ds;
    spec.alpha_enc = uint64_t(std::floor(std::fabs(alpha) * scale * std::ldexp(1.0, (int)ds)));
    spec.beta_enc = uint64_t(std::floor(std::fabs(beta) * scale * std::ldexp(1.0, (int)ds)));
    spec.sparse_offset = n * TridiagonalSpec::slot_count;
    spec.addr_mask = 2 * spec.sparse_offset - 1;
    size_t lg = 0;
    while ((size_t(1) << lg) < n) ++lg;
    spec.k_width = std::max(lg + 1, ds);
    std::vector<uint64_t> elements;
    std::vector<size_t> sparsity;
    elements.reserve(n * TridiagonalSpec::slot_count);
    sparsity.reserve(n * TridiagonalSpec::slot_count);
    for (size_t i = 0; i < n; ++i)
        for (size_t s = 0; s < TridiagonalSpec::slot_count; ++s) {
            elements.push_back(spec.value_of(i, s));
            sparsity.push_back(spec.col_of(i, s));
        }
    return {SparseMatrix(elements, sparsity, ds, TridiagonalSpec::slot_count, n), spec};
}
static void set_tridiag(PointwiseOperator& p, const TridiagonalSpec& spec) {
    if (spec.n_row < 4 || (spec.n_row & (spec.n_row - 1)) || spec.n_row >= (uint64_t(1) << 48))
        throw std::invalid_argument("tridiagonal oracle: bad dimension");
    if (spec.addr_mask != 8 * spec.n_row - 1) throw std::invalid_argument("tridiagonal oracle: address mask must be 8n-1");
    p.c[0] = uint64_t(spec.n_row) | (uint64_t(spec.k_width) << 48);
    p.c[1] = spec.alpha_enc;
    p.c[2] = spec.beta_enc;
}
TridiagOracle1::TridiagOracle1(size_t offset, size_t row, size_t slot, size_t out, const TridiagonalSpec& spec) {
    require_ids({offset, row, slot, out});
    set_op(*this, "TridiagOracle1", SPARQ_PW_TRIDIAG_LOAD, {offset, row, slot, out});
    set_tridiag(*this, spec);
}
TridiagOracle2::TridiagOracle2(size_t offset, size_t row, size_t k, size_t sr, const TridiagonalSpec& spec) {
    require_ids({offset, row, k, sr});
    set_op(*this, "TridiagOracle2", SPARQ_PW_TRIDIAG_COL2SLOT, {offset, row, k, sr}, SPARQ_PW_TRIDIAG_SLOT2COL);
    set_tridiag(*this, spec);
    std::copy(c, c + 3, dag_c);
}
TridiagParityPhase::TridiagParityPhase(size_t reg) {   // hamiltonian_tridiagonal.cpp:90-105
    require_ids({reg});
    set_op(*this, "TridiagParityPhase", SPARQ_PW_NEG_AMP, {});
    extra_bit_regs.emplace_back(reg, 0);
}

void add_systems(SparseState& current, SparseState& new_state, double coef, double eps) {   // system_operations.h:157-181
    current.sync();
    new_state.sync();
    current.flush();
    new_state.flush();
    check(sparq_add_systems(current.handle(), new_state.handle(), coef, eps));
}

SparseState split_systems(SparseState& state, const std::vector<size_t>& nonzeros, const std::vector<size_t>& all_ones,
                          const std::vector<std::pair<size_t, size_t>>& by_bit,
                          const std::vector<std::pair<size_t, uint64_t>>& by_value) {   // system_operations.cpp:53-97
    Controllable c;
    c.condition_variable_nonzeros = nonzeros;
    c.condition_variable_all_ones = all_ones;
    c.condition_variable_by_bit = by_bit;
    c.condition_variable_by_value = by_value;
    const Layout& L = state.sync();
    state.flush();
    const auto conds = c.conditions(L);
    SparseState rest;
    rest.sync();
    check(sparq_partition(state.handle(), conds.data(), (uint32_t)conds.size(), rest.handle()));
    return rest;
}
void combine_systems(SparseState& to, SparseState& from) {   // system_operations.cpp:99-106
    to.sync();
    from.sync();
    to.flush();
    from.flush();
    check(sparq_append(to.handle(), from.handle()));
}
void merge_system(System& a, System& b) {   // basic_components.cpp:517-521
    a.amplitude += b.amplitude;
    b.amplitude = 0;
}
bool remove_system(const System& s) { return std::norm(s.amplitude) < epsilon; }   // :523-526

void reset_systems(SparseState& state) {
    state.sync();
    state.flush();
    check(sparq_state_set_basis(state.handle(), nullptr));
}

}  // namespace sparqsim
