// pybind11 module `sparqsim_amd._core`: the pysparq._core surface for the
// sparse-state operator-apply path (PySparQ/core.cpp:21-180,
// PySparQ/BindUtils.h:117-207).  Device work runs with the GIL released.
#include <pybind11/complex.h>
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "sparq.hpp"

namespace py = pybind11;
using namespace sparqsim;

namespace {

using Release = py::call_guard<py::gil_scoped_release>;

size_t resolve(const py::handle& h) {
    if (py::isinstance<py::str>(h)) return System::get(h.cast<std::string>());
    return h.cast<size_t>();
}

u22_t to_u22(const py::handle& m) {
    auto v = m.cast<std::vector<complex_t>>();
    if (v.size() != 4) throw std::invalid_argument("2x2 matrix must have 4 entries (row-major)");
    return u22_t{v[0], v[1], v[2], v[3]};
}

// conditioned_by_* / clear_control_* (PySparQ/BindUtils.h:117-190)
template <class T, class... Extra>
void bind_controllable(py::class_<T, Extra...>& c) {
    c.def_property_readonly("condition_variable_nonzeros", [](T& o) { return o.condition_variable_nonzeros; })
        .def_property_readonly("condition_variable_all_ones", [](T& o) { return o.condition_variable_all_ones; })
        .def_property_readonly("condition_variable_by_bit", [](T& o) { return o.condition_variable_by_bit; })
        .def_property_readonly("condition_variable_by_clear_bit", [](T& o) { return o.condition_variable_by_clear_bit; })
        .def_property_readonly("condition_variable_by_value", [](T& o) { return o.condition_variable_by_value; })
        .def("conditioned_by_nonzeros", [](T& o, py::object c) -> T& {
            o.condition_variable_nonzeros.clear();
            if (py::isinstance<py::str>(c) || py::isinstance<py::int_>(c)) o.condition_variable_nonzeros.push_back(resolve(c));
            else for (auto it : c) o.condition_variable_nonzeros.push_back(resolve(it));
            return o;
        }, py::arg("cond"), py::return_value_policy::reference_internal)
        .def("conditioned_by_all_ones", [](T& o, py::object c) -> T& {
            o.condition_variable_all_ones.clear();
            if (py::isinstance<py::str>(c) || py::isinstance<py::int_>(c)) o.condition_variable_all_ones.push_back(resolve(c));
            else for (auto it : c) o.condition_variable_all_ones.push_back(resolve(it));
            return o;
        }, py::arg("cond"), py::return_value_policy::reference_internal)
        .def("conditioned_by_bit", [](T& o, py::object c, py::object pos) -> T& {
            o.condition_variable_by_bit.clear();
            if (pos.is_none()) {
                for (auto it : c) {
                    auto t = it.cast<py::tuple>();
                    o.condition_variable_by_bit.emplace_back(resolve(t[0]), t[1].cast<size_t>());
                }
            } else {
                o.condition_variable_by_bit.emplace_back(resolve(c), pos.cast<size_t>());
            }
            return o;
        }, py::arg("cond"), py::arg("pos") = py::none(), py::return_value_policy::reference_internal)
        .def("conditioned_by_clear_bit", [](T& o, py::object c, py::object pos) -> T& {
            o.condition_variable_by_clear_bit.clear();
            if (pos.is_none()) {
                for (auto it : c) {
                    auto t = it.cast<py::tuple>();
                    o.condition_variable_by_clear_bit.emplace_back(resolve(t[0]), t[1].cast<size_t>());
                }
            } else {
                o.condition_variable_by_clear_bit.emplace_back(resolve(c), pos.cast<size_t>());
            }
            return o;
        }, py::arg("cond"), py::arg("pos") = py::none(), py::return_value_policy::reference_internal)
        .def("conditioned_by_value", [](T& o, py::object c, py::object val) -> T& {
            o.condition_variable_by_value.clear();
            if (val.is_none()) {
                for (auto it : c) {
                    auto t = it.cast<py::tuple>();
                    o.condition_variable_by_value.emplace_back(resolve(t[0]), t[1].cast<uint64_t>());
                }
            } else {
                o.condition_variable_by_value.emplace_back(resolve(c), val.cast<uint64_t>());
            }
            return o;
        }, py::arg("cond"), py::arg("value") = py::none(), py::return_value_policy::reference_internal)
        .def("clear_control_nonzeros", [](T& o) { o.clear_control_nonzeros(); })
        .def("clear_control_all_ones", [](T& o) { o.clear_control_all_ones(); })
        .def("clear_control_by_bit", [](T& o) { o.clear_control_by_bit(); })
        .def("clear_control_by_clear_bit", [](T& o) { o.clear_control_by_clear_bit(); })
        .def("clear_control_by_value", [](T& o) { o.clear_control_by_value(); });
}

template <class T>
py::class_<T, PointwiseOperator, std::shared_ptr<T>> pw_class(py::module_& m, const char* name) {
    return py::class_<T, PointwiseOperator, std::shared_ptr<T>>(m, name);
}

}  // namespace

PYBIND11_MODULE(_core, m) {
    m.doc() = "MI355X sparse-state engine: pysparq-compatible register-level API (HIP backend)";

    py::enum_<StateStorageType>(m, "StateStorageType")
        .value("General", General).value("UnsignedInteger", UnsignedInteger)
        .value("SignedInteger", SignedInteger).value("Boolean", Boolean).value("Rational", Rational)
        .export_values();

    // StateStorage / System entries (PySparQ/core.cpp:75-77, 104-170): one host-side
    // entry of a state, as SparseState.basis_states returns them
    py::class_<StateStorage>(m, "StateStorage")
        .def(py::init<>())
        .def_readonly("value", &StateStorage::value)
        .def("__repr__", [](const StateStorage& s) { return "StateStorage(" + std::to_string(s.value) + ")"; });

    py::class_<System>(m, "System")
        .def(py::init<>())
        .def_readwrite("amplitude", &System::amplitude)
        .def("get", [](System& s, py::object r) -> StateStorage& { return s.get(resolve(r)); }, py::arg("reg"),
             py::return_value_policy::reference_internal)
        .def_property_readonly("registers", [](System& s) {
            s.registers.resize(std::max(s.registers.size(), System::name_register_map.size()));
            return s.registers;
        })
        .def("__less__", &System::operator<)
        .def("__lt__", &System::operator<)
        .def("__eq__", &System::operator==)
        .def("__ne__", &System::operator!=)
        .def("__str__", [](const System& s) { return s.to_string(); })
        .def("to_string", [](const System& s) { return s.to_string(); })
        .def_static("clear", &System::clear)
        .def_static("add_register", [](std::string_view n, StateStorageType t, size_t s) { return System::add_register(n, t, s); })
        .def_static("get_id", [](std::string_view n) { return System::get(n); })
        .def_static("size_of", [](py::object r) { return py::isinstance<py::str>(r) ? System::size_of(r.cast<std::string>()) : System::size_of(r.cast<size_t>()); })
        .def_static("type_of", [](py::object r) { return py::isinstance<py::str>(r) ? System::type_of(r.cast<std::string>()) : System::type_of(r.cast<size_t>()); })
        .def_static("status_of", [](py::object r) { return py::isinstance<py::str>(r) ? System::status_of(r.cast<std::string>()) : System::status_of(r.cast<size_t>()); })
        .def_static("name_of", [](size_t id) { return System::name_of(id); })
        .def_static("remove_register", [](py::object r) { System::remove_register(resolve(r)); })
        .def_static("get_qubit_count", &System::get_qubit_count)
        .def_static("get_activated_register_size", &System::get_activated_register_size)
        .def_static("get_register_info", &System::get_register_info)
        .def_static("update_max_size", &System::update_max_size)
        .def_static("max_system_size", []() { return System::max_system_size; })
        .def_static("max_qubit_count", []() { return System::max_qubit_count; })
        .def_static("max_register_count", []() { return System::max_register_count; })
        .def_static("reusable_registers", []() { return System::reusable_registers; })
        .def_static("register_count", []() { return System::name_register_map.size(); });

    py::class_<SparseState>(m, "SparseState")
        .def(py::init<>())
        .def(py::init<const SparseState&>(), py::arg("other"))
        .def("clone", [](const SparseState& s) { return SparseState(s); })
        .def("size", [](SparseState& s) { return s.size(); }, Release())
        // The register's value when every entry holds the same one, else None
        // (one device reduction; no counterpart in the reference).
        .def("uniform_value", [](SparseState& s, py::object reg) -> py::object {
            const size_t id = resolve(reg);
            int uni = 1;
            uint64_t v = 0;
            {
                py::gil_scoped_release r;
                const Layout& L = s.sync();
                s.flush();
                check(sparq_field_uniform(s.handle(), L.at(id), &uni, &v));
            }
            if (!uni) return py::none();
            return py::int_(v);
        }, py::arg("reg"))
        .def("empty", [](SparseState& s) { return s.empty(); }, Release())
        .def("key_words", [](SparseState& s) { return s.sync().words; })
        // basis_states (PySparQ/core.cpp:182): the entries as System objects,
        // downloaded from the device in the engine's entry order (read-only copy)
        .def_property_readonly("basis_states", [](SparseState& s) {
            std::vector<System> entries;
            {
                py::gil_scoped_release r;
                s.to_host(entries);
            }
            return entries;
        })
        .def("_handle", [](SparseState& s) { return reinterpret_cast<std::uintptr_t>(s.handle()); })
        .def("_cpp_ptr", [](SparseState& s) { return reinterpret_cast<std::uintptr_t>(&s); })
        .def("flush", [](SparseState& s) { s.flush(); }, Release())
        // export(): (active ids ascending, values[N, R] uint64, amps[N] complex128)
        .def("export", [](SparseState& s) {
            std::vector<size_t> ids;
            std::vector<uint64_t> vals;
            std::vector<complex_t> amps;
            {
                py::gil_scoped_release r;
                s.export_entries(ids, vals, amps);
            }
            size_t n = amps.size(), R = ids.size();
            py::array_t<uint64_t> v({n, R});
            if (n * R) std::memcpy(v.mutable_data(), vals.data(), n * R * 8);
            py::array_t<std::complex<double>> a(n);
            if (n) std::memcpy(a.mutable_data(), amps.data(), n * 16);
            return py::make_tuple(ids, v, a);
        })
        .def("load", [](SparseState& s, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> vals,
                        py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> amps) {
            std::vector<uint64_t> v(vals.data(), vals.data() + vals.size());
            std::vector<complex_t> a(amps.data(), amps.data() + amps.size());
            py::gil_scoped_release r;
            s.import_entries(v, a);
        })
        .def("post_select_dense", [](SparseState& s, py::object zero, py::object index, uint64_t length) {
            std::vector<size_t> z, ix;
            for (auto it : zero) z.push_back(resolve(it));
            for (auto it : index) ix.push_back(resolve(it));
            std::vector<complex_t> out;
            double w;
            {
                py::gil_scoped_release r;
                w = s.post_select_dense(z, ix, out, length);
            }
            py::array_t<std::complex<double>> a(out.size());
            if (!out.empty()) std::memcpy(a.mutable_data(), out.data(), out.size() * 16);
            return py::make_tuple(a, w);
        }, py::arg("zero_regs"), py::arg("index_regs"), py::arg("length") = 0)
        .def("to_string", [](SparseState& s, int precision) { return s.to_string(precision); }, py::arg("precision") = 6)
        .def("__str__", [](SparseState& s) { return s.to_string(6); })
        .def("__repr__", [](SparseState& s) { return "<SparseState entries=" + std::to_string(s.size()) + ">"; });

    py::class_<BaseOperator, std::shared_ptr<BaseOperator>>(m, "BaseOperator")
        .def("__call__", [](const BaseOperator& o, SparseState& s) { o(s); }, py::arg("state"), Release())
        .def("dag", [](const BaseOperator& o, SparseState& s) { o.dag(s); }, py::arg("state"), Release());
    py::class_<SelfAdjointOperator, BaseOperator, std::shared_ptr<SelfAdjointOperator>>(m, "SelfAdjointOperator");
    py::class_<OperatorSequence, BaseOperator, std::shared_ptr<OperatorSequence>>(m, "OperatorSequence")
        .def(py::init<std::vector<std::shared_ptr<BaseOperator>>>(), py::arg("ops"))
        .def("__len__", [](const OperatorSequence& q) { return q.ops.size(); });

    auto pwo = py::class_<PointwiseOperator, BaseOperator, std::shared_ptr<PointwiseOperator>>(m, "PointwiseOperator");
    pwo.def_readonly("name", &PointwiseOperator::name);
    bind_controllable(pwo);

#define SPQ_BIND3(NAME) pw_class<NAME>(m, #NAME).def(py::init<std::string_view, std::string_view, std::string_view>()).def(py::init<size_t, size_t, size_t>());
#define SPQ_BIND2(NAME) pw_class<NAME>(m, #NAME).def(py::init<std::string_view, std::string_view>()).def(py::init<size_t, size_t>());
#define SPQ_BIND4(NAME) pw_class<NAME>(m, #NAME).def(py::init<std::string_view, std::string_view, std::string_view, std::string_view>()).def(py::init<size_t, size_t, size_t, size_t>());
#define SPQ_BINDRCR(NAME) pw_class<NAME>(m, #NAME).def(py::init<std::string_view, uint64_t, std::string_view>()).def(py::init<size_t, uint64_t, size_t>());
#define SPQ_BINDRC(NAME) pw_class<NAME>(m, #NAME).def(py::init<std::string_view, uint64_t>()).def(py::init<size_t, uint64_t>());
    SPQ_BIND3(Add_UInt_UInt) SPQ_BIND3(Sub_UInt_UInt) SPQ_BIND3(Mul_UInt_UInt) SPQ_BIND3(Div_UInt_UInt)
    SPQ_BIND3(And_UInt_UInt) SPQ_BIND3(Or_UInt_UInt) SPQ_BIND3(Xor_UInt_UInt) SPQ_BIND3(Less_SInt_SInt)
    SPQ_BIND3(Less_UInt_UInt) SPQ_BIND3(GetMid_UInt_UInt) SPQ_BIND3(Div_Sqrt_Arccos_UInt_UInt)
    SPQ_BIND3(Sqrt_Div_Arccos_Int_UInt) SPQ_BIND3(GetRotateAngle_Int_Int)
    SPQ_BIND2(Add_UInt_UInt_InPlace) SPQ_BIND2(Neg_UInt) SPQ_BIND2(Abs_SInt) SPQ_BIND2(Sqrt_UInt) SPQ_BIND2(IsZero_UInt)
    SPQ_BIND2(Negative_SInt) SPQ_BIND2(Add_AnyInt_AnyInt_InPlace) SPQ_BIND2(Assign) SPQ_BIND2(Swap_General_General)
    SPQ_BIND4(Select_Bool_UInt_UInt) SPQ_BIND4(Carry_UInt_UInt) SPQ_BIND4(Overflow_SInt_SInt)
    SPQ_BIND4(MulOverflow_UInt_UInt) SPQ_BIND4(Compare_UInt_UInt)
    SPQ_BINDRCR(Mult_UInt_ConstUInt) SPQ_BINDRCR(Add_Mult_UInt_ConstUInt_InPlace) SPQ_BINDRCR(Add_UInt_ConstUInt)
    SPQ_BINDRC(Add_ConstUInt_InPlace) SPQ_BINDRC(ShiftLeft_InPlace) SPQ_BINDRC(ShiftRight_InPlace) SPQ_BINDRC(Init_Unsafe)
    pw_class<X_Bool>(m, "X_Bool").def(py::init<std::string_view, uint64_t>(), py::arg("reg"), py::arg("digit") = 0)
        .def(py::init<size_t, uint64_t>(), py::arg("reg_id"), py::arg("digit") = 0);
    pw_class<Y_Bool>(m, "Y_Bool").def(py::init<std::string_view, uint64_t>(), py::arg("reg"), py::arg("digit") = 0)
        .def(py::init<size_t, uint64_t>(), py::arg("reg_id"), py::arg("digit") = 0);
    pw_class<FlipBools>(m, "FlipBools").def(py::init<std::string_view>()).def(py::init<size_t>());
    pw_class<Mod_Mult_UInt_ConstUInt_InPlace>(m, "Mod_Mult_UInt_ConstUInt_InPlace")
        .def(py::init<std::string_view, uint64_t, uint64_t, uint64_t>()).def(py::init<size_t, uint64_t, uint64_t, uint64_t>());
    pw_class<Swap_Bool_Bool>(m, "Swap_Bool_Bool")
        .def(py::init<std::string_view, size_t, std::string_view, size_t>()).def(py::init<size_t, size_t, size_t, size_t>());
    pw_class<Phase_Bool>(m, "Phase_Bool")
        .def(py::init<std::string_view, size_t, double>(), py::arg("reg"), py::arg("digit"), py::arg("lambda_"))
        .def(py::init<size_t, size_t, double>(), py::arg("reg_id"), py::arg("digit"), py::arg("lambda_"))
        .def(py::init([](std::string_view r, double l) { return std::make_shared<Phase_Bool>(r, 0, l); }), py::arg("reg"), py::arg("lambda_"));
    py::class_<Z_Bool, Phase_Bool, std::shared_ptr<Z_Bool>>(m, "Z_Bool")
        .def(py::init<std::string_view, size_t>(), py::arg("reg"), py::arg("digit") = 0)
        .def(py::init<size_t, size_t>(), py::arg("reg_id"), py::arg("digit") = 0);
    py::class_<S_Bool, Phase_Bool, std::shared_ptr<S_Bool>>(m, "S_Bool")
        .def(py::init<std::string_view, size_t>(), py::arg("reg"), py::arg("digit") = 0)
        .def(py::init<size_t, size_t>(), py::arg("reg_id"), py::arg("digit") = 0);
    py::class_<T_Bool, Phase_Bool, std::shared_ptr<T_Bool>>(m, "T_Bool")
        .def(py::init<std::string_view, size_t>(), py::arg("reg"), py::arg("digit") = 0)
        .def(py::init<size_t, size_t>(), py::arg("reg_id"), py::arg("digit") = 0);
    pw_class<RZ_Bool>(m, "RZ_Bool").def(py::init<std::string_view, size_t, double>()).def(py::init<size_t, size_t, double>());
    pw_class<ZeroConditionalPhaseFlip>(m, "ZeroConditionalPhaseFlip")
        .def(py::init<const std::vector<std::string>&>(), py::arg("regs"))
        .def(py::init<const std::vector<size_t>&>(), py::arg("reg_ids"));
    pw_class<Reflection_Bool>(m, "Reflection_Bool")
        .def(py::init<std::string_view, bool>(), py::arg("reg"), py::arg("inverse") = false)
        .def(py::init<const std::vector<std::string>&, bool>(), py::arg("regs"), py::arg("inverse") = false)
        .def(py::init<const std::vector<size_t>&, bool>(), py::arg("reg_ids"), py::arg("inverse") = false);
    pw_class<GlobalPhase>(m, "GlobalPhase").def(py::init<complex_t>(), py::arg("phase"));
    pw_class<PlusOneAndOverflow>(m, "PlusOneAndOverflow")
        .def(py::init<std::string_view, std::string_view>(), py::arg("main_reg"), py::arg("overflow"))
        .def(py::init<size_t, size_t>(), py::arg("main_reg"), py::arg("overflow"));
    auto rgu = py::class_<Rot_GeneralUnitary, BaseOperator, std::shared_ptr<Rot_GeneralUnitary>>(m, "Rot_GeneralUnitary")
        .def(py::init([](py::object r, std::vector<std::vector<complex_t>> rows) {
            std::vector<complex_t> flat;
            for (auto& row : rows) flat.insert(flat.end(), row.begin(), row.end());
            return std::make_shared<Rot_GeneralUnitary>(resolve(r), flat);
        }), py::arg("reg"), py::arg("unitary_matrix"));
    bind_controllable(rgu);
    py::class_<Rot_GeneralStatePrep, Rot_GeneralUnitary, std::shared_ptr<Rot_GeneralStatePrep>>(m, "Rot_GeneralStatePrep")
        .def(py::init([](py::object r, std::vector<complex_t> vec) { return std::make_shared<Rot_GeneralStatePrep>(resolve(r), vec); }),
             py::arg("reg"), py::arg("state"));
    m.def("stateprep_unitary", &stateprep_unitary);

    py::class_<QRAMCircuit_qutrit, std::shared_ptr<QRAMCircuit_qutrit>>(m, "QRAMCircuit_qutrit")
        .def(py::init<size_t, size_t, std::vector<uint64_t>>(), py::arg("addr_size"), py::arg("data_size"), py::arg("memory"))
        .def_readonly("address_size", &QRAMCircuit_qutrit::address_size)
        .def_readonly("data_size", &QRAMCircuit_qutrit::data_size);
    pw_class<QRAMLoad>(m, "QRAMLoad")
        .def(py::init<const QRAMCircuit_qutrit*, std::string_view, std::string_view>(), py::keep_alive<1, 2>())
        .def(py::init<const QRAMCircuit_qutrit*, size_t, size_t>(), py::keep_alive<1, 2>());
    m.attr("QRAMLoadFast") = m.attr("QRAMLoad");

    // captured sequences and the table-driven noise channel (no counterpart in the reference)
    py::class_<CapturedSequence>(m, "CapturedSequence")
        .def(py::init([](SparseState& st, py::object body) {
                 return std::make_unique<CapturedSequence>(
                     st, [&body](SparseState& s) { body(py::cast(&s, py::return_value_policy::reference)); });
             }),
             py::arg("state"), py::arg("body"), py::keep_alive<1, 2>())
        .def("launch", &CapturedSequence::launch, py::arg("times") = 1, Release())
        .def("sync", &CapturedSequence::sync, Release())
        .def_property_readonly("capacity", &CapturedSequence::capacity)
        .def_property_readonly("nodes", &CapturedSequence::nodes);
    py::class_<PauliNoiseTable, std::shared_ptr<PauliNoiseTable>>(m, "PauliNoiseTable")
        .def(py::init<size_t, size_t, const std::vector<uint64_t>&, const std::vector<uint64_t>&>(), py::arg("rows"),
             py::arg("iterations"), py::arg("x_masks"), py::arg("z_masks"))
        .def("set_iteration", &PauliNoiseTable::set_iteration, py::arg("k"), Release())
        .def_readonly("rows", &PauliNoiseTable::rows)
        .def_readonly("iterations", &PauliNoiseTable::iterations);
    pw_class<PauliNoise>(m, "PauliNoise")
        .def(py::init<std::string_view, std::string_view, const PauliNoiseTable&>(), py::arg("segment"), py::arg("target"),
             py::arg("table"))
        .def(py::init<size_t, size_t, const PauliNoiseTable&>(), py::arg("segment"), py::arg("target"), py::arg("table"));
    py::class_<AdvancePauliNoise, BaseOperator, std::shared_ptr<AdvancePauliNoise>>(m, "AdvancePauliNoise")
        .def(py::init<const PauliNoiseTable&>(), py::arg("table"));

    // branching
    auto hi = py::class_<Hadamard_Int, SelfAdjointOperator, std::shared_ptr<Hadamard_Int>>(m, "Hadamard_Int")
        .def(py::init<std::string_view, size_t>(), py::arg("reg_in"), py::arg("n_digits"))
        .def(py::init<size_t, size_t>(), py::arg("reg_in"), py::arg("n_digits"));
    bind_controllable(hi);
    auto hf = py::class_<Hadamard_Int_Full, SelfAdjointOperator, std::shared_ptr<Hadamard_Int_Full>>(m, "Hadamard_Int_Full")
        .def(py::init<std::string_view>(), py::arg("reg_in")).def(py::init<size_t>(), py::arg("reg_in"));
    bind_controllable(hf);
    auto hb = py::class_<Hadamard_Bool, SelfAdjointOperator, std::shared_ptr<Hadamard_Bool>>(m, "Hadamard_Bool")
        .def(py::init<std::string_view>(), py::arg("reg_in")).def(py::init<size_t>(), py::arg("reg_in"));
    bind_controllable(hb);
    auto crb = py::class_<CondRot_Rational_Bool, BaseOperator, std::shared_ptr<CondRot_Rational_Bool>>(m, "CondRot_Rational_Bool")
        .def(py::init<std::string_view, std::string_view>()).def(py::init<size_t, size_t>());
    bind_controllable(crb);
    py::class_<CondRot_Fixed_Bool, CondRot_Rational_Bool, std::shared_ptr<CondRot_Fixed_Bool>>(m, "CondRot_Fixed_Bool")
        .def(py::init<std::string_view, std::string_view>()).def(py::init<size_t, size_t>());
    py::class_<CondRot_General_Bool, BaseOperator, std::shared_ptr<CondRot_General_Bool>>(m, "CondRot_General_Bool")
        .def(py::init([](py::object in, py::object out, py::function f) {
            std::function<u22_t(uint64_t)> fn = [f](uint64_t v) { py::gil_scoped_acquire g; return to_u22(f(v)); };
            return std::make_shared<CondRot_General_Bool>(resolve(in), resolve(out), fn);
        }), py::arg("reg_in"), py::arg("reg_out"), py::arg("angle_function"));
    m.def("make_func", [](uint64_t v, size_t n) { auto u = make_func(v, n); return std::vector<complex_t>(u.begin(), u.end()); });
    m.def("make_func_inv", [](uint64_t v, size_t n) { auto u = make_func_inv(v, n); return std::vector<complex_t>(u.begin(), u.end()); });
    auto rb = py::class_<Rot_Bool, BaseOperator, std::shared_ptr<Rot_Bool>>(m, "Rot_Bool")
        .def(py::init([](py::object r, size_t d, py::object mat) { return std::make_shared<Rot_Bool>(resolve(r), d, to_u22(mat)); }),
             py::arg("reg"), py::arg("digit"), py::arg("matrix"));
    bind_controllable(rb);
    py::class_<RX_Bool, Rot_Bool, std::shared_ptr<RX_Bool>>(m, "RX_Bool")
        .def(py::init<std::string_view, size_t, double>()).def(py::init<size_t, size_t, double>());
    py::class_<RY_Bool, Rot_Bool, std::shared_ptr<RY_Bool>>(m, "RY_Bool")
        .def(py::init<std::string_view, size_t, double>()).def(py::init<size_t, size_t, double>());
    py::class_<SX_Bool, Rot_Bool, std::shared_ptr<SX_Bool>>(m, "SX_Bool")
        .def(py::init<std::string_view, size_t>(), py::arg("reg"), py::arg("digit") = 0)
        .def(py::init<size_t, size_t>(), py::arg("reg_id"), py::arg("digit") = 0);
    py::class_<U2_Bool, Rot_Bool, std::shared_ptr<U2_Bool>>(m, "U2_Bool")
        .def(py::init<std::string_view, size_t, double, double>()).def(py::init<size_t, size_t, double, double>());
    py::class_<U3_Bool, Rot_Bool, std::shared_ptr<U3_Bool>>(m, "U3_Bool")
        .def(py::init<std::string_view, size_t, double, double, double>()).def(py::init<size_t, size_t, double, double, double>());

    // structural
    py::class_<ClearZero, SelfAdjointOperator, std::shared_ptr<ClearZero>>(m, "ClearZero")
        .def(py::init<>()).def(py::init<double>(), py::arg("eps"));
    m.def("set_clearzero_default_eps", [](double e) { clearzero_default_eps() = e; });
    m.def("get_clearzero_default_eps", []() { return clearzero_default_eps(); });
    py::class_<Normalize, SelfAdjointOperator, std::shared_ptr<Normalize>>(m, "Normalize").def(py::init<>());
    py::class_<SortUnconditional, SelfAdjointOperator, std::shared_ptr<SortUnconditional>>(m, "SortUnconditional").def(py::init<>());
    py::class_<CheckDuplicateKey, SelfAdjointOperator, std::shared_ptr<CheckDuplicateKey>>(m, "CheckDuplicateKey").def(py::init<>());
    // grouping sorts (PySparQ/core.cpp:1249-1273)
    py::class_<SortExceptKey, SelfAdjointOperator, std::shared_ptr<SortExceptKey>>(m, "SortExceptKey")
        .def(py::init<std::string_view>(), py::arg("key")).def(py::init<size_t>(), py::arg("key_id"));
    py::class_<SortByKey, SelfAdjointOperator, std::shared_ptr<SortByKey>>(m, "SortByKey")
        .def(py::init<std::string_view>(), py::arg("key")).def(py::init<size_t>(), py::arg("key_id"));
    py::class_<SortExceptBit, SelfAdjointOperator, std::shared_ptr<SortExceptBit>>(m, "SortExceptBit")
        .def(py::init<std::string_view, size_t>(), py::arg("key"), py::arg("digit"))
        .def(py::init<size_t, size_t>(), py::arg("key_id"), py::arg("digit"));
    py::class_<SortExceptKeyHadamard, SelfAdjointOperator, std::shared_ptr<SortExceptKeyHadamard>>(m, "SortExceptKeyHadamard")
        .def(py::init([](py::object key, py::iterable q) {
            std::vector<size_t> v;
            for (auto it : q) v.push_back(it.cast<size_t>());
            return std::make_shared<SortExceptKeyHadamard>(resolve(key), v);
        }), py::arg("key"), py::arg("qubit_ids"));
    py::class_<SortByAmplitude, SelfAdjointOperator, std::shared_ptr<SortByAmplitude>>(m, "SortByAmplitude").def(py::init<>());
    py::class_<SortByKey2, SelfAdjointOperator, std::shared_ptr<SortByKey2>>(m, "SortByKey2")
        .def(py::init<std::string_view, std::string_view>(), py::arg("key1"), py::arg("key2"))
        .def(py::init<size_t, size_t>(), py::arg("key1"), py::arg("key2"));
    auto hp = py::class_<Hadamard_Partial, SelfAdjointOperator, std::shared_ptr<Hadamard_Partial>>(m, "Hadamard_Partial")
        .def(py::init([](py::object reg, py::iterable q) {
            std::vector<size_t> v;
            for (auto it : q) v.push_back(it.cast<size_t>());
            return std::make_shared<Hadamard_Partial>(resolve(reg), v);
        }), py::arg("reg_in"), py::arg("qubit_positions"));
    bind_controllable(hp);
    py::class_<MoveBackRegister>(m, "MoveBackRegister")
        .def(py::init<std::string_view>(), py::arg("reg")).def(py::init<size_t>(), py::arg("reg_id"))
        .def("__call__", [](const MoveBackRegister& a, SparseState& s) { a(s); }, Release());
    py::class_<Push, BaseOperator, std::shared_ptr<Push>>(m, "Push")
        .def(py::init<std::string_view, std::string_view>(), py::arg("reg"), py::arg("garbage"))
        .def(py::init<size_t, std::string_view>(), py::arg("reg_id"), py::arg("garbage"));
    py::class_<Pop, BaseOperator, std::shared_ptr<Pop>>(m, "Pop")
        .def(py::init<std::string_view>(), py::arg("reg")).def(py::init<size_t>(), py::arg("reg_id"));
    m.def("split_systems", [](SparseState& s, py::iterable nz, py::iterable ones, py::iterable by_bit, py::iterable by_val) {
        std::vector<size_t> a, b;
        std::vector<std::pair<size_t, size_t>> c;
        std::vector<std::pair<size_t, uint64_t>> d;
        for (auto it : nz) a.push_back(resolve(it));
        for (auto it : ones) b.push_back(resolve(it));
        for (auto it : by_bit) { auto t = it.cast<py::tuple>(); c.emplace_back(resolve(t[0]), t[1].cast<size_t>()); }
        for (auto it : by_val) { auto t = it.cast<py::tuple>(); d.emplace_back(resolve(t[0]), t[1].cast<uint64_t>()); }
        py::gil_scoped_release r;
        return split_systems(s, a, b, c, d);
    }, py::arg("state"), py::arg("condition_variable_nonzeros"), py::arg("condition_variable_all_ones"),
       py::arg("condition_variable_by_bit"), py::arg("condition_variable_by_value"));
    m.def("combine_systems", [](SparseState& to, SparseState& from) { combine_systems(to, from); }, py::arg("to"),
          py::arg("from_"), Release());
    m.def("merge_system", &merge_system);
    m.def("remove_system", &remove_system);
    py::class_<TestRemovable, SelfAdjointOperator, std::shared_ptr<TestRemovable>>(m, "TestRemovable")
        .def(py::init<std::string_view>()).def(py::init<size_t>());
    py::class_<AddRegister>(m, "AddRegister")
        .def(py::init<std::string_view, StateStorageType, size_t>())
        .def("__call__", [](const AddRegister& a, SparseState& s) { return a(s); }, Release());
    py::class_<AddRegisterWithHadamard>(m, "AddRegisterWithHadamard")
        .def(py::init<std::string_view, StateStorageType, size_t>())
        .def("__call__", [](const AddRegisterWithHadamard& a, SparseState& s) { return a(s); }, Release());
    py::class_<RemoveRegister>(m, "RemoveRegister")
        .def(py::init<std::string_view>()).def(py::init<size_t>())
        .def("__call__", [](const RemoveRegister& a, SparseState& s) { a(s); }, Release());
    py::class_<SplitRegister>(m, "SplitRegister")
        .def(py::init<std::string_view, std::string_view, size_t>()).def(py::init<size_t, std::string_view, size_t>())
        .def("__call__", [](const SplitRegister& a, SparseState& s) { return a(s); }, Release());
    py::class_<CombineRegister>(m, "CombineRegister")
        .def(py::init<std::string_view, std::string_view>()).def(py::init<size_t, size_t>())
        .def("__call__", [](const CombineRegister& a, SparseState& s) { return a(s); }, Release());
    m.def("reset_systems", [](SparseState& s) { reset_systems(s); }, Release());
    // eps: merge threshold on |sum|^2 (the reference always uses the library epsilon, 1e-14)
    m.def("add_systems", [](SparseState& cur, SparseState& add, double coef, double eps) { add_systems(cur, add, coef, eps); },
          py::arg("cur"), py::arg("add"), py::arg("coef"), py::arg("eps") = epsilon, Release());

    // debug checks and post-selection (debugger.h:78-160, partial_trace.h:83-145, PySparQ/core.cpp:249-402)
    py::class_<CheckNormalization, SelfAdjointOperator, std::shared_ptr<CheckNormalization>>(m, "CheckNormalization")
        .def(py::init<>()).def(py::init<double>(), py::arg("threshold"));
    py::class_<CheckNan, SelfAdjointOperator, std::shared_ptr<CheckNan>>(m, "CheckNan").def(py::init<>());
    m.def("set_debug_checks", [](bool on) { debug_checks() = on; }, py::arg("on"),
          "Enable CheckNormalization / CheckNan (off by default, as in the release build).");
    m.def("get_debug_checks", []() { return debug_checks(); });
    py::class_<PartialTraceSelect>(m, "PartialTraceSelect")
        .def(py::init<const std::map<std::string, uint64_t>&>(), py::arg("name_value_map"))
        .def(py::init<const std::map<size_t, uint64_t>&>(), py::arg("id_value_map"))
        .def(py::init<const std::vector<size_t>&, const std::vector<uint64_t>&>(), py::arg("reg_ids"), py::arg("select_values"))
        .def(py::init<const std::vector<std::string>&, const std::vector<uint64_t>&>(), py::arg("reg_names"), py::arg("select_values"))
        .def("__call__", [](const PartialTraceSelect& p, SparseState& s) { return p(s); }, py::arg("state"), Release());
    py::class_<PartialTraceSelectRange>(m, "PartialTraceSelectRange")
        .def(py::init<std::string_view, std::pair<size_t, size_t>>(), py::arg("register_name"), py::arg("select_range"))
        .def(py::init<size_t, std::pair<size_t, size_t>>(), py::arg("register_id"), py::arg("select_range"))
        .def("__call__", [](const PartialTraceSelectRange& p, SparseState& s) { return p(s); }, py::arg("state"), Release());

    // measurement with the seedable engine (PySparQ/core.cpp:404-505)
    m.def("set_seed", [](long long seed) { random_engine::set_seed((uint64_t)seed); }, py::arg("seed"));
    m.def("get_seed", []() { return (long long)random_engine::get_seed(); });
    py::class_<MeasureZ>(m, "MeasureZ")
        .def(py::init<const std::vector<std::string>&>(), py::arg("register_names"))
        .def(py::init<const std::vector<size_t>&>(), py::arg("register_ids"))
        .def(py::init<std::string_view>(), py::arg("register_name"))
        .def(py::init<size_t>(), py::arg("register_id"))
        .def_readonly("registers", &MeasureZ::registers)
        .def("__call__", [](const MeasureZ& mz, SparseState& s) { return mz(s); }, py::arg("state"), Release());
    py::class_<Reset>(m, "Reset")
        .def(py::init<const std::vector<std::string>&>(), py::arg("register_names"))
        .def(py::init<const std::vector<std::string>&, const std::vector<uint64_t>&>(), py::arg("register_names"), py::arg("targets"))
        .def(py::init<const std::vector<size_t>&>(), py::arg("register_ids"))
        .def(py::init<const std::vector<size_t>&, const std::vector<uint64_t>&>(), py::arg("register_ids"), py::arg("targets"))
        .def(py::init<std::string_view, uint64_t>(), py::arg("register_name"), py::arg("target") = 0)
        .def(py::init<size_t, uint64_t>(), py::arg("register_id"), py::arg("target") = 0)
        .def_readonly("registers", &Reset::registers)
        .def_readonly("target_values", &Reset::target_values)
        .def("__call__", [](const Reset& r, SparseState& s) { return r(s); }, py::arg("state"), Release());
    py::class_<Probability>(m, "Probability")
        .def(py::init<const std::map<std::string, uint64_t>&>(), py::arg("name_value_map"))
        .def(py::init<const std::map<size_t, uint64_t>&>(), py::arg("id_value_map"))
        .def(py::init<const std::vector<std::string>&, const std::vector<uint64_t>&>(), py::arg("register_names"),
             py::arg("target_values"))
        .def(py::init<const std::vector<size_t>&, const std::vector<uint64_t>&>(), py::arg("register_ids"), py::arg("target_values"))
        .def(py::init<std::string_view, uint64_t>(), py::arg("register_name"), py::arg("value"))
        .def(py::init<size_t, uint64_t>(), py::arg("register_id"), py::arg("value"))
        .def_readonly("registers", &Probability::registers)
        .def_readonly("values", &Probability::values)
        .def("__call__", [](const Probability& p, SparseState& s) { return p(s); }, py::arg("state"), Release())
        .def_static("distribution", [](SparseState& s, py::object r) {
            return py::isinstance<py::str>(r) ? Probability::distribution(s, r.cast<std::string>())
                                              : Probability::distribution(s, r.cast<size_t>());
        }, py::arg("state"), py::arg("register"));

    // CKS quantum-walk primitives (PySparQ/core.cpp:53-65, 1472-1510)
    py::class_<SparseMatrix, std::shared_ptr<SparseMatrix>>(m, "SparseMatrix")
        .def(py::init<>())
        .def(py::init<const std::vector<double>&, const std::vector<size_t>&, size_t, size_t, size_t, bool>(), py::arg("elements"),
             py::arg("sparsity"), py::arg("data_size"), py::arg("nnz_col"), py::arg("n_row"), py::arg("positive_only"))
        .def_static("from_encoded", [](const std::vector<uint64_t>& enc, const std::vector<size_t>& sp, size_t ds, size_t nnz,
                                       size_t nrow, bool pos) { return std::make_shared<SparseMatrix>(enc, sp, ds, nnz, nrow, pos); },
                    py::arg("elements"), py::arg("sparsity"), py::arg("data_size"), py::arg("nnz_col"), py::arg("n_row"),
                    py::arg("positive_only") = true)
        .def("get_data", &SparseMatrix::get_data)
        .def("get_sparsity_offset", &SparseMatrix::get_sparsity_offset)
        .def_readonly("elements", &SparseMatrix::elements)
        .def_readonly("sparsity", &SparseMatrix::sparsity)
        .def_readonly("positive_only", &SparseMatrix::positive_only)
        .def_readonly("data_size", &SparseMatrix::data_size)
        .def_readonly("nnz_col", &SparseMatrix::nnz_col)
        .def_readonly("n_row", &SparseMatrix::n_row);
    pw_class<GetQWRotateAngle_Int_Int_Int>(m, "GetQWRotateAngle_Int_Int_Int")
        .def(py::init<std::string_view, std::string_view, std::string_view, std::string_view, const SparseMatrix*>(), py::arg("data"),
             py::arg("row"), py::arg("col"), py::arg("out"), py::arg("mat"))
        .def(py::init<size_t, size_t, size_t, size_t, const SparseMatrix*>(), py::arg("data"), py::arg("row"), py::arg("col"),
             py::arg("out"), py::arg("mat"));
    pw_class<QuantumBinarySearch_Fast>(m, "QuantumBinarySearch_Fast")
        .def(py::init<const QRAMCircuit_qutrit*, std::string_view, size_t, std::string_view, std::string_view>(), py::arg("qram"),
             py::arg("address_offset_register"), py::arg("total_length"), py::arg("target_register"), py::arg("result_register"),
             py::keep_alive<1, 2>())
        .def(py::init<const QRAMCircuit_qutrit*, size_t, size_t, size_t, size_t>(), py::arg("qram"),
             py::arg("address_offset_register"), py::arg("total_length"), py::arg("target_register"), py::arg("result_register"),
             py::keep_alive<1, 2>());
    pw_class<GetRowAddr>(m, "GetRowAddr")
        .def(py::init<std::string_view, std::string_view, size_t, std::string_view>(), py::arg("reg_offset"), py::arg("reg_row"),
             py::arg("row_size"), py::arg("reg_row_offset"))
        .def(py::init<size_t, size_t, size_t, size_t>(), py::arg("reg_offset"), py::arg("reg_row"), py::arg("row_size"),
             py::arg("reg_row_offset"));
    pw_class<GetDataAddr>(m, "GetDataAddr")
        .def(py::init<std::string_view, std::string_view, std::string_view, size_t, std::string_view>(), py::arg("reg_offset"),
             py::arg("reg_row"), py::arg("reg_col_sparse"), py::arg("row_size"), py::arg("reg_data_offset"))
        .def(py::init<size_t, size_t, size_t, size_t, size_t>(), py::arg("reg_offset"), py::arg("reg_row"), py::arg("reg_col_sparse"),
             py::arg("row_size"), py::arg("reg_data_offset"));
    py::class_<TridiagonalSpec>(m, "TridiagonalSpec")
        .def_readonly("n_row", &TridiagonalSpec::n_row)
        .def_readonly("data_size", &TridiagonalSpec::data_size)
        .def_readonly("alpha_enc", &TridiagonalSpec::alpha_enc)
        .def_readonly("beta_enc", &TridiagonalSpec::beta_enc)
        .def_readonly("sparse_offset", &TridiagonalSpec::sparse_offset)
        .def_readonly("addr_mask", &TridiagonalSpec::addr_mask)
        .def_readonly("k_width", &TridiagonalSpec::k_width)
        .def_property_readonly_static("slot_count", [](py::object) { return TridiagonalSpec::slot_count; })
        .def("window_base", &TridiagonalSpec::window_base)
        .def("col_of", &TridiagonalSpec::col_of)
        .def("value_of", &TridiagonalSpec::value_of)
        .def("combined_mem", &TridiagonalSpec::combined_mem);
    m.def("make_tridiagonal_matrix_spec", [](double alpha, double beta, size_t n, size_t ds) {
        auto [mat, spec] = make_tridiagonal_matrix_spec(alpha, beta, n, ds);
        return py::make_tuple(std::make_shared<SparseMatrix>(std::move(mat)), spec);
    }, py::arg("alpha"), py::arg("beta"), py::arg("n"), py::arg("ds"));
    pw_class<TridiagOracle1>(m, "TridiagOracle1")
        .def(py::init<std::string_view, std::string_view, std::string_view, std::string_view, const TridiagonalSpec&>(),
             py::arg("reg_offset"), py::arg("reg_row"), py::arg("reg_slot"), py::arg("reg_out"), py::arg("spec"))
        .def(py::init<size_t, size_t, size_t, size_t, const TridiagonalSpec&>(), py::arg("reg_offset"), py::arg("reg_row"),
             py::arg("reg_slot"), py::arg("reg_out"), py::arg("spec"));
    pw_class<TridiagOracle2>(m, "TridiagOracle2")
        .def(py::init<std::string_view, std::string_view, std::string_view, std::string_view, const TridiagonalSpec&>(),
             py::arg("reg_offset"), py::arg("reg_row"), py::arg("reg_k"), py::arg("reg_sr"), py::arg("spec"))
        .def(py::init<size_t, size_t, size_t, size_t, const TridiagonalSpec&>(), py::arg("reg_offset"), py::arg("reg_row"),
             py::arg("reg_k"), py::arg("reg_sr"), py::arg("spec"));
    pw_class<TridiagParityPhase>(m, "TridiagParityPhase").def(py::init<std::string_view>(), py::arg("reg")).def(py::init<size_t>());

    m.def("device_name", []() {
        char buf[256] = {0};
        check(sparq_device_name(buf, sizeof(buf)));
        return std::string(buf);
    });
    m.def("memory_stats", []() {
        uint64_t v[6] = {0};
        check(sparq_memory_stats(v));
        py::dict d;
        d["live_bytes"] = v[0];
        d["cached_bytes"] = v[1];
        d["peak_bytes"] = v[2];
        d["hits"] = v[3];
        d["misses"] = v[4];
        d["cached_blocks"] = v[5];
        return d;
    });
    m.def("release_cached_memory", []() { check(sparq_release_cached_memory()); });
    m.def("_take_plugin_error", []() { return take_plugin_error(); });
    m.def(
        "sort_stats",
        [](bool reset) {
            uint64_t v[6] = {0, 0, 0, 0, 0, 0};
            check(sparq_sort_stats(v, 6, reset ? 1 : 0));
            py::dict d;
            d["local_sorts"] = v[0];
            d["local_entries"] = v[1];
            d["local_overflows"] = v[2];
            d["lsd_sorts"] = v[3];
            d["lsd_entries"] = v[4];
            d["presorted_sorts"] = v[5];
            return d;
        },
        py::arg("reset") = false, "Sorts by path since the last reset: prefix + local chunk sort, overflowed local attempts, LSD passes, chunk sorts of states already sorted by their prefix.");
    m.def("set_dense_transform", [](bool on) { check(sparq_set_dense_transform(on ? 1 : 0)); }, py::arg("on"),
          "Hadamard_Int on a small, well filled key domain through a dense amplitude array (default on).");
    m.def("set_canonical_output", [](bool on) { check(sparq_set_canonical_output(on ? 1 : 0)); }, py::arg("on"),
          "Hadamard outputs in canonical (target-major) order where every other varying bit lies below the target "
          "(default on).");
    m.def("set_group_pass_through", [](bool on) { check(sparq_set_group_pass_through(on ? 1 : 0)); }, py::arg("on"),
          "Conditioned group unitaries pass failing groups through inside the transform (on) or partition the state "
          "first (off, the default).");
    m.def("set_sort_mode", [](int mode) { check(sparq_set_sort_mode(mode)); }, py::arg("mode"),
          "Sort strategy: 0 prefix passes + LDS chunk sort where it applies (default), 1 LSD passes only.");
    m.def("set_radix_bits", [](int bits) { check(sparq_set_radix_bits(bits)); }, py::arg("bits"),
          "Digit width of the radix sort: 8 or 11 (forced) or 0 (autotuned, the default).");
    m.def("set_pointwise_items", [](int items) { check(sparq_set_pointwise_items(items)); }, py::arg("items"),
          "Entries per thread of the fused pointwise kernel: 1, 2 or 4 (forced) or 0 (autotuned, the default).");
    m.def("autotune_report", []() {
        size_t need = 0;
        check(sparq_autotune_report(nullptr, 0, &need));
        std::string buf(need, '\0');
        check(sparq_autotune_report(&buf[0], need, &need));
        buf.resize(need ? need - 1 : 0);
        return py::module_::import("json").attr("loads")(buf);
    }, "The autotuners' measurements and choices: pointwise launch shape per class, radix digit width per size class.");
}
