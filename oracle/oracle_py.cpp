// Python binding of the CPU oracle (module `_oracle`).  TEST INFRASTRUCTURE
// ONLY: loaded by tests/, the smoke check and bench.py's CPU-baseline leg.
// Names and argument meaning follow the reference pysparq surface
// (PySparQ/core.cpp, PySparQ/BindUtils.h:117-207) so that a circuit can be
// written once and run on both this oracle and the product module.
#include <pybind11/complex.h>
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <chrono>

#include "sparq_oracle_meas.hpp"

namespace py = pybind11;
using namespace oracle;

static size_t rid(const py::handle& h) {
    if (py::isinstance<py::str>(h)) {
        size_t id = Registry::get(h.cast<std::string>());
        if (id == SIZE_MAX) throw std::invalid_argument("unknown register " + h.cast<std::string>());
        return id;
    }
    return h.cast<size_t>();
}
static std::vector<size_t> rids(const py::handle& h) {
    std::vector<size_t> out;
    if (py::isinstance<py::str>(h) || py::isinstance<py::int_>(h)) { out.push_back(rid(h)); return out; }
    for (auto item : h) out.push_back(rid(item));
    return out;
}
static u22_t to_u22(const py::handle& m) {
    auto v = m.cast<std::vector<complex_t>>();
    if (v.size() != 4) throw std::invalid_argument("2x2 matrix must have 4 entries");
    return u22_t{v[0], v[1], v[2], v[3]};
}

struct PyState { SparseState s; };

// Self-adjoint state-level operator built from a callable.
struct FnOp {
    std::function<void(State&)> f;
    void operator()(State& st) const { f(st); }
    void dag(State& st) const { f(st); }
};
static OpPtr fnop(std::function<void(State&)> f, const char* n) { return wrap(FnOp{std::move(f)}, n); }

// The OpenMP build (-DORACLE_OMP) is the same module under the name _oracle_omp.
#ifdef ORACLE_OMP
PYBIND11_MODULE(_oracle_omp, m) {
#else
PYBIND11_MODULE(_oracle, m) {
#endif
    m.doc() = "CPU oracle of the SparQ sparse-state operator-apply path (test infrastructure)";

    py::enum_<StateStorageType>(m, "StateStorageType")
        .value("General", General).value("UnsignedInteger", UnsignedInteger)
        .value("SignedInteger", SignedInteger).value("Boolean", Boolean).value("Rational", Rational)
        .export_values();

    py::class_<Registry>(m, "System")
        .def_static("clear", &Registry::clear)
        .def_static("add_register", [](const std::string& n, StateStorageType t, size_t s) { return Registry::add_register(n, t, s); })
        .def_static("get_id", [](const std::string& n) { return Registry::get(n); })
        .def_static("size_of", [](py::object r) { return Registry::size_of(rid(r)); })
        .def_static("type_of", [](py::object r) { return Registry::type_of(rid(r)); })
        .def_static("status_of", [](py::object r) { return Registry::status_of(rid(r)); })
        .def_static("name_of", [](size_t id) { return Registry::name_of(id); })
        .def_static("remove_register", [](py::object r) { Registry::remove_register(rid(r)); })
        .def_static("get_qubit_count", &Registry::get_qubit_count)
        .def_static("get_activated_register_size", &Registry::get_activated_register_size)
        .def_static("max_system_size", []() { return Registry::max_system_size(); })
        .def_static("max_qubit_count", []() { return Registry::max_qubit_count(); })
        .def_static("max_register_count", []() { return Registry::max_register_count(); })
        .def_static("reusable_registers", []() { return Registry::reusable(); })
        .def_static("register_count", []() { return Registry::map().size(); });

    py::class_<PyState>(m, "SparseState")
        .def(py::init<>())
        .def("size", [](const PyState& p) { return p.s.size(); })
        .def("clone", [](const PyState& p) { return PyState{p.s}; })
        // export(): (active ids ascending, values[N, R] uint64, amps[N] complex128)
        .def("export", [](const PyState& p) {
            std::vector<size_t> ids;
            for (size_t i = 0; i < Registry::map().size(); ++i) if (Registry::status_of(i)) ids.push_back(i);
            size_t n = p.s.size(), r = ids.size();
            py::array_t<uint64_t> vals({n, r});
            py::array_t<std::complex<double>> amps(n);
            auto v = vals.mutable_unchecked<2>();
            auto a = amps.mutable_unchecked<1>();
            for (size_t i = 0; i < n; ++i) {
                for (size_t j = 0; j < r; ++j) v(i, j) = p.s.basis_states[i].get(ids[j]);
                a(i) = p.s.basis_states[i].amplitude;
            }
            return py::make_tuple(ids, vals, amps);
        })
        // load(values[N, R] for the active ids ascending, amps[N])
        .def("load", [](PyState& p, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> vals,
                        py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> amps) {
            std::vector<size_t> ids;
            for (size_t i = 0; i < Registry::map().size(); ++i) if (Registry::status_of(i)) ids.push_back(i);
            auto v = vals.unchecked<2>();
            auto a = amps.unchecked<1>();
            if ((size_t)v.shape(1) != ids.size() || v.shape(0) != a.shape(0)) throw std::invalid_argument("shape mismatch");
            p.s.basis_states.assign(v.shape(0), System());
            for (py::ssize_t i = 0; i < v.shape(0); ++i) {
                for (size_t j = 0; j < ids.size(); ++j) p.s.basis_states[i].get(ids[j]) = v(i, j);
                p.s.basis_states[i].amplitude = a(i);
            }
        });

    py::class_<QRAMCircuit, std::shared_ptr<QRAMCircuit>>(m, "QRAMCircuit_qutrit")
        .def(py::init([](size_t a, size_t d, std::vector<uint64_t> mem) {
            if (mem.size() < pow2(a)) throw std::invalid_argument("memory smaller than 2^address_size");
            return std::make_shared<QRAMCircuit>(QRAMCircuit{a, d, std::move(mem)});
        }))
        .def_readonly("address_size", &QRAMCircuit::address_size)
        .def_readonly("data_size", &QRAMCircuit::data_size);

    py::class_<Op, OpPtr>(m, "Operator")
        .def("__call__", [](const Op& o, PyState& p) { o.apply(p.s.basis_states); })
        .def("dag", [](const Op& o, PyState& p) { o.apply_dag(p.s.basis_states); })
        .def_readonly("name", &Op::name)
        .def("conditioned_by_nonzeros", [](OpPtr o, py::object c) { o->cond.nonzeros = rids(c); return o; })
        .def("conditioned_by_all_ones", [](OpPtr o, py::object c) { o->cond.all_ones = rids(c); return o; })
        .def("conditioned_by_bit", [](OpPtr o, py::object c, py::object pos) {
            o->cond.by_bit.clear();
            if (pos.is_none()) for (auto item : c) { auto t = item.cast<py::tuple>(); o->cond.by_bit.emplace_back(rid(t[0]), t[1].cast<size_t>()); }
            else o->cond.by_bit.emplace_back(rid(c), pos.cast<size_t>());
            return o;
        }, py::arg("cond"), py::arg("pos") = py::none())
        .def("conditioned_by_value", [](OpPtr o, py::object c, py::object val) {
            o->cond.by_value.clear();
            if (val.is_none()) for (auto item : c) { auto t = item.cast<py::tuple>(); o->cond.by_value.emplace_back(rid(t[0]), t[1].cast<uint64_t>()); }
            else o->cond.by_value.emplace_back(rid(c), val.cast<uint64_t>());
            return o;
        }, py::arg("cond"), py::arg("value") = py::none())
        .def("clear_control_nonzeros", [](OpPtr o) { o->cond.nonzeros.clear(); })
        .def("clear_control_all_ones", [](OpPtr o) { o->cond.all_ones.clear(); })
        .def("clear_control_by_bit", [](OpPtr o) { o->cond.by_bit.clear(); })
        .def("clear_control_by_value", [](OpPtr o) { o->cond.by_value.clear(); });

    // ---- branching
    m.def("Hadamard_Int", [](py::object r, size_t n) { return wrap(Hadamard_Int(rid(r), n), "Hadamard_Int"); });
    m.def("Hadamard_Int_Full", [](py::object r) { return wrap(Hadamard_Int_Full(rid(r)), "Hadamard_Int_Full"); });
    m.def("Hadamard_Bool", [](py::object r) { return wrap(Hadamard_Bool(rid(r)), "Hadamard_Bool"); });
    m.def("CondRot_Rational_Bool", [](py::object i, py::object o) { return wrap(CondRot_Rational_Bool(rid(i), rid(o)), "CondRot_Rational_Bool"); });
    m.def("CondRot_Fixed_Bool", [](py::object i, py::object o) { return wrap(CondRot_Rational_Bool(rid(i), rid(o)), "CondRot_Fixed_Bool"); });
    m.def("CondRot_General_Bool", [](py::object i, py::object o, py::function f) {
        std::function<u22_t(uint64_t)> fn = [f](uint64_t v) { py::gil_scoped_acquire g; return to_u22(f(v)); };
        return wrap(CondRot_General_Bool(rid(i), rid(o), fn), "CondRot_General_Bool");
    });
    m.def("make_func", &make_func);
    m.def("make_func_inv", &make_func_inv);
    m.def("Rot_Bool", [](py::object r, size_t d, py::object mat) { return wrap(Rot_Bool(rid(r), d, to_u22(mat)), "Rot_Bool"); },
          py::arg("reg"), py::arg("digit"), py::arg("matrix"));
    m.def("Rot_GeneralUnitary", [](py::object r, std::vector<std::vector<complex_t>> m) {
        std::vector<complex_t> flat;
        for (auto& row : m) flat.insert(flat.end(), row.begin(), row.end());
        return wrap(Rot_GeneralUnitary(rid(r), flat), "Rot_GeneralUnitary");
    });
    m.def("Rot_GeneralStatePrep", [](py::object r, std::vector<complex_t> vec) {
        size_t id = rid(r);
        if (vec.size() != pow2(Registry::size_of(id))) throw std::invalid_argument("State size does not match the register's size.");
        return wrap(Rot_GeneralUnitary(id, stateprep_unitary(vec)), "Rot_GeneralStatePrep");
    });
    m.def("stateprep_unitary", [](std::vector<complex_t> vec) { return stateprep_unitary(vec); });
    m.def("PlusOneAndOverflow", [](py::object a, py::object o) { return PlusOneAndOverflow(rid(a), rid(o)); });
    // ---- structural
    m.def("ClearZero", [](py::object eps) {
        double e = eps.is_none() ? clearzero_default_eps() : eps.cast<double>();
        return fnop([e](State& st) { ClearZero(st, e); }, "ClearZero");
    }, py::arg("eps") = py::none());
    m.def("set_clearzero_default_eps", [](double e) { clearzero_default_eps() = e; });
    m.def("get_clearzero_default_eps", []() { return clearzero_default_eps(); });
    // fixed-point encoding of a fraction into a w-bit Rational (checked by tests/test_oracle.py)
    m.def("get_rational", [](double x, size_t w) { return get_rational(x, w); });
    m.def("Normalize", []() { return fnop([](State& st) { Normalize(st); }, "Normalize"); });
    m.def("SortUnconditional", []() { return fnop([](State& st) { SortUnconditional(st); }, "SortUnconditional"); });
    m.def("Init_Unsafe", [](py::object r, uint64_t v) {
        size_t id = rid(r);
        return fnop([id, v](State& st) { Init_Unsafe(st, id, v); }, "Init_Unsafe");
    });
    m.def("AddRegister", [](const std::string& n, StateStorageType t, size_t s) {
        return py::cpp_function([n, t, s](PyState& p) { return AddRegister(p.s.basis_states, n, t, s); });
    });
    m.def("AddRegisterWithHadamard", [](const std::string& n, StateStorageType t, size_t s) {
        return py::cpp_function([n, t, s](PyState& p) { return AddRegisterWithHadamard(p.s.basis_states, n, t, s); });
    });
    m.def("RemoveRegister", [](py::object r) {
        size_t id = rid(r);
        return py::cpp_function([id](PyState& p) { RemoveRegister(p.s.basis_states, id); });
    });
    m.def("SplitRegister", [](py::object f, const std::string& second, size_t sz) {
        std::string first = py::isinstance<py::str>(f) ? f.cast<std::string>() : Registry::name_of(f.cast<size_t>());
        return py::cpp_function([first, second, sz](PyState& p) { return SplitRegister(p.s.basis_states, first, second, sz); });
    });
    m.def("CombineRegister", [](py::object f, py::object s) {
        std::string a = py::isinstance<py::str>(f) ? f.cast<std::string>() : Registry::name_of(f.cast<size_t>());
        std::string b = py::isinstance<py::str>(s) ? s.cast<std::string>() : Registry::name_of(s.cast<size_t>());
        return py::cpp_function([a, b](PyState& p) { return CombineRegister(p.s.basis_states, a, b); });
    });
    // eps: merge threshold on |sum|^2 (the reference always uses the library epsilon, 1e-14)
    m.def("add_systems", [](PyState& cur, const PyState& add, double coef, double eps) { add_systems(cur.s.basis_states, add.s.basis_states, coef, eps); },
          py::arg("cur"), py::arg("add"), py::arg("coef"), py::arg("eps") = epsilon);
    m.def("reset_systems", [](PyState& p) { p.s.basis_states.clear(); p.s.basis_states.emplace_back(); });   // system_operations.cpp:108-117
    m.def("CheckDuplicateKey", [](PyState& p) { CheckDuplicateKey(p.s.basis_states); });

    // ---- per-entry ops
    m.def("X_Bool", [](py::object r, size_t d) { return X_Bool(rid(r), d); }, py::arg("reg"), py::arg("digit") = 0);
    m.def("FlipBools", [](py::object r) { return FlipBools(rid(r)); });
    m.def("Swap_Bool_Bool", [](py::object a, size_t d1, py::object b, size_t d2) { return Swap_Bool_Bool(rid(a), d1, rid(b), d2); });
    m.def("Swap_General_General", [](py::object a, py::object b) { return Swap_General_General(rid(a), rid(b)); });
    m.def("ShiftLeft_InPlace", [](py::object r, size_t d) { return ShiftLeft_InPlace(rid(r), d); });
    m.def("ShiftRight_InPlace", [](py::object r, size_t d) { return ShiftRight_InPlace(rid(r), d); });
    m.def("Mult_UInt_ConstUInt", [](py::object a, uint64_t c, py::object b) { return Mult_UInt_ConstUInt(rid(a), c, rid(b)); });
    m.def("Add_Mult_UInt_ConstUInt_InPlace", [](py::object a, uint64_t c, py::object b) { return Add_Mult_UInt_ConstUInt_InPlace(rid(a), c, rid(b)); });
    m.def("Mod_Mult_UInt_ConstUInt_InPlace", [](py::object r, uint64_t a, uint64_t x, uint64_t N) { return Mod_Mult_UInt_ConstUInt_InPlace(rid(r), a, x, N); });
    m.def("Add_UInt_UInt", [](py::object a, py::object b, py::object c) { return Add_UInt_UInt(rid(a), rid(b), rid(c)); });
    m.def("Add_UInt_UInt_InPlace", [](py::object a, py::object b) { return Add_UInt_UInt_InPlace(rid(a), rid(b)); });
    m.def("Add_UInt_ConstUInt", [](py::object a, uint64_t c, py::object b) { return Add_UInt_ConstUInt(rid(a), c, rid(b)); });
    m.def("Add_ConstUInt_InPlace", [](py::object a, uint64_t c) { return Add_ConstUInt_InPlace(rid(a), c); });
    m.def("Div_Sqrt_Arccos_UInt_UInt", [](py::object a, py::object b, py::object c) { return Div_Sqrt_Arccos_UInt_UInt(rid(a), rid(b), rid(c)); });
    m.def("Sqrt_Div_Arccos_Int_UInt", [](py::object a, py::object b, py::object c) { return Sqrt_Div_Arccos_Int_UInt(rid(a), rid(b), rid(c)); });
    m.def("GetRotateAngle_Int_Int", [](py::object a, py::object b, py::object c) { return GetRotateAngle_Int_Int(rid(a), rid(b), rid(c)); });
    m.def("Sub_UInt_UInt", [](py::object a, py::object b, py::object c) { return Sub_UInt_UInt(rid(a), rid(b), rid(c)); });
    m.def("Neg_UInt", [](py::object a, py::object b) { return Neg_UInt(rid(a), rid(b)); });
    m.def("Abs_SInt", [](py::object a, py::object b) { return Abs_SInt(rid(a), rid(b)); });
    m.def("Mul_UInt_UInt", [](py::object a, py::object b, py::object c) { return Mul_UInt_UInt(rid(a), rid(b), rid(c)); });
    m.def("Div_UInt_UInt", [](py::object a, py::object b, py::object c) { return Div_UInt_UInt(rid(a), rid(b), rid(c)); });
    m.def("Sqrt_UInt", [](py::object a, py::object b) { return Sqrt_UInt(rid(a), rid(b)); });
    m.def("Select_Bool_UInt_UInt", [](py::object c, py::object a, py::object b, py::object r) { return Select_Bool_UInt_UInt(rid(c), rid(a), rid(b), rid(r)); });
    m.def("And_UInt_UInt", [](py::object a, py::object b, py::object c) { return And_UInt_UInt(rid(a), rid(b), rid(c)); });
    m.def("Or_UInt_UInt", [](py::object a, py::object b, py::object c) { return Or_UInt_UInt(rid(a), rid(b), rid(c)); });
    m.def("Xor_UInt_UInt", [](py::object a, py::object b, py::object c) { return Xor_UInt_UInt(rid(a), rid(b), rid(c)); });
    m.def("Less_SInt_SInt", [](py::object a, py::object b, py::object f) { return Less_SInt_SInt(rid(a), rid(b), rid(f)); });
    m.def("Carry_UInt_UInt", [](py::object a, py::object b, py::object r, py::object f) { return Carry_UInt_UInt(rid(a), rid(b), rid(r), rid(f)); });
    m.def("Overflow_SInt_SInt", [](py::object a, py::object b, py::object r, py::object f) { return Overflow_SInt_SInt(rid(a), rid(b), rid(r), rid(f)); });
    m.def("MulOverflow_UInt_UInt", [](py::object a, py::object b, py::object r, py::object f) { return MulOverflow_UInt_UInt(rid(a), rid(b), rid(r), rid(f)); });
    m.def("IsZero_UInt", [](py::object a, py::object f) { return IsZero_UInt(rid(a), rid(f)); });
    m.def("Negative_SInt", [](py::object a, py::object f) { return Negative_SInt(rid(a), rid(f)); });
    m.def("Add_AnyInt_AnyInt_InPlace", [](py::object a, py::object b) { return Add_AnyInt_AnyInt_InPlace(rid(a), rid(b)); });
    m.def("Assign", [](py::object a, py::object b) { return Assign(rid(a), rid(b)); });
    m.def("Compare_UInt_UInt", [](py::object a, py::object b, py::object l, py::object e) { return Compare_UInt_UInt(rid(a), rid(b), rid(l), rid(e)); });
    m.def("Less_UInt_UInt", [](py::object a, py::object b, py::object l) { return Less_UInt_UInt(rid(a), rid(b), rid(l)); });
    m.def("GetMid_UInt_UInt", [](py::object a, py::object b, py::object c) { return GetMid_UInt_UInt(rid(a), rid(b), rid(c)); });
    m.def("QRAMLoad", [](std::shared_ptr<QRAMCircuit> q, py::object a, py::object d) { return QRAMLoad(q, rid(a), rid(d)); });
    m.def("QRAMLoadFast", [](std::shared_ptr<QRAMCircuit> q, py::object a, py::object d) { return QRAMLoad(q, rid(a), rid(d)); });
    m.def("Phase_Bool", [](py::object r, size_t d, double l) { return Phase_Bool(rid(r), d, l); }, py::arg("reg"), py::arg("digit"), py::arg("lambda_"));
    m.def("Z_Bool", [](py::object r, size_t d) { return Phase_Bool(rid(r), d, pi); }, py::arg("reg"), py::arg("digit") = 0);
    m.def("S_Bool", [](py::object r, size_t d) { return Phase_Bool(rid(r), d, pi / 2); }, py::arg("reg"), py::arg("digit") = 0);
    m.def("T_Bool", [](py::object r, size_t d) { return Phase_Bool(rid(r), d, pi / 4); }, py::arg("reg"), py::arg("digit") = 0);
    m.def("Y_Bool", [](py::object r, size_t d) { return Y_Bool(rid(r), d); }, py::arg("reg"), py::arg("digit") = 0);
    m.def("RZ_Bool", [](py::object r, size_t d, double a) { return RZ_Bool(rid(r), d, a); });
    m.def("ZeroConditionalPhaseFlip", [](py::object regs) { return ZeroConditionalPhaseFlip(rids(regs)); });
    m.def("Reflection_Bool", [](py::object regs, bool inv) { return Reflection_Bool(rids(regs), inv); }, py::arg("regs"), py::arg("inverse") = false);
    m.def("GlobalPhase", [](complex_t c) { return GlobalPhase(c); });

    // ---- CKS quantum-walk primitives and post-selection
    py::class_<SparseMatrix, std::shared_ptr<SparseMatrix>>(m, "SparseMatrix")
        .def(py::init([](std::vector<double> vals, std::vector<size_t> sp, size_t ds, size_t nnz, size_t nrow, bool pos) {
            if (vals.size() != nnz * nrow || sp.size() != nnz * nrow) throw std::invalid_argument("need nnz_col entries per row");
            auto mat = std::make_shared<SparseMatrix>();
            for (double x : vals) mat->elements.push_back(quantise_element(x, ds, pos));
            mat->sparsity = sp; mat->data_size = ds; mat->nnz_col = nnz; mat->n_row = nrow; mat->positive_only = pos;
            return mat;
        }), py::arg("elements"), py::arg("sparsity"), py::arg("data_size"), py::arg("nnz_col"), py::arg("n_row"), py::arg("positive_only"))
        .def_static("from_encoded", [](std::vector<uint64_t> enc, std::vector<size_t> sp, size_t ds, size_t nnz, size_t nrow, bool pos) {
            if (enc.size() != nnz * nrow || sp.size() != nnz * nrow) throw std::invalid_argument("need nnz_col entries per row");
            auto mat = std::make_shared<SparseMatrix>();
            mat->elements = enc; mat->sparsity = sp; mat->data_size = ds; mat->nnz_col = nnz; mat->n_row = nrow; mat->positive_only = pos;
            return mat;
        }, py::arg("elements"), py::arg("sparsity"), py::arg("data_size"), py::arg("nnz_col"), py::arg("n_row"), py::arg("positive_only") = true)
        .def("get_data", &SparseMatrix::get_data)
        .def("get_sparsity_offset", &SparseMatrix::get_sparsity_offset)
        .def_readonly("elements", &SparseMatrix::elements)
        .def_readonly("sparsity", &SparseMatrix::sparsity)
        .def_readonly("positive_only", &SparseMatrix::positive_only)
        .def_readonly("data_size", &SparseMatrix::data_size)
        .def_readonly("nnz_col", &SparseMatrix::nnz_col)
        .def_readonly("n_row", &SparseMatrix::n_row);
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
    m.def("make_tridiagonal_matrix_spec", [](double a, double b, size_t n, size_t ds) {
        auto [mat, spec] = make_tridiagonal_matrix_spec(a, b, n, ds);
        return py::make_tuple(std::make_shared<SparseMatrix>(std::move(mat)), spec);
    });
    m.def("GetRowAddr", [](py::object o, py::object r, size_t sz, py::object out) { return GetRowAddr(rid(o), rid(r), sz, rid(out)); });
    m.def("GetDataAddr", [](py::object o, py::object r, py::object c, size_t sz, py::object out) {
        return GetDataAddr(rid(o), rid(r), rid(c), sz, rid(out));
    });
    m.def("QuantumBinarySearch_Fast", [](std::shared_ptr<QRAMCircuit> q, py::object o, size_t len, py::object t, py::object r) {
        return QuantumBinarySearch_Fast(q, rid(o), len, rid(t), rid(r));
    });
    m.def("GetQWRotateAngle_Int_Int_Int", [](py::object d, py::object r, py::object c, py::object o, const SparseMatrix& mat) {
        return GetQWRotateAngle_Int_Int_Int(rid(d), rid(r), rid(c), rid(o), mat);
    });
    m.def("TridiagOracle1", [](py::object o, py::object r, py::object s, py::object out, const TridiagonalSpec& sp) {
        return TridiagOracle1(rid(o), rid(r), rid(s), rid(out), sp);
    });
    m.def("TridiagOracle2", [](py::object o, py::object r, py::object k, py::object sr, const TridiagonalSpec& sp) {
        return TridiagOracle2(rid(o), rid(r), rid(k), rid(sr), sp);
    });
    m.def("TridiagParityPhase", [](py::object r) { return TridiagParityPhase(rid(r)); });
    m.def("CheckNan", []() { return fnop([](State& st) { CheckNan(st); }, "CheckNan"); });
    m.def("CheckNormalization", [](double t) { return fnop([t](State& st) { CheckNormalization(st, t); }, "CheckNormalization"); },
          py::arg("threshold") = 1e-5);
    struct Select { std::vector<size_t> regs; std::vector<uint64_t> vals; };
    py::class_<Select>(m, "PartialTraceSelect")
        .def(py::init([](py::dict d) {
            Select s;
            for (auto item : d) { s.regs.push_back(rid(item.first)); s.vals.push_back(item.second.cast<uint64_t>()); }
            return s;
        }))
        .def("__call__", [](const Select& s, PyState& p) { return PartialTraceSelect(p.s.basis_states, s.regs, s.vals); });
    struct SelectRange { size_t reg; size_t lo, hi; };
    py::class_<SelectRange>(m, "PartialTraceSelectRange")
        .def(py::init([](py::object r, std::pair<size_t, size_t> rg) { return SelectRange{rid(r), rg.first, rg.second}; }))
        .def("__call__", [](const SelectRange& s, PyState& p) { return PartialTraceSelectRange(p.s.basis_states, s.reg, s.lo, s.hi); });
    py::class_<ChebyshevPolynomialCoefficient>(m, "ChebyshevPolynomialCoefficient")
        .def(py::init<size_t>())
        .def_readonly("b", &ChebyshevPolynomialCoefficient::b)
        .def("C", &ChebyshevPolynomialCoefficient::C)
        .def("coef", &ChebyshevPolynomialCoefficient::coef)
        .def("sign", &ChebyshevPolynomialCoefficient::sign)
        .def("step", &ChebyshevPolynomialCoefficient::step);

    // ---- measurement
    m.def("set_seed", [](long long s) { RandomEngine::get().set_seed((uint64_t)s); });
    m.def("get_seed", []() { return (long long)RandomEngine::get().seed; });
    struct Meas { std::vector<size_t> regs; };
    py::class_<Meas>(m, "MeasureZ")
        .def(py::init([](py::object r) { return Meas{rids(r)}; }))
        .def("__call__", [](const Meas& mz, PyState& p) { return MeasureZ(p.s.basis_states, mz.regs); });
    struct Rst { std::vector<size_t> regs; std::vector<uint64_t> targets; };
    py::class_<Rst>(m, "Reset")
        .def(py::init([](py::object r, py::object t) {
            Rst x{rids(r), {}};
            if (t.is_none()) x.targets.assign(x.regs.size(), 0);
            else if (py::isinstance<py::int_>(t)) x.targets.assign(1, t.cast<uint64_t>());
            else x.targets = t.cast<std::vector<uint64_t>>();
            if (x.targets.size() != x.regs.size()) throw std::invalid_argument("one target per register");
            return x;
        }), py::arg("registers"), py::arg("targets") = py::none())
        .def("__call__", [](const Rst& r, PyState& p) { return Reset(p.s.basis_states, r.regs, r.targets); });
    struct Prob { std::vector<size_t> regs; std::vector<uint64_t> vals; };
    py::class_<Prob>(m, "Probability")
        .def(py::init([](py::dict d) {
            Prob x;
            for (auto item : d) { x.regs.push_back(rid(item.first)); x.vals.push_back(item.second.cast<uint64_t>()); }
            return x;
        }))
        .def(py::init([](py::object r, uint64_t v) { return Prob{{rid(r)}, {v}}; }))
        .def("__call__", [](const Prob& x, PyState& p) { return Probability(p.s.basis_states, x.regs, x.vals); })
        .def_static("distribution", [](PyState& p, py::object r) { return Distribution(p.s.basis_states, rid(r)); });

    // Timing helper for the CPU baseline: wall seconds of fn() measured in C++.
    m.def("wall_seconds", [](py::function fn) {
        auto t0 = std::chrono::steady_clock::now();
        fn();
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    });
}
