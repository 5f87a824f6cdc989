"""Runtime-compiled operator plugins (the reference's PySparQ/test/test_dynamic_operator.py
cases, plus the plugin ABI and, on the GPU, application to device states).

CPU: code generation, compilation errors, argument validation, caching, the six
exported plugin symbols, proxy construction.  GPU: a plugin's host-side operator
applied to a device state matches the oracle's built-in equivalent; dag undoes
the operator; a plugin calling a built-in engine operator on its host view;
a C++ exception inside a plugin surfaces as RuntimeError.
"""
import ctypes
import math
import shutil
import subprocess

import numpy as np
import pytest

from parity import assert_same_state, random_state

if not shutil.which("g++"):
    pytest.skip("g++ not available", allow_module_level=True)

from sparqsim_amd.dynamic_operator import (CompilationError, clear_cache, compile_operator,  # noqa: E402
                                           get_cache_info)

FLIP = """
class TestFlipOp : public SelfAdjointOperator {
    size_t reg_id;
public:
    TestFlipOp(size_t r) : reg_id(r) {}
    void operator()(std::vector<System>& state) const override {
        for (auto& s : state) s.get(reg_id).value ^= 1;
    }
};
"""

PHASE = """
class TestPhaseOp : public BaseOperator {
    size_t reg_id;
    double phase;
public:
    TestPhaseOp(size_t r, double p) : reg_id(r), phase(p) {}
    void operator()(std::vector<System>& state) const override {
        for (auto& s : state)
            if (s.get(reg_id).value != 0) s.amplitude *= std::exp(std::complex<double>(0, phase));
    }
    void dag(std::vector<System>& state) const override {
        for (auto& s : state)
            if (s.get(reg_id).value != 0) s.amplitude *= std::exp(std::complex<double>(0, -phase));
    }
};
"""

COMPOSITE = """
class TestHadamardThenFlip : public BaseOperator {
    size_t reg_id;
public:
    TestHadamardThenFlip(size_t r) : reg_id(r) {}
    void operator()(std::vector<System>& state) const override {
        (Hadamard_Int_Full(reg_id))(state);
        for (auto& s : state) s.get(reg_id).value ^= 1;
    }
    void dag(std::vector<System>& state) const override {
        for (auto& s : state) s.get(reg_id).value ^= 1;
        (Hadamard_Int_Full(reg_id))(state);
    }
};
"""

OVERFLOW = """
class TestOverflowOp : public SelfAdjointOperator {
    size_t reg_id;
public:
    TestOverflowOp(size_t r) : reg_id(r) {}
    void operator()(std::vector<System>& state) const override {
        for (auto& s : state) s.get(reg_id).value = ~uint64_t(0);
    }
};
"""


@pytest.fixture(scope="module")
def cache(tmp_path_factory):
    return str(tmp_path_factory.mktemp("plugins"))


# ---------------------------------------------------------------- CPU
def test_compile_self_adjoint_and_symbols(cache):
    cls = compile_operator("TestFlipOp", FLIP, "SelfAdjointOperator", constructor_args=[("size_t", "reg_id")],
                           cache_dir=cache)
    assert cls.__name__ == "TestFlipOp" and cls._base_class == "SelfAdjointOperator"
    out = subprocess.run(["nm", "-D", "--defined-only", cls._lib_path], capture_output=True, text=True).stdout
    for sym in ("create_operator", "destroy_operator", "get_operator_name", "apply_operator",
                "apply_operator_dag", "get_base_class"):
        assert f" T {sym}\n" in out, sym
    lib = ctypes.CDLL(cls._lib_path)
    lib.get_operator_name.restype = ctypes.c_char_p
    lib.get_base_class.restype = ctypes.c_char_p
    assert lib.get_operator_name() == b"TestFlipOp"
    assert lib.get_base_class() == b"SelfAdjointOperator"


def test_constructor_args_and_repr(cache):
    code = """
class TestMultiParamOp : public SelfAdjointOperator {
    size_t reg_id; double factor; int offset;
public:
    TestMultiParamOp(size_t r, double f, int o) : reg_id(r), factor(f), offset(o) {}
    void operator()(std::vector<System>& state) const override {
        for (auto& s : state) s.get(reg_id).value = static_cast<size_t>(s.get(reg_id).value * factor + offset);
    }
};
"""
    cls = compile_operator("TestMultiParamOp", code, "SelfAdjointOperator",
                           constructor_args=[("size_t", "reg_id"), ("double", "factor"), ("int", "offset")],
                           cache_dir=cache)
    for arg in ("reg_id", "factor", "offset"):
        assert arg in cls.__doc__
    op = cls(reg_id=0, factor=2.0, offset=1)
    assert repr(op) == "TestMultiParamOp(reg_id=0, factor=2.0, offset=1)"
    with pytest.raises(TypeError):
        cls(reg_id=0)


def test_base_vs_self_adjoint(cache):
    a = compile_operator("TestFlipOp", FLIP, "SelfAdjointOperator", constructor_args=[("size_t", "reg_id")],
                         cache_dir=cache)
    b = compile_operator("TestPhaseOp", PHASE, "BaseOperator",
                         constructor_args=[("size_t", "reg_id"), ("double", "phase")], cache_dir=cache)
    assert a._base_class == "SelfAdjointOperator" and b._base_class == "BaseOperator"


def test_compilation_error(cache):
    bad = """
class BadOp : public BaseOperator {
    void operator()(std::vector<System>& state) const override { undefined_variable = 42; }
};
"""
    with pytest.raises(CompilationError) as e:
        compile_operator("BadOp", bad, "BaseOperator", cache_dir=cache)
    assert "Compilation failed" in str(e.value) and "undefined_variable" in str(e.value)


def test_argument_validation():
    with pytest.raises(ValueError, match="base_class"):
        compile_operator("TestOp", "class TestOp : public InvalidBase {};", "InvalidBase")
    with pytest.raises(ValueError):
        compile_operator("", "class TestOp : public BaseOperator {};")
    with pytest.raises(ValueError):
        compile_operator("TestOp", "")


def test_caching(tmp_path):
    d = str(tmp_path / "cache")
    c1 = compile_operator("TestFlipOp", FLIP, "SelfAdjointOperator", constructor_args=[("size_t", "reg_id")],
                          cache_dir=d)
    c2 = compile_operator("TestFlipOp", FLIP, "SelfAdjointOperator", constructor_args=[("size_t", "reg_id")],
                          cache_dir=d)
    assert c1._lib_path == c2._lib_path
    info = get_cache_info(d)
    assert info["exists"] and info["so_count"] == 1 and info["cpp_count"] == 1
    assert clear_cache(d) == 2
    assert get_cache_info(d)["file_count"] == 0


def test_constructor_parameter_types():
    """C++ parameter spellings map to the ctypes type of the same width and
    signedness (LP64); unknown spellings are refused at load time."""
    from sparqsim_amd.dynamic_operator.operator_wrapper import ctype_of

    want = {"size_t": ctypes.c_size_t, "const size_t": ctypes.c_size_t, "std::size_t": ctypes.c_size_t,
            "int": ctypes.c_int32, "unsigned int": ctypes.c_uint32, "unsigned": ctypes.c_uint32,
            "long": ctypes.c_int64, "unsigned long long": ctypes.c_uint64, "short": ctypes.c_int16,
            "uint8_t": ctypes.c_uint8, "std::int64_t": ctypes.c_int64, "double": ctypes.c_double,
            "float": ctypes.c_float, "bool": ctypes.c_bool, "const char*": ctypes.c_char_p,
            "char *": ctypes.c_char_p, "double*": ctypes.c_void_p, "unsigned char": ctypes.c_ubyte,
            "signed char": ctypes.c_byte}
    for spelling, t in want.items():
        assert ctype_of(spelling) is t, spelling
    from sparqsim_amd.dynamic_operator import DynamicOperatorLoadError

    with pytest.raises(DynamicOperatorLoadError):
        ctype_of("std::vector<int>")


def test_proxy_lifetime_and_arguments(cache):
    """Positional and keyword construction, the errors of a bad call, close()
    and cleanup_all_instances destroy the C++ object once; a missing library
    is a load error."""
    from sparqsim_amd.dynamic_operator import (DynamicOperatorLoadError, cleanup_all_instances,
                                               create_operator_class)

    cls = compile_operator("TestPhaseOp", PHASE, "BaseOperator",
                           constructor_args=[("size_t", "reg_id"), ("double", "phase")], cache_dir=cache)
    a = cls(1, phase=0.5)
    assert repr(a) == "TestPhaseOp(reg_id=1, phase=0.5)"
    with pytest.raises(TypeError):
        cls(1, 0.5, 3)
    with pytest.raises(TypeError):
        cls(1, reg_id=2, phase=0.1)
    with pytest.raises(TypeError):
        cls(reg_id=1, angle=0.1)
    a.close()
    a.close()
    with pytest.raises(RuntimeError):
        a._ptr()
    b = cls(reg_id=0, phase=0.25)
    cleanup_all_instances()
    with pytest.raises(RuntimeError):
        b._ptr()
    with pytest.raises(DynamicOperatorLoadError):
        create_operator_class("Nothing", "/nonexistent/libnothing.so")


# ---------------------------------------------------------------- GPU
def _setup(mod, seed):
    mod.System.clear()
    st = mod.SparseState()
    mod.AddRegister("a", mod.UnsignedInteger, 6)(st)
    mod.AddRegister("b", mod.UnsignedInteger, 5)(st)
    vals, amps = random_state(mod, 40, seed)
    st.load(vals, amps)
    return st


@pytest.mark.gpu
def test_plugin_flip_matches_builtin(ps, oracle, cache):
    cls = compile_operator("TestFlipOp", FLIP, "SelfAdjointOperator", constructor_args=[("size_t", "reg_id")],
                           cache_dir=cache)
    sg = _setup(ps, 1)
    cls(reg_id=ps.System.get_id("b"))(sg)
    so = _setup(oracle, 1)
    oracle.X_Bool("b", 0)(so)
    assert_same_state(sg.export(), so.export(), exact_amps=True, what="plugin flip")
    cls(reg_id=ps.System.get_id("b")).dag(sg)
    oracle.X_Bool("b", 0)(so)
    assert_same_state(sg.export(), so.export(), exact_amps=True, what="plugin flip dag")


@pytest.mark.gpu
def test_plugin_phase_and_dag(ps, oracle, cache):
    cls = compile_operator("TestPhaseOp", PHASE, "BaseOperator",
                           constructor_args=[("size_t", "reg_id"), ("double", "phase")], cache_dir=cache)
    sg = _setup(ps, 2)
    before = sg.export()
    op = cls(reg_id=ps.System.get_id("a"), phase=0.37)
    op(sg)
    ids, vals, amps = before
    expect = np.where(vals[:, 0] != 0, amps * np.exp(1j * 0.37), amps)
    assert_same_state(sg.export(), (ids, vals, expect), what="plugin phase")
    op.dag(sg)
    assert_same_state(sg.export(), before, what="plugin phase dag")


@pytest.mark.gpu
def test_plugin_calls_builtin_on_host_view(ps, oracle, cache):
    cls = compile_operator("TestHadamardThenFlip", COMPOSITE, "BaseOperator",
                           constructor_args=[("size_t", "reg_id")], cache_dir=cache)
    sg = _setup(ps, 3)
    cls(reg_id=ps.System.get_id("a"))(sg)
    so = _setup(oracle, 3)
    oracle.Hadamard_Int_Full("a")(so)
    oracle.X_Bool("a", 0)(so)
    assert_same_state(sg.export(), so.export(), what="plugin composite")
    cls(reg_id=ps.System.get_id("a")).dag(sg)
    assert_same_state(sg.export(), _setup(oracle, 3).export(), what="plugin composite dag")


@pytest.mark.gpu
def test_plugin_exception_becomes_runtime_error(ps, cache):
    cls = compile_operator("TestOverflowOp", OVERFLOW, "SelfAdjointOperator", constructor_args=[("size_t", "reg_id")],
                           cache_dir=cache)
    sg = _setup(ps, 4)
    before = sg.export()
    with pytest.raises(RuntimeError, match="exceeds its width"):
        cls(reg_id=ps.System.get_id("b"))(sg)
    assert_same_state(sg.export(), before, exact_amps=True, what="state untouched")
