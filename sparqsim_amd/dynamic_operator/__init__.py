"""Runtime-compiled C++ operators (the reference's pysparq.dynamic_operator,
PySparQ/pysparq/dynamic_operator/__init__.py).

`compile_operator` wraps a user class deriving from BaseOperator or
SelfAdjointOperator (overriding `operator()(std::vector<System>&)` and, for a
BaseOperator, `dag`) into a plugin library and returns a Python class whose
instances apply it to a SparseState.  The user code runs on the host over a copy
of the device entries, which the engine downloads before and uploads after the
call.  As in the reference, compiling proves nothing about unitarity.
"""
from __future__ import annotations

from typing import List, Optional, Tuple, Type

from .compiler import (CompilationError, CompilerConfig, clear_cache, compile_cpp_code, compute_code_hash,
                       find_project_root, format_compile_error, generate_cpp_source, get_cache_info,
                       quick_compile)
from .operator_wrapper import (CppOperatorWrapper, DynamicOperatorError, DynamicOperatorFactoryError,
                               DynamicOperatorLoadError, cleanup_all_instances, create_operator_class)

__all__ = [
    "CompilerConfig", "CompilationError", "compile_cpp_code", "compute_code_hash", "find_project_root",
    "generate_cpp_source", "format_compile_error", "clear_cache", "get_cache_info", "quick_compile",
    "compile_operator", "CppOperatorWrapper", "DynamicOperatorError", "DynamicOperatorLoadError",
    "DynamicOperatorFactoryError", "create_operator_class", "cleanup_all_instances",
]

_BASES = ("BaseOperator", "SelfAdjointOperator")


def compile_operator(name: str, cpp_code: str, base_class: str = "BaseOperator", extra_includes: List[str] = None,
                     extra_libs: List[str] = None, constructor_args: List[Tuple[str, str]] = None,
                     cache_dir: Optional[str] = None, verbose: bool = False) -> Type:
    """Compile `cpp_code` (defining class `name`) and return its Python proxy class."""
    if not name or not isinstance(name, str):
        raise ValueError("name must be a valid string")
    if not cpp_code or not isinstance(cpp_code, str):
        raise ValueError("cpp_code must be a valid C++ code string")
    if base_class not in _BASES:
        raise ValueError(f"base_class must be one of {list(_BASES)}")
    constructor_args = list(constructor_args or [])
    ctor_params = ", ".join(f"{t} {n}" for t, n in constructor_args)
    ctor_args = ", ".join(n for _, n in constructor_args)
    config = CompilerConfig(include_paths=extra_includes or [], libraries=extra_libs or [],
                            template=CompilerConfig.PYTHON_TEMPLATE.replace("{BASE_CLASS}", base_class))
    lib_path = compile_cpp_code(cpp_code=cpp_code, class_name=name, cache_dir=cache_dir, ctor_params=ctor_params,
                                ctor_args=ctor_args, config=config, verbose=verbose)
    return create_operator_class(name=name, lib_path=lib_path, base_class=base_class,
                                 constructor_args=constructor_args)
