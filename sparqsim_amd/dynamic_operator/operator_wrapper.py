"""ctypes proxies for plugin operator libraries.

Same interface as the reference's PySparQ/pysparq/dynamic_operator/operator_wrapper.py
(CppOperatorWrapper, create_operator_class, cleanup_all_instances and the error
classes): the library is loaded with RTLD_GLOBAL, the factory is typed from the
constructor argument list, and calls pass the C++ SparseState* from
`state._cpp_ptr()`.
"""
from __future__ import annotations

import ctypes
import os
import warnings
import weakref
from typing import List, Tuple, Type


class DynamicOperatorError(Exception):
    pass


class DynamicOperatorLoadError(DynamicOperatorError):
    pass


class DynamicOperatorFactoryError(DynamicOperatorError):
    pass


_active_instances: dict = {}
_instance_counter = [0]

_CTYPES = {
    "int": ctypes.c_int, "size_t": ctypes.c_size_t, "unsigned int": ctypes.c_uint,
    "unsigned long": ctypes.c_ulong, "unsigned long long": ctypes.c_ulonglong, "long": ctypes.c_long,
    "long long": ctypes.c_longlong, "float": ctypes.c_float, "double": ctypes.c_double,
    "bool": ctypes.c_bool, "char": ctypes.c_char, "char*": ctypes.c_char_p, "const char*": ctypes.c_char_p,
    "uint64_t": ctypes.c_uint64, "int64_t": ctypes.c_int64, "uint32_t": ctypes.c_uint32, "int32_t": ctypes.c_int32,
}


class CppOperatorWrapper:
    """One loaded plugin library and its six entry points."""

    def __init__(self, lib_path: str):
        self.lib_path = lib_path
        self._handle = None
        self._create_func = self._destroy_func = None
        self._get_name_func = self._get_base_class_func = None
        self._apply_func = self._apply_dag_func = None
        self._arg_types: List[str] = []

    def load(self, arg_types: List[str] = None):
        if not os.path.exists(self.lib_path):
            raise DynamicOperatorLoadError(f"Dynamic library does not exist: {self.lib_path}")
        try:
            self._handle = ctypes.CDLL(self.lib_path, mode=ctypes.RTLD_GLOBAL)
        except OSError as e:
            raise DynamicOperatorLoadError(f"Failed to load dynamic library: {e}")
        try:
            self._create_func = self._handle.create_operator
            self._destroy_func = self._handle.destroy_operator
            self._get_name_func = self._handle.get_operator_name
            self._get_base_class_func = self._handle.get_base_class
        except AttributeError as e:
            raise DynamicOperatorLoadError(f"Required factory function not found: {e}")
        self._apply_func = getattr(self._handle, "apply_operator", None)
        self._apply_dag_func = getattr(self._handle, "apply_operator_dag", None)
        self._arg_types = list(arg_types or [])
        self._setup_arg_types()

    def _setup_arg_types(self):
        self._create_func.argtypes = [_CTYPES.get(t.strip(), ctypes.c_size_t) for t in self._arg_types]
        self._create_func.restype = ctypes.c_void_p
        self._destroy_func.argtypes = [ctypes.c_void_p]
        self._destroy_func.restype = None
        for f in (self._get_name_func, self._get_base_class_func):
            f.argtypes = []
            f.restype = ctypes.c_char_p
        for f in (self._apply_func, self._apply_dag_func):
            if f is not None:
                f.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
                f.restype = None

    def create(self, *args) -> int:
        if not self._create_func:
            raise DynamicOperatorFactoryError("Factory function not loaded")
        try:
            return self._create_func(*args)
        except Exception as e:
            raise DynamicOperatorFactoryError(f"Failed to create operator: {e}")

    def destroy(self, ptr: int):
        if self._destroy_func and ptr:
            self._destroy_func(ptr)

    def close(self):
        self._create_func = self._destroy_func = None
        self._get_name_func = self._get_base_class_func = None
        self._apply_func = self._apply_dag_func = None
        self._handle = None

    def get_name(self) -> str:
        r = self._get_name_func() if self._get_name_func else None
        return r.decode("utf-8") if r else ""

    def get_base_class(self) -> str:
        r = self._get_base_class_func() if self._get_base_class_func else None
        return r.decode("utf-8") if r else "BaseOperator"

    def apply(self, ptr: int, state_cpp_ptr: int):
        if self._apply_func is None:
            raise DynamicOperatorError("library has no apply_operator")
        if ptr and state_cpp_ptr:
            self._apply_func(ptr, state_cpp_ptr)
            _raise_parked_error()

    def apply_dag(self, ptr: int, state_cpp_ptr: int):
        if self._apply_dag_func is None:
            raise DynamicOperatorError("library has no apply_operator_dag")
        if ptr and state_cpp_ptr:
            self._apply_dag_func(ptr, state_cpp_ptr)
            _raise_parked_error()


def _raise_parked_error():
    """The generated entry points catch C++ exceptions and park the message."""
    from .. import _core

    msg = _core._take_plugin_error()
    if msg:
        raise RuntimeError(msg)


def create_operator_class(name: str, lib_path: str, base_class: str = "BaseOperator",
                          constructor_args: List[Tuple[str, str]] = None) -> Type:
    """A Python class whose instances own one C++ operator of the plugin."""
    constructor_args = list(constructor_args or [])
    wrapper = CppOperatorWrapper(lib_path)
    wrapper.load([t for t, _ in constructor_args])
    detected = wrapper.get_base_class()
    if detected and detected != base_class:
        warnings.warn(f"Detected base class is {detected}, but {base_class} was specified")
        base_class = detected

    def __init__(self, **kwargs):
        args = []
        for _, arg_name in constructor_args:
            if arg_name not in kwargs:
                raise TypeError(f"Missing required argument: {arg_name}")
            args.append(kwargs[arg_name])
        self._args = tuple(args)
        self._instance_id = _instance_counter[0]
        _instance_counter[0] += 1
        _active_instances[self._instance_id] = weakref.ref(self)
        self._cpp_ptr = wrapper.create(*args)

    def __call__(self, state):
        if not self._cpp_ptr:
            raise RuntimeError("Operator not initialized or already destroyed")
        wrapper.apply(self._cpp_ptr, state._cpp_ptr())
        return state

    def dag(self, state):
        if not self._cpp_ptr:
            raise RuntimeError("Operator not initialized or already destroyed")
        if base_class == "SelfAdjointOperator":
            wrapper.apply(self._cpp_ptr, state._cpp_ptr())
        else:
            wrapper.apply_dag(self._cpp_ptr, state._cpp_ptr())
        return state

    def __repr__(self):
        return f"{name}(" + ", ".join(f"{n}={v!r}" for (_, n), v in zip(constructor_args, self._args)) + ")"

    def __del__(self):
        if getattr(self, "_cpp_ptr", 0):
            wrapper.destroy(self._cpp_ptr)
            self._cpp_ptr = 0
        _active_instances.pop(getattr(self, "_instance_id", None), None)

    arg_docs = "\n".join(f"        {n} ({t})" for t, n in constructor_args) or "        (none)"
    cls = type(name, (object,), {
        "__init__": __init__, "__call__": __call__, "dag": dag, "__repr__": __repr__, "__del__": __del__,
        "_is_dynamic_operator": True, "_base_class": base_class, "_lib_path": lib_path, "_wrapper": wrapper,
        "__doc__": f"\nDynamically generated operator class: {name}\nBase class: {base_class}\n"
                   f"Constructor arguments:\n{arg_docs}\n",
    })
    return cls


def cleanup_all_instances():
    """Destroy the C++ objects of every live plugin operator instance."""
    for key, ref in list(_active_instances.items()):
        inst = ref()
        if inst is not None and getattr(inst, "_cpp_ptr", 0):
            inst._wrapper.destroy(inst._cpp_ptr)
            inst._cpp_ptr = 0
        _active_instances.pop(key, None)
