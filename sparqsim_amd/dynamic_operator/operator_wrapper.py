"""Python proxies for runtime-compiled operator plugins.

A plugin library exports six C symbols (include/sparq_plugin.h): a factory
``create_operator(<constructor arguments>)`` returning an opaque operator
pointer, ``destroy_operator``, ``apply_operator`` / ``apply_operator_dag``
taking (operator, SparseState*), and two string getters.  The public names kept
from the reference's PySparQ/pysparq/dynamic_operator/operator_wrapper.py
(``CppOperatorWrapper`` with load/create/destroy/apply/apply_dag/get_name/
get_base_class/close, ``create_operator_class``, ``cleanup_all_instances`` and
the three exception classes) are what callers and the reference's tests use;
the implementation below is this package's own.

Design:
* ``_Symbols`` binds every entry point once, from a table of
  (name, argument types, result type, required); a missing required symbol is
  reported with the library path.
* The factory's argument types come from the C++ spelling of each constructor
  parameter, read by ``ctype_of`` (qualifiers dropped, fixed-width and
  pointer spellings recognised); an unknown spelling is an error at load time
  rather than a silent guess.
* Every proxy instance owns its C++ object through a ``weakref.finalize``, so
  the object is destroyed exactly once: by ``close()``/``cleanup_all_instances``
  or when the proxy is collected, whichever comes first.
* The generated entry points catch C++ exceptions and park the message
  (the host layer's ``_take_plugin_error``); ``apply``/``apply_dag`` raise it
  as ``RuntimeError`` after the call returns.
"""
from __future__ import annotations

import ctypes
import os
import re
import warnings
import weakref
from typing import Callable, Dict, List, Optional, Sequence, Tuple, Type


class DynamicOperatorError(Exception):
    """Base class of every plugin-operator failure."""


class DynamicOperatorLoadError(DynamicOperatorError):
    """The plugin library could not be opened or lacks a required symbol."""


class DynamicOperatorFactoryError(DynamicOperatorError):
    """The plugin's factory could not construct an operator."""


# ------------------------------------------------------------ C++ type spellings
_INTEGER_BITS = {"char": 8, "short": 16, "int": 32, "long": 64, "long long": 64}
_FIXED = re.compile(r"^(u?)int(8|16|32|64)_t$")


def ctype_of(cpp_type: str):
    """ctypes type of one constructor parameter, from its C++ spelling.

    ``const`` and ``std::`` are ignored; ``char*`` (any constness) is a C
    string; other pointers and references are passed as addresses; integers
    follow the LP64 model of the host compiler.
    """
    t = " ".join(cpp_type.replace("std::", "").replace("const", " ").split())
    if not t:
        raise DynamicOperatorLoadError("empty constructor parameter type")
    if t.replace(" ", "") == "char*":
        return ctypes.c_char_p
    if t.endswith("*") or t.endswith("&"):
        return ctypes.c_void_p
    if t in ("double", "float", "bool"):
        return {"double": ctypes.c_double, "float": ctypes.c_float, "bool": ctypes.c_bool}[t]
    if t in ("size_t", "uintptr_t"):
        return ctypes.c_size_t
    if t in ("ptrdiff_t", "ssize_t", "intptr_t"):
        return ctypes.c_ssize_t
    m = _FIXED.match(t)
    if m:
        return getattr(ctypes, f"c_{'u' if m.group(1) else ''}int{m.group(2)}")
    unsigned = t.startswith("unsigned")
    base = t.replace("unsigned", "").replace("signed", "").strip() or "int"
    if base in _INTEGER_BITS:
        bits = _INTEGER_BITS[base]
        if bits == 8:
            if unsigned:
                return ctypes.c_ubyte
            return ctypes.c_byte if "signed" in t else ctypes.c_char
        return getattr(ctypes, f"c_{'u' if unsigned else ''}int{bits}")
    raise DynamicOperatorLoadError(f"cannot pass constructor parameter of type '{cpp_type}' through the plugin factory")


# ------------------------------------------------------------ symbol binding
_VOIDP = ctypes.c_void_p
# (symbol, argument types (None: from the constructor), result type, required)
_SYMBOL_TABLE = (
    ("create_operator", None, _VOIDP, True),
    ("destroy_operator", [_VOIDP], None, True),
    ("get_operator_name", [], ctypes.c_char_p, True),
    ("get_base_class", [], ctypes.c_char_p, True),
    ("apply_operator", [_VOIDP, _VOIDP], None, False),
    ("apply_operator_dag", [_VOIDP, _VOIDP], None, False),
)


class _Symbols:
    """The bound entry points of one opened plugin library."""

    def __init__(self, lib: ctypes.CDLL, path: str, factory_args: Sequence):
        self.lib = lib
        self.fn: Dict[str, Optional[Callable]] = {}
        for name, argtypes, restype, required in _SYMBOL_TABLE:
            f = getattr(lib, name, None)
            if f is None:
                if required:
                    raise DynamicOperatorLoadError(f"{path}: plugin symbol '{name}' is missing")
                self.fn[name] = None
                continue
            f.argtypes = list(factory_args) if argtypes is None else argtypes
            f.restype = restype
            self.fn[name] = f

    def text(self, name: str, default: str) -> str:
        raw = self.fn[name]()
        return raw.decode("utf-8") if raw else default


def _raise_parked_error() -> None:
    """The generated apply entry points catch C++ exceptions and park the message."""
    from .. import _core

    msg = _core._take_plugin_error()
    if msg:
        raise RuntimeError(msg)


class CppOperatorWrapper:
    """One plugin library: opens it, types its factory and forwards calls."""

    def __init__(self, lib_path: str):
        self.lib_path = lib_path
        self._sym: Optional[_Symbols] = None
        self._arg_types: List[str] = []

    # -- lifetime of the library binding
    def load(self, arg_types: List[str] = None) -> None:
        """Open the library (RTLD_GLOBAL, so the plugin resolves the host layer's
        register table) and bind its entry points for these constructor types."""
        if not os.path.isfile(self.lib_path):
            raise DynamicOperatorLoadError(f"no plugin library at {self.lib_path}")
        self._arg_types = list(arg_types or [])
        factory_args = [ctype_of(t) for t in self._arg_types]
        try:
            lib = ctypes.CDLL(self.lib_path, mode=ctypes.RTLD_GLOBAL)
        except OSError as exc:
            raise DynamicOperatorLoadError(f"cannot open plugin library {self.lib_path}: {exc}") from exc
        self._sym = _Symbols(lib, self.lib_path, factory_args)

    def close(self) -> None:
        """Forget the bound entry points (the library stays mapped, as dlclose of a
        library holding live C++ objects is unsafe)."""
        self._sym = None

    def _need(self, name: str) -> Callable:
        if self._sym is None:
            raise DynamicOperatorError(f"plugin library {self.lib_path} is not loaded")
        f = self._sym.fn[name]
        if f is None:
            raise DynamicOperatorError(f"plugin library {self.lib_path} does not export {name}")
        return f

    # -- factory
    def create(self, *args) -> int:
        """Construct one operator; returns its address."""
        if self._sym is None:
            raise DynamicOperatorFactoryError(f"plugin library {self.lib_path} is not loaded")
        if len(args) != len(self._arg_types):
            raise DynamicOperatorFactoryError(
                f"the factory takes {len(self._arg_types)} argument(s), {len(args)} given")
        try:
            ptr = self._sym.fn["create_operator"](*args)
        except (ctypes.ArgumentError, TypeError) as exc:
            raise DynamicOperatorFactoryError(f"bad constructor argument for {self.lib_path}: {exc}") from exc
        if not ptr:
            raise DynamicOperatorFactoryError(f"the factory of {self.lib_path} returned no operator")
        return ptr

    def destroy(self, ptr: int) -> None:
        if ptr and self._sym is not None:
            self._sym.fn["destroy_operator"](ptr)

    # -- queries
    def get_name(self) -> str:
        return self._sym.text("get_operator_name", "") if self._sym else ""

    def get_base_class(self) -> str:
        return self._sym.text("get_base_class", "BaseOperator") if self._sym else "BaseOperator"

    # -- application
    def apply(self, ptr: int, state_cpp_ptr: int) -> None:
        f = self._need("apply_operator")
        if ptr and state_cpp_ptr:
            f(ptr, state_cpp_ptr)
            _raise_parked_error()

    def apply_dag(self, ptr: int, state_cpp_ptr: int) -> None:
        f = self._need("apply_operator_dag")
        if ptr and state_cpp_ptr:
            f(ptr, state_cpp_ptr)
            _raise_parked_error()


# ------------------------------------------------------------ proxy classes
# Proxies whose C++ object is still alive (cleanup_all_instances walks them).
_live_proxies: "weakref.WeakSet" = weakref.WeakSet()


class _OwnedOperator:
    """Owns one C++ operator; destroyed once, explicitly or at collection."""

    def __init__(self, wrapper: CppOperatorWrapper, ptr: int, owner):
        self.ptr = ptr
        self._finalizer = weakref.finalize(owner, wrapper.destroy, ptr)

    def release(self) -> None:
        if self._finalizer.alive:
            self._finalizer()
        self.ptr = 0


def _proxy_doc(name: str, base: str, lib_path: str, params: Sequence[Tuple[str, str]]) -> str:
    lines = [f"{name}: C++ operator ({base}) from {os.path.basename(lib_path)}.", ""]
    if params:
        lines.append("Keyword arguments of the constructor:")
        lines += [f"    {arg} : {typ}" for typ, arg in params]
    else:
        lines.append("The constructor takes no arguments.")
    lines += ["", "Instances are applied as op(state) and op.dag(state)."]
    return "\n".join(lines)


def create_operator_class(name: str, lib_path: str, base_class: str = "BaseOperator",
                          constructor_args: List[Tuple[str, str]] = None) -> Type:
    """Load a plugin library and return a class whose instances each own one of its
    operators.  Instances are constructed with the constructor's parameters as
    keyword (or positional) arguments and applied as ``op(state)`` / ``op.dag(state)``."""
    params = [(str(t), str(n)) for t, n in (constructor_args or [])]
    wrapper = CppOperatorWrapper(lib_path)
    wrapper.load([t for t, _ in params])
    exported = wrapper.get_base_class()
    if exported and exported != base_class:
        warnings.warn(f"{lib_path} implements a {exported}; the declared base {base_class} is replaced by it")
        base_class = exported
    self_adjoint = base_class == "SelfAdjointOperator"
    names = [n for _, n in params]

    def _bind(args, kwargs) -> tuple:
        if len(args) > len(names):
            raise TypeError(f"{name}() takes {len(names)} argument(s), {len(args)} given")
        values = dict(zip(names, args))
        for k, v in kwargs.items():
            if k not in names:
                raise TypeError(f"{name}() has no parameter '{k}'")
            if k in values:
                raise TypeError(f"{name}() got two values for '{k}'")
            values[k] = v
        absent = [n for n in names if n not in values]
        if absent:
            raise TypeError(f"{name}() is missing argument(s): {', '.join(absent)}")
        return tuple(values[n] for n in names)

    def __init__(self, *args, **kwargs):
        self._args = _bind(args, kwargs)
        self._owned = _OwnedOperator(wrapper, wrapper.create(*self._args), self)
        _live_proxies.add(self)

    def _ptr(self) -> int:
        owned = getattr(self, "_owned", None)
        if owned is None or not owned.ptr:
            raise RuntimeError(f"{name}: the C++ operator has been destroyed")
        return owned.ptr

    def __call__(self, state):
        wrapper.apply(self._ptr(), state._cpp_ptr())
        return state

    def dag(self, state):
        # a self-adjoint operator is its own inverse: the forward form is applied
        (wrapper.apply if self_adjoint else wrapper.apply_dag)(self._ptr(), state._cpp_ptr())
        return state

    def close(self) -> None:
        """Destroy the C++ operator now (later calls raise)."""
        owned = getattr(self, "_owned", None)
        if owned is not None:
            owned.release()
        _live_proxies.discard(self)

    def __repr__(self):
        shown = ", ".join(f"{n}={v!r}" for n, v in zip(names, getattr(self, "_args", ())))
        return f"{name}({shown})"

    namespace = {
        "__init__": __init__, "__call__": __call__, "dag": dag, "close": close, "__repr__": __repr__,
        "_ptr": _ptr, "__doc__": _proxy_doc(name, base_class, lib_path, params),
        "_is_dynamic_operator": True, "_base_class": base_class, "_lib_path": lib_path, "_wrapper": wrapper,
        "_constructor_args": tuple(params),
    }
    return type(name, (object,), namespace)


def cleanup_all_instances() -> None:
    """Destroy the C++ operator of every live proxy instance."""
    for proxy in list(_live_proxies):
        proxy.close()
