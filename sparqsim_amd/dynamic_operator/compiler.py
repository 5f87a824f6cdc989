"""Runtime compilation of user C++ operators into plugin libraries.

Same interface as the reference's PySparQ/pysparq/dynamic_operator/compiler.py
(CompilerConfig, compile_cpp_code, compute_code_hash, generate_cpp_source,
format_compile_error, clear_cache, get_cache_info, quick_compile): the user's class
is wrapped in a generated source that exports the six plugin symbols declared in
include/sparq_plugin.h, compiled with g++ against this package's host layer
(libsparq_host.so) and cached by a hash of the code and the build settings.
"""
from __future__ import annotations

import hashlib
import os
import re
import shutil
import subprocess
import tempfile
import warnings
from pathlib import Path
from typing import Optional

PKG = Path(__file__).resolve().parent.parent          # sparqsim_amd/
HOST_INCLUDE = PKG / "csrc" / "host"
ABI_INCLUDE = PKG.parent / "include"
DEFAULT_CACHE = os.path.join(tempfile.gettempdir(), "sparqsim_amd_dynamic_ops")


class CompilationError(Exception):
    """The compiler rejected the generated source; `stderr` holds its condensed output."""

    def __init__(self, message: str, stderr: str = "", returncode: int = -1):
        super().__init__(message)
        self.stderr = stderr
        self.returncode = returncode

    def __str__(self) -> str:
        msg = super().__str__()
        return msg + (f"\n\nCompiler output:\n{self.stderr}" if self.stderr else "")


_FACTORY = """
extern "C" BaseOperator* create_operator({CTOR_PARAMS}) {{
    return new {CLASS_NAME}({CTOR_ARGS});
}}

extern "C" void destroy_operator(BaseOperator* op) {{
    delete op;
}}

extern "C" const char* get_operator_name() {{
    return "{CLASS_NAME}";
}}
"""

_APPLY = """
// op(*state): the engine hands the operator a host copy of the device entries.
// An exception is parked for the caller instead of crossing the C boundary.
extern "C" void apply_operator(BaseOperator* op, SparseState* state) {{
    if (op == nullptr || state == nullptr) return;
    try {{ (*op)(*state); }}
    catch (const std::exception& e) {{ record_plugin_error(e.what()); }}
    catch (...) {{ record_plugin_error("unknown C++ exception"); }}
}}

extern "C" void apply_operator_dag(BaseOperator* op, SparseState* state) {{
    if (op == nullptr || state == nullptr) return;
    try {{ op->dag(*state); }}
    catch (const std::exception& e) {{ record_plugin_error(e.what()); }}
    catch (...) {{ record_plugin_error("unknown C++ exception"); }}
}}

extern "C" const char* get_base_class() {{
    return "{BASE_CLASS}";
}}
"""

_PRELUDE = """#include "basic_components.h"
#include <complex>
#include <vector>

using namespace qram_simulator;

{USER_CPP_CODE}
"""


class CompilerConfig:
    """Compiler command and flags for plugin builds."""

    DEFAULT_TEMPLATE = _PRELUDE + _FACTORY
    PYTHON_TEMPLATE = _PRELUDE + _FACTORY + _APPLY

    def __init__(self, cxx: str = "g++", std: str = "c++17", opt_level: str = "O2",
                 include_paths: Optional[list] = None, lib_paths: Optional[list] = None,
                 libraries: Optional[list] = None, extra_flags: Optional[list] = None,
                 template: Optional[str] = None):
        self.cxx = cxx
        self.std = std
        self.opt_level = opt_level
        self.include_paths = list(include_paths or [])
        self.lib_paths = list(lib_paths or [])
        self.libraries = list(libraries or [])
        self.extra_flags = list(extra_flags or [])
        self.template = template or self.DEFAULT_TEMPLATE

    def get_compile_flags(self) -> list:
        flags = [f"-std={self.std}", f"-{self.opt_level}", "-fPIC", "-shared"]
        flags += [f"-I{p}" for p in self.include_paths]
        flags += [f"-I{HOST_INCLUDE}", f"-I{ABI_INCLUDE}"]
        flags += [f"-L{p}" for p in self.lib_paths]
        flags += [f"-L{PKG}", f"-Wl,-rpath,{PKG}"]
        flags += [f"-l{lib}" for lib in self.libraries]
        flags += ["-lsparq_host", "-lsparq_hip"]
        flags += self.extra_flags
        return flags


def _compiler_version(cxx: str) -> str:
    try:
        r = subprocess.run([cxx, "--version"], capture_output=True, text=True, timeout=5)
        return r.stdout.strip()[:100] if r.returncode == 0 else ""
    except Exception:
        return ""


def compute_code_hash(cpp_code: str, class_name: str, config: CompilerConfig) -> str:
    """16 hex digits over the code, class name, template, flags and compiler version."""
    h = hashlib.sha256()
    for part in (cpp_code, class_name, config.template, config.std, config.opt_level,
                 ",".join(sorted(config.include_paths)), ",".join(config.libraries),
                 ",".join(config.extra_flags), _compiler_version(config.cxx), str(PKG)):
        h.update(part.encode("utf-8"))
        h.update(b"\x00")
    return h.hexdigest()[:16]


def find_project_root() -> Optional[Path]:
    """Directory holding the host headers plugins compile against (this package)."""
    return PKG if (HOST_INCLUDE / "basic_components.h").exists() else None


def generate_cpp_source(cpp_code: str, class_name: str, ctor_params: str = "", ctor_args: str = "",
                        config: Optional[CompilerConfig] = None) -> str:
    cfg = config or CompilerConfig()
    return cfg.template.format(USER_CPP_CODE=cpp_code, CLASS_NAME=class_name,
                               CTOR_PARAMS=ctor_params, CTOR_ARGS=ctor_args)


def format_compile_error(stderr: str, source_path: str) -> str:
    """Condense compiler output: the source path shortened to <source>, the first
    errors and warnings listed."""
    if not stderr:
        return "Unknown compilation error"
    lines = [ln.strip().replace(source_path, "<source>") for ln in stderr.strip().splitlines() if ln.strip()]
    errors = [ln for ln in lines if re.search(r"error:", ln, re.I)]
    warns = [ln for ln in lines if re.search(r"warning:", ln, re.I)]
    out = []
    if errors:
        out.append(f"Errors ({len(errors)}):")
        out += [f"  - {e}" for e in errors[:5]]
        if len(errors) > 5:
            out.append(f"  ... {len(errors) - 5} more error(s)")
    if warns:
        out.append(f"\nWarnings ({len(warns)}):")
        out += [f"  - {w}" for w in warns[:3]]
    if not out:
        out = ["Compiler output:"] + lines[:20]
    return "\n".join(out)


def compile_cpp_code(cpp_code: str, class_name: str, cache_dir: Optional[str] = None, ctor_params: str = "",
                     ctor_args: str = "", config: Optional[CompilerConfig] = None,
                     project_root: Optional[str] = None, verbose: bool = False) -> str:
    """Compile (or reuse from the cache) the plugin library; returns its path."""
    cfg = config or CompilerConfig()
    if find_project_root() is None:
        raise RuntimeError("sparqsim_amd host headers not found")
    if not (PKG / "libsparq_host.so").exists():
        raise RuntimeError("libsparq_host.so is not built (run python sparqsim_amd/build.py)")
    code_hash = compute_code_hash(cpp_code, class_name, cfg)
    cache_dir = cache_dir or DEFAULT_CACHE
    os.makedirs(cache_dir, exist_ok=True)
    lib_path = os.path.join(cache_dir, f"{class_name}_{code_hash}.so")
    if os.path.exists(lib_path):
        if verbose:
            print(f"[compiler] cached: {lib_path}")
        return lib_path
    source_path = os.path.join(cache_dir, f"{class_name}_{code_hash}.cpp")
    with open(source_path, "w", encoding="utf-8") as f:
        f.write(generate_cpp_source(cpp_code, class_name, ctor_params, ctor_args, cfg))
    if not shutil.which(cfg.cxx):
        raise FileNotFoundError(f"C++ compiler not found: {cfg.cxx}")
    tmp_path = lib_path + ".tmp"
    cmd = [cfg.cxx, *cfg.get_compile_flags()[:4], "-o", tmp_path, source_path, *cfg.get_compile_flags()[4:]]
    if verbose:
        print("[compiler] " + " ".join(cmd))
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    except subprocess.TimeoutExpired:
        raise CompilationError("Compilation timed out (over 2 minutes)")
    if r.returncode != 0:
        raise CompilationError(f"Compilation failed (return code: {r.returncode})",
                               stderr=format_compile_error(r.stderr, source_path), returncode=r.returncode)
    os.replace(tmp_path, lib_path)
    return lib_path


def clear_cache(cache_dir: Optional[str] = None) -> int:
    """Delete the cached sources and libraries; returns the number of files removed."""
    cache_dir = cache_dir or DEFAULT_CACHE
    if not os.path.isdir(cache_dir):
        return 0
    count = 0
    for name in os.listdir(cache_dir):
        path = os.path.join(cache_dir, name)
        if os.path.isfile(path):
            try:
                os.remove(path)
                count += 1
            except OSError as e:
                warnings.warn(f"could not remove {path}: {e}")
                raise
    return count


def get_cache_info(cache_dir: Optional[str] = None) -> dict:
    cache_dir = cache_dir or DEFAULT_CACHE
    info = {"cache_dir": cache_dir, "exists": os.path.isdir(cache_dir), "file_count": 0, "so_count": 0,
            "cpp_count": 0, "total_size_mb": 0.0}
    if not info["exists"]:
        return info
    total = 0
    for name in os.listdir(cache_dir):
        path = os.path.join(cache_dir, name)
        if os.path.isfile(path):
            info["file_count"] += 1
            total += os.path.getsize(path)
            info["so_count"] += name.endswith(".so")
            info["cpp_count"] += name.endswith(".cpp")
    info["total_size_mb"] = round(total / (1024 * 1024), 2)
    return info


def quick_compile(class_code: str, class_name: str, verbose: bool = False) -> str:
    return compile_cpp_code(cpp_code=class_code, class_name=class_name, verbose=verbose)
