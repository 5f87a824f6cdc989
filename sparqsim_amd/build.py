"""In-tree build of the native parts of sparqsim_amd.

* ``libsparq_hip.so``  -- the HIP engine (csrc/engine/*.hip) for gfx950, the
  C ABI declared in include/sparq_hip.h.
* ``libsparq_host.so`` -- the host C++ layer (csrc/host/sparq.cpp): register
  table, SparseState, the operator classes.  Shared so that the pybind module
  and runtime-compiled operator plugins see one register table.
* ``_core.<ext>.so``   -- the pybind11 bindings (csrc/host/core_py.cpp).

Every library finds its neighbours through an $ORIGIN rpath.
Run ``python sparqsim_amd/build.py`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
ENGINE_SRC = PKG / "csrc" / "engine"
HOST_SRC = PKG / "csrc" / "host"
BUILD = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPARQ_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off keeps every FP64 product/sum separately rounded, matching
# the reference's std::complex<double> arithmetic; packed FP32 ops are disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd[:4]) + " ...")


def engine_lib() -> Path:
    return PKG / "libsparq_hip.so"


def host_lib() -> Path:
    return PKG / "libsparq_host.so"


def core_ext() -> Path:
    return PKG / ("_core" + sysconfig.get_config_var("EXT_SUFFIX"))


def _link(objs, out: Path, libs: list[str]) -> Path:
    tmp = out.with_suffix(".so.tmp")
    _run(["g++", "-shared", *map(str, objs), "-o", str(tmp), f"-L{PKG}", *libs, "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_engine(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(ENGINE_SRC.glob("*.hip"))
    objs = [BUILD / (s.stem + ".o") for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [HIPCC, *HIP_FLAGS, "-c", str(s), "-o", str(o)]) for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    out = engine_lib()
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_host(jobs: int = 8) -> tuple[Path, Path]:
    import pybind11

    BUILD.mkdir(exist_ok=True)
    inc = [f"-I{sysconfig.get_paths()['include']}", f"-I{pybind11.get_include()}"]
    host_obj = BUILD / "sparq.host.o"
    core_obj = BUILD / "core_py.host.o"
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, ["g++", *HOST_FLAGS, "-c", str(HOST_SRC / "sparq.cpp"), "-o", str(host_obj)]),
                ex.submit(_run, ["g++", *HOST_FLAGS, *inc, "-c", str(HOST_SRC / "core_py.cpp"), "-o", str(core_obj)])]
        for f in futs:
            f.result()
    host = _link([host_obj], host_lib(), ["-lsparq_hip"])
    core = _link([core_obj], core_ext(), ["-lsparq_host", "-lsparq_hip"])
    return host, core


def build_all() -> None:
    build_engine()
    build_host()


if __name__ == "__main__":
    build_all()
    print("built", engine_lib(), host_lib(), core_ext())
