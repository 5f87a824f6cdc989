"""sparqsim_amd -- MI355X-native engine for SparQ's sparse-state operator-apply path.

The public surface mirrors ``pysparq._core`` for the operator-apply path:
``System`` (process-global register table), ``SparseState`` (device-resident,
structure-of-arrays keys + FP64 complex amplitudes) and the operator classes
(``Hadamard_Int``, ``QRAMLoad``, ``Add_UInt_UInt`` ...), each callable as
``op(state)`` / ``op.dag(state)``.  All compute runs in the HIP engine
(``libsparq_hip.so``); there is no CPU fallback.
"""
from __future__ import annotations

import os as _os

_here = _os.path.dirname(_os.path.abspath(__file__))

try:
    from . import _core  # noqa: F401  (native extension, built in-tree)
except ImportError as exc:  # pragma: no cover - exercised only when unbuilt
    raise ImportError(
        "sparqsim_amd native extension is not built or failed to load "
        f"({exc}); run `python sparqsim_amd/build.py`"
    ) from exc

from ._core import *  # noqa: F401,F403,E402
from ._core import System, SparseState  # noqa: E402

ENGINE_LIBRARY = _os.path.join(_here, "libsparq_hip.so")

__all__ = [name for name in dir(_core) if not name.startswith("_")]

from .rir import RIRError, RIRResult, load_rir, run_rir, run_rir_file  # noqa: E402

__all__ += ["RIRError", "RIRResult", "load_rir", "run_rir", "run_rir_file"]
