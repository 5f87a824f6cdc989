"""The engine's own per-kernel profile: launch counts, time (HIP events on each
launch's stream) and algorithmic HBM bytes, as accumulated by the engine while
profiling is on (``sparq_prof_*`` in include/sparq_hip.h), plus the sort-path
counters.  Used by bench.py and by the solver CLIs' ``--kernel-profile``.
"""
from __future__ import annotations

import ctypes
import json

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E


class ProfRecord(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_uint64),
                ("total_ms", ctypes.c_double), ("total_bytes", ctypes.c_double)]


_LIB = None


def engine_library():
    global _LIB
    if _LIB is None:
        from . import ENGINE_LIBRARY

        _LIB = ctypes.CDLL(ENGINE_LIBRARY)
    return _LIB


def _check(rc, what):
    if rc != 0:
        lib = engine_library()
        lib.sparq_last_error.restype = ctypes.c_char_p
        raise RuntimeError(f"{what} failed: {lib.sparq_last_error().decode()}")


def reset() -> None:
    _check(engine_library().sparq_prof_reset(), "sparq_prof_reset")


def enable(on: bool = True) -> None:
    _check(engine_library().sparq_prof_enable(1 if on else 0), "sparq_prof_enable")


def report(capacity: int = 128, host: bool = False) -> list[dict]:
    """One record per kernel name since the last reset, largest total time first:
    launches, ms, bytes (algorithmic), us per launch, GB/s and fraction of the
    HBM peak on those bytes.  host=True returns the host-side waits instead
    (names starting with "host:", e.g. the readbacks: count and blocked ms)."""
    arr = (ProfRecord * capacity)()
    cnt = ctypes.c_uint32(0)
    _check(engine_library().sparq_prof_report(arr, capacity, ctypes.byref(cnt)), "sparq_prof_report")
    out = []
    for i in range(min(cnt.value, capacity)):
        r = arr[i]
        if r.name.decode().startswith("host:") != host:
            continue
        ms, byt, launches = r.total_ms, r.total_bytes, int(r.launches)
        gbs = byt / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        out.append({"name": r.name.decode(), "launches": launches, "ms": ms, "bytes": byt,
                    "us_per_launch": ms / launches * 1e3 if launches else 0.0,
                    "bytes_per_launch": byt / launches if launches else 0.0,
                    "GBs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS})
    return sorted(out, key=lambda r: -r["ms"])


class KernelProfile:
    """Context manager: on entry (or at ``start()`` when ``start_now`` is false)
    resets and enables the profiler and the sort-path counters; on exit
    ``records`` and ``sort_stats`` hold the result."""

    def __init__(self, ps, start_now: bool = True):
        self.ps, self.start_now, self.records, self.host_waits, self.sort_stats = ps, start_now, [], [], {}

    def start(self) -> None:
        reset()
        self.ps.sort_stats(reset=True)
        enable(True)

    def __enter__(self):
        if self.start_now:
            self.start()
        return self

    def __exit__(self, *exc):
        enable(False)
        self.records = report()
        self.host_waits = report(host=True)
        self.sort_stats = self.ps.sort_stats(reset=True)
        return False

    def write(self, path: str, **extra) -> None:
        total_ms = sum(r["ms"] for r in self.records)
        with open(path, "w") as f:
            json.dump({"kernel_ms": total_ms, "sort_stats": self.sort_stats, "host_waits": self.host_waits,
                       "kernels": self.records, **extra}, f, indent=1)
