"""The reference's on-disk formats: text state files and the solver benches' artifacts.

* State lines (System::to_string(precision), SparQ/src/basic_components.cpp:497-515):
  the amplitude as ``"{: 1.<p>f} {:+1.<p>f}i"`` followed by ``|name:value>`` for
  every active register in id order (value = the stored unsigned integer).
  ``print_state_to_file`` writes a ``Print State To File:`` header and one line
  per entry (SparQ/include/system_operations.h:703-725).
* ``StateLoad`` reads such a file back for the QDA registers (main_reg, anc_UA,
  anc_1..anc_4; SparQ/src/system_operations.cpp:388-482): entry lines are the
  ones containing ``|...>``, the amplitude is the first ``re im i`` pair, each
  register value is found by name; the loaded state must be normalised to
  1e-14.  ``load_state_file`` reads every ``|name:value>`` of a line instead.
  Where the reference exits the process (missing file, no entry lines) these
  raise.
* Solver artifacts (PoiseuilleQDABench.cu:170-293, CKSPoiseuilleBench.cu:305-420):
  ``residual_curve.csv`` (``step,res_dir,res_opt,success_prob,fidelity``, %.17g)
  and ``qda_summary_cpp.json``.

Works with the engine module and with the CPU oracle.
"""
from __future__ import annotations

import json
import math
import os
import re

import numpy as np

_AMP = re.compile(r"(-?\d*\.?\d+)\s+([-+]?\d*\.?\d+)i")
_REG = re.compile(r"\|([^:|>]+):([^>]+)>")


def _fmt_amp(a: complex, precision: int) -> str:
    return format(a.real, f" 1.{precision}f") + " " + format(a.imag, f"+1.{precision}f") + "i"


def state_to_lines(mod, state, precision: int = 16) -> list[str]:
    """One line per entry in the state's entry order."""
    if precision < 2 or precision > 18:
        raise ValueError("precision must be in [2, 18]")
    ids, vals, amps = state.export()
    names = [mod.System.name_of(i) for i in ids]
    out = []
    for row, a in zip(vals, amps):
        out.append(_fmt_amp(complex(a), precision) + "".join(f"|{nm}:{int(v)}>" for nm, v in zip(names, row)))
    return out


def print_state_to_file(mod, state, filename: str, precision: int = 16) -> None:
    if precision == 0:
        raise ValueError("precision must not be 0")
    lines = state_to_lines(mod, state, precision)
    with open(filename, "w") as f:
        f.write("Print State To File:\n")
        for line in lines:
            f.write(line + "\n")


def _entry_lines(filename: str):
    if not os.path.exists(filename):
        raise FileNotFoundError(f"file `{filename}` does not exist.")
    branch = False
    out = []
    with open(filename) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line:
                continue
            if not branch:
                branch = re.search(r"\|[^>]*>", line) is not None
            if branch:
                out.append(line)
    if not branch:
        raise ValueError(f"file `{filename}` is not a record of quantum state.")
    return out


def _parse_amp(line: str) -> complex:
    m = _AMP.search(line)
    if m is None:
        raise ValueError(f"no amplitude in line: {line!r}")
    return complex(float(m.group(1)), float(m.group(2)))


def _build(mod, regs_values, amps, check_norm: float | None):
    ids = [i for i in range(mod.System.register_count()) if mod.System.status_of(i)]
    col = {i: c for c, i in enumerate(ids)}
    vals = np.zeros((len(amps), len(ids)), dtype=np.uint64)
    for name, column in regs_values.items():
        vals[:, col[mod.System.get_id(name)]] = column
    amps = np.asarray(amps, dtype=np.complex128)
    if check_norm is not None:
        total = float(np.sum(np.abs(amps) ** 2))
        if not abs(total - 1.0) < check_norm:
            raise RuntimeError(f"State is not normalized (factor {total!r}).")
    st = mod.SparseState()
    st.load(vals, amps)
    return st


class StateLoad:
    """Reader of a printed QDA state (system_operations.h:640-700, .cpp:388-482)."""

    def __init__(self, mod, main_reg="main_reg", anc_UA="anc_UA", anc_1="anc_1", anc_2="anc_2", anc_3="anc_3",
                 anc_4="anc_4"):
        self.mod = mod
        self.regs = [main_reg, anc_UA, anc_1, anc_2, anc_3, anc_4]

    def __call__(self, filename: str):
        lines = _entry_lines(filename)
        amps, cols = [], {r: [] for r in self.regs}
        for line in lines:
            amps.append(_parse_amp(line))
            for r in self.regs:
                m = re.search(r"\|" + re.escape(r) + r":([^>]+)>", line)
                if m is None:
                    raise ValueError(f"register {r} missing in line: {line!r}")
                cols[r].append(int(m.group(1)))
        return _build(self.mod, {r: np.array(v, dtype=np.uint64) for r, v in cols.items()}, amps, 1e-14)


def load_state_file(mod, filename: str, check_norm: float | None = 1e-14):
    """Read every |name:value> of each entry line into the registers of those names."""
    lines = _entry_lines(filename)
    amps, cols = [], {}
    for k, line in enumerate(lines):
        amps.append(_parse_amp(line))
        for name, v in _REG.findall(line):
            cols.setdefault(name, [0] * len(lines))[k] = int(v)
    return _build(mod, {r: np.array(v, dtype=np.uint64) for r, v in cols.items()}, amps, check_norm)


def _num(x):
    return None if x is None or (isinstance(x, float) and math.isnan(x)) else x


def write_residual_curve(path: str, rows) -> None:
    """rows: iterables of (step, res_dir, res_opt, success_prob, fidelity)."""
    with open(path, "w") as f:
        f.write("step,res_dir,res_opt,success_prob,fidelity\n")
        for step, rd, ro, sp, fid in rows:
            f.write(f"{int(step)},{rd:.17g},{ro:.17g},{sp:.17g},{fid:.17g}\n")


def write_summary_json(path: str, summary: dict) -> None:
    with open(path, "w") as f:
        json.dump({k: _num(v) for k, v in summary.items()}, f, indent=2)
        f.write("\n")
