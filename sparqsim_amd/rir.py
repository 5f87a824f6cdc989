"""RIR JSON programs (the register-level IR of QECC.Lang, schema 0.1-0.3) on the engine.

Same entry points as the reference's ``pysparq.rir`` (PySparQ/pysparq/rir.py):
``load_rir`` validates a document, ``run_rir`` / ``run_rir_file`` execute it on
a fresh device state and return an ``RIRResult`` (entry registers and the
amplitude of every basis state, keyed by the registers' integer values).

How a program runs here.  The module graph is flattened on the host into a
stream of bit-level instructions: a ``Call`` binds the callee's registers to
bit views of the caller's (``Ref`` spans), ``Repeat`` replays its body,
``Adjoint`` walks its body backwards with every instruction inverted, and
``Control`` adds a (bits, value) requirement.  Each instruction then becomes
engine operators on one SparseState, with its requirements as per-entry
conditions -- a bit that must be 1 is ``conditioned_by_bit``, a bit that must
be 0 ``conditioned_by_clear_bit`` (SPARQ_COND_BIT_CLEAR), so no control bit is
ever flipped around an operator and a QRAM query under a zero control on one
of its own address bits reads the cell it names.  Whole-register operands use
the register-level operators (``Assign`` for xor, ``Swap_General_General``,
``Add_ConstUInt_InPlace``, ``QRAMLoad``); slices and mixed views go bit by bit
(controlled X, ``Swap_Bool_Bool``, an increment ripple of multi-controlled X,
and a QRAM query through scratch registers).  Single-qubit gates are
``Rot_Bool`` with the gate's 2x2 matrix (the engine runs diagonal and
off-diagonal matrices as per-entry phases).  Consecutive per-entry operators
fuse into one device launch.
"""
from __future__ import annotations

import cmath
import json
import math
from dataclasses import dataclass
from pathlib import Path
from typing import Any, Iterator, Mapping, Sequence

import numpy as np

__all__ = ["RIRError", "RIRResult", "load_rir", "run_rir", "run_rir_file"]

SCHEMA_VERSIONS = ("0.1", "0.2", "0.3")
WORD_BITS = 64
DENSE_QRAM_MAX_ADDRESS = 24      # larger QRAM resources run cell by cell


class RIRError(ValueError):
    """A malformed RIR document or a violated execution limit."""


@dataclass(frozen=True)
class RIRResult:
    """``registers``: the entry module's registers as (name, width), in
    declaration order; ``amplitudes``: {tuple of their integer values: amplitude}."""

    registers: tuple
    amplitudes: dict

    def statevector(self, max_qubits: int = 20) -> list:
        """Dense vector over the entry registers, the first register in the lowest bits."""
        total = sum(w for _, w in self.registers)
        if total > max_qubits:
            raise RIRError(f"dense state conversion exceeds {max_qubits} qubits")
        out = [0j] * (1 << total)
        for values, amp in self.amplitudes.items():
            index, shift = 0, 0
            for (_, w), v in zip(self.registers, values):
                index |= v << shift
                shift += w
            out[index] = amp
        return out


# ---------------------------------------------------------------- documents
def load_rir(source) -> dict:
    """Decode (a mapping, a JSON string or a file path) and check the document's frame."""
    if isinstance(source, Mapping):
        doc = dict(source)
    else:
        text = str(source)
        if not text.lstrip().startswith("{"):
            text = Path(text).read_text(encoding="utf-8")
        try:
            doc = json.loads(text)
        except json.JSONDecodeError as exc:
            raise RIRError(f"invalid RIR JSON: {exc}") from exc
    if not isinstance(doc, dict) or doc.get("tag") != "Program":
        raise RIRError("the root of an RIR document must be a Program node")
    if doc.get("version") not in SCHEMA_VERSIONS:
        raise RIRError(f"unsupported RIR schema version: {doc.get('version')!r}")
    modules = doc.get("modules")
    if not isinstance(modules, list) or not modules:
        raise RIRError("an RIR program declares at least one module")
    names = [m.get("name") if isinstance(m, dict) else None for m in modules]
    if None in names or len(set(names)) != len(names):
        raise RIRError("RIR module names must be present and unique")
    if doc.get("entry") not in names:
        raise RIRError(f"entry module {doc.get('entry')!r} is not declared")
    return doc


# ---------------------------------------------------------------- gates
def gate_matrix(name: str, angle) -> tuple:
    """Row-major 2x2 matrix (u00, u01, u10, u11) of a named single-qubit gate."""
    if name == "h":
        r = 1.0 / math.sqrt(2.0)
        return (r, r, r, -r)
    if name == "x":
        return (0, 1, 1, 0)
    if name == "y":
        return (0, -1j, 1j, 0)
    if name == "z":
        return (1, 0, 0, -1)
    if name in ("s", "t", "phase"):
        lam = {"s": math.pi / 2, "t": math.pi / 4}.get(name, angle)
        if lam is None:
            raise RIRError(f"gate {name!r} needs an angle")
        return (1, 0, 0, cmath.exp(1j * float(lam)))
    if angle is None and name in ("rz", "rx", "ry"):
        raise RIRError(f"gate {name!r} needs an angle")
    if name == "rz":
        a = float(angle)
        return (cmath.exp(-0.5j * a), 0, 0, cmath.exp(0.5j * a))
    if name in ("rx", "ry"):
        c, s = math.cos(float(angle) / 2), math.sin(float(angle) / 2)
        return (c, -1j * s, -1j * s, c) if name == "rx" else (c, -s, s, c)
    raise RIRError(f"unknown single-qubit gate: {name!r}")


def adjoint_matrix(m: tuple) -> tuple:
    c = [complex(v).conjugate() for v in m]
    return (c[0], c[2], c[1], c[3])


# ---------------------------------------------------------------- flattening
# A bit is (native register name, bit index); a view is a list of bits, least
# significant first.  Instructions yielded by Flattener.run:
#   ("enter", native, kind, width)  /  ("exit", native, width)
#   ("gate", matrix, view, ctrl)    ("gphase", theta, ctrl)
#   ("xor", src, dst, ctrl)         ("swap", a, b, ctrl)
#   ("add", view, value, ctrl)      ("load", resource, address, data, ctrl)
# ctrl is a tuple of (bit, required value) pairs (no bit twice); an
# instruction whose controls contradict each other is dropped.
def native_bits(native: str, width: int) -> list:
    if width <= WORD_BITS:
        return [(native, b) for b in range(width)]
    return [(f"{native}__w{b // WORD_BITS}", b % WORD_BITS) for b in range(width)]


def native_words(native: str, width: int) -> list:
    if width <= WORD_BITS:
        return [(native, width)] if width else []
    return [(f"{native}__w{o // WORD_BITS}", min(WORD_BITS, width - o)) for o in range(0, width, WORD_BITS)]


def view_of(ref: Mapping, bound: Mapping) -> list:
    bits = []
    for span in ref["parts"]:
        reg = bound[span["register"]]
        start, width = int(span["start"]), int(span["width"])
        if start < 0 or start + width > len(reg):
            raise RIRError(f"span {span['register']}[{start}:{start + width}] outside its register")
        bits.extend(reg[start:start + width])
    return bits


class Flattener:
    def __init__(self, doc: Mapping, max_steps: int):
        self.modules = {m["name"]: m for m in doc["modules"]}
        self.max_steps = max_steps
        self.locals = 0
        self._cost: dict = {}

    def _body(self, name: str):
        body = self.modules[name].get("body")
        if body is None:
            raise RIRError(f"module {name!r} has no implementation")
        return body

    def cost(self, nodes) -> int:
        """Expanded size: one per primitive or load, one per call / control /
        adjoint plus its body, a repeat's count times its body; stops counting
        past the budget."""
        total = 0
        for node in nodes:
            tag = node["tag"]
            if tag in ("Primitive", "Load"):
                c = 1
            elif tag == "Call":
                name = node["module"]
                if name not in self._cost:
                    self._cost[name] = self.cost(self._body(name))
                c = 1 + self._cost[name]
            elif tag == "Repeat":
                c = int(node["count"]) * self.cost(node["body"])
            elif tag in ("Control", "Adjoint"):
                c = 1 + self.cost(node["body"])
            else:
                raise RIRError(f"unknown RIR node tag: {tag!r}")
            total += c
            if total > self.max_steps:
                return total
        return total

    def check_budget(self, entry: str) -> None:
        if self.cost(self._body(entry)) > self.max_steps:
            raise RIRError(f"RIR expansion exceeds the step budget ({self.max_steps})")

    def run(self, name: str, bound: dict, resources: dict, ctrl: tuple = (), inverse: bool = False) -> Iterator[tuple]:
        module = self.modules[name]
        body = self._body(name)
        bound = dict(bound)
        scratch = []
        for reg in module.get("locals") or []:
            self.locals += 1
            native, width = f"rir_local{self.locals}", int(reg["type"]["width"])
            bound[reg["name"]] = native_bits(native, width)
            scratch.append((native, width))
            yield ("enter", native, reg["type"]["kind"], width)
        yield from self._nodes(body, bound, resources, ctrl, inverse)
        for native, width in reversed(scratch):
            yield ("exit", native, width)

    def _nodes(self, nodes, bound, resources, ctrl, inverse) -> Iterator[tuple]:
        for node in (reversed(nodes) if inverse else nodes):
            tag = node["tag"]
            if tag == "Primitive":
                yield from self._primitive(node, bound, ctrl, inverse)
            elif tag == "Load":      # XOR load: its own inverse
                yield ("load", resources[node["resource"]], view_of(node["address"], bound),
                       view_of(node["data"], bound), ctrl)
            elif tag == "Call":
                callee = self.modules[node["module"]]
                args = {f["name"]: view_of(a, bound) for f, a in zip(callee["registers"], node["arguments"])}
                res = {f["name"]: resources[a] for f, a in zip(callee.get("resources") or [], node.get("resources") or [])}
                yield from self.run(node["module"], args, res, ctrl, inverse)
            elif tag == "Repeat":
                for _ in range(int(node["count"])):
                    yield from self._nodes(node["body"], bound, resources, ctrl, inverse)
            elif tag == "Control":
                bits, value = view_of(node["register"], bound), int(node["value"])
                merged = dict(ctrl)
                dead = False
                for i, bit in enumerate(bits):
                    want = (value >> i) & 1
                    if merged.setdefault(bit, want) != want:
                        dead = True           # contradictory controls: the body never acts
                if not dead:
                    yield from self._nodes(node["body"], bound, resources, tuple(merged.items()), inverse)
            elif tag == "Adjoint":
                yield from self._nodes(node["body"], bound, resources, ctrl, not inverse)
            else:
                raise RIRError(f"unknown RIR node tag: {tag!r}")

    def _primitive(self, node, bound, ctrl, inverse) -> Iterator[tuple]:
        op = node["op"]
        views = [view_of(r, bound) for r in node.get("operands") or []]
        if op == "gphase":
            yield ("gphase", -float(node["angle"]) if inverse else float(node["angle"]), ctrl)
        elif op in ("xor", "swap"):
            if len(views) != 2 or len(views[0]) != len(views[1]):
                raise RIRError(f"{op} needs two operands of equal width")
            yield (op, views[0], views[1], ctrl)
        elif op == "add_const":
            yield ("add", views[0], -int(node["value"]) if inverse else int(node["value"]), ctrl)
        else:
            m = gate_matrix(op, node.get("angle"))
            yield ("gate", adjoint_matrix(m) if inverse else m, views[0], ctrl)


def check_memory(resources: Sequence[Mapping], memory) -> dict:
    """{resource: {address: nonzero word}} from word lists or sparse mappings."""
    memory = {} if memory is None else memory
    if not isinstance(memory, Mapping):
        raise RIRError("QRAM contents must be a mapping keyed by resource name")
    specs = {r["name"]: r["type"] for r in resources}
    if set(memory) != set(specs):
        raise RIRError("memory must give exactly one entry per declared QRAM resource")
    out = {}
    for name, spec in specs.items():
        aw, dw = int(spec["address_width"]), int(spec["data_width"])
        raw = memory[name]
        cells = {}
        for a, v in (raw.items() if isinstance(raw, Mapping) else enumerate(raw)):
            if not isinstance(a, int) or not 0 <= a < (1 << aw):
                raise RIRError(f"QRAM {name!r}: address out of range: {a!r}")
            if not isinstance(v, int) or not 0 <= v < (1 << dw):
                raise RIRError(f"QRAM {name!r}: word out of range: {v!r}")
            if v:
                cells[a] = v
        out[name] = cells
    return out


# ---------------------------------------------------------------- execution
class Executor:
    """Runs flattened instructions on one engine state."""

    def __init__(self, ps, doc: Mapping, memory, max_states: int):
        self.ps = ps
        self.entry = {m["name"]: m for m in doc["modules"]}[doc["entry"]]
        self.max_states = max_states
        self.cells = check_memory(self.entry.get("resources") or [], memory)
        self.specs = {r["name"]: r["type"] for r in self.entry.get("resources") or []}
        self.state = ps.SparseState()
        self.width: dict = {}     # native register -> width
        self.qram: dict = {}
        self.scratch = 0

    # registers
    def add(self, native: str, kind: str, width: int) -> None:
        ps = self.ps
        storage = {"bits": ps.General, "uint": ps.UnsignedInteger, "sint": ps.SignedInteger,
                   "rational": ps.Rational}.get(kind)
        if storage is None:
            raise RIRError(f"unknown register kind: {kind!r}")
        for word, w in native_words(native, width):
            ps.AddRegister(word, storage if width <= WORD_BITS else ps.General, w)(self.state)
            self.width[word] = w

    def remove(self, native: str, width: int, require_zero: bool) -> None:
        for word, _ in reversed(native_words(native, width)):
            if require_zero and self.state.uniform_value(word) != 0:
                raise RIRError(f"local register not uncomputed: {native}")
            self.ps.RemoveRegister(word)(self.state)
            del self.width[word]

    def setup(self) -> dict:
        bound = {}
        for i, reg in enumerate(self.entry["registers"]):
            width = int(reg["type"]["width"])
            native = f"rir_reg{i}"
            if width:
                self.add(native, reg["type"]["kind"], width)
            bound[reg["name"]] = native_bits(native, width)
        for name, spec in self.specs.items():
            aw, dw = int(spec["address_width"]), int(spec["data_width"])
            if aw <= DENSE_QRAM_MAX_ADDRESS and dw <= WORD_BITS:
                words = [0] * (1 << aw)
                for a, v in self.cells[name].items():
                    words[a] = v
                self.qram[name] = self.ps.QRAMCircuit_qutrit(aw, dw, words)
        return bound

    # helpers
    def whole(self, view) -> str | None:
        """The native register a view covers exactly, bit 0 first, else None."""
        if not view:
            return None
        name = view[0][0]
        if self.width.get(name) != len(view) or view != [(name, b) for b in range(len(view))]:
            return None
        return name

    def apply(self, op, ctrl=(), extra_set=()) -> None:
        ones = [bit for bit, v in ctrl if v] + list(extra_set)
        zeros = [bit for bit, v in ctrl if not v]
        if ones:
            op.conditioned_by_bit(ones)
        if zeros:
            op.conditioned_by_clear_bit(zeros)
        op(self.state)

    def check_size(self) -> None:
        if self.state.size() > self.max_states:
            raise RIRError(f"execution exceeds the sparse-state budget ({self.max_states})")

    # instructions
    def run(self, ins: tuple) -> None:
        kind, ps = ins[0], self.ps
        if kind == "enter":
            _, native, rkind, width = ins
            if width:
                self.add(native, rkind, width)
        elif kind == "exit":
            _, native, width = ins
            if width:
                self.remove(native, width, require_zero=True)
        elif kind == "gate":
            _, m, view, ctrl = ins
            for reg, bit in view:
                if m == (0, 1, 1, 0):
                    self.apply(ps.X_Bool(reg, bit), ctrl)
                else:
                    self.apply(ps.Rot_Bool(reg, bit, list(m)), ctrl)
                    self.check_size()
        elif kind == "gphase":
            _, theta, ctrl = ins
            self.apply(ps.GlobalPhase(cmath.exp(1j * theta)), ctrl)
        elif kind == "xor":
            _, src, dst, ctrl = ins
            a, b = self.whole(src), self.whole(dst)
            if a and b and a != b:
                self.apply(ps.Assign(a, b), ctrl)
            else:
                for s, d in zip(src, dst):
                    self.apply(ps.X_Bool(*d), ctrl, extra_set=[s])
        elif kind == "swap":
            _, x, y, ctrl = ins
            a, b = self.whole(x), self.whole(y)
            if a and b and a != b:
                self.apply(ps.Swap_General_General(a, b), ctrl)
            else:
                for p, q in zip(x, y):
                    if p != q:
                        self.apply(ps.Swap_Bool_Bool(p[0], p[1], q[0], q[1]), ctrl)
        elif kind == "add":
            _, view, value, ctrl = ins
            self.add_const(view, value, ctrl)
        elif kind == "load":
            _, res, addr, data, ctrl = ins
            self.load(res, addr, data, ctrl)
        else:
            raise RIRError(f"unknown instruction {kind!r}")

    def add_const(self, view, value: int, ctrl) -> None:
        width = len(view)
        if not width:
            return
        value %= 1 << width
        reg = self.whole(view)
        if reg is not None:
            self.apply(self.ps.Add_ConstUInt_InPlace(reg, value), ctrl)
            return
        # += 2^k for each set bit k: bit i (i > k) flips when bits k..i-1 are all 1,
        # highest i first, then bit k flips
        for k in range(width):
            if (value >> k) & 1:
                for i in range(width - 1, k, -1):
                    self.apply(self.ps.X_Bool(*view[i]), ctrl, extra_set=view[k:i])
                self.apply(self.ps.X_Bool(*view[k]), ctrl)

    def load(self, res: str, addr, data, ctrl) -> None:
        spec = self.specs[res]
        aw, dw = int(spec["address_width"]), int(spec["data_width"])
        if len(addr) != aw or len(data) != dw:
            raise RIRError(f"QRAM {res!r} query: address / data views of {len(addr)} / {len(data)} bits "
                           f"for a {aw}-bit address and {dw}-bit words")
        ps = self.ps
        a_reg, d_reg = self.whole(addr), self.whole(data)
        qram = self.qram.get(res)
        if qram is not None and a_reg and d_reg:
            self.apply(ps.QRAMLoad(qram, a_reg, d_reg), ctrl)
            return
        # through scratch registers: copy the address, query into fresh data bits,
        # xor them onto the data view under the controls, unquery, uncopy
        self.scratch += 1
        ta, td = f"rir_qaddr{self.scratch}", f"rir_qdata{self.scratch}"
        self.add(ta, "uint", aw)
        ta_bits = native_bits(ta, aw)
        for s, d in zip(addr, ta_bits):
            self.apply(ps.X_Bool(*d), (), extra_set=[s])
        if qram is not None:
            self.add(td, "uint", dw)
            td_bits = native_bits(td, dw)
            query = ps.QRAMLoad(qram, ta, td)
            query(self.state)
            for s, d in zip(td_bits, data):
                self.apply(ps.X_Bool(*d), ctrl, extra_set=[s])
            query(self.state)
            self.remove(td, dw, require_zero=False)
        else:   # cell by cell: data ^= word where the copied address equals the cell's
            words = native_words(ta, aw)
            for cell, word in sorted(self.cells[res].items()):
                for j in range(dw):
                    if (word >> j) & 1:
                        op = ps.X_Bool(*data[j])
                        op.conditioned_by_value([(w, (cell >> (WORD_BITS * k)) & ((1 << wb) - 1))
                                                 for k, (w, wb) in enumerate(words)])
                        self.apply(op, ctrl)
        for s, d in zip(reversed(addr), reversed(ta_bits)):
            self.apply(ps.X_Bool(*d), (), extra_set=[s])
        self.remove(ta, aw, require_zero=False)

    def result(self, bound: dict) -> RIRResult:
        regs = tuple((r["name"], int(r["type"]["width"])) for r in self.entry["registers"])
        ids, vals, amps = self.state.export()
        col = {self.ps.System.name_of(i): k for k, i in enumerate(ids)}
        n = len(amps)
        keys = []
        for name, _ in regs:
            v = [0] * n
            for off, (native, bit) in enumerate(bound[name]):
                bits = ((vals[:, col[native]] >> np.uint64(bit)) & np.uint64(1)).astype(np.int64)
                for e in np.flatnonzero(bits):
                    v[e] |= 1 << off
            keys.append(v)
        out: dict = {}
        for e in range(n):
            k = tuple(keys[r][e] for r in range(len(regs)))
            out[k] = out.get(k, 0j) + complex(amps[e])
        return RIRResult(regs, {k: a for k, a in out.items() if abs(a) > 1e-15})


def run_rir(document, memory=None, *, max_steps: int = 1_000_000, max_states: int = 65_536) -> RIRResult:
    """Execute an RIR document on a fresh device state.

    ``memory`` gives each declared QRAM resource's contents, as a word list or a
    sparse {address: word} mapping.  ``max_steps`` bounds the expanded
    instruction count (checked before anything runs), ``max_states`` the
    number of basis states."""
    import sparqsim_amd as ps

    doc = load_rir(document)
    if ps.System.get_activated_register_size():
        raise RIRError("the process-global register table is not empty; finish the other run first")
    flat = Flattener(doc, max_steps)
    flat.check_budget(doc["entry"])
    ex = None
    try:
        ex = Executor(ps, doc, memory, max_states)
        bound = ex.setup()
        resources = {name: name for name in ex.specs}
        for ins in flat.run(doc["entry"], bound, resources):
            ex.run(ins)
        return ex.result(bound)
    finally:
        ex = None              # the state goes before the register table is cleared
        ps.System.clear()


def run_rir_file(path, memory=None, *, max_steps: int = 1_000_000, max_states: int = 65_536) -> RIRResult:
    """Load an RIR JSON file and execute it."""
    return run_rir(load_rir(path), memory, max_steps=max_steps, max_states=max_states)
