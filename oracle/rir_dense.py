"""Dense statevector checker for RIR programs -- TEST INFRASTRUCTURE ONLY:
used by tests/ to check sparqsim_amd.rir, never by the product.

An independent statement of the semantics the reference's RIR interpreter
gives a program (PySparQ/pysparq/rir.py:219-427 expansion, :429-650
execution): every bit of every register is one qubit of a dense vector (bit p
of index i is qubit p); Call binds the callee's registers to bit views of the
caller's, Repeat replays, Adjoint runs its body backwards with each operation
inverted, Control requires bit values (contradictory controls make the body a
no-op); a module's locals start at 0 and must be 0 again at its end.
Primitives: single-qubit gates (h, x, y, z, s, t, phase, rz, rx, ry) on every
bit of their operand, gphase, xor / swap of equal-width views, add_const
(modulo the view's width), and Load: data ^= memory[address].  Controls are
evaluated on the input basis state.  Pinned by the reference's own expected
amplitudes (PySparQ/test/test_rir.py) in tests/test_rir.py.
"""
from __future__ import annotations

import cmath
import math

import numpy as np


class DenseRIR:
    def __init__(self, doc, memory=None):
        self.mods = {m["name"]: m for m in doc["modules"]}
        self.entry = self.mods[doc["entry"]]
        mem = memory or {}
        self.mem = {}
        for r in self.entry.get("resources") or []:
            raw = mem[r["name"]]
            self.mem[r["name"]] = dict(raw.items()) if isinstance(raw, dict) else dict(enumerate(raw))
        self.nq = 0
        self.psi = np.ones(1, dtype=complex)
        self.regs = {}
        for r in self.entry["registers"]:
            self.regs[r["name"]] = self._alloc(int(r["type"]["width"]))

    def _alloc(self, width):
        pos = list(range(self.nq, self.nq + width))
        self.psi = np.concatenate([self.psi, np.zeros(len(self.psi) * ((1 << width) - 1), dtype=complex)])
        self.nq += width
        return pos

    def _free(self, pos):
        # the freed qubits are the highest ones (locals are nested)
        assert pos == list(range(self.nq - len(pos), self.nq))
        keep = 1 << (self.nq - len(pos))
        if np.any(np.abs(self.psi[keep:]) > 1e-12):
            raise ValueError("local register not uncomputed")
        self.psi = self.psi[:keep].copy()
        self.nq -= len(pos)

    # views and controls
    @staticmethod
    def view(ref, env):
        out = []
        for s in ref["parts"]:
            out.extend(env[s["register"]][s["start"]:s["start"] + s["width"]])
        return out

    def _ctl_mask(self, ctl):
        idx = np.arange(len(self.psi))
        ok = np.ones(len(self.psi), dtype=bool)
        for q, v in ctl.items():
            ok &= ((idx >> q) & 1) == v
        return idx, ok

    def _values(self, idx, bits):
        v = np.zeros_like(idx)
        for k, q in enumerate(bits):
            v |= ((idx >> q) & 1) << k
        return v

    def _set(self, idx, bits, v):
        out = idx.copy()
        for k, q in enumerate(bits):
            out = (out & ~(1 << q)) | (((v >> k) & 1) << q)
        return out

    def _permute(self, ok, idx, new):
        dest = np.where(ok, new, idx)
        out = np.zeros_like(self.psi)
        np.add.at(out, dest, self.psi)
        self.psi = out

    # operations
    @staticmethod
    def matrix(op, angle, inverse):
        r = 1 / math.sqrt(2)
        if op == "h":
            m = np.array([[r, r], [r, -r]], dtype=complex)
        elif op == "x":
            m = np.array([[0, 1], [1, 0]], dtype=complex)
        elif op == "y":
            m = np.array([[0, -1j], [1j, 0]])
        elif op == "z":
            m = np.diag([1, -1]).astype(complex)
        elif op in ("s", "t", "phase"):
            lam = {"s": math.pi / 2, "t": math.pi / 4}.get(op, angle)
            m = np.diag([1, cmath.exp(1j * lam)])
        elif op == "rz":
            m = np.diag([cmath.exp(-0.5j * angle), cmath.exp(0.5j * angle)])
        elif op == "rx":
            c, s = math.cos(angle / 2), math.sin(angle / 2)
            m = np.array([[c, -1j * s], [-1j * s, c]])
        elif op == "ry":
            c, s = math.cos(angle / 2), math.sin(angle / 2)
            m = np.array([[c, -s], [s, c]], dtype=complex)
        else:
            raise ValueError(op)
        return m.conj().T if inverse else m

    def gate(self, m, q, ctl):
        idx, ok = self._ctl_mask(ctl)
        lo = ok & (((idx >> q) & 1) == 0)
        i0 = idx[lo]
        i1 = i0 | (1 << q)
        a0, a1 = self.psi[i0].copy(), self.psi[i1].copy()
        self.psi[i0] = m[0, 0] * a0 + m[0, 1] * a1
        self.psi[i1] = m[1, 0] * a0 + m[1, 1] * a1

    def run(self, name, env, res, ctl, inverse):
        mod = self.mods[name]
        env = dict(env)
        allocated = []
        for r in mod.get("locals") or []:
            env[r["name"]] = self._alloc(int(r["type"]["width"]))
            allocated.append(env[r["name"]])
        self.nodes(mod["body"], env, res, ctl, inverse)
        for pos in reversed(allocated):
            self._free(pos)

    def nodes(self, nodes, env, res, ctl, inverse):
        for n in (nodes[::-1] if inverse else nodes):
            t = n["tag"]
            if t == "Call":
                callee = self.mods[n["module"]]
                e2 = {f["name"]: self.view(a, env) for f, a in zip(callee["registers"], n["arguments"])}
                r2 = {f["name"]: res[a] for f, a in zip(callee.get("resources") or [], n.get("resources") or [])}
                self.run(n["module"], e2, r2, ctl, inverse)
            elif t == "Repeat":
                for _ in range(int(n["count"])):
                    self.nodes(n["body"], env, res, ctl, inverse)
            elif t == "Control":
                bits, val = self.view(n["register"], env), int(n["value"])
                c2 = dict(ctl)
                if any(c2.setdefault(q, (val >> k) & 1) != (val >> k) & 1 for k, q in enumerate(bits)):
                    continue
                self.nodes(n["body"], env, res, c2, inverse)
            elif t == "Adjoint":
                self.nodes(n["body"], env, res, ctl, not inverse)
            elif t == "Load":
                a, d = self.view(n["address"], env), self.view(n["data"], env)
                mem = self.mem[res[n["resource"]]]
                idx, ok = self._ctl_mask(ctl)
                av = self._values(idx, a)
                word = np.array([mem.get(int(x), 0) for x in av], dtype=np.int64)
                self._permute(ok, idx, self._set(idx, d, self._values(idx, d) ^ word))
            elif t == "Primitive":
                self.primitive(n, env, ctl, inverse)
            else:
                raise ValueError(t)

    def primitive(self, n, env, ctl, inverse):
        op = n["op"]
        views = [self.view(r, env) for r in n.get("operands") or []]
        if op == "gphase":
            theta = -n["angle"] if inverse else n["angle"]
            _, ok = self._ctl_mask(ctl)
            self.psi[ok] *= cmath.exp(1j * theta)
        elif op == "xor":
            idx, ok = self._ctl_mask(ctl)
            s, d = views
            self._permute(ok, idx, self._set(idx, d, self._values(idx, d) ^ self._values(idx, s)))
        elif op == "swap":
            idx, ok = self._ctl_mask(ctl)
            a, b = views
            va, vb = self._values(idx, a), self._values(idx, b)
            self._permute(ok, idx, self._set(self._set(idx, a, vb), b, va))
        elif op == "add_const":
            idx, ok = self._ctl_mask(ctl)
            v = views[0]
            k = -int(n["value"]) if inverse else int(n["value"])
            self._permute(ok, idx, self._set(idx, v, (self._values(idx, v) + k) % (1 << len(v))))
        else:
            m = self.matrix(op, n.get("angle"), inverse)
            for q in views[0]:
                self.gate(m, q, ctl)

    def result(self):
        self.run(self.entry["name"], self.regs, {r["name"]: r["name"] for r in self.entry.get("resources") or []}, {}, False)
        idx = np.arange(len(self.psi))
        out = {}
        for i in np.flatnonzero(np.abs(self.psi) > 1e-15):
            out[tuple(int(self._values(idx[i:i + 1], self.regs[r["name"]])[0]) for r in self.entry["registers"])] = complex(self.psi[i])
        return out
