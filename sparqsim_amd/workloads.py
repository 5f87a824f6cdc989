"""Seeded synthetic workloads shared by the benchmark, the tests and the CPU baseline.

All randomness comes from splitmix64 over a counter (base seed 0x5EED2026, plus
a per-workload seed offset), so the CPU oracle and the GPU engine see identical
inputs without depending on any library RNG.
"""
from __future__ import annotations

import numpy as np

BASE_SEED = 0x5EED2026

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n outputs of splitmix64 started at `seed` (outputs for states seed+G, seed+2G, ...)."""
    with np.errstate(over="ignore"):
        x = np.arange(1, n + 1, dtype=np.uint64) * _G + np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        z = (x ^ (x >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def qram_memory(address_bits: int, data_bits: int, seed_offset: int) -> list[int]:
    """2^address_bits words, uniform in [0, 2^data_bits)."""
    words = splitmix64(BASE_SEED + seed_offset, 1 << address_bits)
    if data_bits < 64:
        words = words & np.uint64((1 << data_bits) - 1)
    return words.tolist()


class QRAMInterference:
    """Single-state QRAM interference on registers addr (A bits) and data (D bits):
    H(addr); data ^= M[addr]; phase flip on the parity of data (one Z per data
    bit); uncompute the load; H(addr).  The registers must already exist.

    `entry_ops` sums max(N_in, N_out) over the reference operators of one pass
    (2^A entries for each of the 4 + D operators)."""

    def __init__(self, mod, qram, address_bits: int, data_bits: int):
        self.mod = mod
        self.ops = [mod.Hadamard_Int("addr", address_bits), mod.QRAMLoad(qram, "addr", "data")]
        self.ops += [mod.Phase_Bool("data", i, np.pi) for i in range(data_bits)]
        self.ops += [mod.QRAMLoad(qram, "addr", "data"), mod.Hadamard_Int("addr", address_bits)]
        self.entry_ops = (1 << address_bits) * (4 + data_bits)
        # one native call per pass where the module offers it (the engine)
        seq = getattr(mod, "OperatorSequence", None)
        self._seq = seq(self.ops) if seq is not None else None

    def __call__(self, state):
        if self._seq is not None:
            self._seq(state)
            return state
        for op in self.ops:
            op(state)
        return state


def qram_interference(mod, qram, address_bits: int, data_bits: int):
    """Run one pass on a fresh |0> state; returns (state, entry-ops)."""
    circ = QRAMInterference(mod, qram, address_bits, data_bits)
    st = mod.SparseState()
    circ(st)
    return st, circ.entry_ops
