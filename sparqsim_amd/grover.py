"""Grover search composites and the batched noisy-trajectory workload.

The composites follow the reference operator by operator
(SparQ_Algorithm/src/grover.cpp:45-76, PySparQ/pysparq/algorithms/grover.py:26-230)
and run on any module exposing the pysparq operator surface -- the HIP engine
(`sparqsim_amd`) or the CPU oracle.

Batching: many independent trajectories live in ONE sparse state whose first
register, `traj`, holds the trajectory id.  Register ids order the packed key,
so `traj` is the most significant field: every grouping operator
(Hadamard_Int_Full groups by all registers except its target) keeps the
trajectories apart, every pointwise operator touches all of them in one launch,
and a trajectory's own noise is an operator conditioned on `traj == t`.  Each
trajectory starts with amplitude 1, so its amplitudes are exactly those of the
same trajectory run alone (the operators are linear and never mix `traj`).
"""
from __future__ import annotations

import math

import numpy as np

from .workloads import BASE_SEED, splitmix64

GROVER_SEED_OFFSET = 3


# ---------------------------------------------------------------- composites
class GroverOracle:
    """Phase oracle: load, compare with the search register, flip, uncompute
    (grover.cpp:45-62; PySparQ grover.py:26-95).  ZeroConditionalPhaseFlip on
    `compare_equal` flips the NON-matching entries, i.e. minus the textbook oracle."""

    def __init__(self, mod, qram, addr: str, data: str, search: str):
        self.mod, self.qram, self.addr, self.data, self.search = mod, qram, addr, data, search
        self.load = mod.QRAMLoad(qram, addr, data)

    def __call__(self, st):
        m = self.mod
        self.load(st)
        m.AddRegister("compare_less", m.Boolean, 1)(st)
        m.AddRegister("compare_equal", m.Boolean, 1)(st)
        cmp = m.Compare_UInt_UInt(self.data, self.search, "compare_less", "compare_equal")
        cmp(st)
        m.ZeroConditionalPhaseFlip(["compare_equal"])(st)
        cmp(st)
        m.RemoveRegister("compare_equal")(st)
        m.RemoveRegister("compare_less")(st)
        self.load(st)


class DiffusionOperator:
    """HPH: Hadamard_Int_Full, phase flip of |0>, Hadamard_Int_Full (grover.cpp:64-69).
    With `digits`, the Hadamards act on the low `digits` address bits only
    (Hadamard_Int(addr, digits); none for 0): the diffusion of a search over the
    first 2^digits addresses, whose upper address bits stay 0."""

    def __init__(self, mod, addr: str, digits: int | None = None):
        if digits is None:
            self.hs = [mod.Hadamard_Int_Full(addr)]
        else:
            self.hs = [mod.Hadamard_Int(addr, digits)] if digits > 0 else []
        self.flip = mod.ZeroConditionalPhaseFlip([addr])

    def __call__(self, st):
        for h in self.hs:
            h(st)
        self.flip(st)
        for h in self.hs:
            h(st)


class ClassDiffusion(DiffusionOperator):
    """The diffusion of a batch whose trajectories search address ranges of
    different widths: register `cls` holds each trajectory's width c, and the
    Hadamard on the low c address bits is conditioned on cls == c, one launch
    per width present.  The phase flip of addr == 0 is shared (upper address
    bits are 0 in every trajectory)."""

    def __init__(self, mod, addr: str, widths):
        self.hs = [mod.Hadamard_Int(addr, c).conditioned_by_value("cls", c) for c in sorted(set(widths)) if c > 0]
        self.flip = mod.ZeroConditionalPhaseFlip([addr])


class GroverOperator:
    """One Grover iteration: oracle then diffusion (grover.cpp:71-76)."""

    def __init__(self, mod, qram, addr: str, data: str, search: str, diffusion: DiffusionOperator | None = None):
        self.oracle = GroverOracle(mod, qram, addr, data, search)
        self.diffusion = diffusion or DiffusionOperator(mod, addr)

    def __call__(self, st):
        self.oracle(st)
        self.diffusion(st)


# ---------------------------------------------------------------- workload inputs
def grover_memory(address_bits: int, data_bits: int, seed_offset: int = GROVER_SEED_OFFSET):
    """2^address_bits words with exactly one equal to the target; every other word is
    uniform in [0, 2^data_bits) and differs from the target.  Returns (memory, target, marked)."""
    n = 1 << address_bits
    mask = np.uint64((1 << data_bits) - 1)
    r = splitmix64(BASE_SEED + seed_offset, n + 2)
    target = int(r[0] & mask)
    marked = int(r[1] % np.uint64(n))
    words = r[2:] & mask
    words = np.where(words == np.uint64(target), words ^ np.uint64(1), words)
    words[marked] = np.uint64(target)
    return words.tolist(), target, marked


def noise_events(n_traj: int, iterations: int, address_bits: int, p: float,
                 seed_offset: int = GROVER_SEED_OFFSET):
    """Seeded Pauli insertions: after iteration k of trajectory t, for each address qubit
    i in ascending order, X(addr, i) with probability p and then Z(addr, i) with
    probability p.  Trajectory t draws from splitmix64 seeded (BASE_SEED + seed_offset)
    ^ ((t + 1) << 32), two uniforms per (k, i).  Returns events[k] = [(t, i, 'X'|'Z'), ...]
    ordered by t, then i, X before Z."""
    per = iterations * address_bits * 2
    events = [[] for _ in range(iterations)]
    for t in range(n_traj):
        u = (splitmix64((BASE_SEED + seed_offset) ^ ((t + 1) << 32), per) >> np.uint64(11)).astype(np.float64)
        hit = (u * (1.0 / (1 << 53))) < p
        for j in np.flatnonzero(hit):
            j = int(j)
            k, rest = divmod(j, address_bits * 2)
            i, kind = divmod(rest, 2)
            events[k].append((t, i, "Z" if kind else "X"))
    for ev in events:
        ev.sort(key=lambda e: (e[0], e[1], e[2]))
    return events


def grover_success_amplitude(address_bits: int, k: int) -> float:
    """Amplitude of the marked address after k noise-free iterations from the uniform
    superposition: sin((2k+1) theta), sin(theta) = 2^(-address_bits/2)."""
    theta = math.asin(2.0 ** (-address_bits / 2))
    return math.sin((2 * k + 1) * theta)


# ---------------------------------------------------------------- noisy trajectories
def uneven_widths(n_traj: int, address_bits: int) -> list[int]:
    """Search widths of the uneven batch: trajectory t searches the first
    2^(t mod (address_bits + 1)) addresses, so its state holds 1 to
    2^address_bits entries."""
    return [t % (address_bits + 1) for t in range(n_traj)]


class NoisyGrover:
    """n_traj Grover trajectories on (addr, data, search) of the given widths, each with
    its own seeded Pauli X/Z insertions after every iteration.

    `run_batched()` keeps all trajectories in one state keyed by `traj`;
    `run_plain(t)` runs trajectory t alone in its own state.  Both start from
    |0>: search := target (Init_Unsafe), Hadamard_Int_Full(addr), then
    `iterations` x (GroverOperator, noise).

    `widths` (one per trajectory, 0..address_bits) makes the batch uneven:
    trajectory t searches only the first 2^widths[t] addresses (Hadamards on
    the low widths[t] address bits, noise on those bits only), so its state
    holds 2^widths[t] entries; the batch carries a register `cls` with each
    trajectory's width for the per-width diffusion."""

    def __init__(self, mod, n_traj: int, address_bits: int, data_bits: int, iterations: int, p: float,
                 seed_offset: int = GROVER_SEED_OFFSET, widths=None):
        self.mod, self.n_traj, self.A, self.D = mod, n_traj, address_bits, data_bits
        self.iterations, self.p = iterations, p
        self.memory, self.target, self.marked = grover_memory(address_bits, data_bits, seed_offset)
        self.widths = None if widths is None else [int(c) for c in widths]
        if self.widths is not None and (len(self.widths) != n_traj or not all(0 <= c <= address_bits for c in self.widths)):
            raise ValueError("widths needs one value in [0, address_bits] per trajectory")
        self.events = noise_events(n_traj, iterations, address_bits, p, seed_offset)
        if self.widths is not None:
            self.events = [[e for e in ev if e[1] < self.widths[e[0]]] for ev in self.events]
        self.traj_bits = max(1, (n_traj - 1).bit_length())
        self.qram = None

    def _registers(self, batched: bool):
        m = self.mod
        m.System.clear()
        st = m.SparseState()
        if batched:
            m.AddRegister("traj", m.UnsignedInteger, self.traj_bits)(st)
            if self.widths is not None:
                m.AddRegister("cls", m.UnsignedInteger, max(1, self.A.bit_length()))(st)
        m.AddRegister("addr", m.UnsignedInteger, self.A)(st)
        m.AddRegister("data", m.UnsignedInteger, self.D)(st)
        m.AddRegister("search", m.UnsignedInteger, self.D)(st)
        self.qram = m.QRAMCircuit_qutrit(self.A, self.D, self.memory)
        return st

    def _noise_op(self, i, kind):
        m = self.mod
        return m.X_Bool("addr", i) if kind == "X" else m.Z_Bool("addr", i)

    def _start_batched(self):
        """The batch state after the first Hadamard, and the Grover iteration operator."""
        m = self.mod
        st = self._registers(batched=True)
        uneven = self.widths is not None
        vals = np.zeros((self.n_traj, 5 if uneven else 4), dtype=np.uint64)
        vals[:, 0] = np.arange(self.n_traj, dtype=np.uint64)
        if uneven:
            vals[:, 1] = np.asarray(self.widths, dtype=np.uint64)
        st.load(vals, np.ones(self.n_traj, dtype=complex))
        m.Init_Unsafe("search", self.target)(st)
        if uneven:
            diffusion = ClassDiffusion(m, "addr", self.widths)
            for h in diffusion.hs:
                h(st)
        else:
            diffusion = None
            m.Hadamard_Int_Full("addr")(st)
        return st, GroverOperator(m, self.qram, "addr", "data", "search", diffusion)

    def run_batched(self):
        st, g = self._start_batched()
        for k in range(self.iterations):
            g(st)
            for t, i, kind in self.events[k]:
                self._noise_op(i, kind).conditioned_by_value("traj", t)(st)
        return st

    def pauli_masks(self):
        """The noise events as [iteration][trajectory] X and Z bit masks on addr
        (row-major lists for PauliNoiseTable)."""
        x = np.zeros((self.iterations, self.n_traj), dtype=np.uint64)
        z = np.zeros((self.iterations, self.n_traj), dtype=np.uint64)
        for k, ev in enumerate(self.events):
            for t, i, kind in ev:
                (x if kind == "X" else z)[k, t] |= np.uint64(1 << i)
        return x.ravel().tolist(), z.ravel().tolist()

    def capture_batched(self):
        """The batch state after the first Hadamard and a CapturedSequence of one
        whole iteration: the Grover operator, then this iteration's noise from a
        device table of per-(iteration, trajectory) Pauli masks, then the
        table's iteration counter.  The noise is the same X_Bool / Z_Bool
        insertions run_batched applies one conditioned operator at a time."""
        m = self.mod
        st, g = self._start_batched()
        x, z = self.pauli_masks()
        self.noise_table = m.PauliNoiseTable(self.n_traj, self.iterations, x, z)
        noise = m.PauliNoise("traj", "addr", self.noise_table)
        advance = m.AdvancePauliNoise(self.noise_table)

        def iteration(s):
            g(s)
            noise(s)
            advance(s)

        return st, m.CapturedSequence(st, iteration)

    def run_captured(self):
        """run_batched with every iteration replayed from one captured graph: no
        host synchronisation between iterations."""
        st, seq = self.capture_batched()
        seq.launch(self.iterations)
        seq.sync()
        del seq
        return st

    def new_plain_registers(self):
        """Register table of the plain (one trajectory per state) runs."""
        self._registers(batched=False)

    def run_plain(self, t: int):
        """Trajectory t alone; call new_plain_registers() once before a series of runs."""
        m = self.mod
        st = m.SparseState()
        m.Init_Unsafe("search", self.target)(st)
        diffusion = None if self.widths is None else DiffusionOperator(m, "addr", self.widths[t])
        if diffusion is None:
            m.Hadamard_Int_Full("addr")(st)
        else:
            for h in diffusion.hs:
                h(st)
        g = GroverOperator(m, self.qram, "addr", "data", "search", diffusion)
        for k in range(self.iterations):
            g(st)
            for tt, i, kind in self.events[k]:
                if tt == t:
                    self._noise_op(i, kind)(st)
        return st

    def entry_ops(self, sizes_per_iteration: int) -> int:
        """Entry-operations of one iteration, counted as the reference operators applied
        (QRAMLoad x2, AddRegister x2, Compare x2, phase flip, RemoveRegister x2,
        Hadamard_Int_Full x2, phase flip = 12) times the entries they see."""
        return 12 * sizes_per_iteration
