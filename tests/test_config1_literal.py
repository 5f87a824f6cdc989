"""Configuration 1 at its literal size: a, b, d of 16 bits, H(a); b += a;
d ^= M[a]; H(a); ClearZero(1e-14).  b and d determine a, so every group of
the final Hadamard is one entry and the state fans out to 2^32 entries
(96 GiB of keys and amplitudes on the device), none pruned (|a|^2 = 2^-32).
Too large for the oracle or a host copy: checked on the device by its closed
form -- the entry count, the norm (exactly 1: every |a|^2 is 2^-32), and
sampled ranges of entries, each of which must be (a', b, M[b]) at position
a' * 2^16 + b (canonical order: a is the most significant register, and the
one-entry groups' outputs are written target-major) with amplitude exactly
(-1)^popcount(a' & b) 2^-16 (the first H gives 2^-8, the one-entry group
multiplies by 2^-8)."""
import ctypes
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_config1_literal_2p32_final_step(ps):
    w = 16
    rng = np.random.default_rng(0x5EED2026 + 0)
    mem = rng.integers(0, 1 << w, size=1 << w).astype(np.uint64)
    st = ps.SparseState()
    for name in ("a", "b", "d"):
        ps.AddRegister(name, ps.UnsignedInteger, w)(st)
    q = ps.QRAMCircuit_qutrit(w, w, [int(x) for x in mem])
    ps.Hadamard_Int("a", w)(st)
    ps.Add_UInt_UInt_InPlace("a", "b")(st)
    ps.QRAMLoad(q, "a", "d")(st)
    assert st.size() == 1 << w
    t0 = time.perf_counter()
    ps.Hadamard_Int("a", w)(st)
    ps.ClearZero(1e-14)(st)
    n = st.size()
    elapsed = time.perf_counter() - t0
    print(f"config 1 literal final step: {n} entries in {elapsed:.3f} s")
    assert n == 1 << 32
    lib = ctypes.CDLL(ps.ENGINE_LIBRARY)
    h = ctypes.c_void_p(st._handle())
    norm = ctypes.c_double()
    assert lib.sparq_norm2(h, ctypes.byref(norm)) == 0
    assert norm.value == 1.0
    # fields packed most significant first in register order: a, b, d in word 0
    shifts = {"a": 48, "b": 32, "d": 16}
    count = 1 << 14
    starts = [0, (1 << 31) - 5, (1 << 32) - count] + [int(x) for x in rng.integers(0, (1 << 32) - count, size=5)]
    for first in starts:
        keys = np.zeros(count, dtype=np.uint64)
        amps = np.zeros(2 * count, dtype=np.float64)
        rc = lib.sparq_state_read_range(h, ctypes.c_uint64(first), ctypes.c_uint64(count),
                                        keys.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                        amps.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert rc == 0
        pos = np.arange(first, first + count, dtype=np.uint64)
        a = (keys >> np.uint64(shifts["a"])) & np.uint64(0xFFFF)
        b = (keys >> np.uint64(shifts["b"])) & np.uint64(0xFFFF)
        d = (keys >> np.uint64(shifts["d"])) & np.uint64(0xFFFF)
        assert np.array_equal(a, pos >> np.uint64(16)), first
        assert np.array_equal(b, pos & np.uint64(0xFFFF)), first
        assert np.array_equal(d, mem[b.astype(np.int64)]), first
        assert np.all(keys & np.uint64(0xFFFF) == 0), first
        sign = 1.0 - 2.0 * (np.bitwise_count(a & b) & 1)
        assert np.array_equal(amps[0::2], sign * 2.0 ** -16), first
        assert np.all(amps[1::2] == 0.0), first
    del st
    ps.release_cached_memory()
