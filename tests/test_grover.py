"""Grover composites and the batched noisy-trajectory workload.

CPU: the oracle's Grover iteration is pinned by the closed form sin((2k+1) theta)
for the marked amplitude (and the uniform cos((2k+1) theta)/sqrt(N-1) elsewhere),
and the oracle's batched run (one state keyed by `traj`) equals its per-trajectory
runs entry for entry.  GPU: the engine's batched and plain runs against the
oracle's per-trajectory runs (keys bit-exact, amplitudes to 1e-12 relative).
"""
import math

import numpy as np
import pytest

from parity import assert_same_state, canonical
from sparqsim_amd.grover import (NoisyGrover, grover_memory, grover_success_amplitude,
                                 noise_events, uneven_widths)


def trajectory_of(export, t, skip=1):
    """Entries of trajectory t of a batched export, with the traj column (and,
    for an uneven batch, skip=2, the cls column) dropped."""
    ids, vals, amps = export
    sel = vals[:, 0] == t
    return [i - skip for i in ids[skip:]], vals[sel][:, skip:], amps[sel]


def check_uneven_closed_form(g, export, trajectories, atol=1e-12):
    """Noise-free trajectories of an uneven batch: 2^c entries for width c; a
    search range holding the marked address gives sin((2K+1) theta_c) there and
    cos((2K+1) theta_c)/sqrt(2^c - 1) elsewhere (sin theta_c = 2^(-c/2)); a
    range without it keeps every |amplitude| at 2^(-c/2)."""
    K = g.iterations
    for t in trajectories:
        c = g.widths[t]
        _, v, a = trajectory_of(export, t, skip=2)
        assert len(a) == 1 << c, (t, c, len(a))
        assert np.all(v[:, 0] < (1 << c)) and np.all(v[:, 1] == 0) and np.all(v[:, 2] == g.target), t
        if g.marked < (1 << c) and c > 0:
            theta = math.asin(2.0 ** (-c / 2))
            s, co = math.sin((2 * K + 1) * theta), math.cos((2 * K + 1) * theta)
            marked = v[:, 0] == g.marked
            assert np.allclose(a[marked], s, atol=atol, rtol=0), t
            if c > 0 and (1 << c) > 1:
                assert np.allclose(np.abs(a[~marked]), abs(co) / math.sqrt((1 << c) - 1), atol=atol, rtol=0), t
        else:
            assert np.allclose(np.abs(a), 2.0 ** (-c / 2), atol=atol, rtol=0), t


# ---------------------------------------------------------------- inputs
def test_memory_has_one_marked_word():
    mem, target, marked = grover_memory(10, 10)
    mem = np.array(mem)
    assert len(mem) == 1024 and mem.max() < 1024
    assert np.flatnonzero(mem == target).tolist() == [marked]


def test_noise_schedule_deterministic_and_ordered():
    a = noise_events(64, 10, 14, 1e-2)
    b = noise_events(64, 10, 14, 1e-2)
    assert a == b
    n = sum(len(e) for e in a)
    expected = 64 * 10 * 14 * 2 * 1e-2
    assert abs(n - expected) < 5 * math.sqrt(expected)
    for ev in a:
        assert ev == sorted(ev)
        assert all(0 <= t < 64 and 0 <= i < 14 and k in "XZ" for t, i, k in ev)
    assert noise_events(8, 5, 6, 0.0) == [[] for _ in range(5)]


# ---------------------------------------------------------------- oracle pinning
@pytest.mark.parametrize("A,K", [(4, 1), (6, 3), (8, 5), (10, 7)])
def test_oracle_grover_closed_form(oracle, A, K):
    g = NoisyGrover(oracle, 1, A, A, K, 0.0)
    g.new_plain_registers()
    ids, vals, amps = g.run_plain(0).export()
    n = 1 << A
    assert len(amps) == n and sorted(vals[:, 0].tolist()) == list(range(n))
    assert np.all(vals[:, 1] == 0) and np.all(vals[:, 2] == g.target)
    s = grover_success_amplitude(A, K)
    c = math.sqrt(max(0.0, 1 - s * s) / (n - 1))
    theta = math.asin(2.0 ** (-A / 2))
    c_signed = math.cos((2 * K + 1) * theta) / math.sqrt(n - 1)
    assert abs(abs(c_signed) - c) < 1e-12
    marked = vals[:, 0] == g.marked
    assert np.allclose(amps[marked], s, atol=1e-12, rtol=0)
    assert np.allclose(amps[~marked], c_signed, atol=1e-12, rtol=0)


def test_oracle_batched_equals_plain(oracle):
    g = NoisyGrover(oracle, 6, 6, 6, 4, 0.1)
    assert sum(len(e) for e in g.events) > 0
    eb = g.run_batched().export()
    assert sorted(set(eb[1][:, 0].tolist())) == list(range(6))
    g.new_plain_registers()
    for t in range(6):
        assert_same_state(trajectory_of(eb, t), g.run_plain(t).export(), exact_amps=True, what=f"traj {t}")


def test_oracle_uneven_batched_equals_plain(oracle):
    """An uneven batch (search widths 0..5 over 8 trajectories): the batched run
    equals each trajectory run alone, entry for entry."""
    g = NoisyGrover(oracle, 8, 5, 5, 3, 0.1, widths=uneven_widths(8, 5))
    assert g.widths == [0, 1, 2, 3, 4, 5, 0, 1]
    assert all(i < g.widths[t] for ev in g.events for (t, i, _) in ev)
    assert sum(len(e) for e in g.events) > 0
    eb = g.run_batched().export()
    g.new_plain_registers()
    for t in range(8):
        assert_same_state(trajectory_of(eb, t, skip=2), g.run_plain(t).export(), exact_amps=True, what=f"traj {t}")


def test_oracle_uneven_closed_form(oracle):
    """Noise-free uneven batch with the marked address inside some search
    ranges and outside others: entry counts 1..2^A and the closed forms."""
    A = 6
    g = NoisyGrover(oracle, 2 * (A + 1), A, A, 3, 0.0, widths=uneven_widths(2 * (A + 1), A))
    g.marked = 5                    # in range for widths >= 3, out of range below
    g.memory = [w if i != 5 else g.target for i, w in enumerate(g.memory)]
    g.memory = [w ^ 1 if (w == g.target and i != 5) else w for i, w in enumerate(g.memory)]
    eb = g.run_batched().export()
    check_uneven_closed_form(g, eb, range(2 * (A + 1)))


def test_oracle_noise_changes_the_state(oracle):
    """A plant: dropping the noise changes exactly the trajectories that had events,
    so the batched-vs-plain comparison does see the noise operators."""
    g = NoisyGrover(oracle, 6, 6, 6, 4, 0.02)
    touched = {t for ev in g.events for (t, _, _) in ev}
    assert touched and touched != set(range(6))
    g.new_plain_registers()
    noisy = [canonical(g.run_plain(t).export()) for t in range(6)]
    g.events = [[] for _ in g.events]
    clean = [canonical(g.run_plain(t).export()) for t in range(6)]
    for t in range(6):
        same = np.array_equal(noisy[t][1], clean[t][1]) and np.array_equal(noisy[t][2], clean[t][2])
        assert same == (t not in touched), t


# ---------------------------------------------------------------- engine parity
@pytest.mark.gpu
@pytest.mark.parametrize("T,A,D,K,p", [(6, 6, 6, 4, 0.1), (16, 8, 9, 3, 0.05), (3, 10, 10, 2, 0.02)])
def test_engine_batched_vs_oracle(ps, oracle, T, A, D, K, p):
    ge = NoisyGrover(ps, T, A, D, K, p)
    eb = ge.run_batched().export()
    go = NoisyGrover(oracle, T, A, D, K, p)
    go.new_plain_registers()
    for t in range(T):
        assert_same_state(trajectory_of(eb, t), go.run_plain(t).export(), what=f"traj {t}")


@pytest.mark.gpu
def test_engine_plain_vs_oracle(ps, oracle):
    T, A, D, K, p = 4, 7, 7, 3, 0.08
    ge = NoisyGrover(ps, T, A, D, K, p)
    go = NoisyGrover(oracle, T, A, D, K, p)
    ge.new_plain_registers()
    go.new_plain_registers()
    for t in range(T):
        assert_same_state(ge.run_plain(t).export(), go.run_plain(t).export(), what=f"traj {t}")


@pytest.mark.gpu
def test_engine_grover_closed_form(ps):
    A, K = 12, 6
    g = NoisyGrover(ps, 1, A, A, K, 0.0)
    g.new_plain_registers()
    ids, vals, amps = g.run_plain(0).export()
    assert len(amps) == 1 << A
    s = grover_success_amplitude(A, K)
    theta = math.asin(2.0 ** (-A / 2))
    c = math.cos((2 * K + 1) * theta) / math.sqrt((1 << A) - 1)
    marked = vals[:, 0] == g.marked
    assert np.allclose(amps[marked], s, atol=1e-12, rtol=0)
    assert np.allclose(amps[~marked], c, atol=1e-12, rtol=0)


@pytest.mark.gpu
def test_engine_batched_full_size_properties(ps):
    """The benchmarked batch (1 024 trajectories, 14-bit registers, 40 iterations,
    p = 1e-3): every trajectory keeps norm 1 (the circuit is unitary), and every
    trajectory that drew no noise event holds the closed form -- 2^14 entries,
    sin((2K+1) theta) on the marked address and cos((2K+1) theta)/sqrt(N-1)
    elsewhere -- to 1e-12."""
    T, A, K, p = 1024, 14, 40, 1e-3
    g = NoisyGrover(ps, T, A, A, K, p)
    ids, vals, amps = g.run_batched().export()
    traj = vals[:, 0].astype(np.int64)
    norms = np.bincount(traj, weights=np.abs(amps) ** 2, minlength=T)
    assert np.all(np.abs(norms - 1.0) <= 1e-10), float(np.max(np.abs(norms - 1.0)))
    touched = {t for it in g.events for (t, _, _) in it}
    clean = [t for t in range(T) if t not in touched]
    assert len(clean) > 200                            # (1 - p)^(2 A K) of 1 024: about a third
    s = grover_success_amplitude(A, K)
    theta = math.asin(2.0 ** (-A / 2))
    c = math.cos((2 * K + 1) * theta) / math.sqrt((1 << A) - 1)
    counts = np.bincount(traj, minlength=T)
    for t in clean[:64]:
        _, v, a = trajectory_of((ids, vals, amps), t)
        assert counts[t] == 1 << A, t
        assert np.all(v[:, 1] == 0) and np.all(v[:, 2] == g.target), t   # data uncomputed, search holds the target
        marked = v[:, 0] == g.marked
        assert np.allclose(a[marked], s, atol=1e-12, rtol=0), t
        assert np.allclose(a[~marked], c, atol=1e-12, rtol=0), t


@pytest.mark.gpu
def test_engine_uneven_batched_vs_oracle(ps, oracle):
    """The uneven batch on the engine (one conditioned Hadamard per width in each
    diffusion) against the oracle's per-trajectory runs."""
    T, A, K, p = 14, 6, 3, 0.08
    ge = NoisyGrover(ps, T, A, A, K, p, widths=uneven_widths(T, A))
    eb = ge.run_batched().export()
    go = NoisyGrover(oracle, T, A, A, K, p, widths=uneven_widths(T, A))
    go.new_plain_registers()
    for t in range(T):
        assert_same_state(trajectory_of(eb, t, skip=2), go.run_plain(t).export(), what=f"traj {t}")


@pytest.mark.gpu
def test_engine_uneven_full_size_properties(ps):
    """The benchmarked uneven batch (1 024 trajectories searching 2^0 .. 2^14
    addresses, 40 iterations, p = 1e-3): every trajectory keeps norm 1 and every
    trajectory without a noise event holds its closed form."""
    T, A, K, p = 1024, 14, 40, 1e-3
    g = NoisyGrover(ps, T, A, A, K, p, widths=uneven_widths(T, A))
    export = g.run_batched().export()
    traj = export[1][:, 0].astype(np.int64)
    norms = np.bincount(traj, weights=np.abs(export[2]) ** 2, minlength=T)
    assert np.all(np.abs(norms - 1.0) <= 1e-10), float(np.max(np.abs(norms - 1.0)))
    counts = np.bincount(traj, minlength=T)
    assert counts.min() == 1 and counts.max() == 1 << A
    touched = {t for it in g.events for (t, _, _) in it}
    clean = [t for t in range(T) if t not in touched]
    assert len(clean) > 300
    check_uneven_closed_form(g, export, clean[:96])
