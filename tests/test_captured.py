"""Captured operator sequences (HIP graphs with the entry count on the device)
and the table-driven Pauli noise channel.

The captured Grover iteration must give the same state as the same operators
run one call at a time: keys bit-exact and amplitudes bit-identical (the
recorded kernels do the same IEEE operations in the same order), after every
replay, with eager work and synchronisations between replays.  The eager path
is itself held to the oracle by tests/test_grover.py.  A replay that breaks
the plan's assumptions is reported, not silently wrong.
"""
import numpy as np
import pytest

from parity import assert_same_state, canonical, random_state
from sparqsim_amd.grover import NoisyGrover, uneven_widths


def test_pauli_masks_cover_every_event(oracle):
    """The per-(iteration, trajectory) masks hold exactly the noise events."""
    g = NoisyGrover(oracle, 12, 8, 8, 6, 0.05)
    x, z = g.pauli_masks()
    x = np.asarray(x, dtype=np.uint64).reshape(6, 12)
    z = np.asarray(z, dtype=np.uint64).reshape(6, 12)
    events = {(k, t, i, kind) for k, ev in enumerate(g.events) for (t, i, kind) in ev}
    assert events
    rebuilt = {(k, t, i, "X") for k in range(6) for t in range(12) for i in range(8) if (int(x[k, t]) >> i) & 1}
    rebuilt |= {(k, t, i, "Z") for k in range(6) for t in range(12) for i in range(8) if (int(z[k, t]) >> i) & 1}
    assert rebuilt == events


@pytest.mark.gpu
def test_pauli_noise_equals_conditioned_gates(ps):
    """PauliNoise from a table equals the X_Bool / Z_Bool insertions it stands
    for, each conditioned on the segment register, bit for bit; the iteration
    counter selects the table row and stops applying past the last iteration."""
    rng = np.random.default_rng(7)
    R, K, A = 9, 3, 6
    x = rng.integers(0, 1 << A, size=R * K, dtype=np.uint64)
    z = rng.integers(0, 1 << A, size=R * K, dtype=np.uint64)
    x[::4] = 0
    z[1::3] = 0
    table = ps.PauliNoiseTable(R, K, x.tolist(), z.tolist())
    st = ps.SparseState()
    ps.AddRegister("seg", ps.UnsignedInteger, 4)(st)
    ps.AddRegister("addr", ps.UnsignedInteger, A)(st)
    ps.AddRegister("other", ps.UnsignedInteger, 5)(st)
    vals, amps = random_state(ps, 600, seed=3)
    st.load(vals, amps)
    ref = ps.SparseState(st)
    noise = ps.PauliNoise("seg", "addr", table)
    advance = ps.AdvancePauliNoise(table)
    for k in range(K + 1):
        noise(st)
        advance(st)
        if k < K:
            for s in range(R):
                xm, zm = int(x[k * R + s]), int(z[k * R + s])
                for i in range(A):
                    if (xm >> i) & 1:
                        ps.X_Bool("addr", i).conditioned_by_value("seg", s)(ref)
                for i in range(A):
                    if (zm >> i) & 1:
                        ps.Z_Bool("addr", i).conditioned_by_value("seg", s)(ref)
        assert_same_state(st.export(), ref.export(), exact_amps=True, what=f"iteration {k}")


def _check_replays(ps, T, A, K, p, widths=None, exact=True):
    """Eager iterations and captured replays side by side, compared after each."""
    g = NoisyGrover(ps, T, A, A, K, p, widths=widths)
    assert sum(len(e) for e in g.events) > 0
    eager = []
    st, it = g._start_batched()
    for k in range(K):
        it(st)
        for t, i, kind in g.events[k]:
            g._noise_op(i, kind).conditioned_by_value("traj", t)(st)
        eager.append(canonical(st.export()))
    st, seq = g.capture_batched()
    assert seq.nodes > 0 and seq.capacity >= st.size()
    for k in range(K):
        seq.launch(1)
        got = st.export()                       # settles the replay (reads the count, checks the flags)
        assert_same_state(got, eager[k], exact_amps=exact, what=f"replay {k}")
    del seq


@pytest.mark.gpu
@pytest.mark.parametrize("T,A,K,p", [(8, 6, 4, 0.08), (33, 9, 3, 0.03)])
def test_captured_grover_equals_eager(ps, T, A, K, p):
    _check_replays(ps, T, A, K, p)


# A segment left with one entry (a width-1 search whose diffusion cancels to
# |0>) is transformed by the captured path with butterflies, a * (1/sqrt 2)^n,
# where the one-call grouped path multiplies once by 2^(-n/2) (hadamard.cpp:26-40):
# equal to a few ulps, so the uneven batch is held to the parity tolerance
# (1e-12 relative), the uniform batch (no such segment) to bit identity.
@pytest.mark.gpu
def test_captured_uneven_grover_equals_eager(ps):
    """The uneven batch: one conditioned Hadamard per search width in each
    diffusion, every one captured (segments of 2^0 .. 2^7 entries)."""
    T, A, K, p = 20, 7, 3, 0.08
    _check_replays(ps, T, A, K, p, widths=uneven_widths(T, A), exact=False)


@pytest.mark.gpu
def test_captured_full_size_batch_equals_eager(ps):
    """The benchmarked batch (1 024 trajectories, 14-bit registers, 40
    iterations, p = 1e-3) from 40 queued replays with no synchronisation in
    between, against the eager run: bit-identical."""
    T, A, K, p = 1024, 14, 40, 1e-3
    ref = NoisyGrover(ps, T, A, A, K, p).run_batched().export()
    got = NoisyGrover(ps, T, A, A, K, p).run_captured().export()
    assert_same_state(got, ref, exact_amps=True, what="full batch")


@pytest.mark.gpu
def test_captured_uneven_full_size_equals_eager(ps):
    T, A, K, p = 1024, 14, 40, 1e-3
    w = uneven_widths(T, A)
    ref = NoisyGrover(ps, T, A, A, K, p, widths=w).run_batched().export()
    got = NoisyGrover(ps, T, A, A, K, p, widths=w).run_captured().export()
    assert_same_state(got, ref, what="uneven batch")


@pytest.mark.gpu
def test_capture_rejects_unsupported_operator(ps):
    """An operator without a captured form fails the capture with ValueError and
    leaves the state usable."""
    st = ps.SparseState()
    ps.AddRegister("a", ps.UnsignedInteger, 4)(st)
    ps.AddRegister("b", ps.Boolean, 1)(st)
    ps.Hadamard_Int_Full("a")(st)

    def body(s):
        ps.Hadamard_Bool("b")(s)

    with pytest.raises(ValueError):
        ps.CapturedSequence(st, body)
    assert st.size() == 16
    ps.Hadamard_Int_Full("a")(st)
    assert st.size() == 1


@pytest.mark.gpu
def test_capture_reports_a_broken_plan(ps):
    """A body whose branching operator sees key bits the plan did not (the flip
    of `b` before the Hadamard makes a bit vary that was constant when the
    sequence was planned) is reported by the next access to the state."""
    st = ps.SparseState()
    ps.AddRegister("a", ps.UnsignedInteger, 5)(st)
    ps.AddRegister("b", ps.UnsignedInteger, 3)(st)
    ps.Hadamard_Int_Full("a")(st)

    def body(s):
        ps.X_Bool("b", 1)(s)
        ps.Hadamard_Int_Full("a")(s)

    seq = ps.CapturedSequence(st, body)
    seq.launch(1)
    with pytest.raises(RuntimeError, match="broke its plan"):
        st.size()
    del seq


@pytest.mark.gpu
def test_capture_layout_must_be_restored(ps):
    st = ps.SparseState()
    ps.AddRegister("a", ps.UnsignedInteger, 4)(st)
    ps.Hadamard_Int_Full("a")(st)

    def body(s):
        ps.AddRegister("extra", ps.UnsignedInteger, 2)(s)

    with pytest.raises(ValueError, match="layout"):
        ps.CapturedSequence(st, body)
