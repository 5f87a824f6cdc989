"""QDA over the sparse block encoding composed from the CKS oracles: the block
encoding against the matrix it encodes, the engine against the oracle, and the
solve against conjugate gradients."""
from __future__ import annotations

import numpy as np
import pytest

from parity import assert_same_state
from sparqsim_amd import cks, qda_sparse


def _encoded_column(mod, nq, i):
    st, walk, enc, A, kappa = qda_sparse.setup(mod, nq, 30.0, 8)
    st = walk.CreateSys()
    mod.Init_Unsafe(walk.j, i)(st)
    enc.apply(st)
    n = 1 << nq
    zero = [walk.b1, walk.k, walk.b2, walk.j_comp, walk.k_comp, walk.data_offset]
    out, _ = cks.select_dense(mod, st, zero, [walk.j], n)
    q = np.array(enc.matrix.elements, dtype=float).reshape(n, 8)
    cols = np.array(enc.matrix.sparsity).reshape(n, 8)
    ref = np.zeros(n)
    for r in range(n):
        for s, c in enumerate(cols[r]):
            if c == i:
                ref[r] += q[r, s] / (8 * (2.0 ** 32 - 1))
    return out, ref


def test_sparse_block_encoding_extracts_the_matrix(oracle):
    # <r, 0| T^dag S T |i, 0> = A_ri / (nnz (2^ds - 1)), up to the 32-bit rotation angle
    # (cos(2 pi theta_q) differs from sqrt(v/Amax) by at most 2 pi 2^-32 relative)
    for i in (0, 7, 31):
        out, ref = _encoded_column(oracle, 5, i)
        assert np.max(np.abs(out - ref)) < 1e-9
        assert np.max(np.abs(out.imag)) < 1e-9


def test_sparse_qda_short_schedule_converges(oracle):
    met, *_ = qda_sparse.solve(oracle, 5, 200)
    assert met["fidelity"] > 0.99 and met["kappa_encoded"] == pytest.approx(100.6, rel=1e-2)


@pytest.mark.gpu
def test_engine_sparse_block_encoding_matches(ps, oracle):
    for i in (0, 7, 31):
        g, ref = _encoded_column(ps, 5, i)
        o, _ = _encoded_column(oracle, 5, i)
        assert np.max(np.abs(g - o)) <= 1e-12 * np.max(np.abs(o))
        assert np.max(np.abs(g - ref)) < 1e-9


@pytest.mark.gpu
def test_engine_sparse_qda_steps_match_oracle(ps, oracle):
    outs = []
    for mod in (ps, oracle):
        met, st, walk, blk = qda_sparse.solve(mod, 5, 2000, stop_after=6)
        outs.append(st.export())
    assert_same_state(outs[0], outs[1], what="6 QDA steps")


@pytest.mark.gpu
def test_engine_sparse_qda_solve_converges(ps):
    met, *_ = qda_sparse.solve(ps, 5, 2000)
    assert met["fidelity"] > 0.9999 and met["res_dir"] < 5e-4


# ---------------------------------------------------------------- complex Hermitian matrix (phase oracle)
def _encoded_hermitian_column(mod, nq, i, dag=False):
    st, walk, enc, A, kappa = qda_sparse.setup(mod, nq, hermitian=True, kappa_encoded=100.0)
    st = walk.CreateSys()
    mod.Init_Unsafe(walk.j, i)(st)
    (enc.apply_dag if dag else enc.apply)(st)
    n = 1 << nq
    zero = [walk.b1, walk.k, walk.b2, walk.j_comp, walk.k_comp, walk.data_offset]
    out, _ = cks.select_dense(mod, st, zero, [walk.j], n)
    D = A.toarray()
    col = (D.conj().T if dag else D)[:, i]
    # entries enter the walk as floor(|a| 2^32) / (2^32 - 1) of the largest; phases as 32-bit turns
    return out, col * 2.0 ** 32 / (8 * (2.0 ** 32 - 1)), D


def test_hermitian_matrix_construction():
    """8 slots per row at random columns, A = A^H, kappa of the block-encoded matrix ~100."""
    v, s, ph, A = qda_sparse.hermitian_for_encoded_kappa(64, 100.0)
    D = A.toarray()
    assert np.allclose(D, D.conj().T) and np.count_nonzero(D) == 64 * 8
    assert np.all(np.linalg.eigvalsh(D) > 0) and abs(np.max(np.abs(D)) - 1.0) < 1e-12
    assert np.count_nonzero(np.abs(D.imag) > 1e-3) > 64 * 5          # genuinely complex
    assert abs(1.0 / cks.walk_spectrum(A, 32, 8)[0] / 100.0 - 1) < 0.02
    assert np.allclose(np.array(v) * np.exp(1j * ph), D[np.repeat(np.arange(64), 8), s])


def test_hermitian_block_encoding_extracts_the_matrix(oracle):
    """<r, 0| T^dag S P T |i, 0> = A_ri / (8 Amax) for a complex Hermitian A, and
    the adjoint T^dag P^dag S T encodes A^H (= A): 32-bit angle and phase
    quantisation bound the error by ~1e-9 of the largest entry."""
    for i in (0, 7, 31):
        for dag in (False, True):
            out, ref, _ = _encoded_hermitian_column(oracle, 5, i, dag)
            assert np.max(np.abs(out - ref)) < 1e-9, (i, dag)


def test_hermitian_qda_short_schedule(oracle):
    """200 steps move the walk most of the way (fidelity 0.845 against conjugate
    gradients; the complete 2000-step schedule reaches 0.99996, res_dir 3.5e-3,
    on the oracle -- run on the engine in the GPU test below)."""
    met, *_ = qda_sparse.solve(oracle, 5, 200, hermitian=True, kappa_encoded=100.0)
    assert met["fidelity"] > 0.8 and met["hermitian"]


@pytest.mark.gpu
def test_engine_hermitian_block_encoding_matches(ps, oracle):
    for i in (0, 7, 31):
        for dag in (False, True):
            g, ref, _ = _encoded_hermitian_column(ps, 5, i, dag)
            o, _, _ = _encoded_hermitian_column(oracle, 5, i, dag)
            assert np.max(np.abs(g - o)) <= 1e-12 * np.max(np.abs(o))
            assert np.max(np.abs(g - ref)) < 1e-9


@pytest.mark.gpu
def test_engine_hermitian_qda_steps_match_oracle(ps, oracle):
    outs = []
    for mod in (ps, oracle):
        met, st, walk, blk = qda_sparse.solve(mod, 5, 2000, stop_after=4, hermitian=True, kappa_encoded=100.0)
        outs.append(st.export())
    assert_same_state(outs[0], outs[1], what="4 Hermitian QDA steps")


@pytest.mark.gpu
def test_engine_hermitian_qda_solve_converges(ps):
    met, *_ = qda_sparse.solve(ps, 5, 2000, hermitian=True, kappa_encoded=100.0)
    assert met["fidelity"] > 0.9999 and met["res_dir"] < 1e-2, met
