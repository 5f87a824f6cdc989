"""QDA linear solve with a sparse block encoding composed from the CKS sparse-access oracles.

The reference's QDA takes any block encoding of A (Walk_s / Block_Encoding_Hs,
SparQ_Algorithm/include/DiscreteAdiabatic/qda_fundamental.h:53-446); its
general-matrix encoding is a dense QRAM tree over n^2 entries
(Block_Encoding_via_QRAM), which does not reach 2^16.  Here the block encoding is
the Szegedy-walk product U_A = T^dag S T of the CKS walk
(SparQ_Algorithm/include/hamiltonian_simulation.h:1276-1773): with every walk
ancilla (reg_b1, col_id, reg_b2, j_comp, k_comp, data_offset) at 0,
<i, 0| T^dag S T |j, 0> = A_ij / (nnz_col (2^ds - 1)) for a symmetric matrix with
non-negative entries (the walk's rotation encodes |A_ij|), so U_A block-encodes
A / (nnz_col Amax) on the row register row_id.  Walk_s then runs unchanged: its
controlled enc_A is U_A with the two control registers AND-ed into every
operator of T, S and T^dag, and its final reflection covers the walk ancillas
together with anc_2 and anc_3.  The right-hand side is |b> = H|0> on the low
log2(n) bits of row_id (uniform), the state preparation the reference QDA uses.
Runs on the engine or the CPU oracle.
"""
from __future__ import annotations

import math

import numpy as np

from . import cks, qda


class BlockEncodingSparse:
    """U_A = T^dag S T over the CKS walk registers (self-adjoint: S is a swap),
    or U_A = T^dag S P T with the phase oracle P of a complex Hermitian matrix
    (then U_A^dag = T^dag P^dag S T is applied for the adjoint)."""

    def __init__(self, mod, walk: cks.QuantumWalkNSteps, phase: cks.PhaseOracle | None = None):
        self.mod, self.walk, self.phase = mod, walk, phase
        self.ancillas = [walk.b1, walk.k, walk.b2, walk.j_comp, walk.k_comp, walk.data_offset]

    def _swaps(self, st, cond):
        w, m = self.walk, self.mod
        for a, b in ((w.j, w.k), (w.b1, w.b2), (w.j_comp, w.k_comp)):
            qda.ctl(m.Swap_General_General(a, b), cond)(st)

    def apply(self, st, cond: qda.Cond | None = None):
        t = self.walk._T()
        t(st, cond)
        if self.phase is not None:
            self.phase.apply(st, cond)
        self._swaps(st, cond)
        t.dag(st, cond)

    def apply_dag(self, st, cond: qda.Cond | None = None):
        t = self.walk._T()
        t(st, cond)
        self._swaps(st, cond)
        if self.phase is not None:
            self.phase.apply_dag(st, cond)
        t.dag(st, cond)


class WalkSSparse:
    """One step W(s) of the QDA walk (qda_fundamental.h:344-446) over the sparse block encoding."""

    def __init__(self, mod, enc_A: BlockEncodingSparse, anc_1, anc_2, anc_3, anc_4, s, kappa, p):
        self.mod = mod
        self.fs = qda.compute_fs(s, kappa, p)
        main = enc_A.walk.j
        n_digits = enc_A.walk.get_init_size()
        enc_b = lambda st: mod.Hadamard_Int(main, n_digits)(st)   # |b> on the low log2(n) bits of row_id
        self.enc_Hs = qda.BlockEncodingHs(mod, enc_A, main, None, anc_1, anc_2, anc_3, anc_4, self.fs, enc_b=enc_b)
        self.refl = list(enc_A.ancillas) + [anc_2, anc_3]

    def __call__(self, st):
        m = self.mod
        self.enc_Hs.apply(st)
        m.Reflection_Bool(self.refl, False)(st)
        m.GlobalPhase(complex(0, 1))(st)


def hermitian_for_encoded_kappa(n: int, kappa_encoded: float, nnz_col: int = 8, ds: int = 32, seed: int = 0x5EED2026 + 3):
    """cks.hermitian_random_matrix with kappa(A) chosen so that the matrix the walk
    block-encodes, A / (nnz_col Amax), has 1 / lambda_min = kappa_encoded (the
    condition number QDA's schedule runs on); a few secant-free rescalings of
    kappa(A), since kappa_encoded is nearly proportional to it."""
    kappa_A = kappa_encoded / 4.0
    for _ in range(6):
        values, sparsity, phases, A = cks.hermitian_random_matrix(n, kappa_A, seed, nnz_col)
        ke = 1.0 / cks.walk_spectrum(A, ds, nnz_col)[0]
        if abs(ke / kappa_encoded - 1.0) < 0.01:
            break
        kappa_A *= kappa_encoded / ke
    A.kappa_A = kappa_A
    return values, sparsity, phases, A


def setup(mod, nqubit: int, kappa_A: float = 30.0, nnz_col: int = 8, ds: int = 32, seed: int = 0x5EED2026 + 3,
          hermitian: bool = False, kappa_encoded: float | None = None):
    """Registers, the matrix and |b> for the sparse QDA: returns (state, walk, enc_A,
    scipy matrix, kappa of the block-encoded matrix).  hermitian: the complex
    Hermitian matrix of cks.hermitian_random_matrix (|A| through the walk, the
    phases through cks.PhaseOracle), with kappa(A) = kappa_A or, when
    kappa_encoded is given, chosen for that encoded condition number; otherwise
    the real band-plus-random one."""
    mod.System.clear()
    n = 1 << nqubit
    phases = None
    if hermitian and kappa_encoded:
        values, sparsity, phases, A = hermitian_for_encoded_kappa(n, kappa_encoded, nnz_col, ds, seed)
    elif hermitian:
        values, sparsity, phases, A = cks.hermitian_random_matrix(n, kappa_A, seed, nnz_col)
    else:
        values, sparsity, A = cks.band_random_matrix(n, kappa_A, seed, nnz_col)
    mat = mod.SparseMatrix(values, sparsity, ds, nnz_col, n, True)
    walk = cks.QuantumWalkNSteps(mod, mat, None)
    walk.InitEnvironment()
    st = walk.CreateSys()
    for name in ("anc_4", "anc_3", "anc_2", "anc_1"):
        mod.AddRegister(name, mod.Boolean, 1)(st)
    mod.Hadamard_Int(walk.j, nqubit)(st)
    lo, hi = cks.walk_spectrum(A, ds, nnz_col)
    kappa_enc = 1.0 / lo          # eigenvalues of A/(nnz Amax) lie in [lo, hi] with hi < 1
    enc = BlockEncodingSparse(mod, walk, cks.PhaseOracle(mod, walk, phases) if hermitian else None)
    enc.matrix = mat
    return st, walk, enc, A, kappa_enc


def get_solution(mod, st, walk, n: int) -> tuple[np.ndarray, float]:
    """Post-selected solution block: walk ancillas, anc_2, anc_3 = 0, anc_1 = 1,
    anc_4 = 0 (the out[n:2n) block of GetOutput, qda_fundamental.cpp:123-169),
    indexed by row_id; returns (vector, probability of the selected block)."""
    zero = [walk.b1, walk.k, walk.b2, walk.j_comp, walk.k_comp, walk.data_offset, "anc_3", "anc_2"]
    out, w = cks.select_dense(mod, st, zero, ["anc_1", "anc_4", walk.j], 4 * n)
    blk = out[1::4]
    return blk, float(np.sum(np.abs(blk) ** 2)) / max(w, 1e-300) if w > 0 else 0.0


def solution_metrics(blk: np.ndarray, A, xref: np.ndarray) -> dict:
    n = blk.shape[0]
    b = np.full(n, 1.0 / math.sqrt(n))
    met = cks.rhs_metrics(blk, b, 0.0, 0.0, xref, matvec=lambda x: A @ x)
    return met


def solve(mod, nqubit: int = 16, steps: int = 2000, p: float = 1.3, kappa_A: float = 30.0, nnz_col: int = 8,
          cz_eps: float = 1e-14, log_every: int = 0, log=print, stop_after: int | None = None, hermitian: bool = False,
          kappa_encoded: float | None = None):
    """The QDA schedule W(s = i/steps), ClearZero(cz_eps), Normalize per step
    (as PoiseuilleQDABench.cu) with the sparse block encoding; returns metrics."""
    import time

    st, walk, enc, A, kappa = setup(mod, nqubit, kappa_A, nnz_col, hermitian=hermitian, kappa_encoded=kappa_encoded)
    n = 1 << nqubit
    xref = cks._spd_solve(A, np.full(n, 1.0 / math.sqrt(n)))
    xref /= np.linalg.norm(xref)
    old = mod.get_clearzero_default_eps()
    mod.set_clearzero_default_eps(cz_eps)
    clear, norm = mod.ClearZero(cz_eps), mod.Normalize()
    t0 = time.perf_counter()
    peak = 0
    run = steps if stop_after is None else min(steps, stop_after)
    try:
        for i in range(run):
            WalkSSparse(mod, enc, "anc_1", "anc_2", "anc_3", "anc_4", i / steps, kappa, p)(st)
            clear(st)
            norm(st)
            if log_every and ((i + 1) % log_every == 0 or i + 1 == run):
                size = st.size()
                peak = max(peak, size)
                blk, _ = get_solution(mod, st, walk, n)
                met = solution_metrics(blk, A, xref)
                log(f"step {i + 1:6d}/{steps} res_dir={met['res_dir']:.6e} fid={met['fidelity']:.8f} entries={size} "
                    f"wall={time.perf_counter() - t0:.1f}s")
        wall = time.perf_counter() - t0
        size = st.size()
        blk, psel = get_solution(mod, st, walk, n)
    finally:
        mod.set_clearzero_default_eps(old)
    met = solution_metrics(blk, A, xref)
    met.update(steps=steps, steps_run=run, wall_s=wall, ms_per_step=wall / max(run, 1) * 1e3, entries=size,
               peak_logged_entries=max(peak, size), kappa_encoded=kappa, n=n, block_probability=psel,
               kappa_A=float(getattr(A, "kappa_A", kappa_A)), hermitian=hermitian)
    return met, st, walk, blk


def main():
    import argparse
    import contextlib
    import json

    ap = argparse.ArgumentParser(description="QDA solve with the sparse (CKS-oracle) block encoding")
    ap.add_argument("--nqubit", type=int, default=16)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--stop-after", type=int, default=None, help="run only the first N steps of the schedule")
    ap.add_argument("--kappa", type=float, default=30.0, help="condition number of A (the encoded matrix's is ~3x)")
    ap.add_argument("--p", type=float, default=1.3)
    ap.add_argument("--cz-eps", type=float, default=1e-14)
    ap.add_argument("--log-every", type=int, default=0)
    ap.add_argument("--learned", action="store_true", help="size branching operators with the learned sizer")
    ap.add_argument("--hermitian", action="store_true",
                    help="complex Hermitian matrix (N(0,1) + i N(0,1) entries at random columns) through the phase oracle")
    ap.add_argument("--kappa-encoded", type=float, default=None,
                    help="with --hermitian: pick kappa(A) so that the block-encoded matrix has this condition number")
    ap.add_argument("--kernel-profile", default=None,
                    help="write the engine's per-kernel profile and the sort-path counters of the run to this JSON file")
    args = ap.parse_args()
    import sparqsim_amd as ps
    from . import sizer
    from .profiling import KernelProfile

    if args.learned and not sizer.activate_default():
        raise SystemExit("no sizer weights found")
    with KernelProfile(ps) if args.kernel_profile else contextlib.nullcontext() as prof:
        met, *_ = solve(ps, args.nqubit, args.steps, args.p, args.kappa, 8, args.cz_eps, args.log_every,
                        log=lambda s: print(s, flush=True), stop_after=args.stop_after, hermitian=args.hermitian,
                        kappa_encoded=args.kappa_encoded)
    if args.kernel_profile:
        prof.write(args.kernel_profile, steps=met["steps_run"], entries=met["entries"])
    if args.learned:
        met["sizer"] = sizer.stats()
    print(json.dumps(met), flush=True)


if __name__ == "__main__":
    main()
