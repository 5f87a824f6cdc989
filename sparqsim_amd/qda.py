"""Discrete-adiabatic (QDA) linear-solver walk built from the engine's operators.

The composites follow the reference's templates operator by operator
(SparQ_Algorithm/include/DiscreteAdiabatic/qda_fundamental.h:53-480,
qda_tridiagonal.h:43-90, BlockEncoding/block_encoding_tridiagonal.h:37-170)
and run on any module exposing the pysparq operator surface -- the HIP engine
(`sparqsim_amd`) or the CPU oracle.  A composite's own conditions (the
reference's SPLIT_BY_CONDITIONS) are AND-ed into every operator of its body,
which is equivalent because no body operator writes those condition registers.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


# ---------------------------------------------------------------- conditions
@dataclass
class Cond:
    all_ones: list = field(default_factory=list)
    by_value: list = field(default_factory=list)

    def merged(self, other: "Cond | None") -> "Cond":
        if other is None:
            return self
        return Cond(self.all_ones + other.all_ones, self.by_value + other.by_value)


def ctl(op, cond: Cond | None):
    """Apply the (merged) conditions to an operator object and return it."""
    if cond is None:
        return op
    if cond.all_ones:
        op.conditioned_by_all_ones(list(cond.all_ones))
    if cond.by_value:
        op.conditioned_by_value([(r, v) for r, v in cond.by_value])
    return op


def compute_fs(s: float, kappa: float, p: float) -> float:
    """f(s) of the discrete adiabatic schedule (qda_fundamental.h, Walk_s ctor, Eq. (69))."""
    return kappa / (kappa - 1) * (1 - math.pow(1 + s * (math.pow(kappa, p - 1) - 1), 1 / (1 - p)))


# ---------------------------------------------------------------- block encodings
class BlockEncodingTridiagonal:
    """Block encoding of tridiag(beta, alpha, beta) via a 4-branch LCU on anc_UA
    (block_encoding_tridiagonal.h:37-170, .cpp:84-125)."""

    def __init__(self, mod, main_reg: str, anc_UA: str, alpha: float, beta: float):
        self.mod, self.main_reg, self.anc_UA, self.alpha, self.beta = mod, main_reg, anc_UA, alpha, beta
        n = 2 ** mod.System.size_of(main_reg)
        norm_f = math.sqrt(n * abs(alpha) * abs(alpha) + 2 * (n - 1) * abs(beta) * abs(beta))
        self.prep_state = [
            math.sqrt(abs(alpha)) / math.sqrt(norm_f),
            math.sqrt(abs(beta)) / math.sqrt(norm_f),
            math.sqrt(abs(beta)) / math.sqrt(norm_f),
            math.sqrt(1 - (abs(alpha) + 2 * abs(beta)) / norm_f),
        ]

    def _body(self, state, cond: Cond | None, dagger: bool):
        m, A, M = self.mod, self.anc_UA, self.main_reg
        m.SplitRegister(A, "overflow", 1)(state)
        m.SplitRegister(A, "other", 1)(state)
        prep = ctl(m.Rot_GeneralStatePrep(A, self.prep_state), cond)
        prep(state)
        first, second = (2, 1) if dagger else (1, 2)
        ctl(m.PlusOneAndOverflow(M, "overflow"), Cond(by_value=[(A, first)]).merged(cond))(state)
        if self.beta < 0:
            ctl(m.Reflection_Bool([M, "overflow"], False), Cond(by_value=[(A, first)]).merged(cond))(state)
            ctl(m.Reflection_Bool([M, "overflow"], False), Cond(by_value=[(A, second)]).merged(cond))(state)
        ctl(m.PlusOneAndOverflow(M, "overflow"), Cond(by_value=[(A, second)]).merged(cond)).dag(state)
        ctl(m.X_Bool("other", 0), Cond(all_ones=[A]).merged(cond))(state)
        prep.dag(state)
        m.CombineRegister(A, "other")(state)
        m.CombineRegister(A, "overflow")(state)

    def apply(self, state, cond: Cond | None = None):
        self._body(state, cond, dagger=False)

    def apply_dag(self, state, cond: Cond | None = None):
        self._body(state, cond, dagger=True)


class BlockEncodingHs:
    """Block encoding of H(s) = (1-f)H0 + f H1 (qda_fundamental.h:53-162)."""

    def __init__(self, mod, enc_A, main_reg, anc_UA, anc_1, anc_2, anc_3, anc_4, fs, enc_b=None):
        self.mod, self.enc_A, self.enc_b = mod, enc_A, enc_b
        self.main_reg, self.anc_UA = main_reg, anc_UA
        self.a1, self.a2, self.a3, self.a4 = anc_1, anc_2, anc_3, anc_4
        sqrt_n = 1.0 / math.sqrt((1 - fs) * (1 - fs) + fs * fs)
        self.R_s = [sqrt_n * (1 - fs), sqrt_n * fs, sqrt_n * fs, sqrt_n * (fs - 1)]

    def apply(self, st):
        m = self.mod
        a1, a2, a3, a4, main = self.a1, self.a2, self.a3, self.a4, self.main_reg
        # state preparation of b: a Hadamard (self-adjoint) for b = ones/sqrt(n),
        # or an object whose `dag` is its adjoint
        enc_b = self.enc_b if self.enc_b is not None else m.Hadamard_Int_Full(main)
        enc_b_dag = getattr(enc_b, "dag", enc_b)
        m.Hadamard_Bool(a3)(st)
        enc_b_dag(st)
        m.X_Bool(a1, 0)(st)
        m.Reflection_Bool(main, True).conditioned_by_all_ones([a1, a3, a4])(st)
        m.X_Bool(a1, 0)(st)
        enc_b(st)
        m.X_Bool(a4, 0)(st)
        m.Rot_Bool(a2, 0, self.R_s).conditioned_by_all_ones(a4)(st)
        m.X_Bool(a4, 0)(st)
        m.Hadamard_Bool(a2).conditioned_by_all_ones(a4)(st)
        self.enc_A.apply(st, Cond(all_ones=[a1, a2]))
        m.X_Bool(a1, 0).conditioned_by_all_ones(a2)(st)
        m.Reflection_Bool(a2, True).conditioned_by_all_ones(a1)(st)
        self.enc_A.apply_dag(st, Cond(all_ones=[a1, a2]))
        m.X_Bool(a4, 0)(st)
        m.Hadamard_Bool(a2).conditioned_by_all_ones(a4)(st)
        m.X_Bool(a4, 0)(st)
        m.Rot_Bool(a2, 0, self.R_s).conditioned_by_all_ones(a4)(st)
        m.X_Bool(a4, 0)(st)
        enc_b_dag(st)
        m.X_Bool(a1, 0)(st)
        m.Reflection_Bool(main, True).conditioned_by_all_ones([a1, a3, a4])(st)
        m.X_Bool(a1, 0)(st)
        enc_b(st)
        m.Hadamard_Bool(a3)(st)


class WalkSTridiagonal:
    """One step W(s) of the QDA walk with the tridiagonal block encoding
    (qda_fundamental.h:344-446, qda_tridiagonal.h:43-90)."""

    def __init__(self, mod, main_reg, anc_UA, anc_1, anc_2, anc_3, anc_4, s, kappa, p, alpha, beta):
        self.mod = mod
        self.fs = compute_fs(s, kappa, p)
        self.enc_A = BlockEncodingTridiagonal(mod, main_reg, anc_UA, alpha, beta)
        self.enc_Hs = BlockEncodingHs(mod, self.enc_A, main_reg, anc_UA, anc_1, anc_2, anc_3, anc_4, self.fs)
        self.refl = [anc_UA, anc_2, anc_3]

    def __call__(self, st):
        m = self.mod
        self.enc_Hs.apply(st)
        m.Reflection_Bool(self.refl, False)(st)
        m.GlobalPhase(complex(0, 1))(st)


# ---------------------------------------------------------------- classical side
def poiseuille_matrix(n: int, alpha: float, beta: float) -> np.ndarray:
    """tridiag(beta, alpha, beta) (generate_Poiseuille_mat, as documented in
    PySparQ/test/algorithms/test_qda_integration.py:64-80)."""
    a = np.zeros((n, n))
    for i in range(n):
        a[i, i] = alpha
        if i > 0:
            a[i, i - 1] = beta
        if i < n - 1:
            a[i, i + 1] = beta
    return a


def condition_number(a: np.ndarray) -> float:
    ev = np.abs(np.linalg.eigvalsh(a))
    return float(ev.max() / ev.min())


def mid_eigenstate(a: np.ndarray, b: np.ndarray, fs: float) -> np.ndarray:
    """Ideal state of the walk at f(s) (qda_fundamental.cpp:42-112), length 4n."""
    n = len(b)
    out = np.zeros(4 * n)
    if fs < 1e-14:
        out[:n] = b
    elif abs(fs - 1.0) < 1e-14:
        x = np.linalg.solve(a, b)
        out[n:2 * n] = x / np.linalg.norm(x)
    else:
        af = np.zeros((2 * n, 2 * n))
        af[:n, :n] = (1 - fs) * np.eye(n)
        af[:n, n:] = fs * a
        af[n:, :n] = fs * a.T
        af[n:, n:] = -(1 - fs) * np.eye(n)
        rhs = np.zeros(2 * n)
        rhs[:n] = b
        x = np.linalg.solve(af, rhs)
        out[:2 * n] = x / np.linalg.norm(x)
    return out


def get_output(state, mod, main_reg="main_reg", anc_UA="anc_UA", anc_1="anc_1", anc_2="anc_2", anc_3="anc_3", anc_4="anc_4"):
    """Post-selected output vector (qda_fundamental.cpp:123-169): entries with
    anc_UA = anc_3 = anc_2 = 0, indexed by main | anc_1 << n | anc_4 << (n+1),
    renormalised.  Returns (vector, probability).  On the engine the selection
    runs on the device (`post_select_dense`) and only the 4n vector is copied."""
    g = mod.System.get_id
    n = mod.System.size_of(main_reg)
    if hasattr(state, "post_select_dense"):
        vec, prob = state.post_select_dense([anc_UA, anc_3, anc_2], [main_reg, anc_1, anc_4])
    else:
        ids, vals, amps = state.export()
        col = {i: k for k, i in enumerate(ids)}
        sel = (vals[:, col[g(anc_UA)]] == 0) & (vals[:, col[g(anc_3)]] == 0) & (vals[:, col[g(anc_2)]] == 0)
        idx = (vals[sel, col[g(main_reg)]].astype(np.int64)
               | (vals[sel, col[g(anc_1)]].astype(np.int64) << n)
               | (vals[sel, col[g(anc_4)]].astype(np.int64) << (n + 1)))
        vec = np.zeros(1 << (n + 2), dtype=complex)
        vec[idx] = amps[sel]
        prob = float(np.sum(np.abs(amps[sel]) ** 2))
    if abs(prob - 1.0) >= 1e-14:
        vec = vec * (1.0 / math.sqrt(prob) if prob != 0 else 1.0)
    return vec, prob


def fidelity(ideal: np.ndarray, out: np.ndarray) -> float:
    """|<ideal|out>| -- the quantity the reference's golden fidelity lists hold
    (CorrectnessTest_QDA_CompareList.inl:173-176 via get_fidelity)."""
    return float(abs(np.vdot(ideal, out)))


QDA_STEP_CONSTANT = 2305    # test/CPUTest/CommonTest/CorrectnessTest_QDA_CompareList.inl:132


def setup_poiseuille(mod, nqubit: int, ancilla_order=("anc_1", "anc_2", "anc_3", "anc_4")):
    """Registers and |b> = H|0> on main_reg (CorrectnessTest_QDA_CompareList.inl:118-131;
    the production bench, Experiments/QDA/QDA_Poiseuille/PoiseuilleQDABench.cu:149-160,
    adds the Boolean ancillas in the order anc_4, anc_3, anc_2, anc_1)."""
    mod.System.clear()
    st = mod.SparseState()
    mod.AddRegister("main_reg", mod.UnsignedInteger, nqubit)(st)
    mod.AddRegister("anc_UA", mod.UnsignedInteger, 4)(st)
    for name in ancilla_order:
        mod.AddRegister(name, mod.Boolean, 1)(st)
    mod.Hadamard_Int_Full("main_reg")(st)
    return st


def analytic_kappa(alpha: float, beta: float, n: int) -> float:
    """max/min |alpha + 2 beta cos(pi k/(n+1))| over k = 1..n (PoiseuilleQDABench.cu:57-65)."""
    k = np.arange(1, n + 1, dtype=np.float64)
    lam = np.abs(alpha + 2.0 * beta * np.cos(np.pi * k / (n + 1)))
    return float(lam.max() / lam.min())


def thomas_solution(alpha: float, beta: float, n: int) -> np.ndarray:
    """Normalised solution of tridiag(beta, alpha, beta) x = ones/sqrt(n) by the Thomas
    algorithm (PoiseuilleQDABench.cu:67-88)."""
    from scipy.linalg import solve_banded

    ab = np.zeros((3, n))
    ab[0, 1:] = beta
    ab[1, :] = alpha
    ab[2, :-1] = beta
    x = solve_banded((1, 1), ab, np.full(n, 1.0 / math.sqrt(n)))
    return x / np.linalg.norm(x)


def block_metrics(blk: np.ndarray, alpha: float, beta: float, xref: np.ndarray) -> dict:
    """res_dir, res_opt, block weight and fidelity of an unnormalised solution vector
    blk of tridiag(beta, alpha, beta) x = ones/sqrt(n), as the production benches
    compute them (PoiseuilleQDABench.cu:194-229, CKSPoiseuilleBench.cu:95-131):
    y = A x/|x|, res_dir = sqrt(2 - 2<y, b>/|y|), res_opt^2 = |y|^2 - |<y, b>|^2,
    fid = |<xref, x/|x|>|^2."""
    n = blk.shape[0]
    succ = float(np.sum(np.abs(blk) ** 2))
    nrm = math.sqrt(succ)
    if nrm <= 1e-300:
        return {"res_dir": float("nan"), "res_opt": float("nan"), "success_prob": succ, "fidelity": float("nan")}
    x = blk / nrm
    y = alpha * x
    y[:-1] += beta * x[1:]
    y[1:] += beta * x[:-1]
    y_nrm2 = float(np.sum(np.abs(y) ** 2))
    y_nrm = math.sqrt(y_nrm2)
    binv = 1.0 / math.sqrt(n)
    proj_re = binv * float(np.sum(y.real))
    proj_im = binv * float(np.sum(y.imag))
    res_dir = math.sqrt(max(2.0 - 2.0 * proj_re / y_nrm, 0.0))
    res_opt = math.sqrt(max(y_nrm2 - (proj_re * proj_re + proj_im * proj_im), 0.0))
    fr = float(np.dot(xref, x.real))
    fi = float(np.dot(xref, x.imag))
    return {"res_dir": res_dir, "res_opt": res_opt, "success_prob": succ, "fidelity": fr * fr + fi * fi}


def solution_metrics(out: np.ndarray, n: int, alpha: float, beta: float, xref: np.ndarray) -> dict:
    """block_metrics of the solution block out[n:2n) of a GetOutput vector."""
    return block_metrics(out[n:2 * n], alpha, beta, xref)


def qda_schedule(nqubit: int, step_rate: float, alpha: float, beta: float):
    """(normalised A, b, kappa, steps) of the Poiseuille test problem."""
    n = 2 ** nqubit
    a = poiseuille_matrix(n, alpha, beta)
    a = a / np.linalg.norm(a, "fro")
    kappa = condition_number(a)
    b = np.ones(n) / math.sqrt(n)
    steps = int(step_rate * QDA_STEP_CONSTANT * kappa)
    if steps % 2:
        steps += 1
    return a, b, kappa, steps
