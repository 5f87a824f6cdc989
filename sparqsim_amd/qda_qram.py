"""QDA over the QRAM block encoding of a dense matrix.

Restates, operator by operator, the reference's general-matrix encodings
(SparQ_Algorithm/include/BlockEncoding/block_encoding_via_QRAM.h:39-470:
U_R, U_L, Block_Encoding_via_QRAM; state_preparation.h:36-190:
State_Prep_via_QRAM; make_qram.h: the data quantisation and the binary tree
of partial norms the QRAM holds), and the walk that uses them
(DiscreteAdiabatic/qda_via_QRAM.h: Walk_s over Block_Encoding_via_QRAM and
State_Prep_via_QRAM).  Runs on the HIP engine or the CPU oracle.

The QRAM holds a tree over the 2^m leaves v_i (two's complement): level 0 is
the root, node j of level L sits at word 2^L - 1 + j, its children at words
2p + 1 and 2p + 2; the level above the leaves holds v_2j^2 + v_2j+1^2, every
higher node the sum of its children, and a trailing 0 word pads the tree to
2^(m+1) words.  A state preparation walks the tree from the root: at level k
the register's next bit is rotated by arccos(sqrt(child / parent)) (the leaf
level by the angle of (v_2j, v_2j+1), which carries the signs).

As with the other composites here, a composite's own conditions (the
reference's SPLIT_BY_CONDITIONS) are AND-ed into every operator of its body
that changes values; register bookkeeping (add / remove / split / combine)
and ClearZero act on the whole state.
"""
from __future__ import annotations

import math

import numpy as np

from . import qda
from .qda import Cond, ctl


def _clear_zero(m, st):
    """ClearZero at the default threshold (the oracle module takes an explicit None)."""
    (m.ClearZero() if m.__name__.startswith("sparqsim") else m.ClearZero(None))(st)


# ---------------------------------------------------------------- data (make_qram.h)
def column_flatten(mat: np.ndarray) -> np.ndarray:
    """Column-major flattening of a square matrix (get_column_flatten)."""
    return np.asarray(mat, dtype=np.float64).T.reshape(-1)


def scale_and_convert(values, exponent: int, data_size: int) -> list[int]:
    """round(v * 2^exponent), halves away from zero, as data_size-bit two's
    complement words (scaleAndConvertVector)."""
    mask = (1 << data_size) - 1
    out = []
    for v in np.asarray(values, dtype=np.float64).reshape(-1):
        x = v * math.ldexp(1.0, exponent)
        r = int(math.floor(abs(x) + 0.5))
        out.append((-r if x < 0 else r) & mask)
    return out


def signed(word: int, data_size: int) -> int:
    return word - (1 << data_size) if (word >> (data_size - 1)) & 1 else word


def make_vector_tree(leaves: list[int], data_size: int) -> list[int]:
    """The QRAM words of the partial-norm tree over the leaves (make_vector_tree)."""
    n = len(leaves)
    if n < 2 or n & (n - 1):
        raise ValueError("the tree needs a power-of-two number of leaves >= 2")
    levels = [list(leaves)]
    cur = [signed(leaves[2 * j], data_size) ** 2 + signed(leaves[2 * j + 1], data_size) ** 2 for j in range(n // 2)]
    levels.append(cur)
    while len(cur) > 1:
        cur = [cur[2 * j] + cur[2 * j + 1] for j in range(len(cur) // 2)]
        levels.append(cur)
    tree = [w for level in reversed(levels) for w in level] + [0]
    if any(w >> data_size for w in tree):
        raise ValueError("a tree node does not fit the data width")
    return tree


# ---------------------------------------------------------------- state preparation
class StatePrepViaQRAM:
    """|0> -> sum_i v_i / |v| |i> on `work` (State_Prep_via_QRAM,
    state_preparation.h:36-190)."""

    def __init__(self, mod, qram, work: str, data_size: int, rational_size: int):
        self.mod, self.qram, self.work = mod, qram, work
        self.m = mod.System.size_of(work)
        self.ds, self.rs = data_size, rational_size

    def _regs(self, st):
        m = self.mod
        m.AddRegister("addr_parent", m.UnsignedInteger, self.m + 1)(st)
        m.AddRegister("addr_child", m.UnsignedInteger, self.m + 1)(st)
        m.AddRegister("data_parent", m.SignedInteger, self.ds)(st)
        m.AddRegister("data_child", m.SignedInteger, self.ds)(st)
        m.AddRegister("div_result", m.Rational, self.rs)(st)

    def _drop(self, st):
        m = self.mod
        for r in ("addr_parent", "data_parent", "addr_child", "data_child", "div_result"):
            m.RemoveRegister(r)(st)
        _clear_zero(m, st)

    def _level(self, st, level: int, leaf: bool, dagger: bool, cond):
        """Rotate the register's next bit by the level's angle (forward or adjoint)."""
        m, w, q = self.mod, self.work, self.qram
        m.SplitRegister(w, "rotation", 1)(st)
        ctl(m.Add_ConstUInt_InPlace("addr_parent", (1 << level) - 1), cond)(st)
        ctl(m.Add_UInt_UInt_InPlace(w, "addr_parent"), cond)(st)
        ctl(m.Mult_UInt_ConstUInt("addr_parent", 2, "addr_child"), cond)(st)
        ctl(m.X_Bool("addr_child", 0), cond)(st)
        _rotate(m, q, st, leaf, dagger, cond, loads_child_first=False, clear_zero=True)
        ctl(m.X_Bool("addr_child", 0), cond)(st)
        ctl(m.Mult_UInt_ConstUInt("addr_parent", 2, "addr_child"), cond)(st)
        ctl(m.Add_UInt_UInt_InPlace(w, "addr_parent"), cond).dag(st)
        ctl(m.Add_ConstUInt_InPlace("addr_parent", (1 << level) - 1), cond).dag(st)
        m.CombineRegister(w, "rotation")(st)

    def __call__(self, st, cond: Cond | None = None):
        m = self.mod
        self._regs(st)
        for k in range(self.m):
            self._level(st, k, k == self.m - 1, False, cond)
            ctl(m.ShiftLeft_InPlace(self.work, 1), cond)(st)
        ctl(m.ShiftRight_InPlace(self.work, 1), cond)(st)
        self._drop(st)

    def dag(self, st, cond: Cond | None = None):
        m = self.mod
        self._regs(st)
        ctl(m.ShiftLeft_InPlace(self.work, 1), cond)(st)
        for k in range(self.m):
            ctl(m.ShiftRight_InPlace(self.work, 1), cond)(st)
            self._level(st, self.m - 1 - k, k == 0, True, cond)
        self._drop(st)


def _rotate(m, q, st, leaf: bool, dagger: bool, cond, loads_child_first: bool, clear_zero: bool):
    """Load parent and child words, rotate `rotation` by their angle, unload
    (the body shared by the tree walks).  `leaf`: the angle of the signed pair
    (GetRotateAngle_Int_Int) on addr_parent's children instead of
    arccos(sqrt(child / parent)) (Div_Sqrt_Arccos_UInt_UInt).  `clear_zero`:
    a ClearZero right after the rotation (the state preparation and U_L's
    adjoint have it, U_L does not)."""
    def loads():
        pair = [m.QRAMLoad(q, "addr_parent", "data_parent"), m.QRAMLoad(q, "addr_child", "data_child")]
        for op in (pair[::-1] if loads_child_first else pair):
            ctl(op, cond)(st)

    if leaf:
        ctl(m.ShiftLeft_InPlace("addr_parent", 1), cond)(st)
        ctl(m.X_Bool("addr_parent", 0), cond)(st)
        ctl(m.Add_ConstUInt_InPlace("addr_child", 1), cond)(st)
        loads()
        angle = lambda: ctl(m.GetRotateAngle_Int_Int("data_parent", "data_child", "div_result"), cond)(st)
    else:
        loads()
        angle = lambda: ctl(m.Div_Sqrt_Arccos_UInt_UInt("data_child", "data_parent", "div_result"), cond)(st)
    angle()
    rot = ctl(m.CondRot_Fixed_Bool("div_result", "rotation"), cond)
    rot.dag(st) if dagger else rot(st)
    if clear_zero:
        _clear_zero(m, st)
    angle()
    loads()
    if leaf:
        ctl(m.Add_ConstUInt_InPlace("addr_child", 1), cond).dag(st)
        ctl(m.X_Bool("addr_parent", 0), cond)(st)
        ctl(m.ShiftRight_InPlace("addr_parent", 1), cond)(st)


# ---------------------------------------------------------------- block encoding
class BlockEncodingViaQRAM:
    """U_A = SWAP(col, row) U_R^dag(col) U_L(row, col), with
    <i|_col <0|_row U_A |j>_col |0>_row = A_ij / |A|_F
    (Block_Encoding_via_QRAM, block_encoding_via_QRAM.h:408-470)."""

    def __init__(self, mod, qram, column: str, row: str, data_size: int, rational_size: int):
        self.mod, self.qram, self.col, self.row = mod, qram, column, row
        self.m = mod.System.size_of(row)
        self.ds, self.rs = data_size, rational_size

    # U_R (block_encoding_via_QRAM.h:39-190): the column norms on `col`
    def _ur(self, st, dagger: bool, cond):
        m, c, q, n = self.mod, self.col, self.qram, self.m
        m.AddRegister("addr_child", m.UnsignedInteger, n + 1)(st)
        m.AddRegister("data_parent", m.SignedInteger, self.ds)(st)
        m.AddRegister("data_child", m.SignedInteger, self.ds)(st)
        m.AddRegister("temp_bit", m.Boolean, 1)(st)
        m.AddRegister("div_result", m.Rational, self.rs)(st)
        if dagger:
            ctl(m.ShiftLeft_InPlace(c, 1), cond)(st)
        for i in range(n):
            k = n - 1 - i if dagger else i
            if dagger:
                ctl(m.ShiftRight_InPlace(c, 1), cond)(st)
            m.SplitRegister(c, "rotation", 1)(st)
            m.CombineRegister(c, "temp_bit")(st)
            ctl(m.ShiftRight_InPlace(c, 1), cond)(st)
            ctl(m.Add_ConstUInt_InPlace(c, (1 << k) - 1), cond)(st)
            ctl(m.Mult_UInt_ConstUInt(c, 2, "addr_child"), cond)(st)
            ctl(m.X_Bool("addr_child", 0), cond)(st)
            pair = [m.QRAMLoad(q, c, "data_parent"), m.QRAMLoad(q, "addr_child", "data_child")]
            for op in (pair[::-1] if dagger else pair):
                ctl(op, cond)(st)
            ctl(m.Div_Sqrt_Arccos_UInt_UInt("data_child", "data_parent", "div_result"), cond)(st)
            rot = ctl(m.CondRot_Fixed_Bool("div_result", "rotation"), cond)
            rot.dag(st) if dagger else rot(st)
            _clear_zero(m, st)
            ctl(m.Div_Sqrt_Arccos_UInt_UInt("data_child", "data_parent", "div_result"), cond)(st)
            for op in (pair if dagger else pair[::-1]):
                ctl(op, cond)(st)
            ctl(m.X_Bool("addr_child", 0), cond)(st)
            ctl(m.Mult_UInt_ConstUInt(c, 2, "addr_child"), cond)(st)
            ctl(m.Add_ConstUInt_InPlace(c, (1 << n) - (1 << k) + 1), cond)(st)
            ctl(m.ShiftLeft_InPlace(c, 1), cond)(st)
            m.SplitRegister(c, "temp_bit", 1)(st)
            m.CombineRegister(c, "rotation")(st)
            if not dagger:
                ctl(m.ShiftLeft_InPlace(c, 1), cond)(st)
        if not dagger:
            ctl(m.ShiftRight_InPlace(c, 1), cond)(st)
        for r in ("data_parent", "addr_child", "data_child", "temp_bit", "div_result"):
            m.RemoveRegister(r)(st)
        _clear_zero(m, st)

    # U_L (block_encoding_via_QRAM.h:195-400): column col's entries on `row`
    def _ul(self, st, dagger: bool, cond):
        m, r, c, q, n = self.mod, self.row, self.col, self.qram, self.m
        m.AddRegister("addr_parent", m.UnsignedInteger, 2 * n + 1)(st)
        m.AddRegister("addr_child", m.UnsignedInteger, 2 * n + 1)(st)
        m.AddRegister("data_parent", m.SignedInteger, self.ds)(st)
        m.AddRegister("data_child", m.SignedInteger, self.ds)(st)
        m.AddRegister("div_result", m.Rational, self.rs)(st)
        if dagger:
            ctl(m.ShiftLeft_InPlace(r, 1), cond)(st)
        levels = range(2 * n - 1, n - 1, -1) if dagger else range(n, 2 * n)
        for k in levels:
            if dagger:
                ctl(m.ShiftRight_InPlace(r, 1), cond)(st)
            m.SplitRegister(r, "rotation", 1)(st)
            ctl(m.Add_ConstUInt_InPlace("addr_parent", (1 << k) - 1), cond)(st)
            ctl(m.Add_Mult_UInt_ConstUInt_InPlace(c, 1 << (k - n), "addr_parent"), cond)(st)
            ctl(m.Add_UInt_UInt_InPlace(r, "addr_parent"), cond)(st)
            ctl(m.Mult_UInt_ConstUInt("addr_parent", 2, "addr_child"), cond)(st)
            ctl(m.X_Bool("addr_child", 0), cond)(st)
            _rotate(m, q, st, k == 2 * n - 1, dagger, cond, loads_child_first=dagger, clear_zero=dagger)
            ctl(m.X_Bool("addr_child", 0), cond)(st)
            ctl(m.Mult_UInt_ConstUInt("addr_parent", 2, "addr_child"), cond)(st)
            ctl(m.Add_UInt_UInt_InPlace(r, "addr_parent"), cond).dag(st)
            ctl(m.Add_Mult_UInt_ConstUInt_InPlace(c, 1 << (k - n), "addr_parent"), cond).dag(st)
            ctl(m.Add_ConstUInt_InPlace("addr_parent", (1 << k) - 1), cond).dag(st)
            m.CombineRegister(r, "rotation")(st)
            if not dagger:
                ctl(m.ShiftLeft_InPlace(r, 1), cond)(st)
        if not dagger:
            ctl(m.ShiftRight_InPlace(r, 1), cond)(st)
        for name in ("addr_parent", "data_parent", "addr_child", "data_child", "div_result"):
            m.RemoveRegister(name)(st)
        _clear_zero(m, st)

    def apply(self, st, cond: Cond | None = None):
        self._ul(st, False, cond)
        self._ur(st, True, cond)
        ctl(self.mod.Swap_General_General(self.row, self.col), cond)(st)

    def apply_dag(self, st, cond: Cond | None = None):
        ctl(self.mod.Swap_General_General(self.row, self.col), cond)(st)
        self._ur(st, False, cond)
        self._ul(st, True, cond)


# ---------------------------------------------------------------- the walk
class WalkSViaQRAM:
    """One step W(s) of the QDA walk over the QRAM encodings (qda_via_QRAM.h
    Walk_s_via_QRAM; qda_fundamental.h:344-446)."""

    def __init__(self, mod, qram_A, qram_b, main_reg, anc_UA, anc_1, anc_2, anc_3, anc_4, s, kappa, p,
                 data_size, rational_size):
        self.mod = mod
        self.fs = qda.compute_fs(s, kappa, p)
        self.enc_A = BlockEncodingViaQRAM(mod, qram_A, main_reg, anc_UA, data_size, rational_size)
        self.enc_b = StatePrepViaQRAM(mod, qram_b, main_reg, data_size, rational_size)
        self.enc_Hs = qda.BlockEncodingHs(mod, self.enc_A, main_reg, anc_UA, anc_1, anc_2, anc_3, anc_4, self.fs,
                                          enc_b=self.enc_b)
        self.refl = [anc_UA, anc_2, anc_3]

    def __call__(self, st):
        m = self.mod
        self.enc_Hs.apply(st)
        m.Reflection_Bool(self.refl, False)(st)
        m.GlobalPhase(complex(0, 1))(st)


EXPONENT, DATA_SIZE, RATIONAL_SIZE = 20, 50, 51   # CorrectnessTest_QDA_CompareList.inl:12-14


def setup_qram_poiseuille(mod, nqubit: int, alpha: float, beta: float):
    """The reference's QRAM Poiseuille test problem
    (CorrectnessTest_QDA_CompareList.inl:5-60): A = tridiag(beta, alpha, beta) /
    |.|_F and b = ones / sqrt(n) quantised into QRAM trees, registers main_reg,
    anc_UA (nqubit bits each), anc_1..anc_4, and |b> prepared by the tree walk.
    Returns (state, qram_A, qram_b, A, b, kappa)."""
    n = 1 << nqubit
    a = qda.poiseuille_matrix(n, alpha, beta)
    a = a / np.linalg.norm(a, "fro")
    kappa = qda.condition_number(a)
    b = np.ones(n) / math.sqrt(n)
    tree_A = make_vector_tree(scale_and_convert(column_flatten(a), EXPONENT, DATA_SIZE), DATA_SIZE)
    tree_b = make_vector_tree(scale_and_convert(b, EXPONENT, DATA_SIZE), DATA_SIZE)
    mod.System.clear()
    qram_A = mod.QRAMCircuit_qutrit(2 * nqubit + 1, DATA_SIZE, tree_A)
    qram_b = mod.QRAMCircuit_qutrit(nqubit + 1, DATA_SIZE, tree_b)
    st = mod.SparseState()
    mod.AddRegister("main_reg", mod.UnsignedInteger, nqubit)(st)
    mod.AddRegister("anc_UA", mod.UnsignedInteger, nqubit)(st)
    for name in ("anc_1", "anc_2", "anc_3", "anc_4"):
        mod.AddRegister(name, mod.Boolean, 1)(st)
    StatePrepViaQRAM(mod, qram_b, "main_reg", DATA_SIZE, RATIONAL_SIZE)(st)
    return st, qram_A, qram_b, a, b, kappa
