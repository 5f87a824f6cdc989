"""RIR JSON programs on the engine (sparqsim_amd.rir).

Expected amplitudes of each scenario are the ones the reference's own RIR
tests assert (PySparQ/test/test_rir.py: gates and controls, module expansion,
add_const views, QRAM queries under controls that share address bits).  The
dense checker (oracle/rir_dense.py) is held to the same expectations on the
CPU, then the engine is compared with the checker on seeded random programs.
"""
import cmath
import json
import math

import numpy as np
import pytest

from rir_dense import DenseRIR
from sparqsim_amd.rir import RIRError, load_rir, run_rir, run_rir_file

R2 = 1 / math.sqrt(2)


# ---------------------------------------------------------------- document builders
def span(reg, start, width):
    return {"tag": "Ref", "parts": [{"tag": "Span", "register": reg, "start": start, "width": width}],
            "type": {"tag": "RegType", "kind": "uint", "width": width}}


def reg(name, width, kind="uint"):
    return {"tag": "Register", "name": name, "type": {"tag": "RegType", "kind": kind, "width": width}}


def op(name, *views, angle=None, value=None):
    return {"tag": "Primitive", "op": name, "operands": list(views), "angle": angle, "value": value}


def ctrl(view, value, *body):
    return {"tag": "Control", "register": view, "value": value, "body": list(body)}


def adj(*body):
    return {"tag": "Adjoint", "body": list(body)}


def rep(count, *body):
    return {"tag": "Repeat", "count": count, "body": list(body)}


def call(name, views, resources=()):
    return {"tag": "Call", "module": name, "arguments": list(views), "resources": list(resources)}


def load(resource, addr, data):
    return {"tag": "Load", "resource": resource, "address": addr, "data": data}


def qram(name, aw, dw):
    return {"tag": "Resource", "name": name, "type": {"tag": "QRAM", "address_width": aw, "data_width": dw}}


def mod(name, regs, body, locals_=(), resources=()):
    return {"tag": "Module", "name": name, "registers": list(regs), "locals": list(locals_),
            "resources": list(resources), "body": body, "attributes": []}


def prog(*mods, entry="main"):
    return {"tag": "Program", "entry": entry, "modules": list(mods), "version": "0.3"}


# ---------------------------------------------------------------- scenarios (reference expectations)
BANK = {0: 5, 3: 7, 8: 1}
A4D3 = [reg("address", 4, "bits"), reg("data", 3, "bits")]
LOAD43 = load("bank", span("address", 0, 4), span("data", 0, 3))
THETA = 0.7


def _scenarios():
    s = {}
    s["bell"] = (prog(mod("main", [reg("a", 1, "bits"), reg("b", 1, "bits")],
                          [op("h", span("a", 0, 1)), ctrl(span("a", 0, 1), 1, op("x", span("b", 0, 1)))])),
                 None, {(0, 0): R2, (1, 1): R2})
    s["control_on_zero"] = (prog(mod("main", [reg("a", 1, "bits"), reg("b", 1, "bits")],
                                     [ctrl(span("a", 0, 1), 0, op("x", span("b", 0, 1)))])), None, {(0, 1): 1})
    s["gphase"] = (prog(mod("main", [reg("q", 1, "bits")], [op("gphase", angle=math.pi)])), None, {(0,): -1})
    s["controlled_gphase"] = (prog(mod("main", [reg("q", 1, "bits")],
                                       [op("x", span("q", 0, 1)), ctrl(span("q", 0, 1), 1, op("gphase", angle=math.pi / 2))])),
                              None, {(1,): 1j})
    s["rotation_adjoint"] = (prog(mod("main", [reg("q", 1, "bits")],
                                      [op("rx", span("q", 0, 1), angle=THETA), adj(op("rx", span("q", 0, 1), angle=THETA))])),
                             None, {(0,): 1})
    s["xor_swap"] = (prog(mod("main", [reg("a", 2), reg("b", 2), reg("c", 2)],
                              [op("x", span("a", 0, 2)), op("xor", span("a", 0, 2), span("b", 0, 2)),
                               op("swap", span("a", 0, 2), span("c", 0, 2))])), None, {(0, 3, 3): 1})
    inc = mod("inc", [reg("x", 4)], [op("add_const", span("x", 0, 4), value=1)])
    s["call_repeat_adjoint"] = (prog(mod("main", [reg("x", 4)],
                                         [rep(3, call("inc", [span("x", 0, 4)])), adj(call("inc", [span("x", 0, 4)]))]), inc),
                                None, {(2,): 1})
    bump = mod("bump", [reg("lo", 2)], [op("add_const", span("lo", 0, 2), value=1)])
    s["nested_views"] = (prog(mod("main", [reg("x", 4)], [op("x", span("x", 2, 2)), call("bump", [span("x", 0, 2)])]), bump),
                         None, {(13,): 1})
    s["clean_local"] = (prog(mod("main", [reg("out", 1, "bits")],
                                 [op("x", span("w", 0, 1)), ctrl(span("w", 0, 1), 1, op("x", span("out", 0, 1))),
                                  op("x", span("w", 0, 1))], locals_=[reg("w", 1, "bits")])), None, {(1,): 1})
    s["add_full_register"] = (prog(mod("main", [reg("x", 5)], [op("add_const", span("x", 0, 5), value=63)])), None, {(31,): 1})
    s["add_partial_view"] = (prog(mod("main", [reg("x", 4)], [op("x", span("x", 0, 1)), op("x", span("x", 2, 1)),
                                                              op("add_const", span("x", 0, 2), value=1)])), None, {(6,): 1})
    s["add_low_slice_wraps"] = (prog(mod("main", [reg("x", 4)], [op("x", span("x", 0, 1)), op("x", span("x", 1, 1)),
                                                                 op("add_const", span("x", 0, 2), value=1)])), None, {(0,): 1})
    s["add_high_slice"] = (prog(mod("main", [reg("x", 4)], [op("add_const", span("x", 2, 2), value=1)])), None, {(4,): 1})
    s["add_high_slice_wraps"] = (prog(mod("main", [reg("x", 4)], [op("x", span("x", 2, 1)), op("x", span("x", 3, 1)),
                                                                  op("add_const", span("x", 2, 2), value=1)])), None, {(0,): 1})
    s["controlled_add"] = (prog(mod("main", [reg("flag", 1, "bits"), reg("x", 3)],
                                    [op("x", span("flag", 0, 1)), ctrl(span("flag", 0, 1), 1, op("add_const", span("x", 0, 3), value=5))])),
                           None, {(1, 5): 1})
    q23 = mod("main", [reg("address", 2), reg("data", 3)],
              [op("h", span("address", 0, 2)), load("values", span("address", 0, 2), span("data", 0, 3))], resources=[qram("values", 2, 3)])
    s["load_superposition"] = (prog(q23), {"values": [1, 2, 4, 7]}, {(a, w): 0.5 for a, w in enumerate([1, 2, 4, 7])})
    s["load_sparse_memory"] = (prog(q23), {"values": {2: 5}}, {(0, 0): 0.5, (1, 0): 0.5, (2, 5): 0.5, (3, 0): 0.5})
    res43 = [qram("bank", 4, 3)]
    s["zero_control_on_address_bit"] = (prog(mod("main", A4D3, [op("h", span("address", 3, 1)), ctrl(span("address", 3, 1), 0, LOAD43)],
                                                 resources=res43)), {"bank": BANK}, {(0, 5): R2, (8, 0): R2})
    s["one_control_on_address_bit"] = (prog(mod("main", A4D3, [op("h", span("address", 3, 1)), ctrl(span("address", 3, 1), 1, LOAD43)],
                                                resources=res43)), {"bank": BANK}, {(0, 0): R2, (8, 1): R2})
    dense_mem = {a: (3 * a + 1) % 8 for a in range(16)}
    s["dense_zero_control"] = (prog(mod("main", A4D3, [op("h", span("address", 0, 4)), ctrl(span("address", 2, 1), 0, LOAD43)],
                                        resources=res43)), {"bank": dense_mem},
                               {(a, dense_mem[a] if not (a >> 2) & 1 else 0): 0.25 for a in range(16)})
    fad = [reg("flag", 1, "bits")] + A4D3
    s["external_one_control"] = (prog(mod("main", fad, [op("x", span("flag", 0, 1)), ctrl(span("flag", 0, 1), 1, LOAD43)],
                                          resources=res43)), {"bank": BANK}, {(1, 0, 5): 1})
    s["external_zero_control"] = (prog(mod("main", fad, [op("h", span("flag", 0, 1)), ctrl(span("flag", 0, 1), 0, LOAD43)],
                                           resources=res43)), {"bank": BANK}, {(0, 0, 5): R2, (1, 0, 0): R2})
    s["conflicting_cells"] = (prog(mod("main", [reg("address", 2, "bits"), reg("data", 3, "bits")],
                                       [op("h", span("address", 0, 2)),
                                        ctrl(span("address", 1, 1), 0, load("bank", span("address", 0, 2), span("data", 0, 3)))],
                                       resources=[qram("bank", 2, 3)])), {"bank": {0: 5, 1: 3, 2: 7, 3: 1}},
                              {(0, 5): 0.5, (1, 3): 0.5, (2, 0): 0.5, (3, 0): 0.5})
    s["contradictory_controls"] = (prog(mod("main", A4D3, [op("h", span("address", 0, 1)),
                                                           ctrl(span("address", 0, 1), 1, ctrl(span("address", 0, 1), 0, LOAD43))],
                                            resources=res43)), {"bank": {0: 5, 1: 7}}, {(0, 0): R2, (1, 0): R2})
    pre = [op("x", span("data", 0, 1)), op("x", span("data", 1, 1)), op("h", span("address", 3, 1))]
    s["xor_onto_nonzero_data"] = (prog(mod("main", A4D3, pre + [ctrl(span("address", 3, 1), 0, LOAD43)], resources=res43)),
                                  {"bank": BANK}, {(0, 3 ^ 5): R2, (8, 3): R2})
    s["query_twice"] = (prog(mod("main", A4D3, pre + [ctrl(span("address", 3, 1), 0, LOAD43, LOAD43)], resources=res43)),
                        {"bank": BANK}, {(0, 3): R2, (8, 3): R2})
    query = mod("query", [reg("q_address", 4, "bits"), reg("q_data", 3, "bits")],
                [load("mem", span("q_address", 0, 4), span("q_data", 0, 3))], resources=[qram("mem", 4, 3)])
    s["query_through_call"] = (prog(mod("main", A4D3, [op("h", span("address", 3, 1)),
                                                       ctrl(span("address", 3, 1), 0,
                                                            call("query", [span("address", 0, 4), span("data", 0, 3)], ["bank"]))],
                                        resources=res43), query), {"bank": BANK}, {(0, 5): R2, (8, 0): R2})
    s["repeat_two_queries"] = (prog(mod("main", A4D3, pre + [ctrl(span("address", 3, 1), 0, rep(2, LOAD43))], resources=res43)),
                               {"bank": BANK}, {(0, 3): R2, (8, 3): R2})
    s["zero_control_keeps_phase"] = (prog(mod("main", A4D3, [op("h", span("address", 1, 1)), op("rz", span("address", 1, 1), angle=THETA),
                                                             ctrl(span("address", 1, 1), 0, LOAD43)], resources=res43)),
                                     {"bank": BANK}, {(0, 5): R2 * cmath.exp(-0.5j * THETA), (2, 0): R2 * cmath.exp(0.5j * THETA)})
    return s


SCENARIOS = _scenarios()


def close(got: dict, want: dict, tol=1e-12):
    assert set(got) == {k for k, v in want.items() if abs(v) > 1e-15}, (sorted(got), sorted(want))
    for k, v in want.items():
        assert abs(got[k] - v) <= tol, (k, got[k], v)


# ---------------------------------------------------------------- CPU: document checks and the dense checker
def test_load_rejects_malformed_documents():
    with pytest.raises(RIRError, match="version"):
        load_rir(dict(prog(mod("main", [], [])), version="9.9"))
    with pytest.raises(RIRError, match="entry"):
        load_rir(prog(mod("main", [], []), entry="nope"))
    with pytest.raises(RIRError, match="Program"):
        load_rir({"tag": "Module"})
    assert load_rir(json.dumps(prog(mod("main", [], []))))["entry"] == "main"


def test_budget_and_missing_bodies_fail_before_running():
    with pytest.raises(RIRError, match="implementation"):
        run_rir(prog(mod("main", [reg("q", 1)], None)))
    with pytest.raises(RIRError, match="budget"):
        run_rir(prog(mod("main", [reg("x", 4)], [rep(10**9, op("x", span("x", 0, 4)))])), max_steps=1000)


def test_memory_must_match_resources():
    doc = SCENARIOS["load_superposition"][0]
    with pytest.raises(RIRError, match="QRAM"):
        run_rir(doc, memory={})
    with pytest.raises(RIRError, match="out of range"):
        run_rir(doc, memory={"values": [8, 0, 0, 0]})


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_dense_checker_reproduces_reference_expectations(name):
    doc, memory, want = SCENARIOS[name]
    close(DenseRIR(doc, memory).result(), want)


def test_dense_checker_rejects_dirty_local():
    doc = prog(mod("main", [reg("out", 1, "bits")], [op("x", span("w", 0, 1))], locals_=[reg("w", 1, "bits")]))
    with pytest.raises(ValueError, match="uncomputed"):
        DenseRIR(doc).result()


# ---------------------------------------------------------------- engine
@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_engine_reproduces_reference_expectations(ps, name):
    doc, memory, want = SCENARIOS[name]
    close(run_rir(doc, memory).amplitudes, want)


@pytest.mark.gpu
def test_engine_rejects_dirty_local_and_clears_registers(ps):
    doc = prog(mod("main", [reg("out", 1, "bits")], [op("x", span("w", 0, 1))], locals_=[reg("w", 1, "bits")]))
    with pytest.raises(RIRError, match="uncomputed"):
        run_rir(doc)
    assert ps.System.get_activated_register_size() == 0


@pytest.mark.gpu
def test_engine_file_and_statevector(ps, tmp_path):
    path = tmp_path / "h.rir.json"
    path.write_text(json.dumps(prog(mod("main", [reg("q", 1, "bits")], [op("h", span("q", 0, 1))]))))
    close(run_rir_file(path).amplitudes, {(0,): R2, (1,): R2})
    doc, memory, _ = SCENARIOS["load_superposition"]
    vec = run_rir(doc, memory).statevector()
    assert abs(vec[(2 << 2) | 1] - 0.5) < 1e-12 and abs(vec[(7 << 2) | 3] - 0.5) < 1e-12


@pytest.mark.gpu
def test_engine_wide_sparse_qram(ps):
    """A 30-bit address resource runs cell by cell (no device table)."""
    doc = prog(mod("main", [reg("address", 30), reg("data", 5)],
                   [op("h", span("address", 0, 2)), op("x", span("address", 29, 1)),
                    load("big", span("address", 0, 30), span("data", 0, 5))], resources=[qram("big", 30, 5)]))
    top = 1 << 29
    got = run_rir(doc, {"big": {top | 1: 17, top | 3: 9, 5: 31}}).amplitudes
    close(got, {(top, 0): 0.5, (top | 1, 17): 0.5, (top | 2, 0): 0.5, (top | 3, 9): 0.5})


def random_program(rng):
    """Registers of 1-3 bits; a body of gates, phases, xor / swap / add_const on
    random views, controls on a register the body does not touch, adjoint and
    repeat blocks, a call with views and a QRAM query."""
    widths = [int(w) for w in rng.integers(1, 4, size=4)]
    names = ["r0", "r1", "r2", "c"]
    regs = [reg(n, w, "bits") for n, w in zip(names, widths)]

    def rview(n):
        w = widths[names.index(n)]
        a = int(rng.integers(0, w))
        b = int(rng.integers(a + 1, w + 1))
        return span(n, a, b - a)

    def prim():
        k = int(rng.integers(0, 6))
        tgt = names[int(rng.integers(0, 3))]
        if k == 0:
            g = ["h", "x", "y", "z", "s", "t"][int(rng.integers(0, 6))]
            return op(g, rview(tgt))
        if k == 1:
            g = ["rx", "ry", "rz", "phase"][int(rng.integers(0, 4))]
            return op(g, rview(tgt), angle=float(rng.uniform(-3, 3)))
        if k == 2:
            return op("gphase", angle=float(rng.uniform(-3, 3)))
        if k in (3, 4):
            a, b = rng.choice(3, size=2, replace=False)
            w = min(widths[a], widths[b])
            return op("xor" if k == 3 else "swap", span(names[a], 0, w), span(names[b], widths[b] - w, w))
        return op("add_const", rview(tgt), value=int(rng.integers(-9, 10)))

    body = [op("h", span("c", 0, widths[3])), op("h", span("r0", 0, widths[0]))]
    for _ in range(int(rng.integers(6, 12))):
        r = rng.random()
        if r < 0.55:
            body.append(prim())
        elif r < 0.75:
            body.append(ctrl(rview("c"), int(rng.integers(0, 8)), prim(), prim()))
        elif r < 0.85:
            body.append(adj(prim(), prim()))
        else:
            body.append(rep(int(rng.integers(1, 3)), prim()))
    body.append(ctrl(span("c", 0, 1), int(rng.integers(0, 2)), load("m", span("r0", 0, widths[0]), span("r1", 0, widths[1]))))
    body.append(call("sub", [span("r2", 0, widths[2]), span("r0", 0, widths[0])]))
    sub = mod("sub", [reg("p", widths[2], "bits"), reg("q", widths[0], "bits")],
              [op("ry", span("p", 0, widths[2]), angle=0.3), op("add_const", span("q", 0, widths[0]), value=3)])
    mem = {"m": [int(v) for v in rng.integers(0, 1 << widths[1], size=1 << widths[0])]}
    main = mod("main", regs, body, resources=[qram("m", widths[0], widths[1])])
    return prog(main, sub), mem


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(24))
def test_engine_equals_dense_checker_on_random_programs(ps, seed):
    doc, mem = random_program(np.random.default_rng(seed))
    close(run_rir(doc, mem).amplitudes, DenseRIR(doc, mem).result(), tol=1e-11)
