"""CPU checks of the drop-in boundary: the engine library loads and exports every
symbol declared in include/*.h; host-side validation errors surface as Python
exceptions without touching a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"^\s*(?:int|const char\s*\*)\s+(sparq_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_declares_the_engine_api():
    names = declared_functions()
    for required in ("sparq_state_create", "sparq_pointwise", "sparq_hadamard_int", "sparq_sort",
                     "sparq_clear_zero", "sparq_pair_op", "sparq_add_systems", "sparq_prof_report"):
        assert required in names


def test_library_exports_every_declared_symbol():
    import sparqsim_amd

    lib = ctypes.CDLL(sparqsim_amd.ENGINE_LIBRARY)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_no_device_is_reported_not_faked():
    """Without a GPU the engine refuses to create a state (no CPU fallback)."""
    import sparqsim_amd as ps

    try:
        import torch
        has_gpu = torch.cuda.device_count() > 0
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    ps.System.clear()
    with pytest.raises(RuntimeError, match="no HIP device"):
        ps.SparseState()


def test_registry_semantics_host_only():
    """The product's register table follows the reference id rules (no GPU needed)."""
    import sparqsim_amd as ps

    ps.System.clear()
    a = ps.System.add_register("a", ps.UnsignedInteger, 4)
    b = ps.System.add_register("b", ps.UnsignedInteger, 5)
    c = ps.System.add_register("c", ps.Boolean, 1)
    assert (a, b, c) == (0, 1, 2)
    ps.System.remove_register("b")
    assert ps.System.add_register("d", ps.UnsignedInteger, 7) == 1      # LIFO reuse
    assert ps.System.get_id("d") == 1 and ps.System.size_of("d") == 7
    assert ps.System.get_qubit_count() == 12
    ps.System.clear()
    with pytest.raises(ValueError):
        ps.AddRegister("w", ps.UnsignedInteger, 65)


def test_profiler_and_sort_counters_host_only():
    """The engine profiler and the sort-path counters answer without a GPU:
    nothing launched, nothing recorded."""
    import sparqsim_amd as ps
    from sparqsim_amd import profiling

    with profiling.KernelProfile(ps) as prof:
        pass
    assert prof.records == []
    assert prof.sort_stats == {"local_sorts": 0, "local_entries": 0, "local_overflows": 0, "lsd_sorts": 0,
                               "lsd_entries": 0, "presorted_sorts": 0}
    lib = profiling.engine_library()
    out = (ctypes.c_uint64 * 5)()
    assert lib.sparq_sort_stats(out, 5, 0) != 0          # too little room is an error, not an overrun
