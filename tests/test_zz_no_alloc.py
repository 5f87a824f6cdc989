"""sparq_clear_zero_noalloc (the allocation-free ClearZero) against the
existing entry point sparq_clear_zero, both through the C ABI: the same kept
entries, bit for bit, on two calls in a row (the second smaller, so stale
scratch contents from the first would show), on several key widths, on an
empty input, and the size limit refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U64P = ctypes.POINTER(ctypes.c_uint64)
F64P = ctypes.POINTER(ctypes.c_double)


def _lib():
    import sparqsim_amd

    lib = ctypes.CDLL(sparqsim_amd.ENGINE_LIBRARY)
    lib.sparq_last_error.restype = ctypes.c_char_p
    lib.sparq_clear_zero_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64,
                                             ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                             ctypes.c_void_p]
    lib.sparq_state_create.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]
    lib.sparq_state_upload.argtypes = [ctypes.c_void_p, U64P, F64P, ctypes.c_uint64]
    lib.sparq_state_download.argtypes = [ctypes.c_void_p, U64P, F64P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    lib.sparq_clear_zero.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.sparq_state_destroy.argtypes = [ctypes.c_void_p]
    return lib


def _ok(lib, rc):
    assert rc == 0, lib.sparq_last_error().decode()


def _inputs(n, words, seed):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 2 ** 63, size=(n, words), dtype=np.uint64)       # entry-major
    mag = 10.0 ** rng.uniform(-10, 0, n)                                    # |a|^2 from 1e-20 to 1
    amps = mag * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    return keys, amps


def _reference(lib, keys, amps, eps):
    """sparq_clear_zero on a state holding the same entries."""
    n, words = keys.shape
    h = ctypes.c_void_p()
    _ok(lib, lib.sparq_state_create(words, max(n, 1), ctypes.byref(h)))
    k = np.ascontiguousarray(keys)
    a = np.ascontiguousarray(np.stack([amps.real, amps.imag], axis=1).reshape(-1))
    _ok(lib, lib.sparq_state_upload(h, k.ctypes.data_as(U64P), a.ctypes.data_as(F64P), n))
    _ok(lib, lib.sparq_clear_zero(h, eps))
    ko = np.zeros((max(n, 1), words), dtype=np.uint64)
    ao = np.zeros(2 * max(n, 1))
    m = ctypes.c_uint64()
    _ok(lib, lib.sparq_state_download(h, ko.ctypes.data_as(U64P), ao.ctypes.data_as(F64P), max(n, 1), ctypes.byref(m)))
    _ok(lib, lib.sparq_state_destroy(h))
    return ko[: m.value], ao[: 2 * m.value].view(np.complex128)


def _noalloc(lib, torch, keys, amps, eps, pad=0):
    """sparq_clear_zero_noalloc on device arrays the caller owns ([words][cap] key words)."""
    n, words = keys.shape
    cap = n + pad
    kin = torch.zeros((words, max(cap, 1)), dtype=torch.int64, device="cuda")
    ain = torch.zeros((max(cap, 1), 2), dtype=torch.float64, device="cuda")
    if n:
        kin[:, :n] = torch.from_numpy(keys.T.copy().view(np.int64)).cuda()
        ain[:n, 0] = torch.from_numpy(amps.real.copy()).cuda()
        ain[:n, 1] = torch.from_numpy(amps.imag.copy()).cuda()
    kout = torch.full((words, max(cap, 1)), -1, dtype=torch.int64, device="cuda")
    aout = torch.full((max(cap, 1), 2), float("nan"), dtype=torch.float64, device="cuda")
    kept = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    _ok(lib, lib.sparq_clear_zero_noalloc(kin.data_ptr(), ain.data_ptr(), words, n, max(cap, 1), eps, kout.data_ptr(),
                                          aout.data_ptr(), max(cap, 1), kept.data_ptr(), None))
    torch.cuda.synchronize()
    m = int(kept.item())
    ko = kout[:, :m].cpu().numpy().view(np.uint64).T.copy()
    ao = aout[:m].cpu().numpy()
    return ko, ao[:, 0] + 1j * ao[:, 1]


def _same(got, want):
    gk, ga = got
    wk, wa = want
    assert gk.shape == wk.shape
    assert np.array_equal(gk, wk)
    assert np.array_equal(ga.real, wa.real) and np.array_equal(ga.imag, wa.imag)


@pytest.mark.parametrize("words", [1, 3, 5])
def test_noalloc_clear_zero_matches_existing_entry_point(words):
    torch = pytest.importorskip("torch")
    lib = _lib()
    eps = 1e-6
    for n, seed in ((300_000, 1), (70_001, 2)):          # second call smaller than the first
        keys, amps = _inputs(n, words, seed + 10 * words)
        want = _reference(lib, keys, amps, eps)
        got = _noalloc(lib, torch, keys, amps, eps, pad=17)
        assert 0 < want[0].shape[0] < n
        _same(got, want)


def test_noalloc_clear_zero_edges():
    torch = pytest.importorskip("torch")
    lib = _lib()
    keys, amps = _inputs(0, 2, 3)
    k, a = _noalloc(lib, torch, keys, amps, 1e-14)
    assert k.shape == (0, 2) and a.shape == (0,)
    keys, amps = _inputs(1, 1, 4)
    _same(_noalloc(lib, torch, keys, amps, 0.0), _reference(lib, keys, amps, 0.0))   # keeps everything
    _same(_noalloc(lib, torch, keys, amps, 2.0), _reference(lib, keys, amps, 2.0))   # drops everything
    kept = torch.zeros(1, dtype=torch.int64, device="cuda")
    rc = lib.sparq_clear_zero_noalloc(None, None, 1, (1 << 28) + 1, (1 << 28) + 1, 1e-14, None, None, (1 << 28) + 1,
                                      kept.data_ptr(), None)
    assert rc == 3 and b"2^28" in lib.sparq_last_error()
