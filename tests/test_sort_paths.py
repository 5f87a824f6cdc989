"""The two sort strategies give the same stable order: the prefix passes + LDS
chunk sort (default) against the LSD radix passes with an index payload, and
both against numpy's stable lexicographic sort of the same ordering.  Sizes
cover one chunk, the chunk edge, one and two prefix passes, keys of one to
five words, more than 64 ordering bits below the prefix, grouping orders, and
a skewed key set whose bucket overflows a chunk (the fallback)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _regs(mod, widths):
    st = mod.SparseState()
    for name, w in widths:
        mod.AddRegister(name, mod.UnsignedInteger, w)(st)
    return st


def _sorted_both(ps, widths, vals, amps, op_factory):
    out = []
    for mode in (0, 1):
        ps.set_sort_mode(mode)
        try:
            ps.System.clear()
            st = _regs(ps, widths)
            st.load(vals, amps)
            op_factory()(st)
            out.append(st.export())
        finally:
            ps.set_sort_mode(0)
    return out


def _expect_order(vals, key_cols):
    """Stable lexicographic order by the given columns (first most significant)."""
    return np.lexsort(tuple(vals[:, c] for c in reversed(key_cols)))


def _check(out, vals, amps, key_cols):
    (ids0, v0, a0), (ids1, v1, a1) = out
    assert np.array_equal(v0, v1) and np.array_equal(a0.view(np.float64), a1.view(np.float64)), "strategies differ"
    order = _expect_order(vals, key_cols)
    assert np.array_equal(v0, vals[order])
    assert np.array_equal(a0, amps[order])


@pytest.mark.parametrize("n", [1, 2, 1000, 4096, 4097, 9000, 200_000, 2_000_000])
def test_canonical_order_one_word(ps, n):
    widths = [("a", 20), ("b", 28)]
    rng = np.random.default_rng(n)
    vals = np.stack([rng.integers(0, 1 << 20, n), rng.integers(0, 1 << 28, n)], axis=1).astype(np.uint64)
    vals[: n // 7] = vals[n // 3: n // 3 + n // 7]          # duplicated keys: stability decides their order
    amps = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    _check(_sorted_both(ps, widths, vals, amps, ps.SortUnconditional), vals, amps, [0, 1])


@pytest.mark.parametrize("words", [2, 3, 5])
@pytest.mark.parametrize("n", [3000, 150_000])
def test_canonical_order_wide_keys(ps, words, n):
    """Register widths chosen so that the packed key takes `words` words and
    (for 3 and 5 words) more than 64 ordering bits lie below the prefix."""
    widths = [(f"r{i}", 60) for i in range(words - 1)] + [("last", 37)]
    ps.System.clear()
    rng = np.random.default_rng(words * 100 + n)
    vals = np.stack([rng.integers(0, 1 << w, n, dtype=np.uint64) for _, w in widths], axis=1)
    vals[: n // 7] = vals[n // 3: n // 3 + n // 7]          # duplicated keys
    amps = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    _check(_sorted_both(ps, widths, vals, amps, ps.SortUnconditional), vals, amps, list(range(words)))


@pytest.mark.parametrize("n", [2500, 120_000])
def test_grouping_orders(ps, n):
    """SortExceptKey (b last) and SortByKey2 (c, a first) on the prefix + local
    path against the LSD path and numpy."""
    widths = [("a", 9), ("b", 21), ("c", 14)]
    rng = np.random.default_rng(n + 5)
    vals = np.stack([rng.integers(0, 1 << w, n, dtype=np.uint64) for _, w in widths], axis=1)
    amps = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    _check(_sorted_both(ps, widths, vals, amps, lambda: ps.SortExceptKey("b")), vals, amps, [0, 2, 1])
    _check(_sorted_both(ps, widths, vals, amps, lambda: ps.SortByKey2("c", "a")), vals, amps, [2, 0, 1])


def test_skewed_bucket_falls_back(ps):
    """90% of the entries share their top 24 ordering bits: that bucket cannot
    be one chunk, so the LSD path sorts the state -- same result."""
    n = 300_000
    widths = [("hi", 24), ("lo", 30)]
    rng = np.random.default_rng(9)
    hi = rng.integers(0, 1 << 24, n, dtype=np.uint64)
    hi[: int(0.9 * n)] = 12345
    vals = np.stack([hi, rng.integers(0, 1 << 30, n, dtype=np.uint64)], axis=1)
    amps = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    _check(_sorted_both(ps, widths, vals, amps, ps.SortUnconditional), vals, amps, [0, 1])


def test_few_ordering_bits(ps):
    """Only 3 varying bits over 100 000 entries (buckets far above a chunk):
    the LSD path is taken up front."""
    n = 100_000
    widths = [("s", 3), ("z", 20)]
    rng = np.random.default_rng(10)
    vals = np.stack([rng.integers(0, 8, n, dtype=np.uint64), np.full(n, 7, dtype=np.uint64)], axis=1)
    amps = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    _check(_sorted_both(ps, widths, vals, amps, ps.SortUnconditional), vals, amps, [0, 1])
