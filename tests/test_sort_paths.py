"""The two sort strategies give the same stable order: the prefix passes + LDS
chunk sort (default) against the LSD radix passes with an index payload, and
both against numpy's stable lexicographic sort of the same ordering.  Sizes
cover one chunk, the chunk edge, one and two prefix passes, keys of one to
five words, more than 64 ordering bits below the prefix, grouping orders, a
skewed key set that needs a wider prefix, one no prefix of <= 32 bits
splits (the LSD fallback), and states already sorted by a key prefix (the
chunk sort alone), including pair operators on such states against the
oracle."""
import numpy as np
import pytest

from parity import assert_same_state, run_both

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


def _sort_stats_of(ps, widths, vals, amps, op_factory=None):
    """Sort once with the default strategy; return the export and the sort-path counters."""
    ps.sort_stats(reset=True)
    ps.System.clear()
    st = _regs(ps, widths)
    st.load(vals, amps)
    (op_factory or ps.SortUnconditional)()(st)
    return st.export(), ps.sort_stats(reset=True)


def test_skewed_bucket_widens_prefix(ps):
    """90% of the entries share their top 24 ordering bits: buckets of a 16-
    and a 24-bit prefix overflow a chunk; a 32-bit prefix splits the heavy
    bucket 256 ways (~1 050 entries each), which fits once the chunk spacing
    leaves room for such buckets.  The next sort of that ordering starts at the
    level that worked: no overflow, no LSD pass."""
    n = 300_000
    widths = [("hi", 24), ("lo", 30)]
    rng = np.random.default_rng(9)
    hi = rng.integers(0, 1 << 24, n, dtype=np.uint64)
    hi[: int(0.9 * n)] = 12345
    vals = np.stack([hi, rng.integers(0, 1 << 30, n, dtype=np.uint64)], axis=1)
    amps = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    _check(_sorted_both(ps, widths, vals, amps, ps.SortUnconditional), vals, amps, [0, 1])
    perm = rng.permutation(n)
    out, stats = _sort_stats_of(ps, widths, vals[perm], amps[perm])
    assert stats["local_sorts"] == 1 and stats["local_overflows"] == 0 and stats["lsd_sorts"] == 0, stats
    order = _expect_order(vals[perm], [0, 1])
    assert np.array_equal(out[1], vals[perm][order]) and np.array_equal(out[2], amps[perm][order])


def test_skewed_bucket_falls_back(ps):
    """90% of the entries share their top 32 ordering bits: no prefix of at most
    32 bits splits that bucket into chunks, so after five attempts (16, 24, 32
    prefix bits, then 32 bits with two smaller chunk spacings) the LSD path
    sorts the state -- same result; the next sort of the ordering goes there
    directly."""
    n = 600_000   # another size class than the test above: the prefix width is remembered per class
    widths = [("hi", 24), ("lo", 30)]
    rng = np.random.default_rng(11)
    hi = rng.integers(0, 1 << 24, n, dtype=np.uint64)
    lo = rng.integers(0, 1 << 30, n, dtype=np.uint64)
    heavy = int(0.9 * n)
    hi[:heavy] = 777
    lo[:heavy] = (5 << 22) | rng.integers(0, 1 << 22, heavy, dtype=np.uint64)
    vals = np.stack([hi, lo], axis=1)
    amps = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    _, first = _sort_stats_of(ps, widths, vals, amps)
    assert first["local_sorts"] == 0 and first["local_overflows"] == 5 and first["lsd_sorts"] == 1, first
    perm = rng.permutation(n)
    out, again = _sort_stats_of(ps, widths, vals[perm], amps[perm])
    assert again["local_overflows"] == 0 and again["lsd_sorts"] == 1, again
    order = _expect_order(vals[perm], [0, 1])
    assert np.array_equal(out[1], vals[perm][order]) and np.array_equal(out[2], amps[perm][order])
    _check(_sorted_both(ps, widths, vals, amps, ps.SortUnconditional), vals, amps, [0, 1])


def test_few_ordering_bits(ps):
    """Only 3 varying bits over 100 000 entries (buckets far above a chunk):
    one whole-entry pass over the 3 bits sorts the state, no chunk sort."""
    n = 100_000
    widths = [("s", 3), ("z", 20)]
    rng = np.random.default_rng(10)
    vals = np.stack([rng.integers(0, 8, n, dtype=np.uint64), np.full(n, 7, dtype=np.uint64)], axis=1)
    amps = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    _check(_sorted_both(ps, widths, vals, amps, ps.SortUnconditional), vals, amps, [0, 1])
    _, stats = _sort_stats_of(ps, widths, vals, amps)
    assert stats["local_sorts"] == 1 and stats["local_overflows"] == 0 and stats["lsd_sorts"] == 0, stats


@pytest.mark.parametrize("n", [300_000, 1_000_000])
def test_short_ordering_passes_only(ps, n):
    """17 varying ordering bits (a 9-bit and an 8-bit register, the rest
    constant): the 16-bit prefix these sizes call for leaves one bit, so three
    whole-entry passes (8 + 8 + 1 bits) sort the state with no chunk sort; a
    second key word rides along unchanged."""
    widths = [("a", 9), ("pad", 60), ("b", 8)]
    rng = np.random.default_rng(n + 77)
    vals = np.stack([rng.integers(0, 1 << 9, n, dtype=np.uint64), np.full(n, (1 << 59) + 3, dtype=np.uint64),
                     rng.integers(0, 1 << 8, n, dtype=np.uint64)], axis=1)
    amps = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    _check(_sorted_both(ps, widths, vals, amps, ps.SortUnconditional), vals, amps, [0, 1, 2])
    _, stats = _sort_stats_of(ps, widths, vals, amps)
    assert stats["local_sorts"] == 1 and stats["lsd_sorts"] == 0, stats


@pytest.mark.parametrize("hi_bits", [20, 14, 12])
def test_presorted_prefix_skips_passes(ps, hi_bits):
    """Keys sorted by `hi` but not within equal `hi`: the key statistics find
    the sorted prefix and the chunk sort alone orders the state: buckets of a
    few entries (20-bit prefix), ~24 (14 bits) and ~100 (12 bits)."""
    n = 400_000
    widths = [("hi", hi_bits), ("lo", 12)]
    rng = np.random.default_rng(21)
    vals = np.stack([np.sort(rng.integers(0, 1 << hi_bits, n, dtype=np.uint64)),
                     rng.integers(0, 1 << 12, n, dtype=np.uint64)], axis=1)
    amps = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    out, stats = _sort_stats_of(ps, widths, vals, amps)
    assert stats["presorted_sorts"] == 1 and stats["lsd_sorts"] == 0, stats
    order = _expect_order(vals, [0, 1])
    assert np.array_equal(out[1], vals[order]) and np.array_equal(out[2], amps[order])
    _check(_sorted_both(ps, widths, vals, amps, ps.SortUnconditional), vals, amps, [0, 1])


def test_presorted_grouping_sort(ps):
    """A canonically sorted state grouped by SortExceptKey("b") (order a, c, b):
    it is already sorted by `a`, so the chunk sort alone regroups it."""
    n = 250_000
    widths = [("a", 16), ("b", 3), ("c", 10)]
    rng = np.random.default_rng(22)
    vals = np.stack([rng.integers(0, 1 << w, n, dtype=np.uint64) for _, w in widths], axis=1)
    vals = vals[_expect_order(vals, [0, 1, 2])]
    amps = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    out, stats = _sort_stats_of(ps, widths, vals, amps, lambda: ps.SortExceptKey("b"))
    assert stats["presorted_sorts"] == 1 and stats["lsd_sorts"] == 0, stats
    order = _expect_order(vals, [0, 2, 1])
    assert np.array_equal(out[1], vals[order]) and np.array_equal(out[2], amps[order])
    _check(_sorted_both(ps, widths, vals, amps, lambda: ps.SortExceptKey("b")), vals, amps, [0, 2, 1])


def test_pair_ops_on_large_sorted_state(ps, oracle):
    """Hadamard_Bool and Rot_Bool on the Boolean registers of a 2^17-entry
    superposition: each pair operator's grouping sort starts from a state
    sorted by the main register (and the registers above its target), the
    chunk-sort-only case; the result equals the oracle's."""
    mat = [complex(0.6, 0.1), complex(-0.3, 0.2), complex(0.5, -0.4), complex(0.2, 0.7)]

    def build(mod):
        st = mod.SparseState()
        mod.AddRegister("main", mod.UnsignedInteger, 17)(st)
        for q in ("q1", "q2", "q3"):
            mod.AddRegister(q, mod.Boolean, 1)(st)
        mod.Hadamard_Int("main", 17)(st)
        mod.Hadamard_Bool("q1")(st)
        mod.Hadamard_Bool("q3")(st)
        mod.Rot_Bool("q2", 0, mat)(st)
        mod.Hadamard_Bool("q1").conditioned_by_nonzeros("q3")(st)
        mod.Rot_Bool("q3", 0, mat)(st)
        mod.Hadamard_Bool("q2")(st)
        return st

    ps.sort_stats(reset=True)
    eg, eo = run_both(ps, oracle, build)
    stats = ps.sort_stats(reset=True)
    assert stats["presorted_sorts"] >= 1, stats
    assert_same_state(eg, eo, what="pair ops on a 2^17-entry superposition")
