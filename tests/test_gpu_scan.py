"""The single-pass exclusive scan every row offset and compaction index of
the pipeline goes through (broadphase rows, contact compaction, the sweep's
body incidences), called alone through sv_debug_scan and checked against
numpy's cumulative sum (mod 2^32): sizes around the 1024-value tile, long
look-back chains, counts read on the device below the launch size, and runs
of calls of changing size on one context (each call zeroes the status half
the next one uses)."""
import numpy as np
import pytest

from seevr_amd import Context, default_caps, default_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(default_caps(max_bodies=64, max_hull_verts=64, max_vertices=64, max_triangles=64,
                             max_support_entries=64, max_pairs=64, max_contacts=64), default_params())
    yield c
    c.close()


def expect(v):
    c = np.cumsum(v.astype(np.uint64)) & 0xFFFFFFFF
    ex = np.concatenate([[0], c[:-1]]).astype(np.uint32) if v.size else np.zeros(0, np.uint32)
    tot = int(c[-1]) if v.size else 0
    return ex, tot


def check(ctx, v, n_max=None):
    out, tot = ctx.debug_scan(v, n_max)
    ex, et = expect(v)
    assert tot == et
    np.testing.assert_array_equal(out, ex)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 255, 1023, 1024, 1025, 2047, 2048, 2049, 12295, 65536, 300001])
def test_scan_matches_cumsum(ctx, n):
    rng = np.random.default_rng(n)
    check(ctx, rng.integers(0, 8, n, dtype=np.uint32))


def test_scan_long_lookback_and_wraparound(ctx):
    rng = np.random.default_rng(7)
    # 3M values: ~3000 tiles; values up to 2^31 so the sums wrap mod 2^32
    check(ctx, rng.integers(0, 1 << 31, 3_000_001, dtype=np.uint32))
    check(ctx, np.ones(5_000_000, np.uint32))


@pytest.mark.parametrize("n,n_max", [(0, 1), (0, 5000), (1, 1), (1000, 50000), (4096, 4097), (70000, 2_000_000)])
def test_scan_count_on_device(ctx, n, n_max):
    rng = np.random.default_rng(n_max)
    check(ctx, rng.integers(0, 100, n, dtype=np.uint32), n_max)


def test_scan_runs_of_changing_size(ctx):
    rng = np.random.default_rng(3)
    for n, n_max in [(3_000_000, None), (10, None), (1_000_000, None), (2049, None), (0, 5000),
                     (3_000_000, None), (5, 3_000_000), (1024, None), (2_500_000, None)] * 2:
        check(ctx, rng.integers(0, 16, n, dtype=np.uint32), n_max)


def test_scan_rejects_bad_arguments(ctx):
    with pytest.raises(Exception):
        ctx.debug_scan(np.zeros(4, np.uint32), n_max=3)
