"""The broadphase's row sort (pairs placed in their rows in any order, then
each row sorted: by one lane up to 16 partners, by rank in a workgroup's LDS
up to 2048, through an LDS bitmap of the partner range beyond), called alone
through sv_debug_sort_rows and checked against numpy: every path, bitmap
windows one at a time or several, and rows cut at max_pairs."""
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


def make_rows(n, seed, long_rows=()):
    """cnt and shuffled entries: most rows 0..20 partners, some 17..2048, the
    rows in long_rows (i, count) longer."""
    rng = np.random.default_rng(seed)
    cnt = np.minimum(rng.integers(0, 21, n), n - 1 - np.arange(n))
    mid = rng.choice(n // 2, size=min(40, n // 2), replace=False)
    cnt[mid] = np.minimum(rng.integers(17, 2049, mid.size), n - 1 - mid)
    for i, c in long_rows:
        cnt[i] = min(c, n - 1 - i)
    rows = []
    for i in range(n):
        js = rng.choice(np.arange(i + 1, n, dtype=np.uint32), size=cnt[i], replace=False) if cnt[i] else \
            np.zeros(0, np.uint32)
        rows.append(js)
    return cnt.astype(np.uint32), rows


def expect(cnt, rows, max_pairs):
    out = [np.stack([np.full(len(js), i, np.uint32), np.sort(js)], 1) for i, js in enumerate(rows) if len(js)]
    allp = np.concatenate(out) if out else np.zeros((0, 2), np.uint32)
    return allp[:max_pairs]


@pytest.mark.parametrize("n,seed,long_rows,window", [
    (200, 1, (), 0),
    (5000, 2, ((0, 4000), (3, 2049), (7, 2048)), 0),
    (40000, 3, ((0, 30000), (1, 12000), (500, 2500)), 0),
    (40000, 4, ((0, 30000), (1, 12000), (500, 2500)), 256),   # 8192-index windows: several per long row
    (40000, 5, ((0, 39999), (2, 20000)), 512),
])
def test_rows_sorted_like_numpy(ctx, n, seed, long_rows, window):
    cnt, rows = make_rows(n, seed, long_rows)
    entries = np.concatenate([r for r in rows if len(r)]).astype(np.uint32)
    total = int(cnt.sum())
    got = ctx.debug_sort_rows(cnt, entries, total, window)
    np.testing.assert_array_equal(got, expect(cnt, rows, total))


@pytest.mark.parametrize("frac", [0.0, 0.3, 0.5001, 0.97])
def test_rows_cut_at_max_pairs(ctx, frac):
    cnt, rows = make_rows(30000, 11, ((0, 25000), (4, 3000)))
    entries = np.concatenate([r for r in rows if len(r)]).astype(np.uint32)
    mp = max(1, int(frac * int(cnt.sum())))
    got = ctx.debug_sort_rows(cnt, entries, mp, 256)
    np.testing.assert_array_equal(got, expect(cnt, rows, mp))


def test_rows_of_every_short_length(ctx):
    # rows of exactly 0..17 partners around the one-lane limit, plus 2047..2049
    n = 6000
    rng = np.random.default_rng(5)
    cnt = np.zeros(n, np.uint32)
    lens = list(range(18)) + [2047, 2048, 2049]
    for k, c in enumerate(lens):
        cnt[k * 100] = c
    rows = [rng.choice(np.arange(i + 1, n, dtype=np.uint32), size=cnt[i], replace=False) for i in range(n)]
    entries = np.concatenate([r for r in rows if len(r)]).astype(np.uint32)
    got = ctx.debug_sort_rows(cnt, entries, int(cnt.sum()))
    np.testing.assert_array_equal(got, expect(cnt, rows, int(cnt.sum())))
