"""The stable LSD radix sort every sorted order of the pipeline goes through
(the LBVH's Morton order, bvh.comp:170-260; the broadphase pair order; the
sweep's body incidences), called alone through sv_debug_radix_sort and
checked against numpy's stable argsort of the masked keys: the same
permutation, element for element, so equal keys keep their input order.

Sizes cross the tile boundaries of both tile shapes (2048 keys up to 2^20,
8192 above) and the bound of the one-launch form (capacity up to 2^19 keys,
grid barriers between passes), the device count is below the buffer
capacity (the idle workgroups' path), keys carry bits above key_bits, and
one context runs many sorts of both forms in a row (the counters each sort
leaves at zero)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from seevr_amd import Context, default_caps
    c = Context(default_caps(max_bodies=64, max_hull_verts=64, max_vertices=64, max_triangles=64))
    yield c
    c.close()


def check(ctx, keys, bits, n_max=None, vals=None):
    keys = np.asarray(keys, np.uint32)
    vals = np.arange(keys.size, dtype=np.uint32) if vals is None else np.asarray(vals, np.uint32)
    ko, vo = ctx.debug_radix_sort(keys, vals, bits, n_max)
    mask = np.uint32(0xFFFFFFFF if bits == 32 else (1 << bits) - 1)
    order = np.argsort(keys & mask, kind="stable")
    assert np.array_equal(ko, keys[order]), f"keys n={keys.size} bits={bits}"
    assert np.array_equal(vo, vals[order]), f"values n={keys.size} bits={bits}"


@pytest.mark.parametrize("n", [0, 1, 2, 5, 2047, 2048, 2049, 40000, 1 << 18, (1 << 18) + 1, 1 << 19,
                               (1 << 19) + 1, 8192 * 65 + 3])
@pytest.mark.parametrize("bits", [32, 30, 21, 16, 8, 3])
def test_random_keys(ctx, n, bits):
    rng = np.random.default_rng(n * 37 + bits)
    keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    if bits < 32:
        keys &= np.uint32((1 << bits) - 1)
    check(ctx, keys, bits, vals=rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32))


def test_bits_above_key_bits_are_carried_not_sorted(ctx):
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 1 << 32, 300001, dtype=np.uint64).astype(np.uint32)
    check(ctx, keys, 12)


@pytest.mark.parametrize("kind", ["equal", "sorted", "reversed", "few", "clustered"])
def test_structured_keys(ctx, kind):
    n = 1_000_003
    rng = np.random.default_rng(11)
    keys = {"equal": np.full(n, 0xABCDEF12, np.uint64),
            "sorted": np.sort(rng.integers(0, 1 << 30, n)),
            "reversed": np.sort(rng.integers(0, 1 << 30, n))[::-1],
            "few": rng.integers(0, 4, n) * 0x01010101,
            "clustered": np.repeat(rng.integers(0, 1 << 30, n // 257 + 1), 257)[:n]}[kind]
    check(ctx, keys.astype(np.uint32), 32)


@pytest.mark.parametrize("n,n_max", [(0, 5000), (1000, 1_000_000), (300000, 4_000_000), (2049, 2_000_000),
                                     (110000, 395000), (4095, 1 << 19), (1 << 19, 1 << 19)])
def test_device_count_below_capacity(ctx, n, n_max):
    rng = np.random.default_rng(n + n_max)
    keys = rng.integers(0, 1 << 21, n).astype(np.uint32)
    check(ctx, keys, 21, n_max=n_max)


def test_many_sorts_in_a_row(ctx):
    rng = np.random.default_rng(5)
    for i, n in enumerate([3, 9000000, 100, 70000, 2048 * 128, 0, 5_000_000, 17, 600000, 32773, 1 << 19]):
        keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        check(ctx, keys, [32, 30, 16, 24][i % 4] if n else 32)


def test_rejects_bad_arguments(ctx):
    with pytest.raises(Exception):
        ctx.debug_radix_sort(np.zeros(4, np.uint32), np.zeros(4, np.uint32), 33)
    with pytest.raises(Exception):
        ctx.debug_radix_sort(np.zeros(4, np.uint32), np.zeros(4, np.uint32), 8, n_max=3)
