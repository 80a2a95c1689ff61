"""The all-pairs broadphase scratch of a context covers every body count up
to its capacity.  The segment count bp_nseg(n) is not monotonic in n (it
drops from 64 at n = 131072 to 57 at n = 131073), so sizing from the
capacity alone under-allocates; these checks run the library's own sizing
code on the host (no device work)."""
import ctypes as C

import pytest


def scratch(cap: int, n: int):
    from seevr_amd import _lib
    lib = _lib.load()
    out = (C.c_uint64 * 4)()
    assert lib.sv_debug_broadphase_scratch(cap, n, out) == 0
    return tuple(int(v) for v in out)


def nseg_restated(n: int) -> int:
    """bp_nseg restated: segments of max(2048, ceil(n / 64) rounded up to 256)."""
    s = -(-n // 64)
    s = -(-s // 256) * 256
    seg = max(2048, s)
    return -(-n // seg) if n else 1


def test_segment_count_is_not_monotonic():
    # the case the capacity-only sizing missed
    assert nseg_restated(131072) == 64
    assert nseg_restated(131073) == 57
    need, have, _, _ = scratch(131073, 131071)
    assert need == 131071 * 64 + 1 <= have


@pytest.mark.parametrize("cap", [1, 41, 2047, 2048, 2049, 4096, 32773, 131072, 131073, 138240, 140000,
                                 200000, 1 << 18, 1048577])
def test_scratch_covers_every_count_below_capacity(cap):
    ns = set(range(1, min(cap, 3000) + 1))
    ns |= {cap, cap - 1, cap // 2, 196608, 131071, 131072, 131073}
    ns |= set(range(max(1, cap - 5000), cap + 1, 7))
    for n in sorted(x for x in ns if 1 <= x <= cap):
        cells, cells_cap, ovf, ovf_cap = scratch(cap, n)
        assert cells == n * nseg_restated(n) + 1, n
        assert cells <= cells_cap, (cap, n, cells, cells_cap)
        # one count word plus one entry per (row tile of 32, segment)
        assert ovf == 1 + -(-n // 32) * nseg_restated(n), n
        assert ovf <= ovf_cap, (cap, n, ovf, ovf_cap)


def test_small_scene_overflow_list_has_its_count_word():
    # 41 bodies: 2 row tiles in one segment, both may be listed after the count
    cells, cells_cap, ovf, ovf_cap = scratch(41, 41)
    assert ovf == 3 and ovf_cap >= 3


def test_all_pairs_limit():
    # past 2^25 bodies the LBVH broadphase runs; the cells stay 32-bit indexable
    cells, cells_cap, ovf, ovf_cap = scratch(1 << 26, 1 << 26)
    assert cells == 0 and ovf == 0
    assert cells_cap < 2 ** 32
