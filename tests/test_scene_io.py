"""Scene files (sv_scene_serialize / deserialize / save / load): the byte
layout of a one-box scene checked field by field by an independent reader,
round trips of seeded scenes (every staging array bit-identical, the bytes
re-serialize identically), and rejection of damaged or oversized files with
the scene left unchanged."""
import struct

import numpy as np
import pytest

from harness import make_scene

ARRAYS = ("VERTICES", "ORIGINAL_VERTICES", "INDICES", "BODY_POS", "BODY_QUAT", "BODY_VEL", "BODY_OMEGA",
          "BODY_INV_MASS", "BODY_INV_INERTIA", "BODY_VERT_START", "BODY_VERT_COUNT", "GEOM_BODY_IDX", "GEOM_DATA",
          "GEOM_FRICTION", "GEOM_RESTITUTION", "HULL_VERTS", "SUPPORT_LOOKUP_OFFSET", "SUPPORT_LOOKUP",
          "GEOM_HULL_RANGE")
TAGS = [b"VERT", b"OVRT", b"INDX", b"BPOS", b"BQUT", b"BVEL", b"BOMG", b"BIMS", b"BINE", b"BINI", b"GBDY", b"GDAT",
        b"GFRC", b"GRST", b"GVST", b"GVCT", b"HULL", b"SLOF", b"SLUT", b"MDEN"]


def fnv1a(b: bytes) -> int:
    h = 0xCBF29CE484222325
    for x in b:
        h = ((h ^ x) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def parse(data: bytes):
    """Independent reader of the documented layout."""
    assert data[:8] == b"SVSCENE\0"
    version, ns = struct.unpack_from("<II", data, 8)
    counts = struct.unpack_from("<7I", data, 16)
    assert struct.unpack_from("<I", data, 44)[0] == 0
    secs = {}
    for i in range(ns):
        tag, esz, off, ln, ck = struct.unpack_from("<4sIQQQ", data, 48 + 32 * i)
        secs[tag] = (esz, off, ln, ck)
    return version, counts, secs


def empty_scene(**over):
    from seevr_amd import Scene, default_caps
    caps = dict(max_bodies=64, max_hull_verts=20000, max_vertices=20000, max_triangles=40000,
                max_support_entries=64 * 193)
    caps.update(over)
    return Scene(default_caps(**caps))


def test_one_box_layout():
    sc = empty_scene()
    m = sc.add_material(1000.0)
    sc.add_box((1.0, 2.0, 3.0), (0.5, 0.25, 1.0), friction=0.4, restitution=0.2, material=m)
    data = sc.to_bytes()
    version, counts, secs = parse(data)
    assert version == 1
    assert counts == (8, 12, 1, 1, 8, 1, 0)
    assert list(secs) == TAGS
    ends = []
    for tag, (esz, off, ln, ck) in secs.items():
        assert esz == 4 and off % 64 == 0 and off >= 48 + 32 * len(secs)
        assert fnv1a(data[off: off + ln]) == ck, tag
        ends.append(off + ln)
    assert len(data) % 64 == 0 and max(ends) <= len(data)
    f = lambda tag: np.frombuffer(data, np.float32, secs[tag][2] // 4, secs[tag][1])
    u = lambda tag: np.frombuffer(data, np.uint32, secs[tag][2] // 4, secs[tag][1])
    assert f(b"BPOS").tolist() == [1.0, 2.0, 3.0]
    assert f(b"BQUT").tolist() == [0.0, 0.0, 0.0, 1.0]
    mass = np.float32(1000.0) * (np.float32(0.5) * 2 * np.float32(0.25) * 2 * np.float32(1.0) * 2)
    assert f(b"BIMS")[0] == np.float32(1.0) / mass
    assert f(b"GDAT").tolist() == [0.0, 8.0, 0, 0, 0, 0, 0, 0]
    assert f(b"GFRC")[0] == np.float32(0.4) and f(b"GRST")[0] == np.float32(0.2)
    assert u(b"INDX").max() == 7 and u(b"INDX").size == 36
    corners = f(b"HULL").reshape(8, 3)
    assert np.array_equal(np.abs(corners), np.tile([0.5, 0.25, 1.0], (8, 1)).astype(np.float32))
    assert np.array_equal(f(b"VERT").reshape(8, 3), corners + np.float32([1, 2, 3]))
    assert secs[b"SLUT"][2] == 0 and u(b"SLOF").tolist() == [0]
    assert f(b"MDEN").tolist() == [1000.0]


@pytest.mark.parametrize("workload,count", [(1, 50), (4, 64)])
def test_round_trip_bit_identical(workload, count, tmp_path):
    sc = make_scene(workload, count, seed=3)
    data = sc.to_bytes()
    sc2 = make_scene(workload, count, seed=99)  # different contents, replaced by the load
    sc2.from_bytes(data)
    assert sc2.counts() == sc.counts()
    for name in ARRAYS:
        assert np.array_equal(sc2.array(name).view(np.uint32), sc.array(name).view(np.uint32)), name
    assert sc2.to_bytes() == data
    path = tmp_path / "scene.svs"
    sc.save(path)
    assert path.read_bytes() == data
    sc3 = empty_scene(max_bodies=count + 16, max_hull_verts=(count + 16) * 300, max_vertices=(count + 16) * 300,
                      max_triangles=(count + 16) * 600, max_support_entries=(count + 16) * 193 + 1)
    sc3.load(path)
    assert sc3.to_bytes() == data


def test_bodies_added_after_load_match():
    a = empty_scene()
    m = a.add_material(800.0)
    a.add_sphere((0, 3, 0), 0.4, segments=8, material=m)
    b = empty_scene()
    b.from_bytes(a.to_bytes())
    for s in (a, b):
        s.add_box((2, 1, 0), (0.3, 0.3, 0.3), material=m)
        s.add_sphere((0, 5, 0), 0.5, segments=16, material=m)
    assert a.to_bytes() == b.to_bytes()


def damaged(data: bytes, kind: str) -> bytes:
    b = bytearray(data)
    _, counts, secs = parse(data)
    if kind == "magic":
        b[0] ^= 1
    elif kind == "version":
        struct.pack_into("<I", b, 8, 2)
    elif kind == "payload":
        b[secs[b"BPOS"][1]] ^= 0x40
    elif kind == "truncated":
        return bytes(b[: secs[b"HULL"][1] + 8])
    elif kind == "count":
        struct.pack_into("<I", b, 16 + 8, counts[2] + 1)  # one more body than the arrays hold
    elif kind == "missing":
        struct.pack_into("<4s", b, 48 + 32 * 3, b"XXXX")  # BPOS renamed: unknown, skipped, then missing
    elif kind in ("index", "hull", "lookup"):
        tag, val = {"index": (b"INDX", counts[0]), "hull": (b"GDAT", None), "lookup": (b"SLUT", 10 ** 6)}[kind]
        esz, off, ln, _ = secs[tag]
        if kind == "hull":
            struct.pack_into("<f", b, off + 4, float(counts[4] + 1))  # geom 0 hull count past the hull array
        else:
            struct.pack_into("<I", b, off + ln - 4, val)
        i = TAGS.index(tag)
        struct.pack_into("<Q", b, 48 + 32 * i + 24, fnv1a(bytes(b[off: off + ln])))
    return bytes(b)


@pytest.mark.parametrize("kind", ["magic", "version", "payload", "truncated", "count", "missing", "index", "hull",
                                  "lookup"])
def test_damaged_files_rejected_scene_unchanged(kind):
    from seevr_amd import SeevrError as SvError
    sc = make_scene(1, 20, seed=4)
    data = sc.to_bytes()
    other = make_scene(1, 10, seed=5)
    before = other.to_bytes()
    with pytest.raises(SvError):
        other.from_bytes(damaged(data, kind))
    assert other.to_bytes() == before


def test_over_capacity_rejected():
    from seevr_amd import SeevrError as SvError
    sc = make_scene(1, 40, seed=4)
    small = empty_scene(max_bodies=8)
    with pytest.raises(SvError, match="capacit"):
        small.from_bytes(sc.to_bytes())
