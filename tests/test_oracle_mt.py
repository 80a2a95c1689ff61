"""The OpenMP build of the oracle (the benchmark's multi-core CPU baseline)
gives the serial build's results bit for bit: the pile fixture's first
checkpoint (25 frames of the 1,000-body column) reproduces through it, and
from a settled state with truncated pair and contact lists both builds
leave identical buffers."""
import numpy as np

from harness import World, make_scene, oracle
from pile_trajectory import digest, load


def test_mt_build_reproduces_fixture_checkpoint():
    fx = load()
    sc = make_scene(1, 1000, seed=1)
    w = World.from_scene(sc, max_pairs=fx["max_pairs"], max_contacts=fx["max_contacts"],
                         pgs_iterations=fx["pgs_iterations"])
    cp = fx["checkpoints"][0]
    for _ in range(cp["frame"]):
        w.oracle_step(mt=True)
    n, a = w.n, w.a
    assert digest([a["BODY_POS"][: 3 * n], a["BODY_QUAT"][: 4 * n], a["BODY_VEL"][: 3 * n],
                   a["BODY_OMEGA"][: 3 * n]]) == cp["state_sha256"]
    assert [int(a["ATOMIC_COUNTERS"][0]), int(a["ATOMIC_COUNTERS"][1])] == cp["counters"]


def test_mt_build_equals_serial_with_truncated_lists():
    sc = make_scene(1, 300, seed=3)
    w = World.from_scene(sc, max_pairs=4096, max_contacts=2048, pgs_iterations=30)
    for _ in range(100):
        w.oracle_step(mt=True)
    assert int(w.a["ATOMIC_COUNTERS"][1]) > 60
    ser, par = w.copy(), w.copy()
    for x in (ser, par):
        x.max_pairs, x.max_contacts = 97, 41
    ser.oracle_step()
    par.oracle_step(mt=True)
    assert int(par.a["ATOMIC_COUNTERS"][0]) > 97 and int(par.a["ATOMIC_COUNTERS"][1]) > 41
    for k, v in ser.a.items():
        assert np.array_equal(v.view(np.uint8), par.a[k].view(np.uint8)), k


def test_mt_build_reports_its_threads():
    assert oracle(mt=True).threads() >= 1
    assert oracle().threads() == 1
