"""The pile trajectory fixture against the oracle that wrote it: the first
checkpoint (25 frames of the 1,000-body column) reproduces, and the
recorded pile statistics are those of a falling, spreading pile (mean
height decreasing, bodies coming to rest, contacts growing)."""
import numpy as np

from harness import World, make_scene
from pile_trajectory import digest, load


def test_first_checkpoint_reproduces():
    fx = load()
    sc = make_scene(1, 1000, seed=1)
    w = World.from_scene(sc, max_pairs=fx["max_pairs"], max_contacts=fx["max_contacts"],
                         pgs_iterations=fx["pgs_iterations"])
    cp = fx["checkpoints"][0]
    for _ in range(cp["frame"]):
        w.oracle_step()
    n = w.n
    a = w.a
    assert digest([a["BODY_POS"][: 3 * n], a["BODY_QUAT"][: 4 * n], a["BODY_VEL"][: 3 * n],
                   a["BODY_OMEGA"][: 3 * n]]) == cp["state_sha256"]
    assert [int(a["ATOMIC_COUNTERS"][0]), int(a["ATOMIC_COUNTERS"][1])] == cp["counters"]


def test_recorded_statistics_are_a_settling_pile():
    cps = load()["checkpoints"]
    assert [c["frame"] for c in cps] == list(range(25, 301, 25))
    h = np.array([c["mean_height"] for c in cps])
    assert np.all(np.diff(h) < 0)
    rest = np.array([c["resting"] for c in cps])
    assert rest[-1] > 500 and np.all(np.diff(rest) >= 0)
    contacts = np.array([c["counters"][1] for c in cps])
    assert contacts[-1] > 1000 and contacts[-1] > contacts[0]
