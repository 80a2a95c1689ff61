"""Resume from a scene file: frames simulated on the device, the state read
back into the scene (sv_scene_read_state), saved and loaded into a fresh
scene and context, continue bit-identically to the context that kept
running (the step state is bodies and render vertices; pairs, contacts and
impulses are rebuilt every substep)."""
import numpy as np
import pytest

from harness import make_scene

pytestmark = pytest.mark.gpu

STATE = ("BODY_POS", "BODY_QUAT", "BODY_VEL", "BODY_OMEGA", "VERTICES")


def test_resume_from_saved_scene(tmp_path):
    from seevr_amd import Context, Scene, caps_for, default_params
    sc = make_scene(1, 300, seed=21)
    cn = sc.counts()
    caps = caps_for(sc, 40000, 20000)
    args = (cn["num_bodies"], cn["num_vertices"], cn["num_hull_verts"])
    a = Context(caps, default_params(pgs_iterations=30))
    sc.upload(a)
    for _ in range(40):
        a.step(*args)
    sc.read_state(a)
    path = tmp_path / "mid.svs"
    sc.save(path)
    sc2 = Scene(sc.caps)
    sc2.load(path)
    b = Context(caps, default_params(pgs_iterations=30))
    sc2.upload(b)
    for _ in range(20):
        a.step(*args)
        b.step(*args)
    assert a.read_counters() == b.read_counters()
    for name in STATE:
        n = {"BODY_QUAT": 4 * cn["num_bodies"], "VERTICES": 3 * cn["num_vertices"]}.get(name, 3 * cn["num_bodies"])
        x, y = a.download(name, n), b.download(name, n)
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), name
    a.close()
    b.close()
