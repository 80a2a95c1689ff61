"""The FP64 adjoint of the differentiable workload's fixed-contact map
(60 frames x 10 substeps, 65-body scenes) against FP64 central differences
(eps 1e-4) on scenes of the committed selection, and the selection itself
recomputed for a few scenes (tests/golden/make_c4_selection.py)."""
import json
from pathlib import Path

import numpy as np
import pytest

from c4_workload import STATE, FixedContactMap

SEL = json.loads((Path(__file__).resolve().parent / "golden" / "c4_smooth_scenes.json").read_text())


@pytest.mark.parametrize("k", SEL["scenes"][:2])
def test_fp64_adjoint_matches_central_differences(k):
    m = FixedContactMap(k)
    assert m.smoothness() < SEL["threshold"]
    adj, af, ae = m.grad()
    mat = m.inp["mat"]
    for d in m.inp["dirs"]:
        fs, fm = m.fd(d)
        an_s = sum(float(np.dot(adj[key], d[key])) for key in ("vel", "omega"))
        an_m = float(np.dot(np.bincount(mat, af, 2), d["fric_m"]) + np.dot(np.bincount(mat, ae, 2), d["rest_m"]))
        assert abs(an_s - fs) <= 1e-3 * abs(fs), (k, an_s, fs)
        assert abs(an_m - fm) <= 1e-3 * abs(fm), (k, an_m, fm)
    assert sum(c["nc"] for c in m.contacts) > 100  # the map holds real contacts


def test_selection_recomputed():
    th = SEL["threshold"]
    k = SEL["scenes"][0]
    m = FixedContactMap(k)
    assert m.precision_sensitivity() < th
    kinked = [t["scene"] for t in SEL["tried"] if t["smoothness"] >= th][0]
    assert FixedContactMap(kinked).smoothness() >= th
    sensitive = [t["scene"] for t in SEL["tried"] if t["smoothness"] < th <= t["precision_sensitivity"]][0]
    assert FixedContactMap(sensitive).precision_sensitivity() >= th


def test_forward32_mode_reproduces_the_fp32_sweep_bitwise():
    """adj_set_forward32: from each FP32 substep start of a 600-substep
    rollout, the adjoint oracle's forward with the recorded contacts gives
    the FP32 oracle's next substep start bit for bit (gravity, sweep,
    integration), so its branch decisions are the FP32 sweep's."""
    import ctypes as C
    import adjoint_harness as AH
    m = FixedContactMap(SEL["scenes"][0])
    lib = AH.lib()
    lib.adj_set_forward32(1)
    try:
        for s, (st0, c) in enumerate(zip(m.starts32[:-1], m.contacts[:-1])):
            st = {k: v.copy() for k, v in st0.items()}
            sc = m.prob._scene(st, m.statics["fric"], m.statics["rest"], c)
            lib.adj_forward(C.byref(sc), None)
            for k in STATE:
                assert np.array_equal(st[k], m.starts32[s + 1][k]), (s, k)
    finally:
        lib.adj_set_forward32(0)


def test_forward32_same_point_adjoint_close_to_fp64_branches():
    """The FP32-branch adjoint and the FP64-branch adjoint along the same
    FP32 trajectory agree except for the few items whose branch differs."""
    m = FixedContactMap(SEL["scenes"][0])
    a = m.blocks(*m.grad_at_fp32(forward32=True))
    b = m.blocks(*m.grad_at_fp32(forward32=False))
    for key in ("vel", "omega", "restitution"):
        assert np.linalg.norm(a[key] - b[key]) <= 1e-4 * np.linalg.norm(b[key]), key
