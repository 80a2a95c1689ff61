"""Writes c4_smooth_scenes.json: the first 16 scenes, in index order, of the
differentiable workload (tests/c4_workload.py) in a contact-stable regime:
  - the fixed-contact map is smooth along the scene's seeded directions at
    the central-difference step (the differences at eps and eps / 4 agree to
    THRESHOLD), so FP64 central differences there check a gradient, and
  - its gradient does not depend on the working precision: the exact
    derivative of the FP32 computation (FP32 rollout, FP32 branches, FP64
    reverse; adj_set_forward32) and the FP64 adjoint of the FP64 rollout
    agree to THRESHOLD in every block, so an FP32 gradient can be checked
    against FP64 references to that accuracy.
Only the CPU oracle runs (FP32 rollout, FP64 map and adjoints); no device
gradient is looked at.  Usage: python tests/golden/make_c4_selection.py
"""
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parents[1]))

from c4_workload import EPS, SCENES, FixedContactMap  # noqa: E402

THRESHOLD = 5e-4
WANT = 16


def main() -> None:
    picked, tried = [], []
    for k in range(SCENES):
        m = FixedContactMap(k)
        s = m.smoothness()
        t = m.precision_sensitivity()
        tried.append({"scene": k, "smoothness": float(f"{s:.3e}"), "precision_sensitivity": float(f"{t:.3e}")})
        print(f"scene {k}: smoothness {s:.2e}, precision sensitivity {t:.2e}", flush=True)
        if s < THRESHOLD and t < THRESHOLD:
            picked.append(k)
            if len(picked) == WANT:
                break
    out = {"criterion": f"|FD(eps) - FD(eps/4)| / |FD(eps)| < {THRESHOLD} along both seeded directions "
                        f"(state and material parts), eps = {EPS}; the FP32 computation's exact derivative and the FP64 "
                        f"rollout's FP64 adjoint within {THRESHOLD} in every block", "threshold": THRESHOLD, "eps": EPS,
           "scenes": picked, "tried": tried}
    (HERE / "c4_smooth_scenes.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"picked {picked} of {len(tried)} tried")


if __name__ == "__main__":
    main()
