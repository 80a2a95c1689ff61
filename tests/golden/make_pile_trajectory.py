"""Writes pile_1000_300f.json: the oracle's 300-frame trajectory of the
seeded 1,000-body column (scene builder workload 1, seed 1; 10 substeps,
30 iterations, dt 1/60, g = -9.81 y) as SHA-256 digests of the body state
(pos, quat, vel, omega as float32 bytes) every 25 frames, with the raw pair
and contact counters and pile statistics.  Run from the repository root:
    python tests/golden/make_pile_trajectory.py
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parents[1]))

from harness import World, make_scene  # noqa: E402

FRAMES, EVERY = 300, 25


def digest(arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, np.float32).tobytes())
    return h.hexdigest()


def stats(pos, vel, inv_mass):
    dyn = inv_mass > 0
    y = pos.reshape(-1, 3)[dyn, 1].astype(np.float64)
    v = vel.reshape(-1, 3)[dyn].astype(np.float64)
    return {"mean_height": float(y.mean()), "max_height": float(y.max()),
            "sum_v2": float((v * v).sum()), "resting": int((np.linalg.norm(v, axis=1) < 0.05).sum())}


def main():
    sc = make_scene(1, 1000, seed=1)
    w = World.from_scene(sc, max_pairs=200000, max_contacts=100000, pgs_iterations=30)
    n = w.n
    out = {"workload": "1,000 bodies (+1 static ground), scene builder workload 1, seed 1", "bodies": n,
           "substeps": 10, "pgs_iterations": 30, "dt": 1 / 60, "max_pairs": 200000, "max_contacts": 100000,
           "checkpoints": []}
    for f in range(1, FRAMES + 1):
        w.oracle_step()
        if f % EVERY == 0:
            a = w.a
            st = [a["BODY_POS"][: 3 * n], a["BODY_QUAT"][: 4 * n], a["BODY_VEL"][: 3 * n], a["BODY_OMEGA"][: 3 * n]]
            out["checkpoints"].append({"frame": f, "state_sha256": digest(st),
                                       "counters": [int(a["ATOMIC_COUNTERS"][0]), int(a["ATOMIC_COUNTERS"][1])],
                                       **stats(st[0], st[2], a["BODY_INV_MASS"][:n])})
            print(out["checkpoints"][-1], flush=True)
    (HERE / "pile_1000_300f.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
