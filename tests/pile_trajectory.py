"""Shared pieces of the 300-frame pile trajectory checks (fixture
tests/golden/pile_1000_300f.json, written by make_pile_trajectory.py)."""
import hashlib
import json
from pathlib import Path

import numpy as np

FIXTURE = Path(__file__).resolve().parent / "golden" / "pile_1000_300f.json"


def load():
    return json.loads(FIXTURE.read_text())


def digest(arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, np.float32).tobytes())
    return h.hexdigest()
