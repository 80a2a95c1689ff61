"""The differentiable workload (256 scenes x 64 bodies over a ground, 60
frames of 10 substeps): per-scene seeded initial velocities, materials,
targets and finite-difference directions, and its CPU side (FP32 oracle
rollout recording each substep's contacts, FP64 adjoint and FP64 central
differences of the same fixed-contact map).  Used by the selection script
under tests/golden/ and by the GPU test of the differentiable form.

Scene k is scene k of build_batch(0, 256, 64, 64, 16, seed 4, top 6 m); its
body 0 is the static ground (material 0), bodies 1..64 are dynamic
(material 1).  Initial v, w ~ N(0, 1) for dynamic bodies; per-material
friction U[0.2, 0.8] and restitution U[0, 0.5]; the loss is the mean
squared distance of the final positions to targets = final positions of
the scene's own rollout + U[-0.5, 0.5].
"""
from __future__ import annotations

import numpy as np

from adjoint_harness import STATE, Problem, record_fp32
from harness import World

SCENES, BODIES, FRAMES, SUB, SEG, SEED, TOP = 256, 64, 60, 10, 16, 4, 6.0
DT = 1.0 / 60.0
SUB_DT = float(np.float32(np.float32(DT) / np.float32(SUB)))
EPS = 1e-4  # central-difference step


def scene_caps():
    from seevr_amd import default_caps
    m = BODIES + 1
    return default_caps(max_bodies=m, max_hull_verts=m * 290, max_vertices=m * 290, max_triangles=m * 520,
                        max_support_entries=m * 193 + 1)


def scene_inputs(k: int, inv_mass: np.ndarray) -> dict:
    """Seeded inputs of scene k: v and w of dynamic bodies, per-geom friction
    and restitution from the two materials, target offsets, FD directions."""
    rng = np.random.default_rng([SEED, k])
    n = inv_mass.size
    dyn = inv_mass > 0
    vel = np.zeros((n, 3), np.float32)
    omg = np.zeros((n, 3), np.float32)
    vel[dyn] = rng.normal(size=(dyn.sum(), 3))
    omg[dyn] = rng.normal(size=(dyn.sum(), 3))
    mu_m = rng.uniform(0.2, 0.8, 2).astype(np.float32)
    e_m = rng.uniform(0.0, 0.5, 2).astype(np.float32)
    mat = np.where(dyn, 1, 0)
    offs = rng.uniform(-0.5, 0.5, 3 * n)
    drng = np.random.default_rng([SEED + 1, k])
    dirs = []
    for _ in range(2):
        dv = np.zeros(3 * n)
        dw = np.zeros(3 * n)
        dv[np.repeat(dyn, 3)] = drng.normal(size=3 * dyn.sum())
        dw[np.repeat(dyn, 3)] = drng.normal(size=3 * dyn.sum())
        dirs.append({"vel": dv, "omega": dw, "fric_m": drng.normal(size=2), "rest_m": drng.normal(size=2)})
    return {"vel": vel, "omega": omg, "fric": mu_m[mat], "rest": e_m[mat], "mat": mat, "offsets": offs,
            "dirs": dirs}


def cpu_scene(k: int):
    """FP32 oracle world of scene k with its seeded inputs applied."""
    from seevr_amd import Scene
    sc = Scene(scene_caps())
    start, gs = sc.build_batch(k, 1, BODIES, BODIES, SEG, SEED, top=TOP)
    w = World.from_scene(sc, max_pairs=2048, max_contacts=512, pgs_iterations=30)
    n = w.n
    inp = scene_inputs(k, w.a["BODY_INV_MASS"][:n].copy())
    w.a["BODY_VEL"][: 3 * n] = inp["vel"].ravel()
    w.a["BODY_OMEGA"][: 3 * n] = inp["omega"].ravel()
    w.a["GEOM_FRICTION"][:n] = inp["fric"]
    w.a["GEOM_RESTITUTION"][:n] = inp["rest"]
    g = (0.0, float(np.float32(-9.81) * np.float32(gs[0])), 0.0)
    sc.close()
    return w, inp, g


class FixedContactMap:
    """FP64 loss, adjoint and central differences of scene k's fixed-contact map."""

    def __init__(self, k: int):
        self.w, self.inp, g = cpu_scene(k)
        self.x0, contacts, self.statics, self.starts32 = record_fp32(self.w, FRAMES * SUB, SUB_DT, g)
        self.n = self.w.n
        self.final32 = self.w.a["BODY_POS"][: 3 * self.n].copy()  # the FP32 rollout's final positions
        self.target = self.final32.astype(np.float64) + self.inp["offsets"]
        self.prob = Problem(self.statics, contacts, SUB_DT, g, iters=30)
        self.contacts = contacts

    def loss(self, x, fr, re) -> float:
        xf, *_ = self.prob.forward(x, fr, re)
        return float(np.sum((xf["pos"] - self.target) ** 2) / (3 * self.n))

    def grad(self):
        xf, *_ = self.prob.forward(self.x0, self.statics["fric"], self.statics["rest"])
        gout = {k: np.zeros_like(self.x0[k]) for k in STATE}
        gout["pos"] = 2.0 * (xf["pos"] - self.target) / (3 * self.n)
        return self.prob.grad(self.x0, self.statics["fric"], self.statics["rest"], gout)

    def grad_at_fp32(self, forward32: bool = False):
        """The FP64 adjoint of the same map linearised along the FP32 rollout
        (each substep's Jacobian at the FP32 state the device also had, the
        loss seeded with the FP32 final positions).  forward32: the forward
        values inside each substep, and every branch they decide, are the
        FP32 sweep's own, so this is the exact derivative of the device's
        FP32 computation, evaluated in double."""
        gout = {k: np.zeros_like(self.x0[k]) for k in STATE}
        gout["pos"] = 2.0 * (self.final32.astype(np.float64) - self.target) / (3 * self.n)
        return self.prob.grad_along(self.starts32, self.statics["fric"], self.statics["rest"], gout, forward32)

    def blocks(self, adj, af, ae) -> dict:
        """Gradient blocks the workload asks for: initial velocities, angular
        velocities, per-material friction and restitution."""
        mat = self.inp["mat"]
        return {"vel": adj["vel"], "omega": adj["omega"], "friction": np.bincount(mat, af, 2),
                "restitution": np.bincount(mat, ae, 2)}

    def trajectory_sensitivity(self) -> float:
        """Largest relative difference, over the blocks, between the FP64
        adjoint along the FP32 rollout and along the FP64 rollout: small when
        the derivative does not depend on the rounding of the trajectory."""
        a = self.blocks(*self.grad_at_fp32())  # FP64 branches on both sides: only the trajectory differs
        b = self.blocks(*self.grad())
        return max(float(np.linalg.norm(a[k] - b[k]) / max(np.linalg.norm(b[k]), 1e-300)) for k in a)

    def precision_sensitivity(self) -> float:
        """Largest relative difference, over the blocks, between the exact
        derivative of the FP32 computation (FP32 rollout and branches) and
        the FP64 adjoint of the FP64 rollout: small when the gradient does
        not depend on the working precision (neither through the trajectory
        nor through a stick/slip or clamp decision that FP32 and FP64 take
        differently)."""
        a = self.blocks(*self.grad_at_fp32(forward32=True))
        b = self.blocks(*self.grad())
        return max(float(np.linalg.norm(a[k] - b[k]) / max(np.linalg.norm(b[k]), 1e-300)) for k in a)

    def fd(self, d: dict, eps: float = EPS) -> tuple[float, float]:
        """Central differences along direction d: (state part, material part)."""
        x0, fr, re, mat = self.x0, self.statics["fric"], self.statics["rest"], self.inp["mat"]
        xp = {k: x0[k] + eps * d.get(k, 0.0) for k in STATE}
        xm = {k: x0[k] - eps * d.get(k, 0.0) for k in STATE}
        fs = (self.loss(xp, fr, re) - self.loss(xm, fr, re)) / (2 * eps)
        dfr, dre = d["fric_m"][mat], d["rest_m"][mat]
        fm = (self.loss(x0, fr + eps * dfr, re + eps * dre) - self.loss(x0, fr - eps * dfr, re - eps * dre)) / (2 * eps)
        return fs, fm

    def smoothness(self) -> float:
        """Largest relative change of the central differences from eps to
        eps / 4 over the scene's directions: small when no kink of the map
        (an impulse clamp, a stick/slip switch) lies within eps."""
        worst = 0.0
        for d in self.inp["dirs"]:
            a, b = self.fd(d, EPS), self.fd(d, EPS / 4)
            for u, v in zip(a, b):
                worst = max(worst, abs(u - v) / max(abs(u), 1e-30))
        return worst
