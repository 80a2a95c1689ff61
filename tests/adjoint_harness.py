"""FP64 CPU adjoint (oracle/sv_adjoint.c) and finite differences of the
same substep map, for the differentiable-step tests.

The map: substeps of gravity, the Gauss-Seidel contact sweep over a contact
set held fixed per substep (recorded from the oracle's own narrowphase on
the base trajectory), and integration.  Inputs: the state (pos, quat, vel,
omega) and the per-geom friction and restitution.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from harness import ORACLE_DIR, build_oracle

LIB = ORACLE_DIR / "_build" / "libsvadjoint.so"
_lib = None

dp = C.POINTER(C.c_double)
up = C.POINTER(C.c_uint32)


class AdjScene(C.Structure):
    _fields_ = [("n", C.c_uint32), ("pos", dp), ("quat", dp), ("vel", dp), ("omega", dp),
                ("inv_mass", dp), ("inv_i", dp), ("fric", dp), ("rest", dp), ("nc", C.c_uint32),
                ("ca", up), ("cb", up), ("cp", dp), ("cn", dp), ("pen", dp), ("iters", C.c_uint32),
                ("dt", C.c_double), ("g", C.c_double * 3), ("baumgarte", C.c_double), ("slop", C.c_double),
                ("ld", C.c_double), ("ad", C.c_double)]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build_oracle()
        _lib = C.CDLL(str(LIB))
        _lib.adj_tape_doubles.restype = C.c_size_t
        _lib.adj_tape_doubles.argtypes = [C.POINTER(AdjScene)]
        _lib.adj_forward.restype = None
        _lib.adj_forward.argtypes = [C.POINTER(AdjScene), dp]
        _lib.adj_backward.restype = None
        _lib.adj_backward.argtypes = [C.POINTER(AdjScene), dp, dp, dp, dp, dp, dp, dp]
        _lib.adj_set_forward32.restype = None
        _lib.adj_set_forward32.argtypes = [C.c_int]
        _lib.adj_set_implicit.restype = None
        _lib.adj_set_implicit.argtypes = [C.c_int]
    return _lib


def _p(a: np.ndarray, t=dp):
    return a.ctypes.data_as(t)


STATE = ("pos", "quat", "vel", "omega")


class Problem:
    """Substeps of the fixed-contact map from one recorded base trajectory."""

    def __init__(self, statics: dict, contacts: list[dict], dt: float, g, iters: int = 30,
                 baumgarte=0.2, slop=0.005, ld=0.999, ad=0.998):
        self.s = {k: np.ascontiguousarray(v, np.float64) for k, v in statics.items()}
        self.n = self.s["inv_mass"].size
        self.contacts = contacts
        self.dt, self.g, self.iters = dt, tuple(g), iters
        self.consts = (baumgarte, slop, ld, ad)

    def _scene(self, st: dict, fric, rest, c: dict) -> AdjScene:
        s = AdjScene()
        s.n = self.n
        s.pos, s.quat, s.vel, s.omega = (_p(st[k]) for k in STATE)
        s.inv_mass, s.inv_i = _p(self.s["inv_mass"]), _p(self.s["inv_i"])
        s.fric, s.rest = _p(fric), _p(rest)
        s.nc = c["nc"]
        s.ca, s.cb = _p(c["a"], up), _p(c["b"], up)
        s.cp, s.cn, s.pen = _p(c["cp"]), _p(c["n"]), _p(c["pen"])
        s.iters, s.dt = self.iters, self.dt
        s.g = (C.c_double * 3)(*self.g)
        s.baumgarte, s.slop, s.ld, s.ad = self.consts
        return s

    def forward(self, x0: dict, fric, rest, keep=False):
        st = {k: np.array(x0[k], np.float64) for k in STATE}
        fric = np.ascontiguousarray(fric, np.float64)
        rest = np.ascontiguousarray(rest, np.float64)
        saved = []
        for c in self.contacts:
            start = {k: v.copy() for k, v in st.items()}
            sc = self._scene(st, fric, rest, c)
            tape = None
            if keep:
                tape = np.zeros(lib().adj_tape_doubles(C.byref(sc)), np.float64)
            lib().adj_forward(C.byref(sc), _p(tape) if keep else None)
            if keep:
                saved.append((start, tape))
        return st, saved, fric, rest

    def grad(self, x0: dict, fric, rest, gout: dict, implicit: bool = False):
        """Adjoint of the final state gout -> (adjoint of x0, d/dfric, d/drest);
        implicit: the sweep differentiated at its converged solution
        (adj_set_implicit) instead of through every pass."""
        _, saved, fric, rest = self.forward(x0, fric, rest, keep=True)
        adj = {k: np.array(gout[k], np.float64) for k in STATE}
        af = np.zeros(self.n)
        ae = np.zeros(self.n)
        lib().adj_set_implicit(1 if implicit else 0)
        try:
            for (start, tape), c in zip(reversed(saved), reversed(self.contacts)):
                sc = self._scene(start, fric, rest, c)
                lib().adj_backward(C.byref(sc), _p(tape), _p(adj["pos"]), _p(adj["quat"]), _p(adj["vel"]),
                                   _p(adj["omega"]), _p(af), _p(ae))
        finally:
            lib().adj_set_implicit(0)
        return adj, af, ae


    def grad_along(self, starts: list[dict], fric, rest, gout: dict, forward32: bool = True,
                   implicit: bool = False):
        """The adjoint linearised along a given trajectory: starts[s] is the
        state at the start of substep s (for instance an FP32 rollout's, the
        device's own), each substep's Jacobian is taken there, and gout is the
        adjoint of the state after the last substep.  forward32 evaluates the
        forward values and their branches as the FP32 sweep does
        (adj_set_forward32), the reverse in FP64; implicit as in grad."""
        fric = np.ascontiguousarray(fric, np.float64)
        rest = np.ascontiguousarray(rest, np.float64)
        L = lib()
        L.adj_set_forward32(1 if forward32 else 0)
        L.adj_set_implicit(1 if implicit else 0)
        try:
            tapes = []
            for st0, c in zip(starts, self.contacts):
                st = {k: np.array(st0[k], np.float64) for k in STATE}
                sc = self._scene(st, fric, rest, c)
                tape = np.zeros(L.adj_tape_doubles(C.byref(sc)), np.float64)
                L.adj_forward(C.byref(sc), _p(tape))
                tapes.append(tape)
            adj = {k: np.array(gout[k], np.float64) for k in STATE}
            af = np.zeros(self.n)
            ae = np.zeros(self.n)
            for st0, tape, c in zip(reversed(starts), reversed(tapes), reversed(self.contacts)):
                st = {k: np.array(st0[k], np.float64) for k in STATE}
                sc = self._scene(st, fric, rest, c)
                L.adj_backward(C.byref(sc), _p(tape), _p(adj["pos"]), _p(adj["quat"]), _p(adj["vel"]),
                               _p(adj["omega"]), _p(af), _p(ae))
            return adj, af, ae
        finally:
            L.adj_set_forward32(0)
            L.adj_set_implicit(0)


def record_fp32(world, substeps: int, dt_sub: float, g) -> tuple[dict, list[dict], dict, list[dict]]:
    """From an FP32 oracle World (the device's trajectory, bit for bit), run
    `substeps` oracle substeps recording each substep's contacts and start
    state.  Returns (initial state, contacts, statics, starts) in FP64."""
    a = world.a
    n = world.n
    starts, contacts = [], []
    statics = {"inv_mass": a["BODY_INV_MASS"][:n].astype(np.float64),
               "inv_i": a["BODY_INV_INERTIA"][: 3 * n].astype(np.float64),
               "fric": a["GEOM_FRICTION"][:n].astype(np.float64), "rest": a["GEOM_RESTITUTION"][:n].astype(np.float64)}
    for _ in range(substeps):
        starts.append({"pos": a["BODY_POS"][: 3 * n].astype(np.float64), "quat": a["BODY_QUAT"][: 4 * n].astype(np.float64),
                       "vel": a["BODY_VEL"][: 3 * n].astype(np.float64),
                       "omega": a["BODY_OMEGA"][: 3 * n].astype(np.float64)})
        world.oracle_stage(8)
        world.oracle_stage(0, dt=dt_sub, g=g)
        world.oracle_stage(1)
        world.oracle_stage(2)
        world.oracle_stage(3)
        nc = world.ncontacts()
        contacts.append({"nc": nc, "a": a["CONTACT_GEOM_A"][:nc].astype(np.uint32).copy(),
                         "b": a["CONTACT_GEOM_B"][:nc].astype(np.uint32).copy(),
                         "cp": a["CONTACT_POS"][: 3 * nc].astype(np.float64),
                         "n": a["CONTACT_NORMAL"][: 3 * nc].astype(np.float64),
                         "pen": a["CONTACT_PENETRATION"][:nc].astype(np.float64)})
        world.oracle_stage(4)
        world.oracle_stage(5, dt=dt_sub)
        world.oracle_stage(6, dt=dt_sub)
    return dict(starts[0]), contacts, statics, starts


def record_base(world, substeps: int, dt_sub: float, g) -> tuple[dict, list[dict], dict]:
    """From a double-precision oracle World's current state, run `substeps`
    oracle substeps (physics.zig order) and record each substep's contacts.
    Returns (initial state, contacts per substep, statics)."""
    a = world.a
    n = world.n
    x0 = {"pos": a["BODY_POS"][: 3 * n].copy(), "quat": a["BODY_QUAT"][: 4 * n].copy(),
          "vel": a["BODY_VEL"][: 3 * n].copy(), "omega": a["BODY_OMEGA"][: 3 * n].copy()}
    gb = a["GEOM_BODY_IDX"][:n]
    assert np.array_equal(gb, np.arange(n)), "geoms must be 1:1 with bodies"
    statics = {"inv_mass": a["BODY_INV_MASS"][:n].copy(), "inv_i": a["BODY_INV_INERTIA"][: 3 * n].copy(),
               "fric": a["GEOM_FRICTION"][:n].copy(), "rest": a["GEOM_RESTITUTION"][:n].copy()}
    contacts = []
    for _ in range(substeps):
        world.oracle_stage(8)
        world.oracle_stage(0, dt=dt_sub, g=g)
        world.oracle_stage(1)
        world.oracle_stage(2)
        world.oracle_stage(3)
        nc = world.ncontacts()
        contacts.append({"nc": nc, "a": a["CONTACT_GEOM_A"][:nc].astype(np.uint32).copy(),
                         "b": a["CONTACT_GEOM_B"][:nc].astype(np.uint32).copy(),
                         "cp": a["CONTACT_POS"][: 3 * nc].copy(), "n": a["CONTACT_NORMAL"][: 3 * nc].copy(),
                         "pen": a["CONTACT_PENETRATION"][:nc].copy()})
        world.oracle_stage(4)
        world.oracle_stage(5, dt=dt_sub)
        world.oracle_stage(6, dt=dt_sub)
    return x0, contacts, statics
