"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and the CPU baseline.  The
product library never imports this module.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_float, c_int, c_uint32, c_void_p
from pathlib import Path

import numpy as np

from diomedes_amd import records as R

import os

# ORACLE_LIB points the wrapper at another build of the same sources (used to
# check that the tests catch deliberately planted mistakes)
LIB_PATH = Path(os.environ.get("ORACLE_LIB", Path(__file__).resolve().parent / "liboracle.so"))
_lib = None

_D = POINTER(c_double)
_U = POINTER(c_uint32)
_F = POINTER(c_float)


_U64 = POINTER(ctypes.c_uint64)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run `make -C oracle`")
    lib = ctypes.CDLL(str(LIB_PATH))
    sig = {
        "oracle_threads": (c_int, [c_int]),
        "oracle_world_create": (c_void_p, [c_int, c_void_p, c_void_p, c_uint32, c_void_p, c_uint32, c_void_p, _F,
                                           c_uint32]),
        "oracle_world_destroy": (None, [c_void_p]),
        "oracle_world_clone": (c_void_p, [c_void_p]),
        "oracle_world_step": (None, [c_void_p, POINTER(R.Options), c_int, c_uint32]),
        "oracle_world_get_state": (None, [c_void_p, _D, _D, _D, _D, _D]),
        "oracle_world_contact_count": (c_uint32, [c_void_p]),
        "oracle_world_get_contacts": (None, [c_void_p, _U, _D]),
        "oracle_world_pair_count": (c_uint32, [c_void_p]),
        "oracle_world_get_pairs": (None, [c_void_p, _U]),
        "oracle_world_csr_nnz": (c_uint32, [c_void_p]),
        "oracle_world_get_csr": (None, [c_void_p, _U, _U]),
        "oracle_world_times": (None, [c_void_p, _D]),
        "oracle_world_reset_times": (None, [c_void_p]),
        "oracle_aabbs": (None, [c_int, c_void_p, c_uint32, _F, c_uint32, _D]),
        "oracle_broad_pairs": (c_uint32, [c_int, c_void_p, c_uint32, _F, c_uint32, _U, c_uint32]),
        "oracle_narrow_pair": (c_uint32, [c_int, c_void_p, c_void_p, _F, c_uint32, _D]),
        "oracle_gjk": (c_int, [c_int, c_void_p, c_void_p, _F, c_uint32]),
        "oracle_world_set_mode": (None, [c_void_p, c_int]),
        "oracle_world_get_colours": (c_uint32, [c_void_p, _U, _U]),
        "oracle_world_get_joints": (c_uint32, [c_void_p, _D, _D]),
        "oracle_colour_graph": (c_uint32, [c_uint32, ctypes.POINTER(ctypes.c_uint8), _U, c_uint32, _U, _U]),
        "oracle_solve6": (None, [c_int, _D, _D, _D, _D, _D]),
        "oracle_container_contacts": (c_uint32, [c_int, c_void_p, c_void_p, POINTER(R.Container), POINTER(R.Options),
                                                 c_uint32, _F, c_uint32, _U, _D]),
        "oracle_raw_contact": (None, [c_int, c_void_p, c_void_p, c_uint32, c_uint32, _D, c_double, _D, _D,
                                      POINTER(R.Options), _U, _D]),
        "oracle_finalize_contacts": (None, [c_int, c_void_p, c_uint32, c_void_p, c_uint32, c_uint32,
                                            POINTER(R.Options), _U, _D, _U, _U]),
        "oracle_avbd_run": (None, [c_int, c_void_p, _F, _F, c_uint32, c_void_p, c_uint32, c_void_p, c_uint32, _U, _U,
                                   c_uint32, POINTER(R.RunOptions), _D, _D, _D, _D, _D, _D, _D]),
        "oracle_avbd_init": (None, [c_int, c_void_p, _F, _F, c_uint32, POINTER(R.RunOptions), _D, _D, _D, _D]),
        "oracle_qrot": (None, [c_int, _D, _D, _D]),
        "oracle_qmul": (None, [c_int, _D, _D, _D]),
        "oracle_instance_transforms": (None, [c_int, _D, _D, _D, c_uint32, _U, c_uint32, _D, _D]),
        "oracle_support": (None, [c_int, c_int, c_void_p, _D, _F, c_uint32, _D]),
        "oracle_replay": (None, [c_int, c_void_p, c_void_p, c_uint32, c_uint32, c_void_p, _U64, _U, _U, _U64,
                                 POINTER(R.Options), _D, c_double, c_uint32, _D, _D, _D, _D]),
        "oracle_replay_frozen": (None, [c_int, c_int, c_int, c_void_p, c_void_p, c_uint32, c_uint32, c_void_p, _U64, _U,
                                        _U, _U64, POINTER(R.Options), _D, _D, c_double, c_uint32, _D, _D, _D, _D]),
        "oracle_world_keep_presolve": (None, [c_void_p, c_int]),
        "oracle_world_presolve_count": (c_uint32, [c_void_p]),
        "oracle_world_get_presolve": (None, [c_void_p, c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a, t=None):
    if a is None:
        return None
    return a.ctypes.data_as(c_void_p if t is None else POINTER(t))


def threads(n: int = 0) -> int:
    """Set (n > 0) and return the OpenMP thread count of the oracle."""
    return load().oracle_threads(n)


# contacts come back as 25 doubles per contact
CONTACT_FIELDS = {"friction": (0, 1), "normal": (1, 4), "tangent1": (4, 7), "tangent2": (7, 10), "r_a": (10, 13),
                  "r_b": (13, 16), "c0": (16, 19), "penalty": (19, 22), "lambda": (22, 25)}


def split_contacts(ab: np.ndarray, data: np.ndarray) -> dict:
    out = {"a": ab[:, 0].copy(), "b": ab[:, 1].copy()}
    for k, (s, e) in CONTACT_FIELDS.items():
        out[k] = data[:, s:e] if e - s > 1 else data[:, s]
    return out


class OracleWorld:
    """The step protocol on the CPU, in float32 (precision 32) or float64."""

    def __init__(self, scene, precision: int = 32):
        lib = load()
        self.scene = scene
        self.precision = precision
        bodies = np.ascontiguousarray(scene.bodies, R.GPU_BODY)
        state = np.ascontiguousarray(scene.state, R.BODY_STATE)
        cons = np.ascontiguousarray(scene.constraints, R.AVBD_CONSTRAINT)
        hull = np.ascontiguousarray(scene.hull, np.float32).reshape(-1, 3)
        ct = scene.container.to_c() if scene.container is not None else None
        self.n = len(bodies)
        self._h = lib.oracle_world_create(precision, _p(bodies), _p(state), len(bodies), _p(cons), len(cons),
                                          ctypes.byref(ct) if ct is not None else None, _p(hull, c_float), len(hull))

    def clone(self) -> "OracleWorld":
        """A copy in the current state: stepping it continues the same
        trajectory."""
        c = OracleWorld.__new__(OracleWorld)
        c.scene, c.precision, c.n = self.scene, self.precision, self.n
        c._h = load().oracle_world_clone(self._h)
        return c

    def close(self):
        if getattr(self, "_h", None):
            load().oracle_world_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, options=None, nsteps: int = 1, lag: int = 1) -> None:
        o = (options or self.scene.options).to_c()
        load().oracle_world_step(self._h, ctypes.byref(o), lag, nsteps)

    def set_mode(self, mode: int) -> None:
        """0: Jacobi (the reference); 1: coloured Gauss-Seidel."""
        load().oracle_world_set_mode(self._h, int(mode))

    def colours(self) -> tuple:
        """The last coloured step's colour per body (0xFFFFFFFE for static
        bodies), the number of colours and of Jones-Plassmann rounds."""
        col = np.zeros(self.n, np.uint32)
        rounds = ctypes.c_uint32()
        nc = load().oracle_world_get_colours(self._h, _p(col, c_uint32), ctypes.byref(rounds))
        return col, int(nc), rounds.value

    def joints(self) -> tuple:
        """Each constraint's multipliers after the last step: lambda and
        penalty per row, (c, 3) float64 each (zero for soft joints)."""
        c = len(self.scene.constraints)
        lam, pen = np.zeros((c, 3)), np.zeros((c, 3))
        load().oracle_world_get_joints(self._h, _p(lam, c_double), _p(pen, c_double))
        return lam, pen

    def state(self) -> dict:
        n = self.n
        out = {k: np.zeros((n, 3)) for k in ("positions", "velocities", "angular_velocities", "prev_velocities")}
        out["orientations"] = np.zeros((n, 4))
        load().oracle_world_get_state(self._h, _p(out["positions"], c_double), _p(out["orientations"], c_double),
                                      _p(out["velocities"], c_double), _p(out["angular_velocities"], c_double),
                                      _p(out["prev_velocities"], c_double))
        return out

    def contacts(self) -> dict:
        lib = load()
        m = lib.oracle_world_contact_count(self._h)
        ab = np.zeros((m, 2), np.uint32)
        data = np.zeros((m, 25))
        lib.oracle_world_get_contacts(self._h, _p(ab, c_uint32), _p(data, c_double))
        return split_contacts(ab, data)

    def pairs(self) -> np.ndarray:
        lib = load()
        p = lib.oracle_world_pair_count(self._h)
        out = np.zeros((p, 2), np.uint32)
        lib.oracle_world_get_pairs(self._h, _p(out, c_uint32))
        return out

    def csr(self) -> tuple:
        lib = load()
        nnz = lib.oracle_world_csr_nnz(self._h)
        off = np.zeros(self.n + 1, np.uint32)
        idx = np.zeros(nnz, np.uint32)
        lib.oracle_world_get_csr(self._h, _p(off, c_uint32), _p(idx, c_uint32))
        return off, idx

    def times(self) -> dict:
        t = np.zeros(9)
        load().oracle_world_times(self._h, _p(t, c_double))
        keys = ("broad", "narrow", "contacts", "avbd", "total", "init", "primal", "dual", "recover")
        return dict(zip(keys, t.tolist()))

    def reset_times(self) -> None:
        load().oracle_world_reset_times(self._h)

    def keep_presolve(self, on: bool = True) -> None:
        load().oracle_world_keep_presolve(self._h, 1 if on else 0)

    def presolve_contacts(self) -> np.ndarray:
        """The contact records (AVBD_CONTACT) the last solve received."""
        lib = load()
        m = lib.oracle_world_presolve_count(self._h)
        out = np.zeros(m, R.AVBD_CONTACT)
        lib.oracle_world_get_presolve(self._h, _p(out))
        return out


# ------------------------------------------------- frozen-contact replay
class Recording:
    """Per-step frozen contact sets (AVBD_CONTACT records as the solve received
    them) and CSRs of a forward run, packed for ``replay``."""

    def __init__(self, n: int):
        self.n = n
        self.contacts, self.offsets, self.indices = [], [], []

    def add(self, contacts: np.ndarray, offsets: np.ndarray, indices: np.ndarray) -> None:
        self.contacts.append(np.ascontiguousarray(contacts, R.AVBD_CONTACT))
        self.offsets.append(np.ascontiguousarray(offsets, np.uint32))
        self.indices.append(np.ascontiguousarray(indices, np.uint32))

    @property
    def steps(self) -> int:
        return len(self.contacts)

    def packed(self):
        coff = np.zeros(self.steps + 1, np.uint64)
        coff[1:] = np.cumsum([len(c) for c in self.contacts])
        ioff = np.zeros(self.steps + 1, np.uint64)
        ioff[1:] = np.cumsum([len(i) for i in self.indices])
        cat = lambda xs, dt: np.concatenate(xs) if xs else np.zeros(0, dt)
        return (cat(self.contacts, R.AVBD_CONTACT), coff, cat(self.offsets, np.uint32), cat(self.indices, np.uint32),
                ioff)


def record_world(scene, nsteps: int, precision: int = 32, lag: int = 1) -> Recording:
    """Step an oracle world and record each step's frozen contact set."""
    w = OracleWorld(scene, precision)
    w.keep_presolve(True)
    rec = Recording(scene.n)
    for _ in range(nsteps):
        w.step(nsteps=1, lag=lag)
        off, idx = w.csr()
        rec.add(w.presolve_contacts(), off, idx)
    return rec


def replay(bodies, state, rec: Recording, options, friction: float, precision: int = 64, seed_body: int = -1,
           params=None):
    """Replay a recording from (bodies, state).  params: (v0 x, y, z, mass) of
    seed_body and the friction, in float64, overriding the state (None: the
    state's values and `friction`).  precision 32 / 64: final positions
    (n x 3) and orientations (n x 4).  precision 65: also their derivatives
    (n x 3 x 5, n x 4 x 5) with respect to those five parameters."""
    lib = load()
    n = rec.n
    bodies = np.ascontiguousarray(bodies, R.GPU_BODY)
    state = np.ascontiguousarray(state, R.BODY_STATE)
    con, coff, csr_off, csr_idx, ioff = rec.packed()
    pos = np.zeros((n, 3))
    rot = np.zeros((n, 4))
    dual = precision in (33, 65)
    dpos = np.zeros((n, 3, 5)) if dual else None
    drot = np.zeros((n, 4, 5)) if dual else None
    pv = None if params is None else np.ascontiguousarray(params, np.float64)
    o = options.to_c()
    sb = n if seed_body is None or seed_body < 0 else seed_body
    lib.oracle_replay(precision, _p(bodies), _p(state), n, rec.steps, _p(con), _p(coff, ctypes.c_uint64),
                      _p(csr_off, c_uint32), _p(csr_idx, c_uint32), _p(ioff, ctypes.c_uint64), ctypes.byref(o),
                      _p(pv, c_double), float(friction), sb, _p(pos, c_double), _p(rot, c_double),
                      _p(dpos, c_double), _p(drot, c_double))
    return (pos, rot, dpos, drot) if dual else (pos, rot)


def replay_frozen(bodies, state, rec: Recording, options, friction: float, precision: int = 64,
                  seed_body: int = -1, params=None, sched_params=None, sched_precision: int = 32, freeze: int = 1):
    """The frozen-penalty replay: the penalty schedule of the replay in
    sched_precision at sched_params (None: the state's values and `friction`)
    held fixed while the replay in `precision` runs at params (as replay()).
    Its dual numbers (precision 65 / 33) are the frozen-penalty derivatives."""
    lib = load()
    n = rec.n
    bodies = np.ascontiguousarray(bodies, R.GPU_BODY)
    state = np.ascontiguousarray(state, R.BODY_STATE)
    con, coff, csr_off, csr_idx, ioff = rec.packed()
    pos = np.zeros((n, 3))
    rot = np.zeros((n, 4))
    dual = precision in (33, 65)
    dpos = np.zeros((n, 3, 5)) if dual else None
    drot = np.zeros((n, 4, 5)) if dual else None
    pv = None if params is None else np.ascontiguousarray(params, np.float64)
    sv = None if sched_params is None else np.ascontiguousarray(sched_params, np.float64)
    o = options.to_c()
    sb = n if seed_body is None or seed_body < 0 else seed_body
    lib.oracle_replay_frozen(precision, sched_precision, freeze, _p(bodies), _p(state), n, rec.steps, _p(con),
                             _p(coff, ctypes.c_uint64), _p(csr_off, c_uint32), _p(csr_idx, c_uint32),
                             _p(ioff, ctypes.c_uint64), ctypes.byref(o), _p(sv, c_double), _p(pv, c_double),
                             float(friction), sb, _p(pos, c_double), _p(rot, c_double), _p(dpos, c_double),
                             _p(drot, c_double))
    return (pos, rot, dpos, drot) if dual else (pos, rot)


def aabbs(bodies: np.ndarray, hull=None, precision: int = 32) -> np.ndarray:
    bodies = np.ascontiguousarray(bodies, R.GPU_BODY)
    hull = np.zeros((0, 3), np.float32) if hull is None else np.ascontiguousarray(hull, np.float32)
    out = np.zeros((len(bodies), 6))
    load().oracle_aabbs(precision, _p(bodies), len(bodies), _p(hull, c_float), len(hull), _p(out, c_double))
    return out


def broad_pairs(bodies: np.ndarray, hull=None, precision: int = 32) -> np.ndarray:
    bodies = np.ascontiguousarray(bodies, R.GPU_BODY)
    hull = np.zeros((0, 3), np.float32) if hull is None else np.ascontiguousarray(hull, np.float32)
    lib = load()
    n = lib.oracle_broad_pairs(precision, _p(bodies), len(bodies), _p(hull, c_float), len(hull), None, 0)
    out = np.zeros((n, 2), np.uint32)
    lib.oracle_broad_pairs(precision, _p(bodies), len(bodies), _p(hull, c_float), len(hull), _p(out, c_uint32), n)
    return out


def narrow_pair(a, b, hull=None, precision: int = 32) -> list:
    """Up to 4 contacts (normal, depth, point_a, point_b) of one pair."""
    a = np.ascontiguousarray(np.atleast_1d(a), R.GPU_BODY)
    b = np.ascontiguousarray(np.atleast_1d(b), R.GPU_BODY)
    hull = np.zeros((0, 3), np.float32) if hull is None else np.ascontiguousarray(hull, np.float32)
    out = np.zeros(40)
    c = load().oracle_narrow_pair(precision, _p(a), _p(b), _p(hull, c_float), len(hull), _p(out, c_double))
    return [
        {"normal": out[10 * k:10 * k + 3], "depth": out[10 * k + 3], "point_a": out[10 * k + 4:10 * k + 7],
         "point_b": out[10 * k + 7:10 * k + 10]}
        for k in range(c)
    ]


def gjk(a, b, hull=None, precision: int = 32) -> bool:
    a = np.ascontiguousarray(np.atleast_1d(a), R.GPU_BODY)
    b = np.ascontiguousarray(np.atleast_1d(b), R.GPU_BODY)
    hull = np.zeros((0, 3), np.float32) if hull is None else np.ascontiguousarray(hull, np.float32)
    return bool(load().oracle_gjk(precision, _p(a), _p(b), _p(hull, c_float), len(hull)))


def solve6(lin6, ang6, cross9, b6, precision: int = 64) -> np.ndarray:
    arrs = [np.ascontiguousarray(x, np.float64) for x in (lin6, ang6, cross9, b6)]
    x = np.zeros(6)
    load().oracle_solve6(precision, *[_p(a, c_double) for a in arrs], _p(x, c_double))
    return x


def container_contacts(body, state, container, options, index: int = 0, hull=None, precision: int = 32) -> dict:
    body = np.ascontiguousarray(np.atleast_1d(body), R.GPU_BODY)
    state = np.ascontiguousarray(np.atleast_1d(state), R.BODY_STATE)
    hull = np.zeros((0, 3), np.float32) if hull is None else np.ascontiguousarray(hull, np.float32)
    ab = np.zeros((24, 2), np.uint32)
    data = np.zeros((24, 25))
    ct, o = container.to_c(), options.to_c()
    m = load().oracle_container_contacts(precision, _p(body), _p(state), ctypes.byref(ct), ctypes.byref(o), index,
                                         _p(hull, c_float), len(hull), _p(ab, c_uint32), _p(data, c_double))
    return split_contacts(ab[:m], data[:m])


def raw_contact(body_a, body_b, a: int, b: int, normal, depth: float, pa, pb, options, precision: int = 32) -> dict:
    ba = np.ascontiguousarray(np.atleast_1d(body_a), R.GPU_BODY)
    bb = np.ascontiguousarray(np.atleast_1d(body_b), R.GPU_BODY)
    n, pa, pb = (np.ascontiguousarray(x, np.float64) for x in (normal, pa, pb))
    ab = np.zeros((1, 2), np.uint32)
    data = np.zeros((1, 25))
    o = options.to_c()
    load().oracle_raw_contact(precision, _p(ba), _p(bb), a, b, _p(n, c_double), float(depth), _p(pa, c_double),
                              _p(pb, c_double), ctypes.byref(o), _p(ab, c_uint32), _p(data, c_double))
    return split_contacts(ab, data)


def finalize_contacts(old: np.ndarray, new: np.ndarray, n: int, options, precision: int = 32):
    old = np.ascontiguousarray(old, R.AVBD_CONTACT)
    new = np.ascontiguousarray(new, R.AVBD_CONTACT)
    m = len(new)
    ab = np.zeros((m, 2), np.uint32)
    data = np.zeros((m, 25))
    nnz = int(np.sum(new["b"] != R.CONTAINER) + m)
    off = np.zeros(n + 1, np.uint32)
    idx = np.zeros(max(nnz, 1), np.uint32)
    o = options.to_c()
    load().oracle_finalize_contacts(precision, _p(old), len(old), _p(new), m, n, ctypes.byref(o), _p(ab, c_uint32),
                                    _p(data, c_double), _p(off, c_uint32), _p(idx, c_uint32))
    return split_contacts(ab, data), off, idx[:nnz]


def avbd_run(state, positions, orientations, contacts, constraints, offsets, indices, run_options,
             precision: int = 32) -> dict:
    st = np.ascontiguousarray(state, R.BODY_STATE)
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    rot = np.ascontiguousarray(orientations, np.float32).reshape(-1, 4)
    con = np.ascontiguousarray(contacts, R.AVBD_CONTACT)
    cns = np.ascontiguousarray(constraints, R.AVBD_CONSTRAINT)
    off = np.ascontiguousarray(offsets, np.uint32)
    idx = np.ascontiguousarray(indices, np.uint32)
    n, m = len(st), len(con)
    out = {"positions": np.zeros((n, 3)), "orientations": np.zeros((n, 4)), "velocities": np.zeros((n, 3)),
           "angular_velocities": np.zeros((n, 3)), "prev_velocities": np.zeros((n, 3)), "lambda": np.zeros((m, 3)),
           "penalty": np.zeros((m, 3))}
    d = lambda k: _p(out[k], c_double)  # noqa: E731
    load().oracle_avbd_run(precision, _p(st), _p(pos, c_float), _p(rot, c_float), n, _p(con), m, _p(cns), len(cns),
                           _p(off, c_uint32), _p(idx, c_uint32), len(idx), ctypes.byref(run_options),
                           d("positions"), d("orientations"), d("velocities"), d("angular_velocities"),
                           d("prev_velocities"), d("lambda"), d("penalty"))
    return out


def avbd_init(state, positions, orientations, run_options, precision: int = 32) -> dict:
    """avbd_init alone: warm-started and inertial poses."""
    st = np.ascontiguousarray(state, R.BODY_STATE)
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    rot = np.ascontiguousarray(orientations, np.float32).reshape(-1, 4)
    n = len(st)
    out = {"warm_pos": np.zeros((n, 3)), "warm_rot": np.zeros((n, 4)), "inertial_pos": np.zeros((n, 3)),
           "inertial_rot": np.zeros((n, 4))}
    d = lambda k: _p(out[k], c_double)  # noqa: E731
    load().oracle_avbd_init(precision, _p(st), _p(pos, c_float), _p(rot, c_float), n, ctypes.byref(run_options),
                            d("warm_pos"), d("warm_rot"), d("inertial_pos"), d("inertial_rot"))
    return out


def qrot(q, v, precision: int = 32) -> np.ndarray:
    q, v = np.ascontiguousarray(q, np.float64), np.ascontiguousarray(v, np.float64)
    out = np.zeros(3)
    load().oracle_qrot(precision, _p(q, c_double), _p(v, c_double), _p(out, c_double))
    return out


def instance_transforms(pos, rot, scale, view_proj, order=None, precision: int = 32) -> np.ndarray:
    """(view_proj @ model, model) per instance slot, shape (count, 2, 4, 4) with
    [k, i] the column-major 16 scalars of the matrix as rows of columns
    ([k, i, col, row]); order[k] is slot k's body (None: slot k draws body k).
    view_proj is given column-major (16 values, glam's Mat4 storage)."""
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    rot = np.ascontiguousarray(rot, np.float64).reshape(-1, 4)
    scale = np.ascontiguousarray(scale, np.float64).reshape(-1, 3)
    vp = np.ascontiguousarray(view_proj, np.float64).reshape(16)
    n = pos.shape[0]
    if order is None:
        count, optr = n, None
    else:
        order = np.ascontiguousarray(order, np.uint32)
        count, optr = order.size, _p(order, c_uint32)
    out = np.zeros(32 * count)
    load().oracle_instance_transforms(precision, _p(pos, c_double), _p(rot, c_double), _p(scale, c_double), n,
                                      optr, count, _p(vp, c_double), _p(out, c_double))
    return out.reshape(count, 2, 4, 4)


def qmul(a, b, precision: int = 32) -> np.ndarray:
    a, b = np.ascontiguousarray(a, np.float64), np.ascontiguousarray(b, np.float64)
    out = np.zeros(4)
    load().oracle_qmul(precision, _p(a, c_double), _p(b, c_double), _p(out, c_double))
    return out


def support(body, direction, host: bool = False, precision: int = 32, hull=None) -> np.ndarray:
    """Local support point of one body: the kernels' rule (first maximal
    vertex) or, with host=True, Shape::support's (last maximal corner of a
    tetrahedron)."""
    body = np.ascontiguousarray(np.atleast_1d(body), R.GPU_BODY)
    d = np.ascontiguousarray(direction, np.float64)
    hull = np.zeros((0, 3), np.float32) if hull is None else np.ascontiguousarray(hull, np.float32)
    out = np.zeros(3)
    load().oracle_support(precision, 1 if host else 0, _p(body), _p(d, c_double), _p(hull, c_float), len(hull),
                          _p(out, c_double))
    return out


def colour_graph(n: int, dynamic, edges) -> tuple:
    """Jones-Plassmann colouring of an arbitrary graph (the solve's colouring
    rule): colour per vertex, number of colours (None on overflow), rounds."""
    dyn = np.ascontiguousarray(np.asarray(dynamic, np.uint8))
    e = np.ascontiguousarray(np.asarray(edges, np.uint32).reshape(-1, 2))
    col = np.zeros(n, np.uint32)
    rounds = ctypes.c_uint32()
    nc = load().oracle_colour_graph(n, _p(dyn, ctypes.c_uint8), _p(e, c_uint32), len(e), _p(col, c_uint32),
                                    ctypes.byref(rounds))
    return col, (None if nc == 0xFFFFFFFF else int(nc)), rounds.value
