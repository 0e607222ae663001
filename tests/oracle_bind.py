"""ctypes binding of the oracle (oracle/build/liboracle.so).

Test infrastructure: the oracle is what the GPU library is compared
against, never a code path of the library itself.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(REPO, "oracle", "build", "liboracle.so")

NEVER = 1e26


class Params(C.Structure):
    _fields_ = [("res", C.c_double), ("first_screen", C.c_double),
                ("dtime", C.c_double), ("first_thermo", C.c_double),
                ("dtime_thermo", C.c_double), ("noise", C.c_int),
                ("dtnoise", C.c_double), ("temperature", C.c_double),
                ("bucket_count", C.c_int), ("vr", C.c_double), ("einit", C.c_double),
                ("wall_x", C.c_int), ("wall_y", C.c_int), ("res_wall", C.c_double),
                ("res_tan", C.c_double), ("add_field", C.c_int), ("field", C.c_double),
                ("damping", C.c_int), ("gamma", C.c_double), ("wall_circle", C.c_int),
                ("mid_wall", C.c_int), ("type", C.c_void_p), ("sizeratio", C.c_double),
                ("thermal_wall", C.c_int), ("shear_wall", C.c_int), ("rng_seed", C.c_uint32)]


class ThermoRow(C.Structure):
    _fields_ = [("t", C.c_double), ("ncol", C.c_uint64), ("e_per_n", C.c_double),
                ("pressure", C.c_double), ("pxx", C.c_double), ("pyy", C.c_double),
                ("pxy", C.c_double), ("pyx", C.c_double), ("pwall", C.c_double),
                ("pmid", C.c_double), ("dp_lr", C.c_double), ("pack_left", C.c_double),
                ("pack_right", C.c_double), ("temp_left", C.c_double),
                ("temp_right", C.c_double), ("temp_solute", C.c_double)]


TRACE_DT = np.dtype([("type", "<i4"), ("i", "<i4"), ("j", "<i4"),
                     ("valid", "<i4"), ("t", "<f8")])

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_default_params.argtypes = [C.POINTER(Params)]
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_double, C.c_double] + [dp] * 6 + [C.POINTER(Params)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_init_events.argtypes = [C.c_void_p]
        L.orc_run.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(C.c_int64)]
        L.orc_run_until.argtypes = [C.c_void_p, C.c_double, C.c_int64, C.POINTER(C.c_int64)]
        L.orc_run_collisions.argtypes = [C.c_void_p, C.c_uint64, C.c_int64, C.POINTER(C.c_int64)]
        L.orc_n.argtypes = [C.c_void_p]
        L.orc_time.restype = C.c_double
        L.orc_time.argtypes = [C.c_void_p]
        L.orc_status.argtypes = [C.c_void_p]
        L.orc_grid.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), dp, dp]
        L.orc_get_state.argtypes = [C.c_void_p] + [C.c_void_p] * 8
        L.orc_get_events.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        L.orc_observables.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_thermo_rows.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.orc_collision_time.restype = C.c_double
        L.orc_collision_time.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_predict_collision.argtypes = [C.c_void_p, C.c_int, dp, C.POINTER(C.c_int32)]
        L.orc_predict_crossing.argtypes = [C.c_void_p, C.c_int, dp, C.POINTER(C.c_int32)]
        L.orc_sync_all.argtypes = [C.c_void_p, C.c_double]
        L.orc_pcf.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, C.c_double,
                              C.c_double, dp, dp]
        L.orc_get_radii.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_add_event.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64]
        L.orc_remove_event.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_coll_count.restype = C.c_uint64
        L.orc_coll_count.argtypes = [C.c_void_p, C.c_int]
        L.orc_crossing_event.argtypes = [C.c_void_p, C.c_int]
        L.orc_collision_event.argtypes = [C.c_void_p, C.c_int]
        L.orc_free_fly.argtypes = [C.c_void_p, C.c_int]
        L.orc_tape_enable.argtypes = [C.c_void_p, C.c_int64]
        L.orc_tape_count.restype = C.c_int64
        L.orc_tape_count.argtypes = [C.c_void_p]
        L.orc_get_tape.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_tape_clear.argtypes = [C.c_void_p]
        L.orc_tangent.restype = C.c_double
        L.orc_tangent.argtypes = [C.c_int, dp, dp, C.c_void_p, C.c_int64, C.c_double, dp, dp,
                                  C.c_double]
        L.orc_adjoint.argtypes = [C.c_int, dp, dp, C.c_void_p, C.c_int64, C.c_double,
                                  C.c_double, C.c_double, C.c_void_p, dp, C.c_double]
        L.orc_replay_fd.restype = C.c_double
        L.orc_replay_fd.argtypes = [C.c_int, C.c_double, C.c_double, dp, dp, C.c_double,
                                    C.c_void_p, C.c_int64, C.c_double] + [dp] * 6 + [
                                    C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_double,
                                    C.c_double]
        L.orc_exp.restype = C.c_double
        L.orc_exp.argtypes = [C.c_double]
        L.orc_log.restype = C.c_double
        L.orc_log.argtypes = [C.c_double]
        L.orc_sincos.argtypes = [C.c_double, C.c_void_p, C.c_void_p]
        L.orc_atan2.restype = C.c_double
        L.orc_atan2.argtypes = [C.c_double, C.c_double]
        L.orc_smallest_root.restype = C.c_double
        L.orc_smallest_root.argtypes = [C.c_double] * 5
        L.orc_mt_draws.restype = None
        L.orc_mt_draws.argtypes = [C.c_uint32, C.c_int, C.c_void_p]
        L.orc_tape_wsum.restype = C.c_double
        L.orc_tape_wsum.argtypes = [C.c_void_p]
        L.orc_replay.restype = C.c_double
        L.orc_replay.argtypes = [C.c_int, C.c_double, C.c_double, dp, dp, C.c_double,
                                 C.c_void_p, C.c_int64, C.c_double] + [dp] * 5 + [dp, C.c_double,
                                                                                 C.c_double]
        _lib = L
    return _lib


TAPE_DT = np.dtype([("i", "<i4"), ("j", "<i4"), ("t", "<f8"), ("ax", "<f8"),
                    ("ay", "<f8"), ("cx", "<f8"), ("cy", "<f8"), ("ff", "<f8"), ("g", "<f8")])


def mt_draws(seed, n):
    """The first n outputs of the oracle's MT19937 seeded by init_genrand."""
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_mt_draws(int(seed), int(n), out.ctypes.data)
    return out


def tangent(m, tl0, tape, t_end, tan, gamma=0.0):
    """Propagates tan (n x 4: dx dy dvx dvy) in place (gamma: damping rate).

    Returns (d sum w, d t_last)."""
    m = np.ascontiguousarray(m, dtype=np.float64)
    tl0 = np.ascontiguousarray(tl0, dtype=np.float64)
    assert tan.dtype == np.float64 and tan.flags.c_contiguous
    tape = np.ascontiguousarray(tape, dtype=TAPE_DT)
    dtl = C.c_double()
    dS = lib().orc_tangent(len(m), _dp(m), _dp(tl0), tape.ctypes.data, len(tape), t_end,
                           _dp(tan), C.byref(dtl), float(gamma))
    return dS, dtl.value


def adjoint(m, tl0, tape, t_end, lw, lt=0.0, lam_end=None, gamma=0.0):
    """Gradient (n x 4) of lw * sum w + lt * t_last + <lam_end, final state>
    with respect to (x, y, vx, vy) at tl0."""
    m = np.ascontiguousarray(m, dtype=np.float64)
    tl0 = np.ascontiguousarray(tl0, dtype=np.float64)
    tape = np.ascontiguousarray(tape, dtype=TAPE_DT)
    g = np.zeros((len(m), 4))
    le = None
    if lam_end is not None:
        le = np.ascontiguousarray(lam_end, dtype=np.float64).reshape(len(m), 4)
    lib().orc_adjoint(len(m), _dp(m), _dp(tl0), tape.ctypes.data, len(tape), t_end, lw, lt,
                      le.ctypes.data if le is not None else None, _dp(g), float(gamma))
    return g


def replay_fd(st, tape, t_end, e, h=1e-7, lw=1.0, lt=0.0, lam=None, res=1.0, field=0.0,
              gamma=0.0):
    """Extended-precision central difference of lw * sum w + lt * t_last +
    <lam, final state> along e (n x 4) over the fixed-sequence replay
    (flights under the field along y when field != 0)."""
    k = {key: np.ascontiguousarray(st[key], dtype=np.float64)
         for key in ("x", "y", "vx", "vy", "t", "rad", "m")}
    e = np.ascontiguousarray(e, dtype=np.float64).reshape(-1, 4)
    tape = np.ascontiguousarray(tape, dtype=TAPE_DT)
    la = None if lam is None else np.ascontiguousarray(lam, dtype=np.float64).reshape(-1, 4)
    return lib().orc_replay_fd(len(k["m"]), st["Lx"], st["Ly"], _dp(k["rad"]), _dp(k["m"]),
                               res, tape.ctypes.data, len(tape), t_end, _dp(k["x"]),
                               _dp(k["y"]), _dp(k["vx"]), _dp(k["vy"]), _dp(k["t"]), _dp(e),
                               h, lw, lt, la.ctypes.data if la is not None else None,
                               float(field), float(gamma))


def replay(st, tape, t_end, res=1.0, field=0.0, gamma=0.0):
    """Fixed-sequence replay of tape from st (x y vx vy and t per disk).

    Returns (sum w, time of the last collision, final state dict)."""
    out = {k: np.array(st[k], dtype=np.float64, copy=True) for k in ("x", "y", "vx", "vy")}
    tl = np.array(st["t"], dtype=np.float64, copy=True)
    rad = np.ascontiguousarray(st["rad"], dtype=np.float64)
    m = np.ascontiguousarray(st["m"], dtype=np.float64)
    tape = np.ascontiguousarray(tape, dtype=TAPE_DT)
    tk = C.c_double()
    S = lib().orc_replay(len(m), st["Lx"], st["Ly"], _dp(rad), _dp(m), res,
                         tape.ctypes.data, len(tape), t_end, _dp(out["x"]), _dp(out["y"]),
                         _dp(out["vx"]), _dp(out["vy"]), _dp(tl), C.byref(tk), float(field),
                         float(gamma))
    return S, tk.value, out


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_params(**kw):
    p = Params()
    lib().orc_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Oracle:
    """One reference-semantics system on the CPU."""

    def __init__(self, st, init_events=True, **params):
        self._keep = {k: np.ascontiguousarray(st[k], dtype=np.float64)
                      for k in ("x", "y", "vx", "vy", "rad", "m")}
        self.n = len(self._keep["x"])
        if "type" in st:
            self._type = np.ascontiguousarray(st["type"], dtype=np.int32)
            params = dict(params, type=self._type.ctypes.data)
        self.params = default_params(**params)
        k = self._keep
        h = lib().orc_create(self.n, st["Lx"], st["Ly"], _dp(k["x"]), _dp(k["y"]),
                             _dp(k["vx"]), _dp(k["vy"]), _dp(k["rad"]), _dp(k["m"]),
                             C.byref(self.params))
        if not h:
            raise ValueError("orc_create failed")
        self.h = h
        if init_events:
            lib().orc_init_events(self.h)

    @property
    def status(self):
        return lib().orc_status(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def run(self, nevents, trace=True):
        tr = np.zeros(nevents, dtype=TRACE_DT) if trace else None
        done = C.c_int64()
        st = lib().orc_run(self.h, nevents, tr.ctypes.data if trace else None, C.byref(done))
        if trace:
            tr = tr[:done.value]
        return st, tr

    def run_until(self, tmax, max_events=1 << 62):
        done = C.c_int64()
        st = lib().orc_run_until(self.h, tmax, max_events, C.byref(done))
        return st, done.value

    def run_collisions(self, target, max_events=1 << 62):
        done = C.c_int64()
        st = lib().orc_run_collisions(self.h, target, max_events, C.byref(done))
        return st, done.value

    @property
    def t(self):
        return lib().orc_time(self.h)

    def grid(self):
        ncx, ncy = C.c_int(), C.c_int()
        csx, csy = C.c_double(), C.c_double()
        lib().orc_grid(self.h, C.byref(ncx), C.byref(ncy), C.byref(csx), C.byref(csy))
        return ncx.value, ncy.value, csx.value, csy.value

    def radii(self):
        out = np.zeros(self.n)
        lib().orc_get_radii(self.h, out.ctypes.data)
        return out

    def state(self):
        n = self.n
        out = {k: np.zeros(n) for k in ("x", "y", "vx", "vy", "t")}
        out["coll"] = np.zeros(n, dtype=np.uint64)
        out["cx"] = np.zeros(n, dtype=np.int32)
        out["cy"] = np.zeros(n, dtype=np.int32)
        lib().orc_get_state(self.h, *[out[k].ctypes.data for k in
                                      ("x", "y", "vx", "vy", "t", "coll", "cx", "cy")])
        return out

    def events(self):
        n = self.n
        out = {"t_cross": np.zeros(n), "dir": np.zeros(n, dtype=np.int32),
               "t_coll": np.zeros(n), "partner": np.zeros(n, dtype=np.int32),
               "coll_actual": np.zeros(n, dtype=np.uint64)}
        lib().orc_get_events(self.h, *[out[k].ctypes.data for k in
                                       ("t_cross", "dir", "t_coll", "partner", "coll_actual")])
        return out

    def observables(self):
        cnt = np.zeros(3, dtype=np.uint64)
        sums = np.zeros(8)
        lib().orc_observables(self.h, cnt.ctypes.data, sums.ctypes.data)
        names = ["t", "virXX", "virYY", "virXY", "virYX", "E", "px", "py"]
        out = dict(zip(names, sums.tolist()))
        out.update(ncol=int(cnt[0]), ncross=int(cnt[1]), nevents=int(cnt[2]))
        return out

    def thermo_rows(self):
        n = lib().orc_thermo_rows(self.h, None, 0)
        arr = (ThermoRow * max(n, 1))()
        lib().orc_thermo_rows(self.h, arr, n)
        return [dict(t=r.t, ncol=r.ncol, e_per_n=r.e_per_n, pressure=r.pressure,
                     pxx=r.pxx, pyy=r.pyy, pxy=r.pxy, pyx=r.pyx, pwall=r.pwall,
                     pmid=r.pmid, dp_lr=r.dp_lr, pack_left=r.pack_left,
                     pack_right=r.pack_right, temp_left=r.temp_left,
                     temp_right=r.temp_right, temp_solute=r.temp_solute)
                for r in arr[:n]]

    def collision_time(self, i, j):
        return lib().orc_collision_time(self.h, i, j)

    def predict_collision(self, i):
        t = C.c_double()
        p = C.c_int32()
        lib().orc_predict_collision(self.h, i, C.byref(t), C.byref(p))
        return t.value, p.value

    def predict_crossing(self, i):
        t = C.c_double()
        d = C.c_int32()
        lib().orc_predict_crossing(self.h, i, C.byref(t), C.byref(d))
        return t.value, d.value

    def sync_all(self, t):
        lib().orc_sync_all(self.h, t)

    def add_event(self, kind, i, j, t, coll_actual=None):
        """addEventToQueue of the eventList slot of (kind, i); a collision
        takes j's current counter unless coll_actual is given."""
        if coll_actual is None:
            coll_actual = lib().orc_coll_count(self.h, j) if kind == 1 else 0
        assert lib().orc_add_event(self.h, kind, i, j, t, coll_actual) == 0

    def remove_event(self, kind, i):
        assert lib().orc_remove_event(self.h, kind, i) == 0

    def coll_count(self, i):
        return lib().orc_coll_count(self.h, i)

    def crossing_event(self, i):
        lib().orc_crossing_event(self.h, i)

    def collision_event(self, i):
        lib().orc_collision_event(self.h, i)

    def free_fly(self, i):
        lib().orc_free_fly(self.h, i)

    def tape_enable(self, capacity):
        assert lib().orc_tape_enable(self.h, capacity) == 0

    def tape(self):
        k = lib().orc_tape_count(self.h)
        out = np.zeros(k, dtype=TAPE_DT)
        lib().orc_get_tape(self.h, out.ctypes.data)
        return out

    def tape_clear(self):
        lib().orc_tape_clear(self.h)

    def tape_wsum(self):
        return lib().orc_tape_wsum(self.h)


def pcf(x, y, Lx, Ly, dr, max_r):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    nb = int(max_r / dr)
    r = np.zeros(max(nb, 1))
    g = np.zeros(max(nb, 1))
    nb = lib().orc_pcf(len(x), _dp(x), _dp(y), Lx, Ly, dr, max_r, _dp(r), _dp(g))
    return r[:nb], g[:nb]
