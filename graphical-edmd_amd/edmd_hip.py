"""ctypes binding of libedmd_hip.so (include/edmd.h).

Used by the tests and by bench.py.  The library is the product; this module
only marshals numpy arrays across the C ABI.  Loading fails loudly when the
built library is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libedmd_hip.so")
HEADER_PATHS = [os.path.join(os.path.dirname(HERE), "include", "edmd.h"),
                os.path.join(os.path.dirname(HERE), "include", "edmd_compat.h")]

EDMD_OK = 0
EV_CELLCROSS, EV_COLLISION, EV_SCREENSHOT, EV_THERMO, EV_ADDINGNOISE = 0, 1, 2, 3, 4
EV_WALL = 5  # j: 0 = wall normal to x, 1 = normal to y; a wall slot's partner is -5 - j
STOP_NAMES = {0: "none", 1: "events", 2: "collisions", 3: "time", 4: "error", 5: "empty",
              6: "special"}
NEVER = 1e26
INT64_MAX = (1 << 63) - 1
UINT64_MAX = (1 << 64) - 1
RES_AUTO, RES_HBM, RES_LDS = 0, 1, 2  # event-loop array residency (edmd_batch_set_residency)


class EdmdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("edmd error %d: %s" % (code, msg))
        self.code = code


class SystemDesc(C.Structure):
    _fields_ = [("n", C.c_int), ("Lx", C.c_double), ("Ly", C.c_double),
                ("x", C.c_void_p), ("y", C.c_void_p), ("vx", C.c_void_p),
                ("vy", C.c_void_p), ("rad", C.c_void_p), ("m", C.c_void_p),
                ("res", C.c_double), ("first_screen", C.c_double),
                ("dtime", C.c_double), ("first_thermo", C.c_double),
                ("dtime_thermo", C.c_double), ("noise", C.c_int),
                ("dtnoise", C.c_double), ("temperature", C.c_double),
                ("cell_capacity", C.c_int), ("trace_capacity", C.c_int64),
                ("thermo_capacity", C.c_int), ("vr", C.c_double), ("einit", C.c_double),
                ("wall_x", C.c_int), ("wall_y", C.c_int), ("res_wall", C.c_double),
                ("res_tan", C.c_double), ("add_field", C.c_int), ("field", C.c_double),
                ("damping", C.c_int), ("gamma", C.c_double), ("wall_circle", C.c_int),
                ("mid_wall", C.c_int), ("type", C.c_void_p), ("sizeratio", C.c_double),
                ("thermal_wall", C.c_int), ("shear_wall", C.c_int), ("rng_seed", C.c_uint32)]


class Observables(C.Structure):
    _fields_ = [("t", C.c_double), ("ncol", C.c_uint64), ("ncross", C.c_uint64),
                ("nevents", C.c_uint64), ("vir_xx", C.c_double), ("vir_yy", C.c_double),
                ("vir_xy", C.c_double), ("vir_yx", C.c_double), ("E", C.c_double),
                ("px", C.c_double), ("py", C.c_double), ("status", C.c_int32),
                ("stop_reason", C.c_int32), ("err_i", C.c_int32), ("err_j", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


TRACE_DT = np.dtype([("type", "<i4"), ("i", "<i4"), ("j", "<i4"), ("valid", "<i4"),
                     ("t", "<f8")])
THERMO_DT = np.dtype([("t", "<f8"), ("ncol", "<u8"), ("e_per_n", "<f8"), ("pressure", "<f8"),
                      ("pxx", "<f8"), ("pyy", "<f8"), ("pxy", "<f8"), ("pyx", "<f8"),
                      ("pwall", "<f8"), ("pmid", "<f8"), ("dp_lr", "<f8"), ("pack_left", "<f8"),
                      ("pack_right", "<f8"), ("temp_left", "<f8"), ("temp_right", "<f8"),
                      ("temp_solute", "<f8")])

EPOCH_DT = np.dtype([("t", "<f8"), ("ncol", "<u8"), ("ncross", "<u8"), ("nevents", "<u8"),
                     ("status", "<i4"), ("stop_reason", "<i4")])


class TuneReport(C.Structure):
    _fields_ = [("sweep_ms", C.c_double * 8), ("sweep_shape", C.c_int), ("in_lds", C.c_int),
                ("loop_ms_hbm", C.c_double), ("loop_ms_lds", C.c_double),
                ("loop_ms_packed", C.c_double * 2), ("loop_systems_per_group", C.c_int)]

TAPE_DT = np.dtype([("t", "<f8"), ("ax", "<f8"), ("ay", "<f8"), ("cx", "<f8"),
                    ("cy", "<f8"), ("ff", "<f8"), ("i", "<i4"), ("j", "<i4"),
                    ("g", "<f8")])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libedmd_hip.so is not built (%s); run __graft_entry__.build()"
                               % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        vp, i64, u64, dbl, i32 = C.c_void_p, C.c_int64, C.c_uint64, C.c_double, C.c_int
        sig = {
            "edmd_last_error": (C.c_char_p, []),
            "edmd_device_count": (i32, [vp]),
            "edmd_default_desc": (None, [vp]),
            "edmd_ctx_create": (i32, [vp, vp]),
            "edmd_ctx_destroy": (i32, [vp]),
            "edmd_ctx_create_from_dump": (i32, [C.c_char_p, vp, vp]),
            "edmd_run_traced": (i32, [vp, i64, u64, dbl, vp, vp]),
            "edmd_sweep_predict": (i32, [vp]),
            "edmd_run": (i32, [vp, i64, u64, dbl]),
            "edmd_run_until_kind": (i32, [vp, i64, u64, dbl, i32, vp, vp]),
            "edmd_get_state": (i32, [vp] * 9),
            "edmd_get_events": (i32, [vp] * 6),
            "edmd_get_observables": (i32, [vp, vp]),
            "edmd_get_trace": (i32, [vp, vp, i64, vp]),
            "edmd_get_thermo": (i32, [vp, vp, i32, vp]),
            "edmd_get_grid": (i32, [vp] * 6),
            "edmd_get_box": (i32, [vp] * 4),
            "edmd_stream": (vp, [vp]),
            "edmd_peek_event": (i32, [vp] * 6),
            "edmd_collision_time": (i32, [vp, i32, i32, vp]),
            "edmd_repredict": (i32, [vp, i32, i32]),
            "edmd_free_fly": (i32, [vp, i32]),
            "edmd_queue_add": (i32, [vp, i32, i32, i32, dbl, C.c_uint32]),
            "edmd_add_collision_event": (i32, [vp, i32, i32, dbl]),
            "edmd_queue_remove": (i32, [vp, i32, i32]),
            "edmd_batch_create": (i32, [vp, i32, vp]),
            "edmd_batch_destroy": (i32, [vp]),
            "edmd_batch_run": (i32, [vp, i64, u64, dbl]),
            "edmd_batch_sweep_predict": (i32, [vp]),
            "edmd_batch_run_epochs": (i32, [vp, i32, i64, u64, dbl, vp]),
            "edmd_batch_get_observables": (i32, [vp, i32, vp]),
            "edmd_batch_get_state": (i32, [vp, i32] + [vp] * 8),
            "edmd_batch_get_events": (i32, [vp, i32] + [vp] * 4),
            "edmd_batch_get_trace": (i32, [vp, i32, vp, i64, vp]),
            "edmd_batch_get_calendar": (i32, [vp, i32, vp, vp, vp, i64, vp, vp, vp]),
            "edmd_batch_get_thermo": (i32, [vp, i32, vp, i32, vp]),
            "edmd_batch_size": (i32, [vp, vp]),
            "edmd_batch_stream": (vp, [vp]),
            "edmd_batch_set_stream": (i32, [vp, vp]),
            "edmd_batch_set_residency": (i32, [vp, i32]),
            "edmd_batch_get_residency": (i32, [vp, vp, vp, vp]),
            "edmd_batch_set_sweep_shape": (i32, [vp, i32]),
            "edmd_batch_get_sweep_shape": (i32, [vp, vp, vp, vp, vp]),
            "edmd_batch_set_screen": (i32, [vp, i32]),
            "edmd_batch_screen_stats": (i32, [vp, vp, vp, vp]),
            "edmd_batch_autotune": (i32, [vp, i32, i64, vp]),
            "edmd_batch_set_loop_waves": (i32, [vp, i32]),
            "edmd_batch_get_loop_waves": (i32, [vp, vp]),
            "edmd_batch_time_sweep": (i32, [vp, i32, vp]),
            "edmd_init_lattice": (i32, [i32, dbl, dbl, i32, u64, dbl] + [vp] * 6),
            "edmd_init_random": (i32, [i32, dbl, dbl, dbl, dbl, dbl, u64, dbl] + [vp] * 8),
            "edmd_growth_rate": (dbl, [dbl]),
            "edmd_dump_info": (i32, [C.c_char_p, vp, vp, vp]),
            "edmd_dump_read": (i32, [C.c_char_p, i32] + [vp] * 6),
            "edmd_dump_write": (i32, [C.c_char_p, i32, dbl, i32, dbl, dbl] + [vp] * 6 + [i32]),
            "edmd_fp64_check": (i32, [i32, vp, vp, vp, vp]),
            "edmd_pcf": (i32, [i32, vp, vp, dbl, dbl, dbl, dbl, vp, vp, vp]),
            "edmd_batch_pcf": (i32, [vp, i32, i32, dbl, dbl, vp, vp, vp]),
            "edmd_pcf_noalloc": (i32, [i32, vp, vp, dbl, dbl, dbl, dbl, vp, vp, vp]),
            "edmd_batch_deriv_begin": (i32, [vp, i64, i64]),
            "edmd_batch_deriv_run": (i32, [vp, i64, u64, dbl]),
            "edmd_batch_tangent": (i32, [vp] * 5),
            "edmd_batch_adjoint": (i32, [vp] * 5),
            "edmd_batch_deriv_values": (i32, [vp, i32, vp, vp, vp]),
            "edmd_batch_get_tape": (i32, [vp, i32, vp, i64, vp]),
            "edmd_batch_deriv_end": (i32, [vp]),
            "edmd_batch_get_radii": (i32, [vp, i32, vp]),
            "edmd_get_radii": (i32, [vp, vp]),
            "edmd_load_dump": (i32, [C.c_char_p, vp]),
            "edmd_adjoint_begin": (i32, [vp, i64, i64]),
            "edmd_adjoint_run": (i32, [vp, i64, u64, dbl]),
            "edmd_adjoint_tangent": (i32, [vp] * 5),
            "edmd_adjoint_gradient": (i32, [vp, dbl, dbl, vp, vp]),
            "edmd_adjoint_values": (i32, [vp] * 4),
            "edmd_adjoint_end": (i32, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != EDMD_OK:
        raise EdmdError(rc, lib().edmd_last_error().decode())


def _ptr(a):
    return a.ctypes.data if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def make_desc(st, keep, **kw):
    """Fills a SystemDesc from a state dict (x y vx vy rad m Lx Ly); the numpy
    arrays are appended to `keep` so they outlive the call."""
    d = SystemDesc()
    lib().edmd_default_desc(C.byref(d))
    arrs = {k: _f64(st[k]) for k in ("x", "y", "vx", "vy", "rad", "m")}
    keep.append(arrs)
    d.n = len(arrs["x"])
    d.Lx = float(st["Lx"])
    d.Ly = float(st["Ly"])
    for k, a in arrs.items():
        setattr(d, k, a.ctypes.data)
    if "type" in st:
        ty = np.ascontiguousarray(st["type"], dtype=np.int32)
        keep.append(ty)
        d.type = ty.ctypes.data
    for k, v in kw.items():
        setattr(d, k, v)
    return d


class System:
    """One system on the GPU (the plain single-system path)."""

    def __init__(self, st, **kw):
        keep = []
        d = make_desc(st, keep, **kw)
        h = C.c_void_p()
        check(lib().edmd_ctx_create(C.byref(d), C.byref(h)))
        self.h = h
        self.n = d.n
        self.trace_capacity = d.trace_capacity

    def close(self):
        if getattr(self, "h", None):
            lib().edmd_ctx_destroy(self.h)
            self.h = None

    __del__ = close

    @classmethod
    def from_dump(cls, path, **kw):
        """A system loaded from the last frame of a LAMMPS dump; kw: the other
        descriptor fields."""
        self = cls.__new__(cls)
        d = SystemDesc()
        lib().edmd_default_desc(C.byref(d))
        for k, v in kw.items():
            setattr(d, k, v)
        h = C.c_void_p()
        if kw:
            check(lib().edmd_ctx_create_from_dump(str(path).encode(), C.byref(d), C.byref(h)))
        else:
            check(lib().edmd_load_dump(str(path).encode(), C.byref(h)))
        self.h = h
        n, Lx, Ly = C.c_int(), C.c_double(), C.c_double()
        check(lib().edmd_get_box(h, C.byref(n), C.byref(Lx), C.byref(Ly)))
        self.n = n.value
        self.trace_capacity = d.trace_capacity
        return self

    def run(self, max_events=INT64_MAX, max_collisions=UINT64_MAX, tmax=1e300):
        check(lib().edmd_run(self.h, int(max_events), int(max_collisions), float(tmax)))

    def run_traced(self, callback, max_events=INT64_MAX, max_collisions=UINT64_MAX, tmax=1e300):
        """edmd_run_traced: callback(records) receives every handled event in
        order (numpy records of TRACE_DT); a true return stops the run."""
        proto = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.c_void_p)

        def tramp(ptr, n, user):
            buf = (C.c_char * (int(n) * TRACE_DT.itemsize)).from_address(ptr)
            recs = np.frombuffer(bytes(buf), dtype=TRACE_DT)
            return 1 if callback(recs) else 0

        cb = proto(tramp)
        check(lib().edmd_run_traced(self.h, int(max_events), int(max_collisions), float(tmax),
                                    C.cast(cb, C.c_void_p), None))

    def run_until_kind(self, stop_mask, max_events=INT64_MAX, max_collisions=UINT64_MAX,
                       tmax=1e300):
        reason, kind = C.c_int(), C.c_int()
        check(lib().edmd_run_until_kind(self.h, int(max_events), int(max_collisions),
                                        float(tmax), int(stop_mask), C.byref(reason),
                                        C.byref(kind)))
        return reason.value, kind.value

    def sweep_predict(self):
        check(lib().edmd_sweep_predict(self.h))

    def radii(self):
        out = np.zeros(self.n)
        check(lib().edmd_get_radii(self.h, out.ctypes.data))
        return out

    def state(self):
        n = self.n
        out = {k: np.zeros(n) for k in ("x", "y", "vx", "vy", "t")}
        out["coll"] = np.zeros(n, dtype=np.uint32)
        out["cx"] = np.zeros(n, dtype=np.int32)
        out["cy"] = np.zeros(n, dtype=np.int32)
        check(lib().edmd_get_state(self.h, *[_ptr(out[k]) for k in
                                             ("x", "y", "vx", "vy", "t", "coll", "cx", "cy")]))
        return out

    def events(self):
        n = self.n
        out = {"t_cross": np.zeros(n), "dir": np.zeros(n, dtype=np.int32),
               "t_coll": np.zeros(n), "partner": np.zeros(n, dtype=np.int32),
               "coll_actual": np.zeros(n, dtype=np.uint32)}
        check(lib().edmd_get_events(self.h, *[_ptr(out[k]) for k in
                                              ("t_cross", "dir", "t_coll", "partner",
                                               "coll_actual")]))
        return out

    def observables(self):
        o = Observables()
        check(lib().edmd_get_observables(self.h, C.byref(o)))
        return o.as_dict()

    def trace(self, cap=None):
        cap = self.trace_capacity if cap is None else cap
        out = np.zeros(max(int(cap), 1), dtype=TRACE_DT)
        n = C.c_int64()
        check(lib().edmd_get_trace(self.h, out.ctypes.data, int(cap), C.byref(n)))
        return out[:n.value]

    def thermo(self):
        n = C.c_int()
        check(lib().edmd_get_thermo(self.h, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=THERMO_DT)
        check(lib().edmd_get_thermo(self.h, out.ctypes.data, n.value, C.byref(n)))
        return out[:n.value]

    def grid(self):
        ncx, ncy, cap = C.c_int(), C.c_int(), C.c_int()
        csx, csy = C.c_double(), C.c_double()
        check(lib().edmd_get_grid(self.h, C.byref(ncx), C.byref(ncy), C.byref(cap),
                                  C.byref(csx), C.byref(csy)))
        return ncx.value, ncy.value, cap.value, csx.value, csy.value

    def peek(self):
        t, k, i, j, ca = C.c_double(), C.c_int(), C.c_int(), C.c_int(), C.c_uint32()
        check(lib().edmd_peek_event(self.h, C.byref(t), C.byref(k), C.byref(i), C.byref(j),
                                    C.byref(ca)))
        return t.value, k.value, i.value, j.value, ca.value

    def collision_time(self, i, j):
        dt = C.c_double()
        check(lib().edmd_collision_time(self.h, int(i), int(j), C.byref(dt)))
        return dt.value

    def repredict(self, i, which=3):
        check(lib().edmd_repredict(self.h, int(i), int(which)))

    def free_fly(self, i):
        check(lib().edmd_free_fly(self.h, int(i)))

    # ---- derivatives of the event map (edmd_adjoint_*)

    def queue_add(self, kind, i, j, t, coll_actual=0):
        """edmd_queue_add: addEventToQueue of the eventList slot of (kind, i)."""
        check(lib().edmd_queue_add(self.h, int(kind), int(i), int(j), float(t), int(coll_actual)))

    def add_collision_event(self, i, j, t):
        check(lib().edmd_add_collision_event(self.h, int(i), int(j), float(t)))

    def queue_remove(self, kind, i):
        check(lib().edmd_queue_remove(self.h, int(kind), int(i)))

    def adjoint_begin(self, checkpoint_every=256, tape_capacity=0):
        check(lib().edmd_adjoint_begin(self.h, int(checkpoint_every), int(tape_capacity)))

    def adjoint_run(self, max_events=INT64_MAX, max_collisions=UINT64_MAX, tmax=1e300):
        check(lib().edmd_adjoint_run(self.h, int(max_events), int(max_collisions),
                                     float(tmax)))

    def adjoint_tangent(self, tan_in):
        """tan_in: (n, 4) (dx dy dvx dvy per disk). Returns (final tangent,
        dW, d t_last)."""
        t = _f64(np.asarray(tan_in, dtype=np.float64).reshape(self.n, 4))
        out = np.zeros_like(t)
        dw, dtl = C.c_double(), C.c_double()
        check(lib().edmd_adjoint_tangent(self.h, _ptr(t), _ptr(out), C.byref(dw),
                                         C.byref(dtl)))
        return out, dw.value, dtl.value

    def adjoint_gradient(self, lw=0.0, lt=0.0, lam_end=None):
        """Gradient ((n, 4)) of lw W + lt t_last + <lam_end, final state>
        with respect to the taping-origin state."""
        le = None if lam_end is None else _f64(
            np.asarray(lam_end, dtype=np.float64).reshape(self.n, 4))
        g = np.zeros((self.n, 4))
        check(lib().edmd_adjoint_gradient(self.h, float(lw), float(lt), _ptr(le), _ptr(g)))
        return g

    def adjoint_values(self):
        w, tl, k = C.c_double(), C.c_double(), C.c_int64()
        check(lib().edmd_adjoint_values(self.h, C.byref(w), C.byref(tl), C.byref(k)))
        return w.value, tl.value, k.value

    def adjoint_end(self):
        check(lib().edmd_adjoint_end(self.h))


class Batch:
    """Many independent systems, one wavefront each."""

    def __init__(self, states, per_system=None, **kw):
        """kw: descriptor fields for every system; per_system: field -> list
        of per-system values."""
        keep = []
        descs = (SystemDesc * len(states))()
        for k, st in enumerate(states):
            own = dict(kw)
            for f, vals in (per_system or {}).items():
                own[f] = vals[k]
            descs[k] = make_desc(st, keep, **own)
        h = C.c_void_p()
        check(lib().edmd_batch_create(descs, len(states), C.byref(h)))
        self.h = h
        self.ns = [len(st["x"]) for st in states]
        self.trace_capacity = descs[0].trace_capacity

    def close(self):
        if getattr(self, "h", None):
            lib().edmd_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def __len__(self):
        return len(self.ns)

    def run(self, max_events=INT64_MAX, max_collisions=UINT64_MAX, tmax=1e300):
        check(lib().edmd_batch_run(self.h, int(max_events), int(max_collisions), float(tmax)))

    def sweep_predict(self):
        check(lib().edmd_batch_sweep_predict(self.h))

    def run_epochs(self, nepochs, events_per_epoch=INT64_MAX, collisions_per_epoch=UINT64_MAX,
                   tmax=1e300):
        """nepochs epochs in one hipGraph launch; returns an (nepochs, nsys)
        array of EPOCH_DT with every system's counters after each epoch."""
        rows = np.zeros((int(nepochs), len(self.ns)), dtype=EPOCH_DT)
        check(lib().edmd_batch_run_epochs(self.h, int(nepochs), int(events_per_epoch),
                                          int(collisions_per_epoch), float(tmax),
                                          rows.ctypes.data))
        return rows

    def radii(self, s):
        out = np.zeros(self.ns[s])
        check(lib().edmd_batch_get_radii(self.h, int(s), out.ctypes.data))
        return out

    def time_sweep(self, reps):
        ms = C.c_double()
        check(lib().edmd_batch_time_sweep(self.h, int(reps), C.byref(ms)))
        return ms.value

    def observables(self, s):
        o = Observables()
        check(lib().edmd_batch_get_observables(self.h, int(s), C.byref(o)))
        return o.as_dict()

    def state(self, s):
        n = self.ns[s]
        out = {k: np.zeros(n) for k in ("x", "y", "vx", "vy", "t")}
        out["coll"] = np.zeros(n, dtype=np.uint32)
        out["cx"] = np.zeros(n, dtype=np.int32)
        out["cy"] = np.zeros(n, dtype=np.int32)
        check(lib().edmd_batch_get_state(self.h, int(s), *[_ptr(out[k]) for k in
                                                           ("x", "y", "vx", "vy", "t", "coll",
                                                            "cx", "cy")]))
        return out

    def events(self, s):
        n = self.ns[s]
        out = {"t_coll": np.zeros(n), "partner": np.zeros(n, dtype=np.int32),
               "t_cross": np.zeros(n), "dir": np.zeros(n, dtype=np.int32)}
        check(lib().edmd_batch_get_events(self.h, int(s), *[_ptr(out[k]) for k in
                                                            ("t_coll", "partner", "t_cross",
                                                             "dir")]))
        return out

    def calendar(self, s):
        """The event calendar (leaves first, then each level) with the
        insertion ranks of the slots and leaves."""
        n = self.ns[s]
        tot = C.c_int64()
        check(lib().edmd_batch_get_calendar(self.h, int(s), None, None, None, 0, C.byref(tot),
                                            None, None))
        m = tot.value
        out = {"key": np.zeros(m), "idx": np.zeros(m, dtype=np.int32),
               "aux": np.zeros(m, dtype=np.int32),
               "slot_rank": np.zeros(2 * n, dtype=np.int64),
               "leaf_rank": np.zeros(n, dtype=np.int64)}
        check(lib().edmd_batch_get_calendar(self.h, int(s), _ptr(out["key"]), _ptr(out["idx"]),
                                            _ptr(out["aux"]), m, C.byref(tot),
                                            _ptr(out["slot_rank"]), _ptr(out["leaf_rank"])))
        return out

    def trace(self, s, cap=None):
        cap = self.trace_capacity if cap is None else cap
        out = np.zeros(max(int(cap), 1), dtype=TRACE_DT)
        n = C.c_int64()
        check(lib().edmd_batch_get_trace(self.h, int(s), out.ctypes.data, int(cap),
                                         C.byref(n)))
        return out[:n.value]

    def thermo(self, s):
        n = C.c_int()
        check(lib().edmd_batch_get_thermo(self.h, int(s), None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=THERMO_DT)
        check(lib().edmd_batch_get_thermo(self.h, int(s), out.ctypes.data, n.value, C.byref(n)))
        return out[:n.value]

    def stream(self):
        return lib().edmd_batch_stream(self.h)

    def set_residency(self, mode):
        """Where the event loop keeps the systems' arrays: RES_AUTO, RES_HBM
        or RES_LDS (EDMD_ERR_ARG when a system does not fit LDS)."""
        check(lib().edmd_batch_set_residency(self.h, int(mode)))

    def set_sweep_shape(self, shape):
        check(lib().edmd_batch_set_sweep_shape(self.h, int(shape)))

    def sweep_shape(self):
        """{'shape', 'tile_x', 'tile_y', 'threads'} of the prediction sweep."""
        v = [C.c_int() for _ in range(4)]
        check(lib().edmd_batch_get_sweep_shape(self.h, *[C.byref(x) for x in v]))
        return dict(zip(("shape", "tile_x", "tile_y", "threads"), [x.value for x in v]))

    def set_screen(self, mode):
        """The sweep's conservative root screen: 0 off, 1 on, 2 on and
        misses counted (edmd_batch_set_screen); clears the counters."""
        check(lib().edmd_batch_set_screen(self.h, int(mode)))

    def screen_stats(self):
        """(marked candidates, roots solved, misses) since set_screen."""
        m, so, mi = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().edmd_batch_screen_stats(self.h, C.byref(m), C.byref(so), C.byref(mi)))
        return m.value, so.value, mi.value

    def autotune(self, sweep_reps=10, trial_events=2000):
        """Times every sweep shape and every event-loop configuration (arrays
        in HBM with 1, 2, 4 systems per workgroup; arrays in LDS) on this
        batch (state saved and restored) and keeps the fastest; returns the
        report."""
        r = TuneReport()
        check(lib().edmd_batch_autotune(self.h, int(sweep_reps), int(trial_events), C.byref(r)))
        return {"sweep_ms": list(r.sweep_ms), "sweep_shape": r.sweep_shape,
                "in_lds": bool(r.in_lds), "loop_ms_hbm": r.loop_ms_hbm,
                "loop_ms_lds": r.loop_ms_lds, "loop_ms_packed": list(r.loop_ms_packed),
                "loop_systems_per_group": r.loop_systems_per_group}

    def set_loop_waves(self, systems_per_group):
        check(lib().edmd_batch_set_loop_waves(self.h, int(systems_per_group)))

    def loop_waves(self):
        v = C.c_int()
        check(lib().edmd_batch_get_loop_waves(self.h, C.byref(v)))
        return v.value

    def residency(self):
        """{'lds': mode in use is LDS, 'lds_bytes': dynamic LDS per workgroup
        in LDS mode, 'lds_slots': workgroups resident at once in LDS mode}"""
        a, b, c = C.c_int(), C.c_int64(), C.c_int64()
        check(lib().edmd_batch_get_residency(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"lds": bool(a.value), "lds_bytes": b.value, "lds_slots": c.value}

    # ---- derivatives of the event map

    def deriv_begin(self, checkpoint_every=256, tape_capacity=0):
        check(lib().edmd_batch_deriv_begin(self.h, int(checkpoint_every), int(tape_capacity)))

    def deriv_run(self, max_events=INT64_MAX, max_collisions=UINT64_MAX, tmax=1e300):
        check(lib().edmd_batch_deriv_run(self.h, int(max_events), int(max_collisions),
                                         float(tmax)))

    def deriv_end(self):
        check(lib().edmd_batch_deriv_end(self.h))

    def _split(self, flat):
        out, o = [], 0
        for n in self.ns:
            out.append(flat[o:o + n])
            o += n
        return out

    def tangent(self, tan_in):
        """tan_in: list of (n_s, 4) arrays (dx dy dvx dvy per disk).

        Returns (list of final tangents, dW per system, d t_last per system)."""
        flat = _f64(np.concatenate([np.asarray(t, dtype=np.float64).reshape(-1, 4)
                                    for t in tan_in]))
        assert flat.shape == (sum(self.ns), 4)
        out = np.zeros_like(flat)
        dw = np.zeros(len(self.ns))
        dtl = np.zeros(len(self.ns))
        check(lib().edmd_batch_tangent(self.h, _ptr(flat), _ptr(out), _ptr(dw), _ptr(dtl)))
        return self._split(out), dw, dtl

    def adjoint(self, lw=None, lt=None, lam_end=None):
        """Gradient (list of (n_s, 4)) of sum_s lw_s W_s + lt_s t_last_s
        + <lam_end, final state> with respect to the taping-origin state."""
        ns = len(self.ns)
        lw = None if lw is None else _f64(np.broadcast_to(lw, (ns,)))
        lt = None if lt is None else _f64(np.broadcast_to(lt, (ns,)))
        le = None
        if lam_end is not None:
            le = _f64(np.concatenate([np.asarray(a, dtype=np.float64).reshape(-1, 4)
                                      for a in lam_end]))
            assert le.shape == (sum(self.ns), 4)
        g = np.zeros((sum(self.ns), 4))
        check(lib().edmd_batch_adjoint(self.h, _ptr(lw), _ptr(lt), _ptr(le), _ptr(g)))
        return self._split(g)

    def deriv_values(self, s):
        w, tl, k = C.c_double(), C.c_double(), C.c_int64()
        check(lib().edmd_batch_deriv_values(self.h, int(s), C.byref(w), C.byref(tl),
                                            C.byref(k)))
        return w.value, tl.value, k.value

    def tape(self, s):
        k = C.c_int64()
        check(lib().edmd_batch_get_tape(self.h, int(s), None, 0, C.byref(k)))
        out = np.zeros(max(k.value, 1), dtype=TAPE_DT)
        check(lib().edmd_batch_get_tape(self.h, int(s), out.ctypes.data, k.value, C.byref(k)))
        return out[:k.value]

    def pcf(self, s, dr, max_r, extrapolate=False):
        nb = int(max_r / dr)
        r, g = np.zeros(max(nb, 1)), np.zeros(max(nb, 1))
        n = C.c_int(max(nb, 1))
        check(lib().edmd_batch_pcf(self.h, int(s), int(bool(extrapolate)), float(dr),
                                   float(max_r), r.ctypes.data, g.ctypes.data, C.byref(n)))
        return r[:n.value], g[:n.value]


def init_lattice(n, phi, rad=0.5, lattice=0, seed=1, temperature=1.0):
    x, y, vx, vy = (np.zeros(n) for _ in range(4))
    Lx, Ly = C.c_double(), C.c_double()
    check(lib().edmd_init_lattice(int(n), float(phi), float(rad), int(lattice), int(seed),
                                  float(temperature), x.ctypes.data, y.ctypes.data,
                                  vx.ctypes.data, vy.ctypes.data, C.byref(Lx), C.byref(Ly)))
    return {"x": x, "y": y, "vx": vx, "vy": vy, "rad": np.full(n, rad), "m": np.ones(n),
            "Lx": Lx.value, "Ly": Ly.value}


def init_random(n, phi, rad=0.5, poly=0.0, sizeratio=1.0, fsmall=0.0, seed=1, e_per=0.05):
    """Random start for a growth run (pair with vr > 0): target radii in
    "rad", uniform centres, Gaussian velocities with variance e_per / m."""
    a = {k: np.zeros(n) for k in ("x", "y", "vx", "vy", "rad", "m")}
    Lx, Ly = C.c_double(), C.c_double()
    check(lib().edmd_init_random(int(n), float(phi), float(rad), float(poly), float(sizeratio),
                                 float(fsmall), int(seed), float(e_per),
                                 *[a[k].ctypes.data for k in ("x", "y", "vx", "vy", "rad", "m")],
                                 C.byref(Lx), C.byref(Ly)))
    a["Lx"], a["Ly"] = Lx.value, Ly.value
    return a


def growth_rate(phi):
    """The growth rate the reference picks for packing fraction phi."""
    return lib().edmd_growth_rate(float(phi))


def fp64_check(a, b):
    a, b = _f64(a), _f64(b)
    q, s = np.zeros_like(a), np.zeros_like(a)
    check(lib().edmd_fp64_check(len(a), a.ctypes.data, b.ctypes.data, q.ctypes.data,
                                s.ctypes.data))
    return q, s


def pcf(x, y, Lx, Ly, dr, max_r):
    """calculate_pcf on the GPU for host positions."""
    x, y = _f64(x), _f64(y)
    nb = int(max_r / dr)
    r, g = np.zeros(max(nb, 1)), np.zeros(max(nb, 1))
    n = C.c_int(max(nb, 1))
    check(lib().edmd_pcf(len(x), x.ctypes.data, y.ctypes.data, float(Lx), float(Ly), float(dr),
                         float(max_r), r.ctypes.data, g.ctypes.data, C.byref(n)))
    return r[:n.value], g[:n.value]


def device_count():
    n = C.c_int()
    lib().edmd_device_count(C.byref(n))
    return n.value
