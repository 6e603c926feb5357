"""ctypes binding of the C ABI in include/gregnbody.h (libgregnbody.so).

This is plumbing for tests, the benchmark and the autograd harness; the
product's host-side API above the C ABI is the C++ facade include/gtd/nbody.hpp.
The library is loaded from this package directory; a missing build raises
immediately -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgregnbody.so")

NB_OK = 0
NB_ERR_INVALID_ARGUMENT = -1
NB_ERR_NEGATIVE_MASS = -2
NB_ERR_NEGATIVE_RADIUS = -3
NB_ERR_OVERLAPPING_BODIES = -4
NB_ERR_TWO_BODY = -5
NB_ERR_EMPTY_SYSTEM = -6
NB_ERR_INVALID_PLACEMENT = -7
NB_ERR_INDISTINGUISHABLE = -8
NB_ERR_NO_EVOLUTION = -9
NB_ERR_NOT_SUPPORTED = -10
NB_ERR_CAPACITY = -11
NB_ERR_DEVICE = -12
NB_ERR_NSYS_LOAD = -13
NB_ERR_NSYS_FORMAT = -14
NB_ERR_CALLBACK = -15
HOOK_STEP = 1
HOOK_BODIES = 2
HOOK_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int)

TWO_BODY = 1
EULER = 2
MODIFIED_EULER = 4
MIDPOINT = 8
LEAPFROG_KDK = 16
LEAPFROG_DKD = 32
BARNES_HUT = 65536
NO_COLL_CHECK = 0
OVERLAP_COLL_CHECK = 3
PREC_FP64 = 0
PREC_MIXED = 1
PREC_FP64_FAST = 2

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)


class NsysInfo(C.Structure):
    """nb_nsys_info: the .nsys header as the engine reports it."""
    _fields_ = [("d_types", C.c_int32), ("time_size", C.c_int32), ("m_size", C.c_int32), ("r_size", C.c_int32),
                ("t_size", C.c_int32), ("rest_size", C.c_int32), ("g_vals", C.c_uint64 * 6),
                ("num_bodies", C.c_uint64), ("time_elapsed", C.c_double), ("dt", C.c_double)]


_NI = C.POINTER(NsysInfo)

# name -> (restype, argtypes); every symbol declared in include/gregnbody.h
SIGNATURES = {
    "nb_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "nb_destroy": (C.c_int, [_P]),
    "nb_last_error": (C.c_char_p, [_P]),
    "nb_set_bodies": (C.c_int, [_P, C.c_int64, _D, _D, _D, _D, _D, C.c_int]),
    "nb_get_state": (C.c_int, [_P, C.c_int64, _D, _D, _D, _D, _D, _D]),
    "nb_num_bodies": (C.c_int, [_P, _I64]),
    "nb_body_ids": (C.c_int, [_P, C.c_int64, _I64]),
    "nb_set_units": (C.c_int, [_P, C.c_char_p]),
    "nb_set_G": (C.c_int, [_P, C.c_double]),
    "nb_get_G": (C.c_int, [_P, _D]),
    "nb_set_timestep": (C.c_int, [_P, C.c_double]),
    "nb_set_iterations": (C.c_int, [_P, C.c_uint64]),
    "nb_set_softening": (C.c_int, [_P, C.c_double]),
    "nb_set_bh_opening_angle": (C.c_int, [_P, C.c_double]),
    "nb_set_collisions": (C.c_int, [_P, C.c_int, C.c_int, C.c_double]),
    "nb_set_precision": (C.c_int, [_P, C.c_int]),
    "nb_evolve": (C.c_int, [_P, C.c_int, _I64]),
    "nb_advance": (C.c_int, [_P, C.c_uint64, _I64]),
    "nb_advance_noalloc": (C.c_int, [_P, C.c_uint64, _I64]),
    "nb_merge_overlapping": (C.c_int, [_P, _I64]),
    "nb_update_state": (C.c_int, [_P, C.c_int64, _D, _D, _D, _D, _D, _D]),
    "nb_set_step_hook": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_int]),
    "nb_energy": (C.c_int, [_P, _D, _D]),
    "nb_merge_log": (C.c_int, [_P, C.c_int64, _I64, _I64]),
    "nb_set_recording": (C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_int]),
    "nb_track": (C.c_int, [_P, _I64, C.c_int64, _D, _D, _D, _D, _D, _D]),
    "nb_record_count": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "nb_get_record": (C.c_int, [_P, C.c_uint64, _I64, _I64, _D, _D, _I64, _D, _D]),
    "nb_tree_topology": (C.c_int, [_P, C.c_int64, _I32, _I32, _I64, _I64, _D, _D, _D, _I64]),
    "nb_bh_accelerations": (C.c_int, [_P, _D, _I64]),
    "nb_adjoint": (C.c_int, [_P, C.c_uint64, C.c_uint64, _D, _D, _D, _D, _D, _I64, _I64]),
    "nb_nsys_read": (C.c_int, [C.c_char_p, C.c_int, _NI, C.c_int64, _D, _D, _D, _D, _D]),
    "nb_nsys_write": (C.c_int, [C.c_char_p, _NI, C.c_int64, _D, _D, _D, _D, _D, C.c_int, _I64]),
    "nb_load_nsys": (C.c_int, [_P, C.c_char_p, C.c_int, C.c_int]),
    "nb_save_nsys": (C.c_int, [_P, C.c_char_p, C.c_int, _I64]),
    "nb_time_elapsed": (C.c_int, [_P, _D]),
    "nb_set_tuning": (C.c_int, [_P, C.c_char_p, C.c_int64]),
    "nb_get_tuning": (C.c_int, [_P, C.c_char_p, _I64]),
    "nb_autotune": (C.c_int, [_P, C.c_int, C.c_char_p]),
    "nb_profile_enable": (C.c_int, [_P, C.c_int]),
    "nb_profile_read": (C.c_int, [_P, C.c_char_p, _D, _I64]),
    "nb_batch_create": (C.c_int, [C.c_int, C.c_int64, _I64, _D, _D, _D, _D, _D, C.POINTER(_P)]),
    "nb_batch_destroy": (C.c_int, [_P]),
    "nb_batch_last_error": (C.c_char_p, [_P]),
    "nb_batch_set_params": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double,
                                      C.c_int]),
    "nb_batch_evolve": (C.c_int, [_P, C.c_uint64, _I64]),
    "nb_batch_kernel_time": (C.c_int, [_P, _D, _I64]),
    "nb_batch_get_state": (C.c_int, [_P, C.c_int64, _I64, _D, _D, _D, _D, _I64]),
    "nb_batch_merge_log": (C.c_int, [_P, C.c_int64, _I64, _I64]),
    "nb_ic_uniform_ball": (C.c_int, [C.c_uint64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double, _D, _D,
                                     _D]),
    "nb_ic_plummer": (C.c_int, [C.c_uint64, C.c_int64, C.c_double, _D, _D, _D, _D, _D]),
    "nb_ic_comet_ensemble": (C.c_int, [C.c_uint64, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_double,
                                       C.c_double, _I64, _I64, _D, _D, _D, _D, _D]),
}

_lib = None


class NBodyError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """Load libgregnbody.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def dp(a):
    """double* of a C-contiguous float64 array (or None)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def ip(a):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I64)


def i32p(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I32)


def _c64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class System:
    """One gtd::system-shaped context on one GPU (thin wrapper over nb_*)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = _P()
        rc = self._L.nb_create(device, C.byref(h))
        if rc != NB_OK:
            raise NBodyError(rc, self._L.nb_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self._L.nb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        if rc != NB_OK:
            raise NBodyError(rc, self._L.nb_last_error(self.h).decode())
        return rc

    # ---- state ----
    def set_bodies(self, m, pos, vel, r=None, rest_c=None, check_overlap=False):
        m, pos, vel, r, rest_c = map(_c64, (m, pos, vel, r, rest_c))
        n = m.shape[0]
        assert pos.size == 3 * n and vel.size == 3 * n
        self._ck(self._L.nb_set_bodies(self.h, n, dp(m), dp(r), dp(pos), dp(vel), dp(rest_c), int(check_overlap)))

    def merge_overlapping(self) -> int:
        """check_overlap with mergeOverlappingBodies; returns the bodies removed."""
        k = C.c_int64()
        self._ck(self._L.nb_merge_overlapping(self.h, C.byref(k)))
        return k.value

    def update_state(self, m=None, r=None, pos=None, vel=None, acc=None, rest_c=None):
        m, r, pos, vel, acc, rest_c = map(_c64, (m, r, pos, vel, acc, rest_c))
        self._ck(self._L.nb_update_state(self.h, self.num_bodies(), dp(m), dp(r), dp(pos), dp(vel), dp(acc),
                                         dp(rest_c)))

    def set_step_hook(self, fn, mask=HOOK_STEP | HOOK_BODIES):
        """fn(step, where) -> int, called where evolve's callbacks run; None removes it."""
        if fn is None:
            self._hook = None
            self._ck(self._L.nb_set_step_hook(self.h, None, None, 0))
            return
        self._hook = HOOK_FN(lambda user, sys, step, where: int(fn(step, where)))
        self._ck(self._L.nb_set_step_hook(self.h, C.cast(self._hook, C.c_void_p), None, mask))

    def num_bodies(self) -> int:
        n = C.c_int64()
        self._ck(self._L.nb_num_bodies(self.h, C.byref(n)))
        return n.value

    def state(self):
        n = self.num_bodies()
        out = {k: np.zeros(n) for k in ("m", "r", "e")}
        for k in ("pos", "vel", "acc"):
            out[k] = np.zeros((n, 3))
        self._ck(self._L.nb_get_state(self.h, n, dp(out["m"]), dp(out["r"]), dp(out["pos"]), dp(out["vel"]),
                                      dp(out["acc"]), dp(out["e"])))
        return out

    def ids(self):
        n = self.num_bodies()
        a = np.zeros(n, dtype=np.int64)
        self._ck(self._L.nb_body_ids(self.h, n, ip(a)))
        return a

    # ---- parameters ----
    def set_units(self, s: str):
        self._ck(self._L.nb_set_units(self.h, s.encode()))

    def set_G(self, G: float):
        self._ck(self._L.nb_set_G(self.h, G))

    @property
    def G(self) -> float:
        g = C.c_double()
        self._ck(self._L.nb_get_G(self.h, C.byref(g)))
        return g.value

    def set_timestep(self, dt):
        self._ck(self._L.nb_set_timestep(self.h, dt))

    def set_iterations(self, it):
        self._ck(self._L.nb_set_iterations(self.h, it))

    def set_softening(self, eps):
        self._ck(self._L.nb_set_softening(self.h, eps))

    def set_bh_opening_angle(self, th):
        self._ck(self._L.nb_set_bh_opening_angle(self.h, th))

    def set_collisions(self, mode, mergers=False, min_tot_com_mom=1e10):
        self._ck(self._L.nb_set_collisions(self.h, mode, int(mergers), min_tot_com_mom))

    def set_precision(self, p):
        self._ck(self._L.nb_set_precision(self.h, p))

    # ---- stepping ----
    def evolve(self, method=LEAPFROG_KDK) -> int:
        ns = C.c_int64()
        self._ck(self._L.nb_evolve(self.h, method, C.byref(ns)))
        return ns.value

    def advance(self, steps) -> int:
        ns = C.c_int64()
        self._ck(self._L.nb_advance(self.h, steps, C.byref(ns)))
        return ns.value

    def advance_noalloc(self, steps) -> int:
        """nb_advance with the force pass's split partials in the library's
        static device scratch (allocates nothing)."""
        ns = C.c_int64()
        self._ck(self._L.nb_advance_noalloc(self.h, steps, C.byref(ns)))
        return ns.value

    def energy(self):
        ke, pe = C.c_double(), C.c_double()
        self._ck(self._L.nb_energy(self.h, C.byref(ke), C.byref(pe)))
        return ke.value, pe.value

    def track(self, idx, point=None):
        """body_tracker over slots idx: dict(mass, com, com_vel, mean_dist, mean_sep)."""
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        mass, md, ms = C.c_double(), C.c_double(), C.c_double()
        com, cv = np.zeros(3), np.zeros(3)
        pt = None if point is None else np.ascontiguousarray(point, dtype=np.float64)
        self._ck(self._L.nb_track(self.h, ip(idx), len(idx), None if pt is None else dp(pt), C.byref(mass), dp(com),
                                  dp(cv), C.byref(md), C.byref(ms)))
        return dict(mass=mass.value, com=com, com_vel=cv, mean_dist=md.value, mean_sep=ms.value)

    def set_recording(self, every, capacity, what=3):
        """what: 1 state, 2 energies, 3 both (NB_RECORD_STATE | NB_RECORD_ENERGY)."""
        self._ck(self._L.nb_set_recording(self.h, every, capacity, what))

    def record_count(self):
        taken, stored = C.c_uint64(), C.c_uint64()
        self._ck(self._L.nb_record_count(self.h, C.byref(taken), C.byref(stored)))
        return taken.value, stored.value

    def record(self, k):
        """Stored record k (0 = oldest kept): dict(step, pos, vel, ids, ke, pe)."""
        step, n = C.c_int64(), C.c_int64()
        self._ck(self._L.nb_get_record(self.h, k, C.byref(step), C.byref(n), None, None, None, None, None))
        nb = n.value
        pos, vel = np.zeros((nb, 3)), np.zeros((nb, 3))
        ids = np.zeros(max(nb, 1), np.int64)
        ke, pe = C.c_double(), C.c_double()
        self._ck(self._L.nb_get_record(self.h, k, None, None, dp(pos), dp(vel), ip(ids), C.byref(ke), C.byref(pe)))
        return dict(step=step.value, pos=pos, vel=vel, ids=ids[:nb], ke=ke.value, pe=pe.value)

    def merge_log(self):
        cnt = C.c_int64()
        self._ck(self._L.nb_merge_log(self.h, 0, None, C.byref(cnt)))
        rec = np.zeros((max(cnt.value, 1), 4), dtype=np.int64)
        self._ck(self._L.nb_merge_log(self.h, cnt.value, ip(rec), C.byref(cnt)))
        return rec[: cnt.value]

    def tree(self, cap=None):
        n = self.num_bodies()
        cap = cap or 4 * n + 64
        while True:
            depth = np.zeros(cap, np.int32)
            octant = np.zeros(cap, np.int32)
            body = np.zeros(cap, np.int64)
            parent = np.zeros(cap, np.int64)
            mass = np.zeros(cap)
            com = np.zeros((cap, 3))
            side = np.zeros(cap)
            cnt = C.c_int64()
            rc = self._L.nb_tree_topology(self.h, cap, i32p(depth), i32p(octant), ip(body), ip(parent), dp(mass),
                                          dp(com), dp(side), C.byref(cnt))
            if rc == NB_ERR_CAPACITY and cnt.value > cap:
                cap = cnt.value
                continue
            self._ck(rc)
            k = cnt.value
            return dict(depth=depth[:k], octant=octant[:k], body=body[:k], parent=parent[:k], mass=mass[:k],
                        com=com[:k], side=side[:k])

    def tree_count(self):
        """Node count of the octree of the current positions (builds it)."""
        cnt = C.c_int64()
        rc = self._L.nb_tree_topology(self.h, 0, None, None, None, None, None, None, None, C.byref(cnt))
        if rc not in (NB_OK, NB_ERR_CAPACITY):
            self._ck(rc)
        return cnt.value

    def bh_accelerations(self):
        n = self.num_bodies()
        acc = np.zeros((n, 3))
        cnt = C.c_int64()
        self._ck(self._L.nb_bh_accelerations(self.h, dp(acc), C.byref(cnt)))
        return acc, cnt.value

    def adjoint(self, steps, ckpt, gxT, gvT):
        n = self.num_bodies()
        gxT, gvT = _c64(gxT), _c64(gvT)
        gx0, gv0, gm = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
        f, b = C.c_int64(), C.c_int64()
        self._ck(self._L.nb_adjoint(self.h, steps, ckpt, dp(gxT), dp(gvT), dp(gx0), dp(gv0), dp(gm), C.byref(f),
                                    C.byref(b)))
        return gx0, gv0, gm, f.value, b.value

    def load_nsys(self, path, check_nan_inf=True, allow_overlapping=True):
        """Replace bodies, dt, time elapsed and G ratios from an .nsys snapshot."""
        self._ck(self._L.nb_load_nsys(self.h, os.fsencode(path), int(check_nan_inf), int(allow_overlapping)))

    def save_nsys(self, path, value_size=16) -> int:
        """Write the current state as an .nsys snapshot; returns the file size."""
        b = C.c_int64()
        self._ck(self._L.nb_save_nsys(self.h, os.fsencode(path), int(value_size), C.byref(b)))
        return b.value

    def time_elapsed(self) -> float:
        t = C.c_double()
        self._ck(self._L.nb_time_elapsed(self.h, C.byref(t)))
        return t.value

    def set_tuning(self, key, value):
        self._ck(self._L.nb_set_tuning(self.h, key.encode(), int(value)))

    def tuning(self, key):
        v = C.c_int64()
        self._ck(self._L.nb_get_tuning(self.h, key.encode(), C.byref(v)))
        return v.value

    def autotune(self, method=LEAPFROG_KDK, cache_path=None):
        self._ck(self._L.nb_autotune(self.h, method, None if cache_path is None else cache_path.encode()))

    def profile(self, on=True):
        self._ck(self._L.nb_profile_enable(self.h, int(on)))

    def profile_read(self, name):
        ms, cnt = C.c_double(), C.c_int64()
        self._ck(self._L.nb_profile_read(self.h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


class Batch:
    """Ensemble of independent small systems (nb_batch_*)."""

    def __init__(self, offsets, m, r, pos, vel, rest_c, device=0):
        self._L = lib()
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        m, r, pos, vel, rest_c = map(_c64, (m, r, pos, vel, rest_c))
        h = _P()
        nsys = offsets.shape[0] - 1
        rc = self._L.nb_batch_create(device, nsys, ip(offsets), dp(m), dp(r), dp(pos), dp(vel), dp(rest_c),
                                     C.byref(h))
        if rc != NB_OK:
            raise NBodyError(rc, self._L.nb_batch_last_error(None).decode())
        self.h = h
        self.nsys = nsys
        self.total = int(offsets[-1])

    def close(self):
        if getattr(self, "h", None):
            self._L.nb_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        if rc != NB_OK:
            raise NBodyError(rc, self._L.nb_batch_last_error(self.h).decode())

    def set_params(self, G, dt, eps=0.0, mode=OVERLAP_COLL_CHECK, mergers=True, min_tot_com_mom=1e10,
                   precision=PREC_MIXED):
        self._ck(self._L.nb_batch_set_params(self.h, G, dt, eps, mode, int(mergers), min_tot_com_mom, precision))

    def evolve(self, steps) -> int:
        ns = C.c_int64()
        self._ck(self._L.nb_batch_evolve(self.h, steps, C.byref(ns)))
        return ns.value

    def kernel_time(self):
        """(total ms, launches) of the last evolve's stepping launches (HIP events)."""
        ms, n = C.c_double(), C.c_int64()
        self._ck(self._L.nb_batch_kernel_time(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def state(self):
        cap = self.total
        counts = np.zeros(self.nsys, np.int64)
        m, r = np.zeros(cap), np.zeros(cap)
        pos, vel = np.zeros((cap, 3)), np.zeros((cap, 3))
        ids = np.zeros(cap, np.int64)
        self._ck(self._L.nb_batch_get_state(self.h, cap, ip(counts), dp(m), dp(r), dp(pos), dp(vel), ip(ids)))
        k = int(counts.sum())
        return dict(counts=counts, m=m[:k], r=r[:k], pos=pos[:k], vel=vel[:k], ids=ids[:k])

    def merge_log(self):
        cnt = C.c_int64()
        self._ck(self._L.nb_batch_merge_log(self.h, 0, None, C.byref(cnt)))
        rec = np.zeros((max(cnt.value, 1), 5), dtype=np.int64)
        self._ck(self._L.nb_batch_merge_log(self.h, cnt.value, ip(rec), C.byref(cnt)))
        return rec[: cnt.value]


# ---- .nsys snapshots without a context (host only) ----
def _static_ck(rc):
    if rc != NB_OK:
        raise NBodyError(rc, lib().nb_last_error(None).decode())


def nsys_read(path, check_nan_inf=True):
    """Parse an .nsys file: dict(info=NsysInfo, m, r, pos, vel, rest_c)."""
    L = lib()
    info = NsysInfo()
    rc = L.nb_nsys_read(os.fsencode(path), int(check_nan_inf), C.byref(info), 0, None, None, None, None, None)
    if rc not in (NB_OK, NB_ERR_CAPACITY):
        _static_ck(rc)
    n = int(info.num_bodies)
    m, r, rest = np.zeros(n), np.zeros(n), np.zeros(n)
    pos, vel = np.zeros((n, 3)), np.zeros((n, 3))
    _static_ck(L.nb_nsys_read(os.fsencode(path), int(check_nan_inf), C.byref(info), n, dp(m), dp(r), dp(pos),
                              dp(vel), dp(rest)))
    return dict(info=info, m=m, r=r, pos=pos, vel=vel, rest_c=rest)


def nsys_write(path, m, r, pos, vel, rest_c, time_elapsed=0.0, dt=1.0, g_vals=(1, 1, 1, 1, 1, 1),
               value_size=16) -> int:
    """Write an .nsys file (M = R = T of value_size bytes); returns its size."""
    m, r, pos, vel, rest_c = map(_c64, (m, r, pos, vel, rest_c))
    info = NsysInfo()
    info.time_elapsed, info.dt = time_elapsed, dt
    for k in range(6):
        info.g_vals[k] = int(g_vals[k])
    b = C.c_int64()
    _static_ck(lib().nb_nsys_write(os.fsencode(path), C.byref(info), len(m), dp(m), dp(r), dp(pos), dp(vel),
                                   dp(rest_c), int(value_size), C.byref(b)))
    return b.value


# ---- seeded initial conditions ----
def ic_uniform_ball(seed, n, mlo, mhi, radius, sigma):
    m, pos, vel = np.zeros(n), np.zeros((n, 3)), np.zeros((n, 3))
    rc = lib().nb_ic_uniform_ball(seed, n, mlo, mhi, radius, sigma, dp(m), dp(pos), dp(vel))
    if rc != NB_OK:
        raise NBodyError(rc, "nb_ic_uniform_ball")
    return m, pos, vel


def ic_plummer(seed, n, a=1.0, offset=None, vel_offset=None):
    m, pos, vel = np.zeros(n), np.zeros((n, 3)), np.zeros((n, 3))
    off = _c64(offset)
    voff = _c64(vel_offset)
    rc = lib().nb_ic_plummer(seed, n, a, dp(off), dp(voff), dp(m), dp(pos), dp(vel))
    if rc != NB_OK:
        raise NBodyError(rc, "nb_ic_plummer")
    return m, pos, vel


def ic_two_plummer(seed, n_total, a=1.0, sep=10.0, vrel=0.5):
    """Two Plummer spheres at x = -/+sep approaching at relative speed vrel."""
    h = n_total // 2
    m1, p1, v1 = ic_plummer(seed, h, a, np.array([-sep, 0, 0.0]), np.array([vrel / 2, 0, 0.0]))
    m2, p2, v2 = ic_plummer(seed + 0x5DEECE66D, n_total - h, a, np.array([sep, 0, 0.0]),
                            np.array([-vrel / 2, 0, 0.0]))
    m = np.concatenate([m1, m2]) * 0.5  # total mass 1
    return m, np.concatenate([p1, p2]), np.concatenate([v1, v2])


def ic_comet_ensemble(seed, nsys, nmin, nmax, central_mass=1.9e27, central_radius=6.9911e7, G=6.6743e-11):
    L = lib()
    counts = np.zeros(nsys, np.int64)
    offsets = np.zeros(nsys + 1, np.int64)
    rc = L.nb_ic_comet_ensemble(seed, nsys, nmin, nmax, central_mass, central_radius, G, ip(counts), ip(offsets),
                                None, None, None, None, None)
    if rc != NB_OK:
        raise NBodyError(rc, "nb_ic_comet_ensemble")
    tot = int(offsets[-1])
    m, r, e = np.zeros(tot), np.zeros(tot), np.zeros(tot)
    pos, vel = np.zeros((tot, 3)), np.zeros((tot, 3))
    rc = L.nb_ic_comet_ensemble(seed, nsys, nmin, nmax, central_mass, central_radius, G, ip(counts), ip(offsets),
                                dp(m), dp(r), dp(pos), dp(vel), dp(e))
    if rc != NB_OK:
        raise NBodyError(rc, "nb_ic_comet_ensemble")
    return offsets, m, r, pos, vel, e
