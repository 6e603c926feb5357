"""ctypes binding of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of the reference's force-and-step path
(nbody_oracle.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, and only as the reference to compare against.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

LD = np.longdouble
_LD = C.c_longdouble
_LDP = C.POINTER(C.c_longdouble)
_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)

FORM_NEWTON = 0
FORM_SOFT = 1
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_direct_acc_ld.argtypes = [C.c_int64, C.c_int, _LDP, _LDP, _LDP, _LDP, _LDP]
        L.orc_direct_acc_d.argtypes = [C.c_int64, C.c_int, C.c_double, C.c_double, _D, _D, _D]
        L.orc_direct_acc_rows_d.argtypes = [C.c_int64, C.c_int, C.c_double, C.c_double, _D, _D, _D]
        L.orc_energy_ld.argtypes = [C.c_int64, C.c_int, _LDP, _LDP, _LDP, _LDP, _LDP, _LDP, _LDP]
        L.orc_energy_d.argtypes = [C.c_int64, C.c_int, C.c_double, C.c_double, _D, _D, _D, _D, _D]
        L.orc_kdk_ld.argtypes = [_I64, C.c_int, _LDP, _LDP, _LDP, C.c_int64, C.c_int, _LDP, _LDP, _LDP, _LDP, _LDP,
                                 _LDP, _LDP, _I64, _I64, C.c_int64, _I64]
        L.orc_kdk_d.argtypes = [_I64, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int64, C.c_int, C.c_double,
                                _D, _D, _D, _D, _D, _D, _I64, _I64, C.c_int64, _I64]
        L.orc_check_overlap_d.argtypes = [_I64, C.c_int, _D, _D, _D, _D, _D, _I64, _I64]
        L.orc_check_overlap_ld.argtypes = [_I64, C.c_int, _LDP, _LDP, _LDP, _LDP, _LDP, _I64, _I64]
        L.orc_s_coll_d.argtypes = [_I64, C.c_int, C.c_double, C.c_int64, _D, _D, _D, _D, _D, _D, _I64, _I64, _I64,
                                   C.c_int64, _I64]
        L.orc_bh_tree_d.restype = C.c_int64
        L.orc_bh_tree_d.argtypes = [C.c_int64, _D, _D, C.c_int64, _I32, _I32, _I64, _I64, _D, _D, _D, _D]
        L.orc_bh_acc_d.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_double, _D, _D, _D, _I64, _I64]
        L.orc_bh_accepted_d.restype = C.c_int64
        L.orc_bh_accepted_d.argtypes = [C.c_int64, C.c_double, _D, _D, C.c_int64, C.c_int64, _I64]
        L.orc_bh_kdk_d.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64, _D, _D,
                                   _D, _D]
        L.orc_adjoint_d.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int64, _D, _D, _D, _D, _D, _D,
                                    _D, _D]
        L.orc_forward_d.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int64, _D, _D, _D]
        L.orc_direct_acc_mt.argtypes = [C.c_int64, C.c_double, C.c_double, _D, _D, _D, C.c_int]
        L.orc_bh_acc_mt.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_double, _D, _D, _D, _I64, C.c_int]
        L.orc_max_threads.restype = C.c_int
        L.orc_integrate_ld.argtypes = [C.c_int] + L.orc_kdk_ld.argtypes
        L.orc_integrate_d.argtypes = [C.c_int] + L.orc_kdk_d.argtypes
        L.orc_bh_s_coll_ld.argtypes = [C.c_int64, _LDP, _LDP, _LDP, _LDP, _LDP]
        L.orc_bh_s_coll_d.argtypes = [C.c_int64, _D, _D, _D, _D, _D]
        L.orc_bh_neighbours_d.restype = C.c_int64
        L.orc_bh_neighbours_d.argtypes = [C.c_int64, _D, _D, C.c_int64, C.c_int64, _I64]
        L.orc_bh_kdk_coll_d.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64,
                                        C.c_int, _D, _D, _D, _D, _D, _D]
        L.orc_track_ld.argtypes = [C.c_int64, _I64, _LDP, _LDP, _LDP, _LDP, _LDP]
        L.orc_track_d.argtypes = [C.c_int64, _I64, _D, _D, _D, _D, _D]
        L.orc_trace.argtypes = [C.c_int]
        L.orc_trace_len.restype = C.c_int64
        L.orc_trace_copy.argtypes = [_I64, _LDP]
        _lib = L
    return _lib


def _ld(a):
    return np.ascontiguousarray(a, dtype=LD)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


_keep = []


def _sc(x):
    """Pointer to a long double scalar (kept alive until the next call)."""
    a = np.array([x], dtype=LD)
    _keep.append(a)
    del _keep[:-16]
    return a.ctypes.data_as(_LDP)


def _lp(a):
    return None if a is None else a.ctypes.data_as(_LDP)


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _ip(a):
    return None if a is None else a.ctypes.data_as(_I64)


def direct_acc(m, pos, G, eps=0.0, form=FORM_NEWTON, dtype=LD):
    n = len(m)
    if dtype == LD:
        m, pos = _ld(m), _ld(pos)
        acc = np.zeros((n, 3), LD)
        lib().orc_direct_acc_ld(n, form, _sc(G), _sc(eps), _lp(m), _lp(pos), _lp(acc))
    else:
        m, pos = _d(m), _d(pos)
        acc = np.zeros((n, 3))
        lib().orc_direct_acc_d(n, form, G, eps, _dp(m), _dp(pos), _dp(acc))
    return acc


def direct_acc_rows(m, pos, G, eps=0.0, form=FORM_NEWTON):
    """FP64 direct_acc with one body per thread (the adjoint's forward)."""
    m, pos = _d(m), _d(pos)
    acc = np.zeros((len(m), 3))
    lib().orc_direct_acc_rows_d(len(m), form, G, eps, _dp(m), _dp(pos), _dp(acc))
    return acc


def energy(m, pos, vel, G, eps=0.0, form=FORM_NEWTON, dtype=LD):
    n = len(m)
    if dtype == LD:
        m, pos, vel = _ld(m), _ld(pos), _ld(vel)
        out = np.zeros(2, LD)
        lib().orc_energy_ld(n, form, _sc(G), _sc(eps), _lp(m), _lp(pos), _lp(vel), _lp(out[0:1]), _lp(out[1:2]))
        return out[0], out[1]
    m, pos, vel = _d(m), _d(pos), _d(vel)
    ke, pe = C.c_double(), C.c_double()
    lib().orc_energy_d(n, form, G, eps, _dp(m), _dp(pos), _dp(vel), C.byref(ke), C.byref(pe))
    return ke.value, pe.value


METHOD_EULER, METHOD_MODIFIED_EULER, METHOD_MIDPOINT, METHOD_KDK, METHOD_DKD = 2, 4, 8, 16, 32


def kdk(m, pos, vel, G, dt, steps, eps=0.0, form=FORM_NEWTON, dtype=LD, coll=0, r=None, e=None, min_mom=1e10,
        labels=None, log_cap=100000, method=METHOD_KDK):
    """Leapfrog KDK, or another direct integrator of evolve via `method` (the
    reference's flag), + optional s_coll: coll=1 bounce, coll=2 mergers."""
    n = len(m)
    conv = _ld if dtype == LD else _d
    ptr = _lp if dtype == LD else _dp
    m, pos, vel = conv(m).copy(), conv(pos).copy(), conv(vel).copy()
    r = conv(r if r is not None else np.zeros(n)).copy()
    e = conv(e if e is not None else np.ones(n)).copy()
    acc = np.zeros((n, 3), dtype=m.dtype)
    lab = np.ascontiguousarray(np.arange(n) if labels is None else labels, dtype=np.int64).copy()
    log = np.zeros((log_cap, 4), dtype=np.int64)
    nn = C.c_int64(n)
    loglen = C.c_int64()
    if dtype == LD:
        lib().orc_integrate_ld(method, C.byref(nn), form, _sc(G), _sc(dt), _sc(eps), steps, coll, _sc(min_mom),
                               ptr(m), ptr(r), ptr(e), ptr(pos), ptr(vel), ptr(acc), _ip(lab), _ip(log), log_cap,
                               C.byref(loglen))
    else:
        lib().orc_integrate_d(method, C.byref(nn), form, float(G), float(dt), float(eps), steps, coll,
                              float(min_mom), ptr(m), ptr(r), ptr(e), ptr(pos), ptr(vel), ptr(acc), _ip(lab), _ip(log),
                              log_cap, C.byref(loglen))
    k = nn.value
    return dict(n=k, m=m[:k], r=r[:k], e=e[:k], pos=pos[:k], vel=vel[:k], acc=acc[:k], labels=lab[:k],
                log=log[: min(loglen.value, log_cap)])


def kdk_trace(*args, **kw):
    """kdk() that also returns every body's state after every step:
    (steps[k], labels[k], vals[k, 8] = m, r, x, y, z, vx, vy, vz)."""
    L = lib()
    L.orc_trace(1)
    try:
        res = kdk(*args, **kw)
        k = L.orc_trace_len()
        hdr = np.zeros((k, 2), np.int64)
        vals = np.zeros((k, 8), LD)
        L.orc_trace_copy(_ip(hdr), _lp(vals))
    finally:
        L.orc_trace(0)
    return res, hdr[:, 0], hdr[:, 1], vals


def s_coll(m, r, e, pos, vel, acc, labels, next_label, merge=True, min_mom=1e10, step=1, log_cap=10000):
    n = len(m)
    m, r, e = _d(m).copy(), _d(r).copy(), _d(e).copy()
    pos, vel, acc = _d(pos).copy(), _d(vel).copy(), _d(acc).copy()
    lab = np.ascontiguousarray(labels, dtype=np.int64).copy()
    nl = C.c_int64(next_label)
    nn = C.c_int64(n)
    log = np.zeros((log_cap, 4), dtype=np.int64)
    loglen = C.c_int64()
    lib().orc_s_coll_d(C.byref(nn), int(merge), min_mom, step, _dp(m), _dp(r), _dp(e), _dp(pos), _dp(vel), _dp(acc),
                       _ip(lab), C.byref(nl), _ip(log), log_cap, C.byref(loglen))
    k = nn.value
    return dict(n=k, m=m[:k], r=r[:k], e=e[:k], pos=pos[:k], vel=vel[:k], acc=acc[:k], labels=lab[:k],
                next_label=nl.value, log=log[: loglen.value])


def check_overlap(m, r, pos, vel, e=None, merge=False, dtype=LD):
    """system::check_overlap (gregsys.hpp:307-342).  merge=False: (outer,
    inner) of the first overlap, or None.  merge=True: the bodies after the
    construction-time merges, dict(n, m, r, e, pos, vel, labels)."""
    n = len(m)
    conv = _ld if dtype == LD else _d
    ptr = _lp if dtype == LD else _dp
    m, r, pos, vel = conv(m).copy(), conv(r).copy(), conv(pos).copy(), conv(vel).copy()
    e = conv(e if e is not None else np.ones(n)).copy()
    lab = np.arange(n, dtype=np.int64)
    first = np.zeros(2, np.int64)
    nn = C.c_int64(n)
    fn = lib().orc_check_overlap_ld if dtype == LD else lib().orc_check_overlap_d
    fn(C.byref(nn), int(merge), ptr(m), ptr(r), ptr(e), ptr(pos), ptr(vel), _ip(lab), _ip(first))
    if not merge:
        return None if first[0] < 0 else (int(first[0]), int(first[1]))
    k = nn.value
    return dict(n=k, m=m[:k], r=r[:k], e=e[:k], pos=pos[:k], vel=vel[:k], labels=lab[:k])


def bh_tree(m, pos, cap=None):
    n = len(m)
    m, pos = _d(m), _d(pos)
    cap = cap or 4 * n + 64
    while True:
        depth, octant = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        body, parent = np.zeros(cap, np.int64), np.zeros(cap, np.int64)
        mass, com, side = np.zeros(cap), np.zeros((cap, 3)), np.zeros(cap)
        box = np.zeros(6)
        k = lib().orc_bh_tree_d(n, _dp(m), _dp(pos), cap, depth.ctypes.data_as(_I32), octant.ctypes.data_as(_I32),
                                _ip(body), _ip(parent), _dp(mass), _dp(com), _dp(side), _dp(box))
        if k == -1:
            cap *= 2
            continue
        if k < 0:
            raise RuntimeError(f"oracle tree build failed: {k}")
        return dict(depth=depth[:k], octant=octant[:k], body=body[:k], parent=parent[:k], mass=mass[:k],
                    com=com[:k], side=side[:k], box=box)


def bh_acc(m, pos, G, eps, theta, per_body=False):
    n = len(m)
    m, pos = _d(m), _d(pos)
    acc = np.zeros((n, 3))
    cnt = C.c_int64()
    pb = np.zeros(n, np.int64) if per_body else None
    rc = lib().orc_bh_acc_d(n, G, eps, theta, _dp(m), _dp(pos), _dp(acc), C.byref(cnt), _ip(pb))
    if rc:
        raise RuntimeError(f"oracle bh_acc failed: {rc}")
    return (acc, cnt.value, pb) if per_body else (acc, cnt.value)


def direct_acc_mt(m, pos, G, eps, threads=0):
    """All-core FP64 softened direct sum (timing baseline, not a parity reference)."""
    n = len(m)
    m, pos = _d(m), _d(pos)
    acc = np.zeros((n, 3))
    lib().orc_direct_acc_mt(n, G, eps, _dp(m), _dp(pos), _dp(acc), threads)
    return acc


def bh_acc_mt(m, pos, G, eps, theta, threads=0):
    """All-core corrected BH (serial insertion tree, parallel walk); timing baseline."""
    n = len(m)
    m, pos = _d(m), _d(pos)
    acc = np.zeros((n, 3))
    cnt = C.c_int64()
    rc = lib().orc_bh_acc_mt(n, G, eps, theta, _dp(m), _dp(pos), _dp(acc), C.byref(cnt), threads)
    if rc:
        raise RuntimeError(f"oracle bh_acc_mt failed: {rc}")
    return acc, cnt.value


def max_threads():
    return lib().orc_max_threads()


def bh_accepted(m, pos, theta, body):
    n = len(m)
    out = np.zeros(8 * n + 64, np.int64)
    k = lib().orc_bh_accepted_d(n, theta, _dp(_d(m)), _dp(_d(pos)), body, out.size, _ip(out))
    return out[:k]


def bh_kdk(m, pos, vel, G, dt, eps, theta, steps):
    n = len(m)
    m, pos, vel = _d(m), _d(pos).copy(), _d(vel).copy()
    acc = np.zeros((n, 3))
    rc = lib().orc_bh_kdk_d(n, G, dt, eps, theta, steps, _dp(m), _dp(pos), _dp(vel), _dp(acc))
    if rc:
        raise RuntimeError(f"oracle bh_kdk failed: {rc}")
    return pos, vel, acc


def bh_s_coll(m, r, e, pos, vel, dtype=LD):
    """One bh_s_coll pass (touching-cell neighbours, bounces) on a state;
    returns the new velocities."""
    n = len(m)
    conv, ptr = (_ld, _lp) if dtype == LD else (_d, _dp)
    m, r, e, pos, vel = conv(m), conv(r), conv(e), conv(pos), conv(vel).copy()
    f = lib().orc_bh_s_coll_ld if dtype == LD else lib().orc_bh_s_coll_d
    rc = f(n, ptr(m), ptr(r), ptr(e), ptr(pos), ptr(vel))
    if rc:
        raise RuntimeError(f"oracle bh_s_coll failed: {rc}")
    return vel


def track(m, pos, vel, idx, point, dtype=LD):
    """body_tracker over bodies idx: dict(mass, com, com_vel, mean_dist, mean_sep)."""
    conv, ptr = (_ld, _lp) if dtype == LD else (_d, _dp)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    m, pos, vel, point = conv(m), conv(pos), conv(vel), conv(point)
    out = np.zeros(9, dtype=m.dtype)
    f = lib().orc_track_ld if dtype == LD else lib().orc_track_d
    f(len(idx), _ip(idx), ptr(m), ptr(pos), ptr(vel), ptr(point), ptr(out))
    return dict(mass=out[0], com=out[1:4], com_vel=out[4:7], mean_dist=out[7], mean_sep=out[8])


def bh_neighbours(m, pos, body):
    """The bodies nn_iterator yields for `body`, in its order."""
    n = len(m)
    out = np.zeros(n, np.int64)
    k = lib().orc_bh_neighbours_d(n, _dp(_d(m)), _dp(_d(pos)), body, out.size, _ip(out))
    if k < 0:
        raise RuntimeError(f"oracle bh_neighbours failed: {k}")
    return out[:k]


def bh_kdk_coll(m, r, e, pos, vel, G, dt, eps, theta, steps, coll=True):
    n = len(m)
    m, r, e, pos, vel = _d(m), _d(r), _d(e), _d(pos).copy(), _d(vel).copy()
    acc = np.zeros((n, 3))
    rc = lib().orc_bh_kdk_coll_d(n, G, dt, eps, theta, steps, int(coll), _dp(m), _dp(r), _dp(e), _dp(pos), _dp(vel),
                                 _dp(acc))
    if rc:
        raise RuntimeError(f"oracle bh_kdk_coll failed: {rc}")
    return pos, vel, acc


def adjoint(m, x0, v0, G, dt, eps, steps, gxT, gvT):
    n = len(m)
    gx0, gv0, gm = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    lib().orc_adjoint_d(n, G, dt, eps, steps, _dp(_d(m)), _dp(_d(x0)), _dp(_d(v0)), _dp(_d(gxT)), _dp(_d(gvT)),
                        _dp(gx0), _dp(gv0), _dp(gm))
    return gx0, gv0, gm


def forward(m, x0, v0, G, dt, eps, steps):
    x, v = _d(x0).copy(), _d(v0).copy()
    lib().orc_forward_d(len(m), G, dt, eps, steps, _dp(_d(m)), _dp(x), _dp(v))
    return x, v


# ---------------- compiled reference (oracle/_ref, this container only) ----------------
def ref_available():
    return os.path.exists(os.path.join(REF_DIR, "ref_harness")) and os.path.exists(
        os.path.join(REF_DIR, "ref_harness_merge"))


def write_ref_input(path, m, pos, vel, r=None, e=None, steps=0, dt=1.0):
    n = len(m)
    r = np.zeros(n) if r is None else r
    e = np.ones(n) if e is None else e
    rec = np.zeros((n, 9))
    rec[:, 0], rec[:, 1], rec[:, 2] = m, r, e
    rec[:, 3:6], rec[:, 6:9] = np.asarray(pos).reshape(n, 3), np.asarray(vel).reshape(n, 3)
    with open(path, "wb") as f:
        f.write(np.array([n, steps], dtype=np.int64).tobytes())
        f.write(np.array([dt], dtype=np.float64).tobytes())
        f.write(rec.astype(np.float64).tobytes())


def ref_bhcoll(m, r, e, pos, vel, tmpdir):
    """The compiled reference's bh_s_coll on one state: velocities (long double)."""
    inp, out = os.path.join(tmpdir, "bhc_in.bin"), os.path.join(tmpdir, "bhc_out.bin")
    write_ref_input(inp, m, pos, vel, r=r, e=e, steps=1, dt=1.0)
    subprocess.run([os.path.join(REF_DIR, "ref_bhcoll"), inp, out], check=True, timeout=600)
    return np.fromfile(out, dtype=LD).reshape(len(m), 3)


TRACK_POINT = (1.5e8, -2.5e8, 0.5e8)  # the point ref_harness's track mode measures distances to


def ref_track(m, pos, vel, tmpdir):
    """The compiled reference's body_tracker over the bodies with x > 0:
    (count, [mass, com xyz, com_vel xyz, mean distance to TRACK_POINT, mean separation]) in long double."""
    inp, out = os.path.join(tmpdir, "trk_in.bin"), os.path.join(tmpdir, "trk_out.bin")
    write_ref_input(inp, m, pos, vel, steps=1, dt=1.0)
    run_ref("track", inp, out)
    raw = open(out, "rb").read()
    cnt = int(np.frombuffer(raw[:8], np.int64)[0])
    return cnt, np.frombuffer(raw[8:], LD).copy()


def run_ref(mode, inp, out, merge=False, timeout=3600):
    exe = os.path.join(REF_DIR, "ref_harness_merge" if merge else "ref_harness")
    subprocess.run([exe, mode, inp, out], check=True, timeout=timeout)


def read_ref_kdk(path, n):
    a = np.fromfile(path, dtype=LD)
    return a.reshape(n, 6)


def read_ref_coll(path):
    raw = open(path, "rb").read()
    n0 = int(np.frombuffer(raw[:8], np.int64)[0])
    ids0 = np.frombuffer(raw[8:8 + 8 * n0], np.int64).copy()
    off = 8 + 8 * n0
    rec_sz = 16 + 8 * 16
    k = (len(raw) - off) // rec_sz
    steps = np.zeros(k, np.int64)
    ids = np.zeros(k, np.int64)
    vals = np.zeros((k, 8), LD)
    for q in range(k):
        b = raw[off + q * rec_sz: off + (q + 1) * rec_sz]
        hdr = np.frombuffer(b[:16], np.int64)
        steps[q], ids[q] = hdr
        vals[q] = np.frombuffer(b[16:], LD)
    return ids0, steps, ids, vals


def read_ref_tree(path):
    raw = open(path, "rb").read()
    cnt = int(np.frombuffer(raw[:8], np.int64)[0])
    rec = np.frombuffer(raw[8:], dtype=np.dtype([("d", "<i8"), ("o", "<i8"), ("b", "<i8"), ("m", "<f8")]))
    assert rec.shape[0] == cnt
    return rec


# ---------------- .nsys snapshot layout (restatement) ----------------
# nsys_file_header / nsys_file_chunk (gregsys.hpp:1444-1477), written by
# system::to_nsys (:3214-3258): 96-byte header, then per body a 16-byte
# restitution field and mass, radius, x, y, z, vx, vy, vz in their declared
# types with natural alignment, file padded to a multiple of 4 bytes.  Only
# floating-point M/R/T (4, 8, 16 bytes) are restated here.
_NSYS_FLOAT = {4: np.float32, 8: np.float64, 16: LD}


def parse_nsys(path):
    raw = open(path, "rb").read()
    h = raw[:96]
    out = {"signature": h[:2], "d_types": h[2], "time_size": h[3],
           "m_size": h[36], "r_size": h[37], "t_size": h[38], "rest_size": h[39]}
    out["g_vals"] = np.frombuffer(h[40:88], np.uint64).copy()
    n = int(np.frombuffer(h[88:96], np.uint64)[0])
    out["num_bodies"] = n
    tt = _NSYS_FLOAT[h[3]]
    out["time_elapsed"] = LD(np.frombuffer(h[4:4 + h[3]], tt)[0])
    out["dt"] = LD(np.frombuffer(h[20:20 + h[3]], tt)[0])
    sizes = [h[36], h[37], h[38]]
    offs, off = [], 16
    for sz in sizes:
        off = (off + sz - 1) // sz * sz
        offs.append(off)
        off += sz if len(offs) < 3 else 6 * sz
    chunk = (off + max(sizes) - 1) // max(sizes) * max(sizes)
    body = n * chunk
    out["expected_size"] = 96 + body + ((4 - body % 4) % 4)
    out["file_size"] = len(raw)
    rest, m, r = np.zeros(n, LD), np.zeros(n, LD), np.zeros(n, LD)
    pos, vel = np.zeros((n, 3), LD), np.zeros((n, 3), LD)
    rt = _NSYS_FLOAT[h[39]]
    for i in range(n):
        c = raw[96 + i * chunk: 96 + (i + 1) * chunk]
        rest[i] = np.frombuffer(c[:h[39]], rt)[0]
        m[i] = np.frombuffer(c[offs[0]:offs[0] + sizes[0]], _NSYS_FLOAT[sizes[0]])[0]
        r[i] = np.frombuffer(c[offs[1]:offs[1] + sizes[1]], _NSYS_FLOAT[sizes[1]])[0]
        t = np.frombuffer(c[offs[2]:offs[2] + 6 * sizes[2]], _NSYS_FLOAT[sizes[2]])
        pos[i], vel[i] = t[:3], t[3:]
    out.update(rest_c=rest, m=m, r=r, pos=pos, vel=vel, chunk=chunk)
    return out


def ref_nsys_copy(src, dst, value_type="ld"):
    """The compiled reference reads src with its .nsys constructor and writes
    it back with to_nsys (system<T,T,T>, T = long double or double)."""
    run_ref("nsys_copy" if value_type == "ld" else "nsys_copy_d", src, dst)
