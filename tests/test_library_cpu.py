"""CPU tests of the library boundary and host-side code (no GPU needed):
the C ABI exports every symbol the header declares, the ctypes binding covers
them, the seeded IC generators behave, and the merged-radius cube root is
faithful."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from gregpackagescpp_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gregnbody.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nb_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(L, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (nb_\w+)", out))
    assert set(syms) <= exported


def test_ctypes_binding_covers_header():
    assert set(declared_symbols()) == set(capi.SIGNATURES)


def test_facade_header_compiles():
    src = os.path.join(ROOT, "gregpackagescpp_amd", "examples", "facade_demo.cpp")
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-I", os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_ic_determinism_and_com_frame():
    a = capi.ic_uniform_ball(5, 1000, 1e20, 1e22, 1e9, 1e3)
    b = capi.ic_uniform_ball(5, 1000, 1e20, 1e22, 1e9, 1e3)
    c = capi.ic_uniform_ball(6, 1000, 1e20, 1e22, 1e9, 1e3)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[1], c[1])
    m, pos, vel = a
    assert np.all((m >= 1e20) & (m <= 1e22))
    assert np.abs((m[:, None] * pos).sum(0) / m.sum()).max() < 1e-6 * 1e9
    assert np.abs((m[:, None] * vel).sum(0) / m.sum()).max() < 1e-9 * 1e3
    # the ball of radius R is shifted to its centre of mass (a few % of R at N = 1000)
    assert np.linalg.norm(pos, axis=1).max() <= 1.1e9
    assert abs(vel.std() - 1e3) < 50


def test_plummer_virial_and_profile():
    m, pos, vel = capi.ic_plummer(1, 20000)
    assert np.allclose(m, 1 / 20000)
    r = np.linalg.norm(pos, axis=1)
    # half-mass radius of a Plummer sphere: a / sqrt(2^(2/3) - 1) = 1.305 a
    assert abs(np.median(r) - 1.305) < 0.05
    ke = 0.5 * (m * (vel ** 2).sum(1)).sum()
    pe = float(O.energy(m[:4000] * 5, pos[:4000], vel[:4000], 1.0, 0.0, dtype=np.float64)[1])
    assert abs(ke / abs(pe) - 0.5) < 0.06  # virial equilibrium (PE from a subsample)


def test_two_plummer_setup():
    m, pos, vel = capi.ic_two_plummer(2, 4096)
    assert abs(m.sum() - 1.0) < 1e-12
    left = pos[:, 0] < 0
    assert abs(np.median(pos[left, 0]) + 10) < 0.5 and abs(np.median(pos[~left, 0]) - 10) < 0.5
    assert abs(vel[:2048, 0].mean() - vel[2048:, 0].mean() - 0.5) < 0.05


def test_comet_ensemble_shape():
    offsets, m, r, pos, vel, e = capi.ic_comet_ensemble(3, 64, 8, 512)
    counts = np.diff(offsets)
    assert np.all((counts >= 9) & (counts <= 513))
    central = offsets[:-1]
    assert np.all(m[central] == 1.9e27) and np.all(r[central] == 6.9911e7)
    frag = np.ones(len(m), bool)
    frag[central] = False
    assert np.all((m[frag] >= 1e9) & (m[frag] <= 1e13))
    assert np.allclose(r[frag], np.cbrt(3 * m[frag] / (4 * np.pi * 500)), rtol=1e-14)
    # fragments start outside the central body, on bound-to-parabolic speeds
    d = np.linalg.norm(pos[frag], axis=1)
    assert np.all(d > 6.9911e7)
    v2 = (vel[frag] ** 2).sum(1)
    esc2 = 2 * 6.6743e-11 * 1.9e27 / d
    assert np.all(np.abs(v2 / esc2 - 1) < 1e-3)


def test_newton_cbrt_faithful():
    """The double restatement's merged radius (fixed Newton cube root shared
    with the GPU) is within 1 ulp of the true cube root."""
    rng = np.random.default_rng(0)
    worst = 0.0
    for _ in range(500):
        r = rng.uniform(1, 2000, 2)
        m = np.array([1e13, 1e12])
        pos = np.array([[0, 0, 0], [0.5 * (r[0] + r[1]), 0, 0]])
        vel = np.array([[1.0, 0, 0], [-1.0, 0, 0]])
        res = O.s_coll(m, r, np.ones(2), pos, vel, np.zeros((2, 3)), np.arange(2), 2, merge=True)
        assert res["n"] == 1
        want = np.cbrt(np.longdouble(r[0]) ** 3 + np.longdouble(r[1]) ** 3)
        worst = max(worst, abs(float(res["r"][0]) - float(want)) / np.spacing(float(want)))
    assert worst <= 1.0


def test_bench_rejects_multi_gpu():
    r = subprocess.run(["python", os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one MI355X" in r.stderr
