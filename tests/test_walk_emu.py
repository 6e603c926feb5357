"""The FP32 Barnes-Hut walks' device source (nb_walk_dev.hpp) built for the
host and run under the wave emulator (tests/emu/walk_emu.cpp): 64 lanes per
wave in lockstep at every intrinsic, guard pages around every buffer.

Bars, on octrees built from the oracle's tree with the walk-record
arithmetic of the GPU build:
  * the depth-first wave walk and the frontier walk accept the same cells:
    identical interaction counts, for several bucket sizes, with far
    outliers (64-body groups whose box spans the domain), with eps = 0 and
    with a ragged last group;
  * their accelerations differ only by FP32 summation order;
  * the depth-first walk reading the grouped layout (as the frontier walk's
    overflow fallback does) gives the same counts and accelerations, bit for
    bit, as reading its own layout;
  * no access leaves a buffer, no intrinsic is reached by part of a wave,
    and every launch ends;
  * a frontier stack too small for the tree reports the overflow.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "emu"))

import make_walk_input as M  # noqa: E402
from gregpackagescpp_amd import capi  # noqa: E402

SRC = os.path.join(HERE, "emu", "walk_emu.cpp")
INC = os.path.join(HERE, "..", "gregpackagescpp_amd", "csrc")


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("emu") / "walk_emu")  # one build per test process
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-U_FORTIFY_SOURCE", "-D_FORTIFY_SOURCE=0",
                           "-ffp-contract=off", "-Wno-unknown-pragmas", "-I", INC, "-o", exe, SRC])
    return exe


def run(emu, tmp_path, m, pos, bucket, eps=0.01, theta=0.5, cap=None):
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    M.write(inp, m, pos, theta, eps, bucket)
    cmd = [emu, inp, out, "--max-rounds", "200000000"]
    if cap:
        cmd += ["--cap", str(cap)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    wave, front, grouped = M.read_output(out, len(m))
    # the depth-first walk reading the grouped layout (the frontier walk's
    # fallback) repeats the one reading its own layout bit for bit
    assert grouped[0] == wave[0] and np.array_equal(grouped[2], wave[2])
    return wave, front


def rel(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


@pytest.mark.parametrize("bucket", [1, 8, 32])
def test_walks_accept_same_cells(emu, tmp_path, bucket):
    m, pos, _ = capi.ic_two_plummer(51, 4096)
    (cw, _, aw), (cf, ov, af) = run(emu, tmp_path, m, pos, bucket)
    assert ov == 0
    assert cw == cf and cw > 4096
    assert rel(af, aw).max() < 1e-4


def test_far_outliers_ragged_and_unsoftened(emu, tmp_path):
    m, pos, _ = capi.ic_plummer(52, 3000)
    rng = np.random.default_rng(52)
    far = rng.choice(3000, 30, replace=False)
    pos[far] = rng.uniform(-500, 500, (30, 3))
    for eps in (0.01, 0.0):
        (cw, _, aw), (cf, ov, af) = run(emu, tmp_path, m, pos, 8, eps=eps)
        assert ov == 0 and cw == cf
        assert np.all(np.isfinite(af))
        assert rel(af, aw).max() < 1e-4


def test_small_stack_reports_overflow(emu, tmp_path):
    m, pos, _ = capi.ic_two_plummer(53, 4096)
    (cw, _, _), (cf, ov, _) = run(emu, tmp_path, m, pos, 1, cap=64)
    assert ov == 1


@pytest.fixture(scope="module")
def emu_small_lds(tmp_path_factory):
    """The same walks with 16 frontier stack entries in LDS, so the stack
    moves to global memory and back many times per group."""
    exe = str(tmp_path_factory.mktemp("emu16") / "walk_emu")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-U_FORTIFY_SOURCE", "-D_FORTIFY_SOURCE=0",
                           "-ffp-contract=off", "-Wno-unknown-pragmas", "-DNB_FRONTIER_LDS_STACK=16", "-I", INC,
                           "-o", exe, SRC])
    return exe


@pytest.mark.parametrize("bucket,theta", [(1, 0.5), (8, 0.3)])
def test_frontier_lds_stack_split_is_invisible(emu, emu_small_lds, tmp_path, bucket, theta):
    """Each pop takes the top of the whole stack (global part below, LDS part
    above), so where the entries live changes no sum: a 16-entry LDS part
    gives bit-identical accelerations and the same counts."""
    m, pos, _ = capi.ic_two_plummer(54, 4096)
    (cw, _, aw), (cf, ov, af) = run(emu, tmp_path, m, pos, bucket, theta=theta)
    (cw2, _, aw2), (cf2, ov2, af2) = run(emu_small_lds, tmp_path, m, pos, bucket, theta=theta)
    assert ov == 0 and ov2 == 0
    assert cf == cw and cf2 == cf and cw2 == cw
    assert np.array_equal(af2, af)
    assert np.array_equal(aw2, aw)
