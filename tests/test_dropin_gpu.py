"""examples/drop_in_driver.cpp, one reference-style driver built unchanged
against the reference (tests/golden fixtures, made with oracle/_ref) and
against this engine's gtd::system facade (run here on the GPU).  It uses
long double M, R and T and the reference's evolve callbacks.

Bars (FP64 engine against the reference's long double):
  * cor / cor_pred: collisions = 3, for_all_bods rewriting every body's
    restitution each step as random_comet does, for_each_bod recording every
    body after each step's collision pass (every 25th step compared): the
    same bodies and step numbers, masses and radii exact, positions and
    velocities within 1e-12 relative (max-norm; only the arithmetic type
    differs, measured 1.5e-15 and 2.6e-15), restitution within 1e-14;
  * merge: the bodies surviving construction with mergeOverlappingBodies,
    the same count and values within 1e-13 relative;
  * check: the plain overlap check throws exactly when the reference does.
"""
import glob
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "gregpackagescpp_amd", "drop_in_driver")


def run(mode, f, steps, dt, td):
    inp, out = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
    O.write_ref_input(inp, f["m"], f["pos"], f["vel"], r=f["r"], e=f["e"], steps=steps, dt=dt)
    r = subprocess.run([EXE, mode, inp, out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.fromfile(out, dtype=np.float64)


@pytest.mark.parametrize("mode", ["cor", "cor_pred"])
def test_callback_driver_matches_reference(mode):
    assert os.path.exists(EXE), "build the package first (make -C gregpackagescpp_amd)"
    f = np.load(os.path.join(GOLD, f"ref_dropin_{mode}_seed1.npz"))
    print(f"case drop-in {mode}", flush=True)
    with tempfile.TemporaryDirectory() as td:
        rec = run(mode, f, int(f["steps"]), float(f["dt"]), td).reshape(-1, 11)
    rec = rec[rec[:, 0] % 25 == 0]
    ref = f["rec"]
    assert rec.shape == ref.shape
    assert np.array_equal(rec[:, :4], ref[:, :4])  # step, slot, mass, radius
    px = np.abs(rec[:, 4:7] - ref[:, 4:7]).max() / np.abs(ref[:, 4:7]).max()
    pv = np.abs(rec[:, 7:10] - ref[:, 7:10]).max() / np.abs(ref[:, 7:10]).max()
    pe = np.abs(rec[:, 10] - ref[:, 10]).max()
    print(f"positions {px:.2e} velocities {pv:.2e} restitution {pe:.2e}")
    assert px < 1e-12 and pv < 1e-12 and pe < 1e-14
    assert len(np.unique(ref[:, 10])) > 10  # the callback rewrote the coefficients


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ref_dropin_merge_*.npz"))),
                         ids=os.path.basename)
def test_merge_driver_matches_reference(path):
    f = np.load(path)
    print(f"case drop-in merge {os.path.basename(path)}", flush=True)
    with tempfile.TemporaryDirectory() as td:
        res = run("merge", f, 1, 1.0, td)
        k = int(res[0])
        bodies = res[1:].reshape(k, 9)
        ref = f["bodies"]
        assert bodies.shape == ref.shape
        scale = np.abs(ref).max(axis=0) + 1e-300
        assert (np.abs(bodies - ref) <= 1e-13 * scale).all()
        g = np.load(path.replace("_merge_", "_check_"))
        flag = run("check", f, 1, 1.0, td)[0]
        assert flag == g["throws"]
