"""The batched ensemble kernel and its collision pass, run on the CPU.

tests/emu/batch_emu.cpp compiles the device source (nb_coll_dev.hpp) for the
host and runs each workgroup's 256 threads as cooperatively scheduled
contexts, with every global buffer against guard pages and a check that all
threads of a workgroup meet at the same barrier.  The host-side driver is
nb_batch_evolve's: a zero-dt force-only launch, then launches of `chunk`
steps.

Bars:
  * FP64 path (reference order, contraction off): every system bit-identical
    to the double oracle (states, labels, merge records) -- the same bar as
    the GPU test, here without a GPU;
  * mixed path (FP32 pair forces, prefilter, skipped replays) on comet-string
    systems of up to 513 bodies, whose passes span several 256-thread chunks:
    no guard-page access and no divergent barrier with the buffers placed
    against either edge, and a consistent result (each system's body count
    equals its initial count minus its merges; labels unique).  The mixed
    trajectories are not compared with the GPU's: rsqrt rounds differently
    here (1/sqrtf).
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle as O
from gregpackagescpp_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G_SI = float(np.longdouble(66743) / np.longdouble(10 ** 15))


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("emu") / "batch_emu")
    cmd = ["g++", "-std=c++17", "-O3", "-U_FORTIFY_SOURCE", "-D_FORTIFY_SOURCE=0", "-ffp-contract=off",
           "-Wall", "-Wno-unknown-pragmas", "-I", os.path.join(ROOT, "gregpackagescpp_amd", "csrc"),
           os.path.join(ROOT, "tests", "emu", "batch_emu.cpp"), "-o", out, "-lpthread"]
    subprocess.run(cmd, check=True)
    return out


def run_emu(exe, tmp, offsets, m, r, e, pos, vel, steps, *, mixed, merge=True, chunk=1000, guard="end",
            dt=5.0, min_mom=1e10):
    nsys = len(offsets) - 1
    inp, outp = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
    with open(inp, "wb") as f:
        f.write(struct.pack("<qqq", nsys, steps, chunk))
        f.write(struct.pack("<dddd", G_SI, dt, 0.0, min_mom))
        f.write(struct.pack("<qqq", 1, int(merge), int(mixed)))
        np.asarray(offsets, np.int64).tofile(f)
        for a in (m, r, e, pos, vel):
            np.ascontiguousarray(a, np.float64).reshape(-1).tofile(f)
    res = subprocess.run([exe, inp, outp, "--guard", guard, "--threads", "8"], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    with open(outp, "rb") as f:
        err = struct.unpack("<q", f.read(8))[0]
        counts = np.frombuffer(f.read(8 * nsys), np.int64).copy()
        nb = int(counts.sum())
        rec = np.frombuffer(f.read(72 * nb), np.uint8).reshape(nb, 72)
        vals = rec[:, :64].copy().view(np.float64)
        ids = rec[:, 64:].copy().view(np.int64).ravel()
        nlog = struct.unpack("<q", f.read(8))[0]
        log = np.frombuffer(f.read(40 * nlog), np.int64).reshape(nlog, 5).copy()
    return dict(err=err, counts=counts, m=vals[:, 0], r=vals[:, 1], pos=vals[:, 2:5], vel=vals[:, 5:8], ids=ids,
                log=log)


def rubble(n, seed):
    rng = np.random.default_rng(seed)
    m = 10 ** rng.uniform(11, 13, n)
    r = (3 * m / (4 * np.pi * 500)) ** (1 / 3)
    pos = rng.uniform(-6e3, 6e3, (n, 3))
    vel = rng.normal(0, 0.5, (n, 3))
    e = rng.uniform(0.1, 0.9, n)
    return m, r, e, pos, vel


@pytest.mark.parametrize("merge", [True, False])
def test_emu_fp64_bit_exact_vs_oracle(emu, tmp_path, merge):
    systems = [rubble(n, 100 + k) for k, n in enumerate([8, 13, 24, 40, 24, 64, 9, 31])]
    offsets = np.concatenate([[0], np.cumsum([len(s[0]) for s in systems])])
    m, r, e, pos, vel = (np.concatenate([s[k] for s in systems]) for k in range(5))
    # chunk 150: state carried across launches as on the GPU
    st = run_emu(emu, str(tmp_path), offsets, m, r, e, pos, vel, 400, mixed=False, merge=merge, chunk=150)
    assert st["err"] == 0
    k = 0
    for s, (ms, rs, es, ps, vs) in enumerate(systems):
        ref = O.kdk(ms, ps, vs, G_SI, 5.0, 400, coll=2 if merge else 1, r=rs, e=es, dtype=np.float64)
        c = int(st["counts"][s])
        assert c == ref["n"]
        assert np.array_equal(st["pos"][k:k + c], ref["pos"])
        assert np.array_equal(st["vel"][k:k + c], ref["vel"])
        assert np.array_equal(st["m"][k:k + c], ref["m"]) and np.array_equal(st["r"][k:k + c], ref["r"])
        assert np.array_equal(st["ids"][k:k + c], ref["labels"])
        assert np.array_equal(st["log"][st["log"][:, 0] == s][:, 1:], ref["log"])
        k += c
    if merge:
        assert len(st["log"]) > 0


@pytest.mark.parametrize("guard", ["end", "start"])
def test_emu_mixed_comet_systems_clean(emu, tmp_path, guard):
    offsets, m, r, pos, vel, e = capi.ic_comet_ensemble(3, 4096, 8, 512)
    sizes = np.diff(offsets)
    order = np.argsort(-sizes, kind="stable")
    keep = np.sort(np.concatenate([order[:6], order[-6:], order[2000:2004]]))  # largest, smallest, middle
    idx = np.concatenate([np.arange(offsets[s], offsets[s + 1]) for s in keep])
    offs = np.concatenate([[0], np.cumsum(sizes[keep])])
    st = run_emu(emu, str(tmp_path), offs, m[idx], r[idx], e[idx], pos[idx], vel[idx], 300, mixed=True,
                 guard=guard)
    assert st["err"] == 0
    assert sizes[keep].max() > 2 * 256  # passes span three 256-thread chunks
    merges = np.bincount(st["log"][:, 0], minlength=len(keep)) if len(st["log"]) else np.zeros(len(keep), int)
    assert np.array_equal(st["counts"], sizes[keep] - merges)
    k = 0
    for c in st["counts"]:
        assert len(np.unique(st["ids"][k:k + c])) == c
        k += c
