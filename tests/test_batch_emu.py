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


def build_emu(out, *defines):
    cmd = ["g++", "-std=c++17", "-O3", "-U_FORTIFY_SOURCE", "-D_FORTIFY_SOURCE=0", "-ffp-contract=off",
           "-Wall", "-Wno-unknown-pragmas", *defines, "-I", os.path.join(ROOT, "gregpackagescpp_amd", "csrc"),
           os.path.join(ROOT, "tests", "emu", "batch_emu.cpp"), "-o", out, "-lpthread"]
    subprocess.run(cmd, check=True)
    return out


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    return build_emu(str(tmp_path_factory.mktemp("emu") / "batch_emu"))


@pytest.fixture(scope="module")
def emu_every_step(tmp_path_factory):
    """The same kernel with the exact collision replay after every mixed step."""
    return build_emu(str(tmp_path_factory.mktemp("emu") / "batch_emu_every"), "-DNB_BATCH_PREFILTER=0")


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


def _concat(systems):
    offsets = np.concatenate([[0], np.cumsum([len(s[0]) for s in systems])])
    m, r, e, pos, vel = (np.concatenate([s[k] for s in systems]) for k in range(5))
    return offsets, m, r, e, pos, vel


@pytest.mark.parametrize("sizes", [[300, 520, 9], [800, 64], [1, 2, 65, 128, 129, 257, 3]])
def test_emu_fp64_large_systems_bit_exact(emu, tmp_path, sizes):
    """Systems that give a wave two, three (up to 768 bodies) or four (above
    768, the larger instance) 64-body blocks, and sizes at the edges (1 and
    2 bodies, one past a block), with merges, bit for bit."""
    systems = [rubble(n, 300 + k) for k, n in enumerate(sizes)]
    offsets, m, r, e, pos, vel = _concat(systems)
    st = run_emu(emu, str(tmp_path), offsets, m, r, e, pos, vel, 12, mixed=False, merge=True, chunk=5)
    assert st["err"] == 0
    k = 0
    for s, (ms, rs, es, ps, vs) in enumerate(systems):
        ref = O.kdk(ms, ps, vs, G_SI, 5.0, 12, coll=2, r=rs, e=es, dtype=np.float64)
        c = int(st["counts"][s])
        assert c == ref["n"]
        assert np.array_equal(st["pos"][k:k + c], ref["pos"])
        assert np.array_equal(st["vel"][k:k + c], ref["vel"])
        assert np.array_equal(st["ids"][k:k + c], ref["labels"])
        assert np.array_equal(st["log"][st["log"][:, 0] == s][:, 1:], ref["log"])
        k += c
    assert len(st["log"]) > 0


def test_emu_mixed_prefilter_skips_only_clean_steps(emu, emu_every_step, tmp_path):
    """The FP32 prefilter lets a step skip the exact replay only when no body
    overlaps another: with the replay after every step as well, every state,
    id and merge record is bit-identical.  Dense rubble (merges and bounces
    from the first steps), comet systems (the largest, 513 bodies, above the
    512 of two blocks per wave) and a system of 900 bodies (the larger
    instance)."""
    offsets, m, r, pos, vel, e = capi.ic_comet_ensemble(3, 4096, 8, 512)
    sizes = np.diff(offsets)
    order = np.argsort(-sizes, kind="stable")
    systems = []
    for s in np.concatenate([order[:3], order[1000:1003]]):
        sl = slice(offsets[s], offsets[s + 1])
        systems.append((m[sl], r[sl], e[sl], pos[sl], vel[sl]))
    systems += [rubble(n, 400 + k) for k, n in enumerate([40, 200, 900])]
    offs, m2, r2, e2, p2, v2 = _concat(systems)
    for steps in (60, 400):
        a = run_emu(emu, str(tmp_path), offs, m2, r2, e2, p2, v2, steps, mixed=True, chunk=45)
        b = run_emu(emu_every_step, str(tmp_path), offs, m2, r2, e2, p2, v2, steps, mixed=True, chunk=45)
        assert a["err"] == 0 and b["err"] == 0
        for key in ("counts", "m", "r", "pos", "vel", "ids", "log"):
            assert np.array_equal(a[key], b[key]), key
        assert len(a["log"]) > 0
