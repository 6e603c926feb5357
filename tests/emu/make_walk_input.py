"""Input file of tests/emu/walk_emu.cpp: the sorted FP32 bodies and both walk
record layouts of a system's octree, built on the host from the oracle's tree
(whose topology the GPU build reproduces; tests/test_bh_gpu.py) with the
record arithmetic of k_walk_prep and k_sorted_pos (G = 1)."""
import struct

import numpy as np

import oracle as O

WALKNODE = np.dtype([("q", "<f4", 4), ("r", "<i4", 4)])


def records(m, pos, theta):
    t = O.bh_tree(m, pos)
    body, parent, mass, com, side = t["body"], t["parent"], t["mass"], t["com"], t["side"]
    total = len(body)
    size = np.ones(total, np.int64)
    leaf = (body >= 0).astype(np.int64)
    nleaf = leaf.copy()
    nchild = np.zeros(total, np.int64)
    rank = np.zeros(total, np.int64)
    for x in range(1, total):
        p = parent[x]
        rank[x] = nchild[p]
        nchild[p] += 1
    for x in range(total - 1, 0, -1):
        size[parent[x]] += size[x]
        nleaf[parent[x]] += nleaf[x]
    skip = np.arange(total) + size
    first = np.concatenate([[0], np.cumsum(leaf)[:-1]])
    count = nleaf
    cbase = np.concatenate([[0], np.cumsum(nchild)[:-1]])
    r2o = np.where(body >= 0, np.float32(-1.0), ((side / theta) ** 2).astype(np.float32)).astype(np.float32)
    q = np.stack([com[:, 0], com[:, 1], com[:, 2], mass], axis=1).astype(np.float32)
    wnode = np.zeros(total, WALKNODE)
    wnode["q"] = q
    wnode["r"][:, 0] = skip
    wnode["r"][:, 1] = first
    wnode["r"][:, 2] = count
    wnode["r"][:, 3] = r2o.view(np.int32)
    at = np.zeros(total, np.int64)
    at[1:] = 1 + cbase[parent[1:]] + rank[1:]
    wgrp = np.zeros(total, WALKNODE)
    wgrp["q"][at] = q
    wgrp["r"][at, 0] = np.where(nchild > 0, 1 + cbase, 0)
    wgrp["r"][at, 1] = first
    wgrp["r"][at, 2] = (np.minimum(count, (1 << 27) - 1) << 4) | nchild
    wgrp["r"][at, 3] = r2o.view(np.int32)
    order = body[body >= 0]
    pos32 = np.concatenate([pos[order], m[order, None]], axis=1).astype(np.float32)
    return pos32, wnode, wgrp


def write(path, m, pos, theta, eps, bucket):
    pos32, wnode, wgrp = records(m, pos, theta)
    with open(path, "wb") as f:
        f.write(struct.pack("<qqqff", len(m), len(wnode), bucket, np.float32(eps * eps), 0.0))
        f.write(pos32.tobytes())
        f.write(wnode.tobytes())
        f.write(wgrp.tobytes())
    return len(wnode)


def read_output(path, n, runs=3):
    out = []
    with open(path, "rb") as f:
        for _ in range(runs):
            cnt, flag = struct.unpack("<qq", f.read(16))
            acc = np.frombuffer(f.read(24 * n), "<f8").reshape(3, n).T.copy()
            out.append((cnt, flag, acc))
    return out
