"""Test helpers: the hierarchy file reader (format "SC23AMG1" written by
sc23ad_hier_export and by the reference harness), hashing, and seeded
systems."""
from __future__ import annotations

import hashlib
import os
import struct
import tempfile
from dataclasses import dataclass

import numpy as np

import sc23ad as S


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def file_sha(path: str) -> str:
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


@dataclass
class Mat:
    n: int
    m: int
    bit: int
    dbit: int
    rptr: np.ndarray
    cind: np.ndarray
    val: np.ndarray

    def dense(self) -> np.ndarray:
        d = np.zeros((self.n, self.m), dtype=np.float64)
        for i in range(self.n):
            for p in range(self.rptr[i], self.rptr[i + 1]):
                d[i, self.cind[p]] += self.val[p]
        return d


@dataclass
class HierFile:
    precision: int
    depth: int
    smoother: int
    iter: int
    b_num: int
    max_levels: int
    coarse_n: int
    n_coarse: int
    omega: float
    theta: float
    levels: list  # per level: dict A, LD, U, R, P
    lc: Mat
    uc: Mat


_VAL = {0: np.float64, 1: np.float32, 2: np.int32}


def read_hier(path: str) -> HierFile:
    with open(path, "rb") as f:
        buf = f.read()
    assert buf[:8] == b"SC23AMG1"
    off = 8
    hdr = struct.unpack_from("<8i2d", buf, off)
    off += struct.calcsize("<8i2d")
    prec = hdr[0]
    vt = _VAL[prec]

    def mat():
        nonlocal off
        n, m, nnz, bit, dbit = struct.unpack_from("<5i", buf, off)
        off += 20
        rptr = np.frombuffer(buf, np.int32, n + 1, off).copy()
        off += 4 * (n + 1)
        cind = np.frombuffer(buf, np.int32, nnz, off).copy()
        off += 4 * nnz
        val = np.frombuffer(buf, vt, nnz, off).copy()
        off += np.dtype(vt).itemsize * nnz
        return Mat(n, m, bit, dbit, rptr, cind, val)

    levels = []
    for _ in range(hdr[1]):
        levels.append({k: mat() for k in ("A", "LD", "U", "R", "P")})
    lc, uc = mat(), mat()
    assert off == len(buf)
    return HierFile(prec, hdr[1], hdr[2], hdr[3], hdr[4], hdr[5], hdr[6], hdr[7],
                    hdr[8], hdr[9], levels, lc, uc)


def export(h: "S.Hierarchy") -> str:
    fd, path = tempfile.mkstemp(suffix=".bin")
    os.close(fd)
    h.export(path)
    return path


def system(case: dict) -> "S.System":
    return S.make_system(case["kind"], tuple(case["params"]), case["seed"])
