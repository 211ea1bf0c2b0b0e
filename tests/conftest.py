import ctypes as C
import gzip
import json
import os
import shutil
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def _cases(mode):
    out = []
    for d in sorted(os.listdir(GOLDEN)):
        f = os.path.join(GOLDEN, d, "case.json")
        if os.path.isfile(f) and json.load(open(f)).get("mode", "se") == mode:
            out.append(d)
    return out


CASES = _cases("se")
PAIR_CASES = _cases("pe")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


class Case:
    def __init__(self, name, tmpdir):
        d = os.path.join(GOLDEN, name)
        self.name = name
        self.meta = json.load(open(os.path.join(d, "case.json")))
        idir = os.path.join(GOLDEN, self.meta["index_dir"])
        self.index = os.path.join(tmpdir, "index.bin")
        for f in ("index.bin", "index.bin_strands"):
            with gzip.open(os.path.join(idir, f + ".gz"), "rb") as a, open(os.path.join(tmpdir, f), "wb") as b:
                shutil.copyfileobj(a, b)
        r = np.load(os.path.join(d, "reads.npz"))
        self.seqs, self.offs, self.lens = r["seqs"], r["offs"], r["lens"]
        g = np.load(os.path.join(d, "golden.npz"))
        self.valid, self.mat, self.meth = g["valid"], g["mat"], g["meth"]
        with gzip.open(os.path.join(d, "cpg.tsv.gz"), "rb") as f:
            self.tsv = f.read()
        self.chunk = self.meta["chunk"]
        self.unord = self.meta["unord"]
        self.ref = self.meta["reference"]


@pytest.fixture(params=CASES)
def golden_case(request, tmp_path):
    return Case(request.param, str(tmp_path))


def _unpack_index(meta, tmpdir):
    idir = os.path.join(GOLDEN, meta["index_dir"])
    path = os.path.join(tmpdir, "index.bin")
    for f in ("index.bin", "index.bin_strands"):
        with gzip.open(os.path.join(idir, f + ".gz"), "rb") as a, open(os.path.join(tmpdir, f), "wb") as b:
            shutil.copyfileobj(a, b)
    return path


class PairCase:
    """A paired-end fixture: two mates per pair, five counters."""

    def __init__(self, name, tmpdir):
        d = os.path.join(GOLDEN, name)
        self.name = name
        self.meta = json.load(open(os.path.join(d, "case.json")))
        self.index = _unpack_index(self.meta, tmpdir)
        r = np.load(os.path.join(d, "reads.npz"))
        self.r1 = (r["seqs1"], r["offs1"], r["lens1"])
        self.r2 = (r["seqs2"], r["offs2"], r["lens2"])
        g = np.load(os.path.join(d, "golden.npz"))
        self.valid1, self.valid2 = g["valid1"], g["valid2"]
        self.mat1, self.mat2, self.meth = g["mat1"], g["mat2"], g["meth"]
        with gzip.open(os.path.join(d, "cpg.tsv.gz"), "rb") as f:
            self.tsv = f.read()
        self.chunk = self.meta["chunk"]
        self.unord = self.meta["unord"]
        self.minpd, self.maxpd = self.meta["minpd"], self.meta["maxpd"]
        self.ref = self.meta["reference"]
        self.counters = [self.ref[k] for k in ("succ", "nonunique", "unsucc", "succ_paired",
                                               "too_short")]


@pytest.fixture(params=PAIR_CASES)
def golden_pair(request, tmp_path):
    return PairCase(request.param, str(tmp_path))


SC_CASES = _cases("sc")


class ScCase:
    """A single-cell fixture: per-cell FASTQs, the cell list, FAME's matrix."""

    def __init__(self, name, tmpdir):
        self.dir = os.path.join(GOLDEN, name)
        self.name = name
        self.meta = json.load(open(os.path.join(self.dir, "case.json")))
        self.index = _unpack_index(self.meta, tmpdir)
        self.expected = gzip.open(os.path.join(self.dir, "sc_out.txt.gz")).read()
        self.cells = []  # (id, [fastq paths])
        for line in open(os.path.join(self.dir, "cells.txt")).read().split("\n"):
            parts = line.split(" ")
            if len(parts) > 1:
                self.cells.append((parts[0], [os.path.join(self.dir, p) for p in parts[1:]]))


@pytest.fixture(params=SC_CASES)
def golden_sc(request, tmp_path):
    return ScCase(request.param, str(tmp_path))


def read_fastq_gz(path):
    """(seqs, offs, lens) of a gzipped FASTQ, every fourth line from the second."""
    lines = gzip.open(path).read().split(b"\n")
    seqs = [lines[i] for i in range(1, len(lines), 4)]
    lens = np.array([len(x) for x in seqs], np.uint32)
    offs = np.zeros(len(seqs), np.uint64)
    if len(seqs) > 1:
        offs[1:] = np.cumsum(lens[:-1].astype(np.uint64))
    return np.frombuffer(b"".join(seqs), np.uint8), offs, lens


H2D, D2H = 1, 2


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libfamehip.so links"""

    def __init__(self):
        from fame_amd import _lib

        _lib.load()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.h = h
        self.live = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(int(nbytes), 1)) == 0
        assert self.h.hipMemset(p, 0, max(int(nbytes), 1)) == 0
        self.live.append(p.value)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, n, dtype):
        out = np.zeros(n, dtype)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


