"""The native index builder against indexes built by FAME itself.

The golden fixtures carry reference-built indexes; their chromosome
sequences and names are read back, rebuilt with fame_amd.index and every
section of the file is compared (filteredKmers and chrMap as set / map: the
reference writes them in hash-container order)."""
import gzip
import json
import os
import shutil
import struct

import numpy as np
import pytest

from fame_amd import index as fidx

from conftest import CASES, GOLDEN


def parse(path):
    b = open(path, "rb").read()
    o = 0

    def u64():
        nonlocal o
        v = struct.unpack_from("<Q", b, o)[0]
        o += 8
        return v

    def u32():
        nonlocal o
        v = struct.unpack_from("<I", b, o)[0]
        o += 4
        return v

    d = {"hdr": (u64(), u32(), u32(), u32(), u64(), u32())}
    for key in ("cpg", "cpgs"):
        n = u64()
        a = np.frombuffer(b, np.uint8, 8 * n, o).reshape(-1, 8)
        o += 8 * n
        d[key] = (a[:, 0].copy(), a[:, 4:8].copy().view(np.uint32).ravel())
    nc = u64()
    d["chroms"] = []
    for _ in range(nc):
        L = u64()
        d["chroms"].append(b[o:o + L])
        o += L
    tl = u64()
    d["tab"] = np.frombuffer(b, np.uint64, tl, o)
    o += 8 * tl
    nk = u64()
    k = np.frombuffer(b, np.uint32, 2 * nk, o).reshape(-1, 2)
    o += 8 * nk
    d["kmeta"], d["ktmask"] = k[:, 0], k[:, 1]
    for _ in range(2):  # metaCpGs, metaStartCpGs
        n = u64()
        o += 8 * n
    nw = u64()
    w = np.frombuffer(b, np.uint32, 4 * nw, o).reshape(-1, 4).copy()
    o += 16 * nw
    w[:, 1] &= 0xFF
    d["win"] = w
    nf = u64()
    d["filt"] = set(np.frombuffer(b, np.uint64, nf, o).tolist())
    o += 8 * nf
    m = {}
    nm = u64()
    for _ in range(nm):
        i = b[o]
        o += 1
        L = u64()
        m[i] = b[o:o + L].decode()
        o += L
    d["chrmap"] = m
    s = open(path + "_strands", "rb").read()
    ns = struct.unpack_from("<Q", s, 0)[0]
    d["strand"] = np.unpackbits(np.frombuffer(s, np.uint8, (ns + 7) // 8, 8), bitorder="little")[:ns]
    return d


@pytest.mark.parametrize("name", [c for c in CASES if not os.path.isfile(
    os.path.join(GOLDEN, c, "case.json")) or '"share"' not in open(os.path.join(GOLDEN, c, "case.json")).read()])
def test_builder_reproduces_reference_index(name, tmp_path):
    src = os.path.join(GOLDEN, name)
    if not os.path.exists(os.path.join(src, "index.bin.gz")):
        pytest.skip("case shares another case's index")
    ref = os.path.join(str(tmp_path), "ref.idx")
    for f, dst in (("index.bin", ref), ("index.bin_strands", ref + "_strands")):
        with gzip.open(os.path.join(src, f + ".gz"), "rb") as a, open(dst, "wb") as b:
            shutil.copyfileobj(a, b)
    a = parse(ref)
    htab, readlen, winlen, kmerlen, cutoff, seedbits = a["hdr"]
    chroms = [np.frombuffer(c, np.uint8) for c in a["chroms"]]
    names = [a["chrmap"][i] for i in range(len(chroms))]
    mine = os.path.join(str(tmp_path), "mine.idx")
    noloss = bool(json.load(open(os.path.join(src, "case.json"))).get("noloss"))
    fidx.build_index(chroms, mine, names=names, readlen=readlen, htab=htab, kmercutoff=cutoff, noloss=noloss)
    assert_same_index(a, parse(mine), noloss)


def assert_same_index(a, b, noloss):
    """Two parsed indexes (parse()) equal section by section: a is FAME's."""
    htab, readlen, winlen, kmerlen, cutoff, seedbits = a["hdr"]
    for key in ("hdr", "chrmap", "filt", "chroms", "win"):
        if key == "win":
            assert np.array_equal(a[key], b[key]), key
        else:
            assert a[key] == b[key], key
    if not noloss and cutoff < 1500:
        assert len(a["filt"]) > 0, "fixture meant to exercise the blacklist has none"
    for key in ("cpg", "cpgs"):
        assert np.array_equal(a[key][0], b[key][0]) and np.array_equal(a[key][1], b[key][1]), key
    # Reverse-strand entries of a chromosome's last, shorter window read
    # past the chromosome in the reference (getTMask RefGenome.h:84-101 and
    # reproduceTMask / reproduceKmerSeq RefGenome.cpp:2028-2190 take 32
    # bytes at an offset computed for a full window): heap bytes, undefined.
    # Their tmask decides filterRedundancyInHashTable's choice among them
    # (:2191-2264), so their number per bucket may differ too.  Every other
    # entry must be identical, in order, bucket by bucket.
    clen = np.array([len(c) for c in a["chroms"]], np.int64)

    def defined(d):
        w = d["win"][d["kmeta"] & 0x7FFFFFFF]
        short = w[:, 0].astype(np.int64) + winlen > clen[w[:, 1]]
        return ~(short & (d["strand"] == 0))

    da, db = defined(a), defined(b)
    # (at most one window per chromosome, about WINLEN / SKIPMOD entries)
    lim = len(clen) * (winlen // 2 + 32)
    assert (~da).sum() <= lim and (~db).sum() <= lim
    bka = np.repeat(np.arange(htab), np.diff(a["tab"].astype(np.int64)))
    bkb = np.repeat(np.arange(htab), np.diff(b["tab"].astype(np.int64)))
    assert np.array_equal(bka[da], bkb[db]), "defined entries fall into other buckets"
    for key in ("kmeta", "strand", "ktmask"):
        assert np.array_equal(a[key][da], b[key][db]), key
