"""FASTQ ingestion (fame_fastq_*, ReadQueue.parseChunk) against what FAME's
own parseChunk / parseChunkGZ (ReadQueue.cpp:218-373, CHUNKSIZE 700) read from
the same files: tests/golden/fastq/<name>.expect.npz.  The parser runs on the
host, so these tests need no GPU."""
import gzip
import json
import os

import numpy as np
import pytest

from fame_amd import _lib
from fame_amd.readqueue import ReadQueue

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fastq")
META = json.load(open(os.path.join(HERE, "case.json")))
NAMES = sorted(f[:-len(".expect.npz")] for f in os.listdir(HERE) if f.endswith(".expect.npz"))


def _input(name, tmp_path):
    """Path of the case's input as the parser reads it (gz cases stay gzipped)."""
    src = os.path.join(HERE, name + ".fq.gz")
    if name in META["gz"]:
        return src, True
    dst = tmp_path / (name + ".fq")
    dst.write_bytes(gzip.open(src).read())
    return str(dst), False


def _drain(path, gz, k):
    """All records, read k at a time, plus each call's (n, n >= k)."""
    rd = _lib.FastqReader(path, gz)
    seqs, lens, calls = [], [], []
    while True:
        s, o, ln = rd.next(k)
        assert len(o) == len(ln)
        if len(ln):
            assert o[0] == 0 and np.array_equal(o[1:], np.cumsum(ln[:-1].astype(np.uint64)))
            seqs.append(s)
            lens.append(ln)
        calls.append((len(ln), int(len(ln) >= k)))
        if len(ln) < k:
            break
    rd.close()
    cat = lambda xs, dt: np.concatenate(xs) if xs else np.zeros(0, dt)  # noqa: E731
    return cat(seqs, np.uint8), cat(lens, np.uint32), calls


@pytest.mark.parametrize("name", NAMES)
def test_parse_matches_reference(name, tmp_path):
    exp = np.load(os.path.join(HERE, name + ".expect.npz"))
    path, gz = _input(name, tmp_path)
    seqs, lens, calls = _drain(path, gz, META["chunksize"])
    assert np.array_equal(lens, exp["lens"])
    assert np.array_equal(seqs, exp["seqs"])
    # per call: procReads and whether parseChunk reported a full chunk
    assert calls == [tuple(int(x) for x in c) for c in exp["calls"]]


@pytest.mark.parametrize("k", [1, 3, 699, 1 << 20])
@pytest.mark.parametrize("name", ["lf", "crlf", "trunc_plus", "gz_notrail"])
def test_parse_independent_of_chunk(name, k, tmp_path):
    exp = np.load(os.path.join(HERE, name + ".expect.npz"))
    path, gz = _input(name, tmp_path)
    seqs, lens, _ = _drain(path, gz, k)
    assert np.array_equal(lens, exp["lens"]) and np.array_equal(seqs, exp["seqs"])


def test_readqueue_parse_chunk(tmp_path):
    """ReadQueue.parseChunk returns (more, procReads) per call as the reference does."""
    exp = np.load(os.path.join(HERE, "gz_lf.expect.npz"))
    rq = ReadQueue.__new__(ReadQueue)
    rq._f = _lib.FastqReader(os.path.join(HERE, "gz_lf.fq.gz"), True)
    rq._f2 = None
    rq.is_paired = False
    rq.chunk_size = META["chunksize"]
    got = []
    while True:
        more, n = rq.parseChunk()
        got.append([n, int(more)])
        if not more:
            break
    rq._f.close()
    assert got == exp["calls"].tolist()


def test_large_records_across_refills(tmp_path):
    """Records spanning the parser's input refills (64 MiB blocks): 600k reads
    of 100-300 bases, read in chunks of 300000 (the reference's CHUNKSIZE)."""
    rng = np.random.default_rng(5)
    n = 600_000
    lens = rng.integers(100, 301, n).astype(np.uint32)
    alpha = np.frombuffer(b"ACGT", np.uint8)
    flat = alpha[rng.integers(0, 4, int(lens.sum()))]
    offs = np.zeros(n, np.uint64)
    offs[1:] = np.cumsum(lens[:-1].astype(np.uint64))
    p = tmp_path / "big.fq"
    with open(p, "wb") as f:
        for i in range(n):
            s = flat[int(offs[i]):int(offs[i]) + int(lens[i])].tobytes()
            f.write(b"@read_%d some description\n%s\n+\n%s\n" % (i, s, b"F" * len(s)))
    seqs, got_lens, calls = _drain(str(p), False, 300_000)
    assert calls == [(300_000, 1), (300_000, 1), (0, 0)]
    assert np.array_equal(got_lens, lens) and np.array_equal(seqs, flat)


def test_missing_file_raises(tmp_path):
    with pytest.raises(_lib.FameError):
        _lib.FastqReader(str(tmp_path / "absent.fq"))


def test_readqueue_pairs_count_mismatch(tmp_path):
    """Paired files with different record counts: the reference prints a
    message and exits (ReadQueue.cpp:280-286); the mirror raises."""
    a, b = tmp_path / "a.fq", tmp_path / "b.fq"
    a.write_bytes(b"@a\nACGT\n+\nIIII\n" * 3)
    b.write_bytes(b"@b\nACGT\n+\nIIII\n" * 2)
    rq = ReadQueue.__new__(ReadQueue)
    rq._f, rq._f2 = _lib.FastqReader(str(a)), _lib.FastqReader(str(b))
    rq.is_paired = True
    rq.chunk_size = 700
    with pytest.raises(_lib.FameError):
        rq.parseChunk()
    rq._f.close()
    rq._f2.close()

