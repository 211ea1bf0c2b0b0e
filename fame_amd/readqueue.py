"""Python mirror of FAME's matching boundary (ReadQueue / RefGenome,
ReadQueue.h:44-129, RefGenome.h:55) over the C ABI of libfamehip.so.

Names, argument meaning and counter conventions follow the reference:
counters are accumulated into the caller's list, the first chunk of a
stranded run decides the orientation of read 1, and the report is written to
``<filename>_cpg.tsv``.
"""
import gzip

import numpy as np

from . import _lib

CHUNKSIZE = 300000  # CONST.h:69


class RefGenome:
    """RefGenome(std::string filePath) RefGenome.cpp:77-80: the loaded index.
    The device copy is created by the ReadQueue that borrows it."""

    def __init__(self, file_path):
        self.path = file_path


class ReadQueue:
    """Single-end ReadQueue (ReadQueue.cpp:25-78)."""

    def __init__(self, file_path, ref, is_gz=False, bs_flag=False, chunk_size=CHUNKSIZE):
        self.ctx = _lib.Context(ref.path, both_strands=bs_flag)
        self.bs_flag = bs_flag
        self.chunk_size = chunk_size
        opener = gzip.open if is_gz else open
        self._f = opener(file_path, "rb") if file_path else None
        self.buf = None  # (seqs, offs, lens) of the parsed chunk
        self.results = None

    # parseChunk ReadQueue.cpp:218-296
    def parseChunk(self):
        seqs = []
        n = 0
        while n < self.chunk_size:
            head = self._f.readline()
            if not head:
                break
            seq = self._f.readline().rstrip(b"\n")
            self._f.readline()
            self._f.readline()
            seqs.append(seq)
            n += 1
        lens = np.array([len(s) for s in seqs], np.uint32)
        offs = np.zeros(len(seqs), np.uint64)
        if len(seqs) > 1:
            offs[1:] = np.cumsum(lens.astype(np.uint64))[:-1]
        self.buf = (np.frombuffer(b"".join(seqs), np.uint8), offs, lens)
        return n >= self.chunk_size, n

    def set_chunk(self, seqs, offs, lens):
        """Places an in-memory chunk in the read buffer (instead of parsing)."""
        self.buf = (seqs, offs, lens)

    # matchReads ReadQueue.cpp:404-933
    def matchReads(self, procReads, counters, getStranded):
        seqs, offs, lens = self.buf
        c = np.array(counters[:3], np.uint64)
        valid, mat = self.ctx.match(seqs, offs[:procReads], lens[:procReads], getStranded, c)
        counters[:3] = [int(x) for x in c]
        self.results = (valid, mat)
        return True

    # decideStrand ReadQueue.cpp:376-400
    def decideStrand(self):
        fwd, f, r = self.ctx.decide_strand()
        return fwd

    # printMethylationLevels ReadQueue.cpp:1860-1893
    def printMethylationLevels(self, filename):
        self.ctx.print_methylation_levels(filename)

    def close(self):
        if self._f:
            self._f.close()
        self.ctx.close()


def query_routine(rque, both_strands, on_chunk=None):
    """queryRoutine main.cpp:363-398; returns [succ, nonunique, unsucc]."""
    counters = [0, 0, 0]
    if not both_strands:
        _, n = rque.parseChunk()
        rque.matchReads(n, counters, True)
        if on_chunk:
            on_chunk(rque.results)
        rque.decideStrand()
    while True:
        more, n = rque.parseChunk()
        if not more:
            break
        rque.matchReads(n, counters, False)
        if on_chunk:
            on_chunk(rque.results)
    rque.matchReads(n, counters, False)
    if on_chunk:
        on_chunk(rque.results)
    return counters
