"""Python mirror of FAME's matching boundary (ReadQueue / RefGenome,
ReadQueue.h:44-129, RefGenome.h:55) over the C ABI of libfamehip.so.

Names, argument meaning and counter conventions follow the reference:
counters are accumulated into the caller's list, the first chunk of a
stranded run decides the orientation of read 1, and the report is written to
``<filename>_cpg.tsv``.
"""
import numpy as np

from . import _lib

CHUNKSIZE = 300000  # CONST.h:69


class RefGenome:
    """RefGenome(std::string filePath) RefGenome.cpp:77-80: the loaded index.
    The device copy is created by the ReadQueue that borrows it."""

    def __init__(self, file_path):
        self.path = file_path


class ReadQueue:
    """ReadQueue: single-end (ReadQueue.cpp:25-78) or, with file_path2, the
    paired-end constructor (:79-180)."""

    def __init__(self, file_path, ref, is_gz=False, bs_flag=False, chunk_size=CHUNKSIZE, file_path2=None,
                 minpd=50, maxpd=550):
        self.ctx = _lib.Context(ref.path, both_strands=bs_flag)
        self.bs_flag = bs_flag
        self.chunk_size = chunk_size
        # the native parser (fame_fastq_*), which reads as parseChunk /
        # parseChunkGZ do: a record every four lines, its second line the
        # sequence
        self._f = _lib.FastqReader(file_path, is_gz) if file_path else None
        self._f2 = _lib.FastqReader(file_path2, is_gz) if file_path2 else None
        self.is_paired = file_path2 is not None
        if self.is_paired:
            self.ctx.set_pair_dist(minpd, maxpd)
        self.buf = None   # (seqs, offs, lens) of the parsed chunk
        self.buf2 = None  # the same for the mates of a paired run
        self.results = None

    # parseChunk ReadQueue.cpp:218-296
    def parseChunk(self):
        self.buf = self._f.next(self.chunk_size)
        n = len(self.buf[2])
        if self.is_paired:
            self.buf2 = self._f2.next(self.chunk_size)
            if len(self.buf2[2]) != n:
                # the reference prints a message and exits (:280-286)
                raise _lib.FameError("Not the same number of reads available in the paired read files")
        return n >= self.chunk_size, n

    def set_chunk(self, seqs, offs, lens):
        """Places an in-memory chunk in the read buffer (instead of parsing)."""
        self.buf = (seqs, offs, lens)

    @classmethod
    def single_cell(cls, sc_output_path, ref, is_gz=False, bs_flag=False, is_paired=False, chunk_size=CHUNKSIZE,
                    minpd=50, maxpd=550):
        """The single-cell constructor (ReadQueue.cpp:181-216): no read file
        yet, the count matrix at sc_output_path with its header rows."""
        rq = cls(None, ref, is_gz, bs_flag, chunk_size, None, minpd, maxpd)
        rq.is_paired = is_paired
        if is_paired:
            rq.ctx.set_pair_dist(minpd, maxpd)
        rq.sc = _lib.ScMatrix(rq.ctx, sc_output_path)
        return rq

    # matchSCBatch ReadQueue.cpp:1660-1731
    def matchSCBatch(self, scFile, scId, isGZ):
        """One cell: strand tallies and counts from zero, the chunk loop of
        queryRoutine over scFile, then the cell's rows of the matrix.
        Returns [succ, nonunique, unsucc]."""
        self.ctx.reset()
        self._f = _lib.FastqReader(scFile, isGZ)
        try:
            counters = query_routine(self, self.bs_flag)
        finally:
            self._f.close()
            self._f = None
        self.sc.write_cell(scId)
        return counters

    # matchSCBatchPaired ReadQueue.cpp:1732-1854
    def matchSCBatchPaired(self, scFile1, scFile2, scId, isGZ):
        """The same for one cell's paired files; returns the five paired
        counters."""
        self.ctx.reset()
        self._f = _lib.FastqReader(scFile1, isGZ)
        self._f2 = _lib.FastqReader(scFile2, isGZ)
        try:
            counters = query_routine_paired(self, self.bs_flag)
        finally:
            self._f.close()
            self._f2.close()
            self._f = self._f2 = None
        self.sc.write_cell(scId)
        return counters

    # matchReads ReadQueue.cpp:404-933
    def matchReads(self, procReads, counters, getStranded):
        seqs, offs, lens = self.buf
        c = np.array(counters[:3], np.uint64)
        valid, mat = self.ctx.match(seqs, offs[:procReads], lens[:procReads], getStranded, c)
        counters[:3] = [int(x) for x in c]
        self.results = (valid, mat)
        return True

    def set_pairs(self, r1, r2):
        """Places an in-memory chunk of pairs in the read buffers."""
        self.buf, self.buf2 = r1, r2

    # matchPairedReads ReadQueue.cpp:935-1655
    def matchPairedReads(self, procReads, counters, getStranded):
        """counters: [succ, nonunique, unsucc, succPaired, tooShort], accumulated."""
        r1 = (self.buf[0], self.buf[1][:procReads], self.buf[2][:procReads])
        r2 = (self.buf2[0], self.buf2[1][:procReads], self.buf2[2][:procReads])
        c = np.array(counters[:5], np.uint64)
        self.results = self.ctx.match_pairs(r1, r2, getStranded, c)
        counters[:5] = [int(x) for x in c]
        return True

    # decideStrand ReadQueue.cpp:376-400
    def decideStrand(self):
        fwd, f, r = self.ctx.decide_strand()
        return fwd

    # printMethylationLevels ReadQueue.cpp:1860-1893
    def printMethylationLevels(self, filename):
        self.ctx.print_methylation_levels(filename)

    def close(self):
        for f in (self._f, self._f2):
            if f:
                f.close()
        if getattr(self, "sc", None):
            self.sc.close()
            self.sc = None
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


def query_routine_paired(rque, both_strands, on_chunk=None):
    """queryRoutinePaired main.cpp:400-441; returns
    [succ, nonunique, unsucc, succPaired, tooShort]."""
    counters = [0, 0, 0, 0, 0]
    if not both_strands:
        _, n = rque.parseChunk()
        rque.matchPairedReads(n, counters, True)
        if on_chunk:
            on_chunk(rque.results)
        rque.decideStrand()
    while True:
        more, n = rque.parseChunk()
        if not more:
            break
        rque.matchPairedReads(n, counters, False)
        if on_chunk:
            on_chunk(rque.results)
    rque.matchPairedReads(n, counters, False)
    if on_chunk:
        on_chunk(rque.results)
    return counters


def _sc_cells(sc_meta_file, nfiles):
    """Lines "<id> <file> [<file2>]" of the cell list, split at single spaces
    as main.cpp:451-457 and :481-491 split them."""
    with open(sc_meta_file) as f:
        for line in f.read().split("\n"):
            parts = line.split(" ")
            if len(parts) >= 1 + nfiles and parts[0]:
                yield parts[0], parts[1:1 + nfiles]


def query_routine_sc(rque, is_gz, sc_meta_file):
    """queryRoutineSC main.cpp:443-468: matchSCBatch per listed cell (file
    paths as listed, relative ones from the working directory)."""
    for sc_id, (path,) in _sc_cells(sc_meta_file, 1):
        rque.matchSCBatch(path, sc_id, is_gz)


def query_routine_sc_paired(rque, is_gz, sc_meta_file):
    """queryRoutineSCPaired main.cpp:469-499: matchSCBatchPaired per cell."""
    for sc_id, (p1, p2) in _sc_cells(sc_meta_file, 2):
        rque.matchSCBatchPaired(p1, p2, sc_id, is_gz)
