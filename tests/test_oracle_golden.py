"""The oracle (CPU restatement) against FAME's own outputs on the golden
fixtures: per-read validity and MATCH words, counters, strand tallies and the
CpG report byte for byte."""
import os

import numpy as np

import orc


def test_oracle_matches_reference(golden_case, tmp_path):
    c = golden_case
    ix = orc.Index(c.index)
    m, valid, mat = orc.match_stream(ix, c.seqs, c.offs, c.lens, chunk=c.chunk,
                                     both_strands=c.unord)
    assert np.array_equal(valid, c.valid)
    assert np.array_equal(mat, c.mat)
    assert m.counters.tolist() == [c.ref["succ"], c.ref["nonunique"], c.ref["unsucc"]]
    assert np.array_equal(m.meth(), c.meth)
    tsv = os.path.join(str(tmp_path), "o_cpg.tsv")
    m.write_tsv(tsv)
    assert open(tsv, "rb").read() == c.tsv
    # reads whose DP looked up to 8 bases before a chromosome start read
    # bytes outside the sequence in the reference (undefined behaviour; in
    # practice the zero upper bytes of the allocator's chunk header).  The
    # oracle and the HIP path read 0 there, and the fixtures still agree.
    assert m.ub_reads() <= 2
    if not c.unord:
        fwd, f, r = m.decide_strand()
        assert (f, r, fwd) == (c.ref["r1_fwd"], c.ref["r1_rev"], c.ref["match_r1_fwd"])


def test_oracle_pairs_match_reference(golden_pair, tmp_path):
    c = golden_pair
    ix = orc.Index(c.index)
    m, v1, m1, v2, m2 = orc.match_pairs_stream(ix, c.r1, c.r2, chunk=c.chunk, both_strands=c.unord,
                                               minpd=c.minpd, maxpd=c.maxpd)
    assert np.array_equal(v1, c.valid1)
    assert np.array_equal(v2, c.valid2)
    assert np.array_equal(m1, c.mat1)
    assert np.array_equal(m2, c.mat2)
    assert m.counters5.tolist() == c.counters
    assert np.array_equal(m.meth(), c.meth)
    tsv = os.path.join(str(tmp_path), "o_cpg.tsv")
    m.write_tsv(tsv)
    assert open(tsv, "rb").read() == c.tsv
    if not c.unord:
        fwd, f, r = m.decide_strand()
        assert (f, r, fwd) == (c.ref["r1_fwd"], c.ref["r1_rev"], c.ref["match_r1_fwd"])
