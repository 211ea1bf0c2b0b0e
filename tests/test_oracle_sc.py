"""Single-cell driver against FAME's own (matchSCBatch / matchSCBatchPaired
via the reference harness's `sc` mode, tests/golden/sc_*): the oracle run
cell by cell (queryRoutine's chunk loop from zero strand tallies), its
counts written as printSCMethylationLevels writes them (ReadQueue.cpp:
1894-1921: four rows per cell, the fourth labelled "unmethFwd"), under the
header of the single-cell ReadQueue (:181-216), equal the reference's matrix
byte for byte."""
import numpy as np

import orc
from conftest import read_fastq_gz


def test_oracle_sc_matrix_matches_reference(golden_sc):
    c = golden_sc
    ix = orc.Index(c.index)
    head = c.expected.split(b"\n")
    assert head[0] == b"" and head[1].startswith(b"SC_ID\tCount_Type\t") and head[2].startswith(b"SC_ID\tCount_Type\t")
    out = [b"", head[1], head[2]]
    unord = c.meta["unord"]
    for cid, files in c.cells:
        if c.meta["pe"]:
            r1, r2 = read_fastq_gz(files[0]), read_fastq_gz(files[1])
            m = orc.match_pairs_stream(ix, r1, r2, chunk=c.meta["chunk"], both_strands=unord,
                                       minpd=c.meta["minpd"], maxpd=c.meta["maxpd"])[0]
        else:
            s, o, ln = read_fastq_gz(files[0])
            m = orc.match_stream(ix, s, o, ln, chunk=c.meta["chunk"], both_strands=unord)[0]
        meth = np.asarray(m.meth(), np.uint16).reshape(-1, 4)
        for k, lab in enumerate((b"methFwd", b"unmethFwd", b"methRev", b"unmethFwd")):
            out.append(cid.encode() + b"\t" + lab + b"\t" + b"".join(b"%d\t" % v for v in meth[:, k]))
        m.close()
    assert b"\n".join(out) == c.expected
