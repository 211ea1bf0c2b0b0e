"""Single-cell driver on the GPU (ReadQueue.single_cell, matchSCBatch /
matchSCBatchPaired, query_routine_sc*): the count matrix written through
fame_sc_* equals the one FAME's own single-cell driver wrote for the same
cells (tests/golden/sc_*), byte for byte.  The cell list names its files
relative to the working directory, as main.cpp:443-499 reads it."""
import os

import pytest

from fame_amd.readqueue import ReadQueue, RefGenome, query_routine_sc, query_routine_sc_paired

pytestmark = pytest.mark.gpu


def test_sc_matrix_matches_reference(golden_sc, tmp_path, monkeypatch):
    c = golden_sc
    print("case", c.name, flush=True)
    out = str(tmp_path / "sc_out.txt")
    rq = ReadQueue.single_cell(out, RefGenome(c.index), is_gz=True, bs_flag=c.meta["unord"],
                               is_paired=c.meta["pe"], chunk_size=c.meta["chunk"],
                               minpd=c.meta.get("minpd", 50), maxpd=c.meta.get("maxpd", 550))
    monkeypatch.chdir(c.dir)
    if c.meta["pe"]:
        query_routine_sc_paired(rq, True, "cells.txt")
    else:
        query_routine_sc(rq, True, "cells.txt")
    rq.close()
    got = open(out, "rb").read()
    assert got == c.expected
