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

CASES = sorted(d for d in os.listdir(GOLDEN) if os.path.isfile(os.path.join(GOLDEN, d, "case.json")))


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
