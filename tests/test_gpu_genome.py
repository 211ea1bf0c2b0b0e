"""Genome-scale constants on the GPU against FAME itself: HTABSIZE 2^30 (the
8 GiB slot table and the tabIndex narrowed to 32 bits), 24 chromosomes
(CHROMNUM 24), 400 000 reads in FAME's chunk loop (a first chunk of 300 000
in both orientations, then a stranded one), in the plain form, in the
batched form with both chunks split over the two pipelined lanes (the first
one's strand tallies summed over the lanes), and
with the 64-bit tabIndex kept (FAME_TAB64, test-only).  The reference is
FAME's own CPU path (main.cpp queryRoutine over ReadQueue::matchReads),
built from its sources into oracle/_ref for these constants; its per-read
(isInvalid, MATCH), counters and CpG report are compared byte for byte."""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from fame_amd import _lib, index, synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "fame_ref_h1073741824_r100_c16_n24_p50-550_k300000")
HTAB = 1 << 30
CHUNK = 300_000
NREADS = 400_000


def workdir(need):
    for d in ("/dev/shm", tempfile.gettempdir()):
        try:
            if shutil.disk_usage(d).free > need:
                return tempfile.mkdtemp(prefix="fame_genome_", dir=d)
        except OSError:
            pass
    pytest.skip(f"no folder with {need / 1e9:.0f} GB free")


@pytest.fixture(scope="module")
def genome_case():
    assert os.path.exists(REF), f"{REF} is not built (build() makes it)"
    wd = workdir(24e9)
    try:
        lens_c = synth.config_genome("genome", scale=0.12, seed=11)
        chroms = synth.make_genome(lens_c, seed=11)
        idx = os.path.join(wd, "g.idx")
        print(f"building a {sum(lens_c) / 1e6:.0f} Mbp index of {len(lens_c)} chromosomes, HTABSIZE 2^30",
              flush=True)
        index.build_index(chroms, idx, readlen=100, htab=HTAB, kmercutoff=1500)
        reads = synth.make_reads(chroms, NREADS, 100, seed=12)
        seqs, offs, lens = reads.flat()
        fq = os.path.join(wd, "r.fq")
        with open(fq, "wb") as f:
            q = b"I" * 100
            for i in range(len(lens)):
                L = int(lens[i])
                f.write(b"@r%d\n" % i + reads.seqs[i, :L].tobytes() + b"\n+\n" + q[:L] + b"\n")
        print("FAME's CPU path on the reads", flush=True)
        out = os.path.join(wd, "fame")
        r = subprocess.run([REF, "match", idx, fq, out], cwd=wd, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        raw = open(out + "_reads.bin", "rb").read()
        n = int(np.frombuffer(raw[8:16], np.uint64)[0])
        ref = {
            "valid": np.frombuffer(raw[16:16 + n], np.uint8).copy(),
            "mat": np.frombuffer(raw[16 + n:16 + 9 * n], np.uint64).copy(),
            "stats": json.load(open(out + "_stats.json")),
            "tsv": open(out + "_cpg.tsv", "rb").read(),
        }
        assert n == NREADS
        yield {"wd": wd, "idx": idx, "seqs": seqs, "offs": offs, "lens": lens, "ref": ref}
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def run_loop(ctx, g):
    seqs, offs, lens = g["seqs"], g["offs"], g["lens"]
    valid = np.zeros(NREADS, np.uint8)
    mat = np.zeros(NREADS, np.uint64)
    cnt = np.zeros(3, np.uint64)
    valid[:CHUNK], mat[:CHUNK] = ctx.match(seqs, offs[:CHUNK], lens[:CHUNK], True, cnt)
    ctx.decide_strand()
    valid[CHUNK:], mat[CHUNK:] = ctx.match(seqs, offs[CHUNK:], lens[CHUNK:], False, cnt)
    return valid, mat, cnt


def check(ctx, g, valid, mat, cnt, tag):
    ref = g["ref"]
    bad = np.nonzero((valid != ref["valid"]) | (mat != ref["mat"]))[0]
    assert bad.size == 0, f"{tag}: {bad.size} reads differ, first {bad[:10].tolist()}"
    st = ref["stats"]
    assert cnt.tolist() == [st["succ"], st["nonunique"], st["unsucc"]], tag
    base = os.path.join(g["wd"], "hip_" + tag)
    ctx.print_methylation_levels(base)
    assert open(base + "_cpg.tsv", "rb").read() == ref["tsv"], tag


def test_genome_scale_plain_and_batched(genome_case):
    g = genome_case
    ctx = _lib.Context(g["idx"])
    assert ctx.info["htab"] == HTAB and ctx.info["nchrom"] == 24
    print("plain form", flush=True)
    check(ctx, g, *run_loop(ctx, g), "plain")
    # the stranded chunk (100 000 reads, one orientation): most k-mers are
    # read from the genome-ordered slot copy at 2^30 keys
    st = ctx.stats()
    print(f"slots: {st}", flush=True)
    assert st["ordered_slots"] > 0.6 * st["kmers"], st
    assert st["random_slots"] < 0.4 * st["kmers"], st
    ctx.reset()
    print("batched form, pipelined lanes", flush=True)
    ctx.set_form("batched")
    ctx.set_pipeline(25_000)
    valid, mat, cnt = run_loop(ctx, g)
    info = ctx.form_info()
    assert info["pipelined_calls"] >= 1 and info["graph_launches"] >= 4
    check(ctx, g, valid, mat, cnt, "batched")
    ctx.close()


def test_genome_scale_without_genome_order(genome_case, monkeypatch):
    monkeypatch.setenv("FAME_NO_GORDER", "1")
    g = genome_case
    print("slot table read at random", flush=True)
    ctx = _lib.Context(g["idx"])
    check(ctx, g, *run_loop(ctx, g), "no_gorder")
    ctx.close()


def test_genome_scale_tab64(genome_case, monkeypatch):
    monkeypatch.setenv("FAME_TAB64", "1")
    g = genome_case
    print("64-bit tabIndex kept", flush=True)
    ctx = _lib.Context(g["idx"])
    check(ctx, g, *run_loop(ctx, g), "tab64")
    ctx.close()
