"""Paired-end HIP path (fame_match_pairs) against FAME's own outputs on the
paired golden fixtures and against the oracle, through the C ABI.
Bit-exact: per-mate validity and MATCH words, the five paired counters,
strand tallies, CpG counts and the CpG report."""
import os

import numpy as np
import pytest

import orc
from fame_amd import _lib

pytestmark = pytest.mark.gpu


def run_hip_pairs(c, chunk=None):
    """queryRoutinePaired main.cpp:400-441 over the case's pairs."""
    chunk = chunk or c.chunk
    ctx = _lib.Context(c.index, both_strands=c.unord)
    ctx.set_pair_dist(c.minpd, c.maxpd)
    n = len(c.r1[2])
    v1 = np.zeros(n, np.uint8)
    v2 = np.zeros(n, np.uint8)
    m1 = np.zeros(n, np.uint64)
    m2 = np.zeros(n, np.uint64)
    cnt = np.zeros(5, np.uint64)
    tally = None

    def run(s, e, stranded):
        a = (c.r1[0], c.r1[1][s:e], c.r1[2][s:e])
        b = (c.r2[0], c.r2[1][s:e], c.r2[2][s:e])
        v1[s:e], m1[s:e], v2[s:e], m2[s:e] = ctx.match_pairs(a, b, stranded, cnt)

    s = 0
    if not c.unord:
        e = min(n, chunk)
        run(s, e, True)
        tally = ctx.decide_strand()
        s = e
    while s < n:
        e = min(n, s + chunk)
        run(s, e, False)
        s = e
    return ctx, v1, m1, v2, m2, cnt, tally


def test_hip_pairs_match_reference(golden_pair, tmp_path):
    c = golden_pair
    print(f"case {c.name}: {len(c.r1[2])} pairs", flush=True)
    ctx, v1, m1, v2, m2, cnt, tally = run_hip_pairs(c)
    bad = np.nonzero((v1 != c.valid1) | (v2 != c.valid2) | (m1 != c.mat1) | (m2 != c.mat2))[0]
    assert bad.size == 0, f"{bad.size} pairs differ, first {bad[:10].tolist()}"
    assert cnt.tolist() == c.counters
    if tally is not None:
        assert (tally[1], tally[2], tally[0]) == (c.ref["r1_fwd"], c.ref["r1_rev"], c.ref["match_r1_fwd"])
    assert np.array_equal(ctx.meth(), c.meth)
    base = os.path.join(str(tmp_path), "hip")
    ctx.print_methylation_levels(base)
    assert open(base + "_cpg.tsv", "rb").read() == c.tsv
    ctx.close()


def test_hip_pairs_match_oracle_other_chunking(golden_pair):
    """Chunks of a third of the pairs: per-pair results do not depend on the
    batch they are matched in (the stranded first chunk aside)."""
    c = golden_pair
    chunk = max(len(c.r1[2]) // 3, 1)
    ctx, v1, m1, v2, m2, cnt, _ = run_hip_pairs(c, chunk=chunk)
    ix = orc.Index(c.index)
    m, ov1, om1, ov2, om2 = orc.match_pairs_stream(ix, c.r1, c.r2, chunk=chunk, both_strands=c.unord,
                                                   minpd=c.minpd, maxpd=c.maxpd)
    assert np.array_equal(v1, ov1) and np.array_equal(v2, ov2)
    assert np.array_equal(m1, om1) and np.array_equal(m2, om2)
    assert cnt.tolist() == m.counters5.tolist()
    assert np.array_equal(ctx.meth(), m.meth())
    ctx.close()


@pytest.mark.parametrize("walk_env", [{"FAME_PE_WALK_LDS": "0"},
                                      {"FAME_PE_WALK_LDS": "0", "FAME_PE_WALK_BIG_LDS": "0"}],
                         ids=["wave-walk-lds", "wave-walk-map-pool"])
def test_hip_pairs_wave_walks_match_reference(golden_pair, monkeypatch, walk_env):
    """Every item through the one-wave walk kernel (k_pe_walk1_big), with its
    windows in LDS or in the map pool: the same per-pair results, counters
    and CpG counts as FAME."""
    for k, v in walk_env.items():
        monkeypatch.setenv(k, v)
    c = golden_pair
    print(f"case {c.name}: {len(c.r1[2])} pairs, {walk_env}", flush=True)
    ctx, v1, m1, v2, m2, cnt, _ = run_hip_pairs(c)
    bad = np.nonzero((v1 != c.valid1) | (v2 != c.valid2) | (m1 != c.mat1) | (m2 != c.mat2))[0]
    assert bad.size == 0, f"{bad.size} pairs differ, first {bad[:10].tolist()}"
    assert cnt.tolist() == c.counters
    assert np.array_equal(ctx.meth(), c.meth)
    ctx.close()


def test_hip_pairs_verify_queue_full_match_reference(golden_pair, monkeypatch):
    """Verify's run queue held to one run per round of 256 jobs
    (FAME_VERIFY_QUEUE_LIMIT, test-only): mates' jobs re-verified in place
    give FAME's paired results."""
    monkeypatch.setenv("FAME_VERIFY_QUEUE_LIMIT", "1")
    c = golden_pair
    print(f"case {c.name}: verify queue of 1 run", flush=True)
    ctx, v1, m1, v2, m2, cnt, _ = run_hip_pairs(c)
    bad = np.nonzero((v1 != c.valid1) | (v2 != c.valid2) | (m1 != c.mat1) | (m2 != c.mat2))[0]
    assert bad.size == 0, f"{bad.size} pairs differ, first {bad[:10].tolist()}"
    assert cnt.tolist() == c.counters
    assert np.array_equal(ctx.meth(), c.meth)
    ctx.close()


def test_hip_pairs_tab64_match_reference(golden_pair, monkeypatch):
    """Paired path with the 64-bit tabIndex kept (FAME_TAB64, test-only: the
    kernels of indexes of 2^32 k-mers or more) gives FAME's results."""
    monkeypatch.setenv("FAME_TAB64", "1")
    c = golden_pair
    print(f"case {c.name}: 64-bit tabIndex", flush=True)
    ctx, v1, m1, v2, m2, cnt, _ = run_hip_pairs(c)
    bad = np.nonzero((v1 != c.valid1) | (v2 != c.valid2) | (m1 != c.mat1) | (m2 != c.mat2))[0]
    assert bad.size == 0, f"{bad.size} pairs differ, first {bad[:10].tolist()}"
    assert cnt.tolist() == c.counters
    assert np.array_equal(ctx.meth(), c.meth)
    ctx.close()


def test_hip_pairs_batched_lanes_match_oracle(golden_pair):
    """Batched form, paired: after a stranded first chunk of a third of the
    pairs, the rest in one call that splits into sub-batches of 300 pairs
    alternating between two lanes (the context and a worker context, each
    with its own buffers, stream and host thread); the oracle on the same
    chunking, per pair, by counters and CpG counts."""
    c = golden_pair
    print(f"case {c.name}: paired lanes", flush=True)
    ctx = _lib.Context(c.index, both_strands=c.unord)
    ctx.set_pair_dist(c.minpd, c.maxpd)
    ctx.set_form("batched")
    ctx.set_pipeline(600)
    n = len(c.r1[2])
    first = 0 if c.unord else max(n // 3, 1)
    v1 = np.zeros(n, np.uint8)
    v2 = np.zeros(n, np.uint8)
    m1 = np.zeros(n, np.uint64)
    m2 = np.zeros(n, np.uint64)
    cnt = np.zeros(5, np.uint64)
    if first:
        a = (c.r1[0], c.r1[1][:first], c.r1[2][:first])
        b = (c.r2[0], c.r2[1][:first], c.r2[2][:first])
        v1[:first], m1[:first], v2[:first], m2[:first] = ctx.match_pairs(a, b, True, cnt)
        ctx.decide_strand()
    a = (c.r1[0], c.r1[1][first:], c.r1[2][first:])
    b = (c.r2[0], c.r2[1][first:], c.r2[2][first:])
    v1[first:], m1[first:], v2[first:], m2[first:] = ctx.match_pairs(a, b, False, cnt)
    assert ctx.form_info()["pipelined_calls"] >= 1
    ix = orc.Index(c.index)
    m, ov1, om1, ov2, om2 = orc.match_pairs_stream(ix, c.r1, c.r2, chunk=first or n, both_strands=c.unord,
                                                   minpd=c.minpd, maxpd=c.maxpd)
    bad = np.nonzero((v1 != ov1) | (v2 != ov2) | (m1 != om1) | (m2 != om2))[0]
    assert bad.size == 0, f"{bad.size} pairs differ, first {bad[:10].tolist()}"
    assert cnt.tolist() == m.counters5.tolist()
    assert np.array_equal(ctx.meth(), m.meth())
    ctx.close()
