"""Batched form (one hipGraph launch per batch, counts kept on the device):
bit-exact against FAME's outputs on the golden fixtures, and every replay of
a captured graph equals the plain form's result for the same chunk, with
eager work of the other form between replays."""
import os

import numpy as np
import pytest

from fame_amd import _lib

pytestmark = pytest.mark.gpu


def stream(ctx, c, chunk):
    n = len(c.lens)
    valid = np.zeros(n, np.uint8)
    mat = np.zeros(n, np.uint64)
    cnt = np.zeros(3, np.uint64)
    s = 0
    if not c.unord:
        e = min(n, chunk)
        valid[s:e], mat[s:e] = ctx.match(c.seqs, c.offs[s:e], c.lens[s:e], True, cnt)
        ctx.decide_strand()
        s = e
    while s < n:
        e = min(n, s + chunk)
        valid[s:e], mat[s:e] = ctx.match(c.seqs, c.offs[s:e], c.lens[s:e], False, cnt)
        s = e
    return valid, mat, cnt


def test_batched_matches_reference(golden_case, tmp_path):
    c = golden_case
    print(f"case {c.name}", flush=True)
    ctx = _lib.Context(c.index, both_strands=c.unord)
    ctx.set_form("batched")
    valid, mat, cnt = stream(ctx, c, max(len(c.lens) // 3, 1))
    info = ctx.form_info()
    assert info["graph_launches"] >= 1
    assert np.array_equal(valid, c.valid) and np.array_equal(mat, c.mat)
    assert cnt.tolist() == [c.ref["succ"], c.ref["nonunique"], c.ref["unsucc"]]
    assert np.array_equal(ctx.meth(), c.meth)
    base = os.path.join(str(tmp_path), "b")
    ctx.print_methylation_levels(base)
    assert open(base + "_cpg.tsv", "rb").read() == c.tsv
    ctx.close()


@pytest.mark.parametrize("env", [{"FAME_NO_MID_TIER": "1"}, {"FAME_MID_SLOTS": "64"}],
                         ids=["heavy_only", "mid_overflow"])
def test_batched_heavy_tiers_match_reference(golden_case, monkeypatch, env):
    """The batched form's in-graph tiers for deferred pairs: the global
    tiers alone, or the LDS tier with a 64-slot counting table handing reads
    to the first global tier (test-only switches): FAME's results."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    c = golden_case
    print(f"case {c.name}: {env}", flush=True)
    ctx = _lib.Context(c.index, both_strands=c.unord)
    ctx.set_form("batched")
    valid, mat, cnt = stream(ctx, c, max(len(c.lens) // 3, 1))
    if int(c.lens.max()) <= 255:  # longer reads always take the plain form
        assert ctx.form_info()["fallbacks"] == 0
    assert np.array_equal(valid, c.valid) and np.array_equal(mat, c.mat)
    assert cnt.tolist() == [c.ref["succ"], c.ref["nonunique"], c.ref["unsucc"]]
    assert np.array_equal(ctx.meth(), c.meth)
    ctx.close()


def test_graph_replays_equal_plain(golden_case):
    """The same chunk six times through its graph (captured once per
    autotuner candidate, then replayed), a plain call of another chunk
    between replays; every replay equals the plain result."""
    c = golden_case
    ctx = _lib.Context(c.index, both_strands=c.unord)
    n = len(c.lens)
    half = n // 2
    cnt = np.zeros(3, np.uint64)
    if not c.unord:
        ctx.match(c.seqs, c.offs[:half], c.lens[:half], True, cnt)
        ctx.decide_strand()
    a = (c.seqs, c.offs[half:], c.lens[half:])
    b = (c.seqs, c.offs[:half], c.lens[:half])
    ctx.set_form("plain")
    pv, pm = ctx.match(*a, False, cnt)
    pv, pm = pv.copy(), pm.copy()
    for k in range(6):
        ctx.set_form("batched")
        cb = np.zeros(3, np.uint64)
        v, m = ctx.match(*a, False, cb)
        assert np.array_equal(v, pv) and np.array_equal(m, pm), f"replay {k}"
        ctx.set_form("plain")
        ctx.match(*b, False, np.zeros(3, np.uint64))
    info = ctx.form_info()
    # a batch holding a read over 255 letters runs on the plain path
    long_reads = bool((c.lens[half:] > _lib.MAX_MATE).any())
    assert info["graph_launches"] == 6 and info["fallbacks"] == (6 if long_reads else 0)
    ctx.close()


def test_pipelined_lanes_match_reference(golden_case, tmp_path):
    """Sub-batches of 700 reads alternating between the two lanes (own
    buffers and streams, work in flight together): still bit-exact."""
    c = golden_case
    ctx = _lib.Context(c.index, both_strands=c.unord)
    ctx.set_form("batched")
    ctx.set_pipeline(700)
    valid, mat, cnt = stream(ctx, c, max(len(c.lens) // 2, 1))
    info = ctx.form_info()
    assert info["pipelined_calls"] >= 1
    assert np.array_equal(valid, c.valid) and np.array_equal(mat, c.mat)
    assert cnt.tolist() == [c.ref["succ"], c.ref["nonunique"], c.ref["unsucc"]]
    assert np.array_equal(ctx.meth(), c.meth)
    ctx.close()


def test_batched_overflow_reruns_plain(golden_case, monkeypatch):
    """Verify pools capped at 256 slots (FAME_POOL_CAP_LIMIT, a test-only
    limit): batches outgrow the graphs' fixed capacities, are detected on
    the device before anything is counted and run again on the plain path;
    per-read results, counters and CpG counts stay FAME's (nothing counted
    twice), in single calls and in the pipelined lanes."""
    monkeypatch.setenv("FAME_POOL_CAP_LIMIT", "256")
    c = golden_case
    print(f"case {c.name}: pools capped", flush=True)
    for pipe in (0, 700):
        ctx = _lib.Context(c.index, both_strands=c.unord)
        ctx.set_form("batched")
        if pipe:
            ctx.set_pipeline(pipe)
        valid, mat, cnt = stream(ctx, c, max(len(c.lens) // 2, 1))
        info = ctx.form_info()
        assert info["fallbacks"] > 0
        assert np.array_equal(valid, c.valid) and np.array_equal(mat, c.mat)
        assert cnt.tolist() == [c.ref["succ"], c.ref["nonunique"], c.ref["unsucc"]]
        assert np.array_equal(ctx.meth(), c.meth)
        ctx.close()


def test_graph_cache_stays_bounded(monkeypatch):
    """Device-path calls with new input buffers every call (70 of them, one
    capacity class): each captures a graph, the plan keeps at most 64 (the
    least recently used destroyed), and every call equals the plain form."""
    from conftest import Case, Hip
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        c = Case("se_basic", td)
        hip = Hip()
        ctx = _lib.Context(c.index)
        cnt = np.zeros(3, np.uint64)
        ctx.match(c.seqs, c.offs[:700], c.lens[:700], True, cnt)
        ctx.decide_strand()
        n = 300
        plain = _lib.Context(c.index)
        plain.match(c.seqs, c.offs[:700], c.lens[:700], True, np.zeros(3, np.uint64))
        plain.decide_strand()
        ctx.set_form("batched")
        for k in range(70):
            s = 700 + (k * 37) % (len(c.lens) - 700 - n)
            offs = c.offs[s:s + n] - c.offs[s]
            lens = c.lens[s:s + n]
            seqs = c.seqs[int(c.offs[s]):int(c.offs[s + n - 1] + lens[-1])]
            ds, do, dl = hip.put(seqs), hip.put(offs.astype(np.uint64)), hip.put(lens.astype(np.uint32))
            dv, dm = hip.alloc(n), hip.alloc(8 * n)
            cb = np.zeros(3, np.uint64)
            ctx.match_device(ds, do, dl, n, False, cb, dv, dm)
            v, m = hip.get(dv, n, np.uint8), hip.get(dm, n, np.uint64)
            cp = np.zeros(3, np.uint64)
            pv, pm = plain.match(c.seqs, c.offs[s:s + n], lens, False, cp)
            assert np.array_equal(v, pv) and np.array_equal(m, pm), f"call {k}"
            assert cb.tolist() == cp.tolist(), f"call {k} counters"
        info = ctx.form_info()
        assert info["graphs"] >= 70 and info["graphs_held"] <= 64 and info["fallbacks"] == 0
        ctx.close()
        plain.close()
        hip.free()
