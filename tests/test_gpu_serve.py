"""Serving loop (fame_serve_*): chunks submitted together are merged into
device batches, and every chunk's per-read results and counters equal those
of matching it alone (fame_match_reads after the same strand decision)."""
import numpy as np
import pytest

from fame_amd import _lib

pytestmark = pytest.mark.gpu


def chunks_of(c, first, sizes):
    out, s = [], first
    n = len(c.lens)
    for k in sizes:
        e = min(n, s + k)
        if e <= s:
            break
        out.append((s, e))
        s = e
    if s < n:
        out.append((s, n))
    return out


def decided_context(c, first):
    ctx = _lib.Context(c.index, both_strands=c.unord)
    cnt = np.zeros(3, np.uint64)
    if not c.unord:
        ctx.match(c.seqs, c.offs[:first], c.lens[:first], True, cnt)
        ctx.decide_strand()
    return ctx


@pytest.mark.parametrize("name", ["se_repeats", "se_long", "se_unord", "se_xlong"])
def test_serve_results_equal_alone_and_in_the_mix(name, tmp_path):
    from conftest import Case
    c = Case(name, str(tmp_path))
    first = 700 if not c.unord else 0
    sizes = [1, 37, 500, 3, 1200, 64, 999, 250, 2, 411]
    parts = chunks_of(c, first, sizes)
    print(f"case {name}: {len(parts)} chunks", flush=True)
    # alone
    a = decided_context(c, first)
    alone = []
    for s, e in parts:
        cnt = np.zeros(3, np.uint64)
        v, m = a.match(c.seqs, c.offs[s:e], c.lens[s:e], False, cnt)
        alone.append((v.copy(), m.copy(), cnt.copy()))
    meth_alone = a.meth()
    a.close()
    # in the mix: every chunk queued at once, batched by the server
    b = decided_context(c, first)
    # batches of 2048 reads, letters budgeted at 255 per read (more room for
    # the chunks of reads of up to 5 000 letters)
    mb = 16384 if (c.lens > 255).any() else 2048
    srv = _lib.Server(b, max_batch_reads=mb)
    tickets = [srv.submit(c.seqs, c.offs[s:e], c.lens[s:e]) for s, e in parts]
    mixed = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    for k, ((v, m, cnt), (v2, m2, cnt2, tm)) in enumerate(zip(alone, mixed)):
        assert np.array_equal(v, v2) and np.array_equal(m, m2), f"chunk {k} differs"
        assert cnt.tolist() == cnt2.tolist(), f"chunk {k} counters"
    assert np.array_equal(b.meth(), meth_alone)
    # chunks shared batches, and no batch exceeded the limit
    assert st["batches"] < len(parts)
    assert max(x[3]["batch_reads"] for x in mixed) <= mb
    assert max(x[3]["batch_reads"] - (e - s) for x, (s, e) in zip(mixed, parts)) > 0
    # and the oracle on the same chunking (the first chunk stranded, the rest
    # in the decided orientation): per-read results and CpG counts
    import orc
    ix = orc.Index(c.index)
    om, ov, omat = orc.match_stream(ix, c.seqs, c.offs, c.lens, chunk=first or len(c.lens),
                                    both_strands=c.unord)
    for k, ((s, e), (v2, m2, cnt2, tm)) in enumerate(zip(parts, mixed)):
        assert np.array_equal(v2, ov[s:e]) and np.array_equal(m2, omat[s:e]), f"chunk {k} against the oracle"
    assert np.array_equal(b.meth(), om.meth())
    b.close()
