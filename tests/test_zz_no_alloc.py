"""fame_match_reads_noalloc (the allocation-free entry point) against
fame_match_reads_device, through the C ABI: per-read validity and MATCH,
counters, strand tallies and CpG counts, bit for bit, on two calls in a row
(the second smaller), on an index with large buckets (deferred reads, both
heavy tiers) and with the verify pools capped so that batches are split.
Device buffers for the inputs come from the HIP runtime the library itself
loaded (hipMalloc through ctypes)."""
import ctypes as C

import numpy as np
import pytest

from conftest import Case, Hip
from fame_amd import _lib

pytestmark = pytest.mark.gpu

MAX_READS = 32768
MAX_BYTES = 8 << 20


def noalloc(ctx, d_seqs, d_offs, d_lens, n, stranded, cnt, d_valid, d_mat):
    return _lib.load().fame_match_reads_noalloc(ctx.h, d_seqs, d_offs, d_lens, int(n), int(stranded),
                                                cnt.ctypes.data_as(C.POINTER(C.c_uint64)), d_valid or None,
                                                d_mat or None, None)


def compare_calls(c, calls):
    """calls: [(start, end, stranded)], run on a context per entry point"""
    hip = Hip()
    ds, do, dl = hip.put(c.seqs), hip.put(c.offs.astype(np.uint64)), hip.put(c.lens.astype(np.uint32))
    n = len(c.lens)
    out = {}
    for form in ("device", "noalloc"):
        ctx = _lib.Context(c.index, both_strands=c.unord)
        v, m = hip.alloc(n), hip.alloc(8 * n)
        res = []
        for s, e, stranded in calls:
            cnt = np.zeros(3, np.uint64)
            args = (ds, do + 8 * s, dl + 4 * s, e - s, stranded, cnt, v + s, m + 8 * s)
            if form == "device":
                ctx.match_device(*args)
            else:
                ctx._check(noalloc(ctx, *args))
            tally = ctx.decide_strand() if stranded else None
            res.append((hip.get(v + s, e - s, np.uint8), hip.get(m + 8 * s, e - s, np.uint64), cnt.copy(), tally,
                        ctx.meth().copy()))
        out[form] = (res, ctx.form_info())
        ctx.close()
    hip.free()
    return out


def check_equal(out):
    for k, (a, b) in enumerate(zip(out["device"][0], out["noalloc"][0])):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), f"call {k}: per-read results"
        assert a[2].tolist() == b[2].tolist(), f"call {k}: counters"
        assert a[3] == b[3], f"call {k}: strand tallies"
        assert np.array_equal(a[4], b[4]), f"call {k}: CpG counts"


@pytest.mark.parametrize("name", ["se_basic", "se_noloss", "se_long"])
def test_noalloc_equals_device_entry(name, tmp_path):
    c = Case(name, str(tmp_path))
    n = len(c.lens)
    print(f"case {name}: {n} reads, calls of {n - 600} then 600", flush=True)
    # a stranded call, then a smaller one in the decided orientation
    out = compare_calls(c, [(0, n - 600, not c.unord), (n - 600, n, False)])
    check_equal(out)
    assert out["noalloc"][1]["fallbacks"] == 0


def test_noalloc_splits_batches_that_outgrow_it(tmp_path, monkeypatch):
    """Verify pools capped at 256 slots (a test-only limit): the batch aborts
    before anything is counted and is matched again in halves; the results
    are those of the device entry point."""
    monkeypatch.setenv("FAME_POOL_CAP_LIMIT", "256")
    c = Case("se_repeats", str(tmp_path))
    n = len(c.lens)
    print(f"case se_repeats: pools capped, {n} reads", flush=True)
    out = compare_calls(c, [(0, 4000, True), (4000, n, False)])
    check_equal(out)
    assert out["noalloc"][1]["fallbacks"] > 0


def test_noalloc_rejects_sizes_beyond_its_limits(tmp_path):
    c = Case("se_basic", str(tmp_path))
    hip = Hip()
    ctx = _lib.Context(c.index)
    k = MAX_READS + 1
    seqs = hip.put(np.full(k * 100, ord("A"), np.uint8))
    offs = hip.put(np.arange(k, dtype=np.uint64) * 100)
    lens = hip.put(np.full(k, 100, np.uint32))
    cnt = np.zeros(3, np.uint64)
    rc = noalloc(ctx, seqs, offs, lens, k, 1, cnt, 0, 0)
    assert rc == -4 and "FAME_NOALLOC_MAX_READS" in _lib.load().fame_last_error(ctx.h).decode()
    # 100 reads, the last one ending past FAME_NOALLOC_MAX_BYTES
    big = hip.put(np.full(MAX_BYTES + 200, ord("A"), np.uint8))
    o2 = np.arange(100, dtype=np.uint64) * 100
    o2[99] = MAX_BYTES + 50
    rc = noalloc(ctx, big, hip.put(o2), lens, 100, 1, cnt, 0, 0)
    assert rc == -4 and "FAME_NOALLOC_MAX_BYTES" in _lib.load().fame_last_error(ctx.h).decode()
    # a read over 255 letters: the allocating entry points serve it
    o3 = np.arange(100, dtype=np.uint64) * 100
    l3 = np.full(100, 100, np.uint32)
    l3[7] = 300
    o3[8:] += 200
    rc = noalloc(ctx, big, hip.put(o3), hip.put(l3), 100, 1, cnt, 0, 0)
    assert rc == -4 and "FAME_MAX_MATE" in _lib.load().fame_last_error(ctx.h).decode()
    assert cnt.tolist() == [0, 0, 0]
    ctx.close()
    hip.free()
