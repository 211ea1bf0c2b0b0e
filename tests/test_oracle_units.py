"""The oracle's building blocks, one at a time, against

* vectors written by FAME's own code (tests/golden/units, made by
  tests/golden/make_units.py through the reference harness's `units` mode):
  ntHash::NTPS64, ShiftAnd<E>, LevenshtDP, tsl::hopscotch_map order;
* the known answers of the reference's unit tests: LevenshtDP_unittest.cpp
  (all six cases) and ShiftAnd_unittest.cpp (masks, reset, matches; its match
  positions are shifted by MISCOUNT = 2 as ShiftAnd.h:164/171 report them
  today, and its mask tests read only the words the current layout has);
* what the mathematics fixes: rolling hash == direct hash, ShiftAnd ==
  brute-force bisulfite approximate matching (Sellers) on small inputs, the
  DP's edit distance == a textbook banded Levenshtein, hopscotch keys and
  counts == the multiset inserted.
"""
import os
import struct
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import orc  # noqa: E402

UNITS = os.path.join(ROOT, "tests", "golden", "units")
SEEDBITS = 0b11011110111111011111111111111101  # CONST.h:58
MISCOUNT = 2
COMP = bytes.maketrans(b"ACGTN", b"TGCAN")
FULL = 0xFFFFFFFFFFFFFFFF


class _Rd:
    def __init__(self, b):
        self.b = bytes(b)
        self.p = 0

    def get(self, fmt):
        v = struct.unpack_from("<" + fmt, self.b, self.p)
        self.p += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    def bytes(self, n):
        v = self.b[self.p:self.p + n]
        self.p += n
        return v

    def done(self):
        return self.p == len(self.b)


def _load(kind):
    z = np.load(os.path.join(UNITS, kind + ".npz"))
    return _Rd(z["input"]), _Rd(z["output"])


# ---- ntHash ------------------------------------------------------------------

def test_ntps64_matches_reference():
    i, o = _load("ntps64")
    for _ in range(i.get("I")):
        seq = i.bytes(i.get("I"))
        nk = len(seq) - 31
        direct = np.array(o.get(f"{nk}Q") if nk > 1 else [o.get("Q")], np.uint64)
        rolling = np.array(o.get(f"{nk}Q") if nk > 1 else [o.get("Q")], np.uint64)
        assert np.array_equal(direct, rolling)  # the reference's own two forms agree
        assert [orc.ntps64(seq[k:k + 32], SEEDBITS) for k in range(nk)] == direct.tolist()
        assert np.array_equal(orc.ntps64_rolling(seq, SEEDBITS), rolling)
    assert o.done()


# ---- ShiftAnd -----------------------------------------------------------------

def test_shiftand_matches_reference():
    i, o = _load("shiftand")
    n = i.get("I")
    total = 0
    for _ in range(n):
        E, rev, plen = i.get("BBI")
        pat = i.bytes(plen)
        text = i.bytes(i.get("I"))
        want = []
        for _ in range(o.get("I")):
            want.append((o.get("Q"), o.get("B")))
        total += len(want)
        assert orc.shiftand(pat, text, rev=bool(rev), E=E) == want
    assert o.done() and total > 300


def _m(pat, E):
    return orc.shiftand_state(pat, E)


def test_shiftand_known_answer_masks():
    """ShiftAnd_unittest.cpp:87-211 (mask words of the current layout, W = plen/64 + 1)."""
    A, C, G, T = 0, 1, 2, 3
    masks, acc, _ = _m(b"ACACACCCC", 1)                       # simpleBitmasks :87-114
    assert [int(x) for x in masks[:, 0]] == [0xFFFFFFFFFFFFFC2B, 0xFFFFFFFFFFFFFFD5, 0xFFFFFFFFFFFFFC01,
                                             0xFFFFFFFFFFFFFC01]
    assert acc.tolist() == [0x200]
    masks, acc, _ = _m(b"A" * 64 + b"CG", 1)                  # simpleBitmasksOverflow :119-150
    assert [int(masks[k, 0]) for k in (A, C, G, T)] == [FULL, 1, 1, 1]
    assert [int(masks[k, 1]) for k in (A, C, G, T)] == [0xFFFFFFFFFFFFFFF9, 0xFFFFFFFFFFFFFFFA,
                                                         0xFFFFFFFFFFFFFFFC, 0xFFFFFFFFFFFFFFF8]
    assert acc.tolist() == [0, 4]
    masks, acc, _ = _m(b"G" * 64 + b"C" * 64 + b"AAA", 2)     # bitmaskLongPattern :153-190
    assert [int(masks[k, 0]) for k in (A, C, G, T)] == [1, 1, FULL, 1]
    assert [int(masks[k, 1]) for k in (A, C, G, T)] == [0, 0xFFFFFFFFFFFFFFFE, 1, 0]
    assert acc.tolist() == [0, 0, 8]  # the test's {0, 1<<63} predates the 3-word layout
    masks, _, _ = _m(b"ATTATTTCCC", 1)                        # bitmaskBisulfite :195-211
    assert [int(x) for x in masks[:, 0]] == [0xFFFFFFFFFFFFF813, 0xFFFFFFFFFFFFFFED, 0xFFFFFFFFFFFFF801,
                                             0xFFFFFFFFFFFFF8ED]


def test_shiftand_known_answer_reset():
    """ShiftAnd_unittest.cpp:47-82: reset() sets layer i to 2^(i+1) - 1."""
    _, _, act = _m(b"ACCATGTGACTGCATG", 0)
    assert act[:, 0].tolist() == [1]
    _, _, act = _m(b"ACCATGTGACTGCATG", 2)
    assert act[:, 0].tolist() == [1, 3, 7]


def test_shiftand_known_answer_matches():
    """ShiftAnd_unittest.cpp:216-347, positions + MISCOUNT."""
    seq = b"A" * 120 + b"TTTTTTT"                              # matching_same :216-275
    assert orc.shiftand(seq, seq, E=0) == [(126 + MISCOUNT, 0)]
    assert orc.shiftand(seq, seq, E=1) == [(126 + MISCOUNT, 0)]
    t = bytearray(seq)
    t[5] = ord("C")
    assert orc.shiftand(seq, bytes(t), E=0) == []
    assert orc.shiftand(seq, bytes(t), E=1) == [(126 + MISCOUNT, 1)]
    p, t = b"AGGCGAGGC", b"AGGCGAGGCGAAGCGAGGC"                # matching_smaller :279-314
    assert orc.shiftand(p, t, E=1) == [(8 + MISCOUNT, 0), (13 + MISCOUNT, 1), (18 + MISCOUNT, 1)]
    assert orc.shiftand(p, t[11:], E=1)[0] == (7 + MISCOUNT, 1)
    p, t = b"AGGTTATTC", bytearray(b"AGGTCACCCAAT")          # matching_bisulfite :317-347
    assert orc.shiftand(p, bytes(t), E=1) == [(8 + MISCOUNT, 0)]
    t[4] = ord("G")
    assert orc.shiftand(p, bytes(t), E=1) == [(8 + MISCOUNT, 1)]


def _bs_eq(pc, tc):
    # pattern T also matches text C (bisulfite); otherwise letters must agree
    return pc == tc or (pc == ord("T") and tc == ord("C"))


def _sellers(pat, text, E):
    """Per text position: the least edit distance of the whole pattern to a
    text substring ending there that crosses no N (None where > E)."""
    m = len(pat)
    out = []
    col = list(range(m + 1))
    for c in text:
        if c == ord("N"):
            col = list(range(m + 1))
            out.append(None)
            continue
        new = [0] * (m + 1)
        for j in range(1, m + 1):
            new[j] = min(col[j] + 1, new[j - 1] + 1, col[j - 1] + (0 if _bs_eq(pat[j - 1], c) else 1))
        col = new
        out.append(col[m] if col[m] <= E else None)
    return out


def _report(dist, plen, E, pos_of):
    """Runs of accepting positions merge; a later one replaces the entry when
    its error count is not larger (ShiftAnd.h:155-176)."""
    res, was, prev, ncomp = [], False, None, 0
    for k, d in enumerate(dist):
        ncomp += 1
        if ncomp < plen - E:
            continue
        if d is None:
            was = False
            continue
        if was:
            if d <= prev:
                res[-1] = (pos_of(k), d)
                prev = d
        else:
            res.append((pos_of(k), d))
            was, prev = True, d
    return res


@pytest.mark.parametrize("seed", range(6))
def test_shiftand_equals_bruteforce(seed):
    rng = np.random.default_rng(seed)
    for _ in range(150):
        plen = int(rng.integers(6, 14))
        E = int(rng.integers(0, 4))
        pat = bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, plen)])
        text = bytearray(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, int(rng.integers(10, 60)))])
        at = int(rng.integers(0, len(text) - 5))
        text[at:at + plen] = pat.replace(b"T", b"C") if rng.random() < 0.5 else pat
        if rng.random() < 0.2:
            text[int(rng.integers(0, len(text)))] = ord("N")
        text = bytes(text)
        # N letters do not advance the letter count (ShiftAnd.h:142-146), so
        # drop them from the gate by counting non-N letters only
        dist = _sellers(pat, text, E)
        nn = [k for k, c in enumerate(text) if c != ord("N")]
        fwd = _report([dist[k] for k in nn], plen, E, lambda q: nn[q] + MISCOUNT)
        assert orc.shiftand(pat, text, E=E) == fwd
        u = text.translate(COMP)[::-1]
        dist = _sellers(pat, u, E)
        nn = [k for k, c in enumerate(u) if c != ord("N")]
        T = len(text)
        rev = _report([dist[k] for k in nn], plen, E, lambda q: (T - 1 - nn[q]) + 1 + plen - 1 + MISCOUNT)
        assert orc.shiftand(pat, text, rev=True, E=E) == rev


# ---- LevenshtDP ---------------------------------------------------------------

MATCHING, MISMATCH, INSERTION, DELETION = 0, 1, 2, 3


def test_dp_matches_reference():
    i, o = _load("dp")
    n = i.get("I")
    for _ in range(n):
        band, rev, cmp, L = i.get("BBBI")
        row = i.bytes(L)
        ref = i.bytes(i.get("I"))
        end = i.get("Q")
        dist = o.get("H")
        trace = list(o.bytes(o.get("I")))
        got, ub, d = orc.dp_align(row, ref, end, rev=bool(rev), band=band, cmp=cmp)
        assert not ub
        assert (got, d) == (trace, dist)
    assert o.done()


def test_dp_known_answers():
    """LevenshtDP_unittest.cpp:26-243 (colPat = s2 + offset; rows read as the reference does)."""
    eq, wc = orc.CMP_EQ, orc.CMP_WC
    # init_test.twoLetterMismatches :26-88
    for band in (0, 1, 2):
        assert orc.dp_align(b"AG", b"CTAG", 3, band=band, cmp=eq)[2] == 0
        assert orc.dp_align(b"AG", b"CCTCG", 4, band=band, cmp=eq)[2] == 1
    assert orc.dp_align(b"AG", b"CTAG", 3, band=2, cmp=eq)[0] == [MATCHING, MATCHING]
    assert orc.dp_align(b"AG", b"CCTCG", 4, band=2, cmp=eq)[0] == [INSERTION, MATCHING]
    assert orc.dp_align(b"AG", b"CCTCT", 4, band=2, cmp=eq)[2] == 2
    # matching_test.simpleSeq :92-117
    tr, _, d = orc.dp_align(b"ACTTTGACCG", b"AAACTGTGACTG", 11, band=2, cmp=eq)
    assert d == 2 and tr == [0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    # matching_test.simpleSeqRev :119-162
    tr, _, d = orc.dp_align(b"TCCAG", b"GACTGAA", 6, rev=True, band=2, cmp=wc)
    assert d == 1 and tr == [0, 0, 0, 1, 0]
    # matching_test.complexSeqIns :165-188
    tr, _, d = orc.dp_align(b"ACTAAGTGACTG", b"TTTTACTGTGACTG", 13, band=2, cmp=eq)
    assert d == 2 and tr == [0, 0, 0, 2, 2, 0, 0, 0, 0, 0, 0, 0]
    # matching_test.complexSeqDel :191-214
    tr, _, d = orc.dp_align(b"TGTGACTT", b"TGTGACTAAT", 9, band=2, cmp=eq)
    assert d == 2 and tr == [0, 0, 0, 0, 0, 0, 0, 3, 3, 0]
    # matching_test.complexSeqAll :217-241
    tr, _, d = orc.dp_align(b"ACGTGAACCTG", b"AAACGGTGAACTG", 12, band=2, cmp=eq)
    assert d == 2 and tr == [0, 0, 0, 3, 0, 0, 0, 0, 0, 2, 0, 0]


def _banded(row, ref, end, band, rev, cmp):
    """Textbook banded Levenshtein: row letters against the reference read
    backwards from ref[end]; |i - j| <= band; least value in the last row."""
    L = len(row)
    INF = 1 << 30
    D = [[INF] * (L + band + 2) for _ in range(L + 1)]
    for c in range(band + 1):
        D[0][c] = c
    for r in range(1, L + 1):
        if r <= band:
            D[r][0] = r
        for j in range(max(1, r - band), r + band + 1):
            rc = row[r - 1] if rev else row[L - r]
            gc = ref[end - j + 1]
            mm = {orc.CMP_FWD: lambda a, b: 0 if a == b or (a == 84 and b == 67) else 1,
                  orc.CMP_EQ: lambda a, b: 0 if a == b else 1}[cmp](rc, gc)
            D[r][j] = min(D[r - 1][j] + 1, D[r][j - 1] + 1, D[r - 1][j - 1] + mm)
    return min(D[L][j] for j in range(L - band, L + band + 1))


def test_dp_distance_equals_textbook():
    rng = np.random.default_rng(11)
    for _ in range(300):
        band = int(rng.choice([1, 2, 8]))
        L = int(rng.integers(band + 1, 40))
        row = bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, L)])
        seg = bytearray(row)
        for _ in range(int(rng.integers(0, band + 1))):
            k = int(rng.integers(0, len(seg)))
            op = rng.integers(0, 3)
            if op == 0:
                seg[k] = ord("ACGT"[rng.integers(0, 4)])
            elif op == 1 and len(seg) > 1:
                del seg[k]
            else:
                seg.insert(k, ord("ACGT"[rng.integers(0, 4)]))
        pad = band + 4
        ref = bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, pad + L + 2 * band)]) + bytes(seg) + b"A" * pad
        end = pad + L + 2 * band + len(seg) - 1
        cmp = orc.CMP_FWD if band == 8 else orc.CMP_EQ
        _, ub, d = orc.dp_align(row, ref, end, band=band, cmp=cmp)
        assert not ub
        assert d == _banded(row, ref, end, band, False, cmp)


# ---- hopscotch order -----------------------------------------------------------

def test_hopscotch_matches_reference():
    i, o = _load("hopscotch")
    for _ in range(i.get("I")):
        bc, n = i.get("QI")
        keys = np.frombuffer(i.bytes(4 * n), "<u4")
        m = o.get("I")
        pairs = np.frombuffer(o.bytes(8 * m), "<u4").reshape(m, 2)
        final_bc = o.get("Q")
        got_k, got_v, got_bc = orc.hopscotch_order(keys, bc)
        assert np.array_equal(got_k, pairs[:, 0]) and np.array_equal(got_v, pairs[:, 1])
        assert got_bc == final_bc
        # the map holds the inserted multiset
        u, c = np.unique(keys, return_counts=True)
        order = np.argsort(got_k)
        assert np.array_equal(got_k[order], u) and np.array_equal(got_v[order], c % 65536)
    assert o.done()
