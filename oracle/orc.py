"""ctypes binding of the oracle library (oracle/_ref/liboracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the product.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "liboracle.so")

_u8p = C.POINTER(C.c_uint8)
_u16p = C.POINTER(C.c_uint16)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


def _ptr(a, t):
    return a.ctypes.data_as(t)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run `make -C oracle oracle`)")
        L = C.CDLL(LIB_PATH)
        L.orc_load_index.restype = C.c_void_p
        L.orc_load_index.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t]
        L.orc_free_index.argtypes = [C.c_void_p]
        L.orc_index_info.argtypes = [C.c_void_p, _u64p]
        L.orc_ctx_new.restype = C.c_void_p
        L.orc_ctx_new.argtypes = [C.c_void_p, C.c_int]
        L.orc_ctx_free.argtypes = [C.c_void_p]
        L.orc_match_reads.argtypes = [C.c_void_p, _u8p, _u64p, _u32p, C.c_uint32, C.c_int, _u64p,
                                      _u8p, _u64p, C.c_int]
        L.orc_decide_strand.argtypes = [C.c_void_p, _u64p, _u64p]
        L.orc_get_meth.argtypes = [C.c_void_p, _u16p]
        L.orc_write_tsv.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_ub_reads.restype = C.c_uint64
        L.orc_ub_reads.argtypes = [C.c_void_p]
        L.orc_ntps64.restype = C.c_uint64
        L.orc_ntps64.argtypes = [_u8p, C.c_uint32]
        L.orc_ntps64_rolling.argtypes = [_u8p, C.c_uint32, C.c_uint32, _u64p]
        L.orc_shiftand.argtypes = [_u8p, C.c_uint32, _u8p, C.c_uint32, C.c_int, _u64p, _u8p, C.c_int,
                                   C.POINTER(C.c_int)]
        L.orc_dp_align.argtypes = [_u8p, C.c_uint32, _u8p, C.c_uint64, C.c_int64, C.c_int, _u8p,
                                   C.c_int, C.POINTER(C.c_int)]
        L.orc_hopscotch_order.argtypes = [_u32p, C.c_uint32, C.c_uint64, _u32p, _u32p, _u64p]
        L.orc_shiftand_e.argtypes = [_u8p, C.c_uint32, _u8p, C.c_uint32, C.c_int, C.c_uint, _u64p, _u8p,
                                     C.c_int, C.POINTER(C.c_int)]
        L.orc_shiftand_state.argtypes = [_u8p, C.c_uint32, C.c_uint, _u64p, _u64p, _u64p]
        L.orc_dp_align_e.argtypes = [_u8p, C.c_uint32, _u8p, C.c_uint64, C.c_int64, C.c_int, C.c_uint, C.c_int,
                                     _u8p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_set_pdist.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
        L.orc_match_pairs.argtypes = [C.c_void_p, _u8p, _u64p, _u32p, _u8p, _u64p, _u32p, C.c_uint32,
                                      C.c_int, _u64p, _u8p, _u64p, _u8p, _u64p, C.c_int]
        L.orc_std_sort_counts.argtypes = [_u32p, _u32p, C.c_uint32, C.c_int]
        L.orc_seed_maps.argtypes = [C.c_void_p, _u8p, C.c_uint32, _u32p, _u16p, C.POINTER(C.c_int),
                                    _u32p, _u16p, C.POINTER(C.c_int), C.c_int, _u32p]
        _lib = L
    return _lib


def _bytes(s):
    if isinstance(s, (bytes, bytearray)):
        return np.frombuffer(bytes(s), dtype=np.uint8).copy()
    return np.ascontiguousarray(s, dtype=np.uint8)


class Index:
    FIELDS = ("htab", "readlen", "winlen", "kmerlen", "kmercutoff", "seedbits", "ncpg", "nchrom",
              "nkmer", "nwin", "nfiltered")

    def __init__(self, path):
        err = C.create_string_buffer(256)
        self.h = lib().orc_load_index(path.encode(), err, 256)
        if not self.h:
            raise RuntimeError(err.value.decode())
        info = np.zeros(11, dtype=np.uint64)
        lib().orc_index_info(self.h, _ptr(info, _u64p))
        self.info = dict(zip(self.FIELDS, (int(x) for x in info)))

    def close(self):
        if self.h:
            lib().orc_free_index(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Matcher:
    """The reference's ReadQueue matching state (single-end) on the CPU."""

    def __init__(self, index, both_strands=False):
        self.index = index
        self.h = lib().orc_ctx_new(index.h, int(both_strands))
        self.counters = np.zeros(3, dtype=np.uint64)

    def match(self, seqs, offs, lens, get_stranded, nthreads=0):
        n = int(len(lens))
        valid = np.zeros(n, dtype=np.uint8)
        mat = np.zeros(n, dtype=np.uint64)
        seqs = np.ascontiguousarray(seqs, dtype=np.uint8)
        offs = np.ascontiguousarray(offs, dtype=np.uint64)
        lens = np.ascontiguousarray(lens, dtype=np.uint32)
        if seqs.size == 0:
            seqs = np.zeros(1, dtype=np.uint8)
        lib().orc_match_reads(self.h, _ptr(seqs, _u8p), _ptr(offs, _u64p), _ptr(lens, _u32p), n,
                              int(get_stranded), _ptr(self.counters, _u64p), _ptr(valid, _u8p),
                              _ptr(mat, _u64p), int(nthreads))
        return valid, mat

    def set_pdist(self, minpd, maxpd):
        lib().orc_set_pdist(self.h, int(minpd), int(maxpd))

    def match_pairs(self, r1, r2, get_stranded, nthreads=0):
        """matchPairedReads over pairs given as two (seqs, offs, lens) triples;
        counters5 accumulates succ, nonunique, unsucc, succPaired, tooShort."""
        n = int(len(r1[2]))
        v1 = np.zeros(n, np.uint8)
        v2 = np.zeros(n, np.uint8)
        m1 = np.zeros(n, np.uint64)
        m2 = np.zeros(n, np.uint64)
        a = [np.ascontiguousarray(x, dtype=t) for x, t in zip(r1, (np.uint8, np.uint64, np.uint32))]
        b = [np.ascontiguousarray(x, dtype=t) for x, t in zip(r2, (np.uint8, np.uint64, np.uint32))]
        for arr in (a, b):
            if arr[0].size == 0:
                arr[0] = np.zeros(1, np.uint8)
        if not hasattr(self, "counters5"):
            self.counters5 = np.zeros(5, dtype=np.uint64)
        lib().orc_match_pairs(self.h, _ptr(a[0], _u8p), _ptr(a[1], _u64p), _ptr(a[2], _u32p),
                              _ptr(b[0], _u8p), _ptr(b[1], _u64p), _ptr(b[2], _u32p), n, int(get_stranded),
                              _ptr(self.counters5, _u64p), _ptr(v1, _u8p), _ptr(m1, _u64p), _ptr(v2, _u8p),
                              _ptr(m2, _u64p), int(nthreads))
        return v1, m1, v2, m2

    def decide_strand(self):
        f = C.c_uint64()
        r = C.c_uint64()
        fwd = lib().orc_decide_strand(self.h, C.byref(f), C.byref(r))
        return bool(fwd), int(f.value), int(r.value)

    def meth(self):
        out = np.zeros((self.index.info["ncpg"], 4), dtype=np.uint16)
        lib().orc_get_meth(self.h, _ptr(out, _u16p))
        return out

    def write_tsv(self, path):
        if lib().orc_write_tsv(self.h, path.encode()) != 0:
            raise RuntimeError("cannot write " + path)

    def ub_reads(self):
        return int(lib().orc_ub_reads(self.h))

    def seed_maps(self, seq, cap=1 << 16):
        seq = _bytes(seq)
        fk = np.zeros(cap, np.uint32)
        fc = np.zeros(cap, np.uint16)
        rk = np.zeros(cap, np.uint32)
        rc = np.zeros(cap, np.uint16)
        nf = C.c_int()
        nr = C.c_int()
        bc = np.zeros(1, np.uint32)
        lib().orc_seed_maps(self.h, _ptr(seq, _u8p), seq.size, _ptr(fk, _u32p), _ptr(fc, _u16p),
                            C.byref(nf), _ptr(rk, _u32p), _ptr(rc, _u16p), C.byref(nr), cap,
                            _ptr(bc, _u32p))
        return (fk[:nf.value], fc[:nf.value]), (rk[:nr.value], rc[:nr.value]), int(bc[0])

    def close(self):
        if self.h:
            lib().orc_ctx_free(self.h)
            self.h = None


def match_stream(index, seqs, offs, lens, chunk=300000, both_strands=False, nthreads=0):
    """queryRoutine main.cpp:363-398 over an in-memory read set."""
    m = Matcher(index, both_strands)
    n = len(lens)
    valid = np.zeros(n, np.uint8)
    mat = np.zeros(n, np.uint64)
    s = 0
    if not both_strands:
        e = min(n, chunk)
        valid[s:e], mat[s:e] = m.match(seqs, offs[s:e], lens[s:e], True, nthreads)
        m.decide_strand()
        s = e
    while s < n:
        e = min(n, s + chunk)
        valid[s:e], mat[s:e] = m.match(seqs, offs[s:e], lens[s:e], False, nthreads)
        s = e
    return m, valid, mat


def match_pairs_stream(index, r1, r2, chunk=300000, both_strands=False, minpd=50, maxpd=550, nthreads=0):
    """queryRoutinePaired main.cpp:400-441 over in-memory pairs; r1 / r2 are
    (seqs, offs, lens) triples of the first and second mates."""
    m = Matcher(index, both_strands)
    m.set_pdist(minpd, maxpd)
    n = len(r1[2])
    v1 = np.zeros(n, np.uint8)
    v2 = np.zeros(n, np.uint8)
    m1 = np.zeros(n, np.uint64)
    m2 = np.zeros(n, np.uint64)

    def run(s, e, stranded):
        sub1 = (r1[0], r1[1][s:e], r1[2][s:e])
        sub2 = (r2[0], r2[1][s:e], r2[2][s:e])
        v1[s:e], m1[s:e], v2[s:e], m2[s:e] = m.match_pairs(sub1, sub2, stranded, nthreads)

    s = 0
    if not both_strands:
        e = min(n, chunk)
        run(s, e, True)
        m.decide_strand()
        s = e
    while s < n:
        e = min(n, s + chunk)
        run(s, e, False)
        s = e
    return m, v1, m1, v2, m2


def std_sort_counts(keys, vals, second):
    """std::sort (libstdc++) of (key, packed count) pairs, count greater."""
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    lib().orc_std_sort_counts(_ptr(k, _u32p), _ptr(v, _u32p), k.size, int(second))
    return k, v


def ntps64(kmer, seedbits):
    k = _bytes(kmer)
    return int(lib().orc_ntps64(_ptr(k, _u8p), seedbits))


def ntps64_rolling(seq, seedbits):
    s = _bytes(seq)
    out = np.zeros(max(0, s.size - 31), dtype=np.uint64)
    lib().orc_ntps64_rolling(_ptr(s, _u8p), s.size, seedbits, _ptr(out, _u64p))
    return out


def shiftand(pattern, text, rev=False, cap=4096, E=6):
    """ShiftAnd<E> query: [(position, errors)] as querySeq / queryRevSeq report them."""
    p = _bytes(pattern)
    t = _bytes(text)
    if t.size == 0:
        t = np.zeros(1, np.uint8)
        tl = 0
    else:
        tl = t.size
    pos = np.zeros(cap, np.uint64)
    err = np.zeros(cap, np.uint8)
    tot = C.c_int()
    n = lib().orc_shiftand_e(_ptr(p, _u8p), p.size, _ptr(t, _u8p), tl, int(rev), int(E), _ptr(pos, _u64p),
                             _ptr(err, _u8p), cap, C.byref(tot))
    if n < 0:
        raise ValueError("pattern too long or E > 6")
    return [(int(a), int(b)) for a, b in zip(pos[:n], err[:n])]


def shiftand_state(pattern, E):
    """(masks[4][W], accepted[W], active[E+1][W]) after construction and reset()."""
    p = _bytes(pattern)
    W = p.size // 64 + 1
    masks = np.zeros(4 * W, np.uint64)
    acc = np.zeros(W, np.uint64)
    act = np.zeros((E + 1) * W, np.uint64)
    w = lib().orc_shiftand_state(_ptr(p, _u8p), p.size, int(E), _ptr(masks, _u64p), _ptr(acc, _u64p),
                                 _ptr(act, _u64p))
    if w < 0:
        raise ValueError("pattern too long or E > 6")
    return masks.reshape(4, W), acc, act.reshape(E + 1, W)


CMP_FWD, CMP_REV, CMP_EQ, CMP_WC = 0, 1, 2, 3


def dp_align(row, ref, ref_end, rev=False, cap=4096, band=8, cmp=None):
    """LevenshtDP fill + backtrack: (trace, undefined-read flag, edit distance).
    cmp defaults to CompiRev for rev, CompiFwd otherwise (the matching path)."""
    r = _bytes(row)
    g = _bytes(ref)
    tr = np.zeros(cap, np.uint8)
    ub = C.c_int()
    dist = C.c_int()
    if cmp is None:
        cmp = CMP_REV if rev else CMP_FWD
    n = lib().orc_dp_align_e(_ptr(r, _u8p), r.size, _ptr(g, _u8p), g.size, int(ref_end), int(rev), int(band),
                             int(cmp), _ptr(tr, _u8p), cap, C.byref(ub), C.byref(dist))
    return tr[:n].tolist(), bool(ub.value), int(dist.value)


def hopscotch_order(keys, bucket_count):
    """(keys in iteration order, their counts, final bucket_count())."""
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    out = np.zeros(max(1, k.size), np.uint32)
    vals = np.zeros(max(1, k.size), np.uint32)
    bc = np.zeros(1, np.uint64)
    n = lib().orc_hopscotch_order(_ptr(k, _u32p), k.size, int(bucket_count), _ptr(out, _u32p),
                                  _ptr(vals, _u32p), _ptr(bc, _u64p))
    return out[:n], vals[:n], int(bc[0])
