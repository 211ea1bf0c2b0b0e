"""ctypes binding of libfamehip.so (the C ABI of include/fame_hip.h).

The library must have been built in-tree (python -m fame_amd.build); a missing
library is an error, never a silent fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfamehip.so")

FAME_OK = 0
ERRORS = {-1: "FAME_ERR_IO", -2: "FAME_ERR_FORMAT", -3: "FAME_ERR_HIP", -4: "FAME_ERR_ARG",
          -5: "FAME_ERR_CAPACITY"}
INFO_FIELDS = ("htab", "readlen", "winlen", "kmerlen", "kmercutoff", "seedbits", "ncpg", "nchrom",
               "nkmer", "nwin", "nfiltered")
MAX_READ = 4096  # single-end (FAME_MAX_READ)
MAX_MATE = 255  # paired mates, the allocation-free entry point (FAME_MAX_MATE)

# every entry point declared in include/fame_hip.h
EXPORTS = ("fame_open", "fame_close", "fame_last_error", "fame_index_info", "fame_match_reads",
           "fame_match_reads_device", "fame_match_reads_noalloc", "fame_set_pair_dist", "fame_match_pairs", "fame_match_pairs_device",
           "fame_decide_strand", "fame_get_meth", "fame_reset", "fame_set_form", "fame_set_pipeline", "fame_form_info", "fame_serve_open", "fame_serve_submit",
           "fame_serve_wait", "fame_serve_stats", "fame_serve_close", "fame_fastq_open", "fame_fastq_next",
           "fame_fastq_close", "fame_sc_open", "fame_sc_write_cell", "fame_sc_close",
           "fame_print_methylation_levels", "fame_last_timings", "fame_last_stats")

_u8p = C.POINTER(C.c_uint8)
_u16p = C.POINTER(C.c_uint16)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_dbl = C.POINTER(C.c_double)

_lib = None


class FameError(RuntimeError):
    pass


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FameError(f"{LIB_PATH} is not built; run `python -m fame_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.fame_open.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]
        L.fame_close.argtypes = [C.c_void_p]
        L.fame_last_error.restype = C.c_char_p
        L.fame_last_error.argtypes = [C.c_void_p]
        L.fame_index_info.argtypes = [C.c_void_p, _u64p]
        L.fame_match_reads.argtypes = [C.c_void_p, _u8p, _u64p, _u32p, C.c_uint32, C.c_int, _u64p,
                                       _u8p, _u64p]
        L.fame_match_reads_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_uint32, C.c_int, _u64p, C.c_void_p, C.c_void_p,
                                              C.c_void_p]
        L.fame_match_reads_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.c_uint32, C.c_int, _u64p, C.c_void_p, C.c_void_p,
                                               C.c_void_p]
        L.fame_set_pair_dist.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
        L.fame_match_pairs.argtypes = [C.c_void_p, _u8p, _u64p, _u32p, _u8p, _u64p, _u32p, C.c_uint32,
                                       C.c_int, _u64p, _u8p, _u64p, _u8p, _u64p]
        L.fame_match_pairs_device.argtypes = [C.c_void_p] + [C.c_void_p] * 5 + [C.c_uint32, C.c_int, _u64p] + \
            [C.c_void_p] * 5
        L.fame_decide_strand.argtypes = [C.c_void_p, _u64p, _u64p, C.POINTER(C.c_int)]
        L.fame_set_form.argtypes = [C.c_void_p, C.c_int]
        L.fame_form_info.argtypes = [C.c_void_p, _u64p]
        L.fame_set_pipeline.argtypes = [C.c_void_p, C.c_uint32]
        L.fame_serve_open.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)]
        L.fame_serve_submit.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                        C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]
        L.fame_serve_wait.argtypes = [C.c_void_p, C.c_uint64, _u64p, _dbl]
        L.fame_serve_stats.argtypes = [C.c_void_p, _u64p, _dbl]
        L.fame_serve_close.argtypes = [C.c_void_p]
        L.fame_fastq_open.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]
        L.fame_fastq_next.argtypes = [C.c_void_p, C.c_uint32, _u32p, C.POINTER(_u8p), C.POINTER(_u64p),
                                      C.POINTER(_u32p)]
        L.fame_fastq_close.argtypes = [C.c_void_p]
        L.fame_sc_open.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p)]
        L.fame_sc_write_cell.argtypes = [C.c_void_p, C.c_char_p]
        L.fame_sc_close.argtypes = [C.c_void_p]
        L.fame_get_meth.argtypes = [C.c_void_p, _u16p, C.c_uint64]
        L.fame_reset.argtypes = [C.c_void_p]
        L.fame_print_methylation_levels.argtypes = [C.c_void_p, C.c_char_p]
        L.fame_last_timings.argtypes = [C.c_void_p, _dbl]
        L.fame_last_stats.argtypes = [C.c_void_p, _u64p]
        _lib = L
    return _lib


def ptr(a, t):
    return a.ctypes.data_as(t)


class Context:
    """An index resident in HBM plus the matching state of one ReadQueue."""

    def __init__(self, index_path, both_strands=False):
        L = load()
        h = C.c_void_p()
        rc = L.fame_open(index_path.encode(), int(both_strands), C.byref(h))
        self.h = h.value
        if rc != FAME_OK:
            msg = L.fame_last_error(self.h).decode() if self.h else "fame_open failed"
            self.close()
            raise FameError(f"{ERRORS.get(rc, rc)}: {msg}")
        info = np.zeros(11, np.uint64)
        self._check(L.fame_index_info(self.h, ptr(info, _u64p)))
        self.info = dict(zip(INFO_FIELDS, (int(x) for x in info)))

    def _check(self, rc):
        if rc != FAME_OK:
            msg = load().fame_last_error(self.h).decode()
            raise FameError(f"{ERRORS.get(rc, rc)}: {msg}")

    def match(self, seqs, offs, lens, get_stranded, counters):
        n = int(len(lens))
        valid = np.zeros(max(n, 1), np.uint8)
        mat = np.zeros(max(n, 1), np.uint64)
        seqs = np.ascontiguousarray(seqs, np.uint8)
        if seqs.size == 0:
            seqs = np.zeros(1, np.uint8)
        offs = np.ascontiguousarray(offs, np.uint64)
        lens = np.ascontiguousarray(lens, np.uint32)
        self._check(load().fame_match_reads(self.h, ptr(seqs, _u8p), ptr(offs, _u64p),
                                            ptr(lens, _u32p), n, int(get_stranded),
                                            ptr(counters, _u64p), ptr(valid, _u8p), ptr(mat, _u64p)))
        return valid[:n], mat[:n]

    def match_device(self, d_seqs, d_offs, d_lens, n, get_stranded, counters, d_valid=0, d_mat=0,
                     stream=0):
        self._check(load().fame_match_reads_device(self.h, d_seqs, d_offs, d_lens, int(n),
                                                   int(get_stranded), ptr(counters, _u64p),
                                                   d_valid or None, d_mat or None, stream or None))

    def match_noalloc(self, d_seqs, d_offs, d_lens, n, get_stranded, counters, d_valid=0, d_mat=0, stream=0):
        """fame_match_reads_noalloc: match_device without any device allocation
        (at most 32768 reads per call, read ends within 8 MiB of d_seqs)"""
        self._check(load().fame_match_reads_noalloc(self.h, d_seqs, d_offs, d_lens, int(n), int(get_stranded),
                                                    ptr(counters, _u64p), d_valid or None, d_mat or None,
                                                    stream or None))

    def set_pair_dist(self, minpd, maxpd):
        self._check(load().fame_set_pair_dist(self.h, int(minpd), int(maxpd)))

    def match_pairs(self, r1, r2, get_stranded, counters):
        """matchPairedReads over pairs (mate 1 triple r1, mate 2 triple r2 of
        (seqs, offs, lens)); counters (5 x u64) accumulate."""
        n = int(len(r1[2]))
        if len(r2[2]) != n:
            raise FameError("mate files hold different numbers of reads")
        v1 = np.zeros(max(n, 1), np.uint8)
        v2 = np.zeros(max(n, 1), np.uint8)
        m1 = np.zeros(max(n, 1), np.uint64)
        m2 = np.zeros(max(n, 1), np.uint64)
        a = []
        for seqs, offs, lens in (r1, r2):
            seqs = np.ascontiguousarray(seqs, np.uint8)
            if seqs.size == 0:
                seqs = np.zeros(1, np.uint8)
            a.append((seqs, np.ascontiguousarray(offs, np.uint64), np.ascontiguousarray(lens, np.uint32)))
        self._check(load().fame_match_pairs(self.h, ptr(a[0][0], _u8p), ptr(a[0][1], _u64p),
                                            ptr(a[0][2], _u32p), ptr(a[1][0], _u8p), ptr(a[1][1], _u64p),
                                            ptr(a[1][2], _u32p), n, int(get_stranded),
                                            ptr(counters, _u64p), ptr(v1, _u8p), ptr(m1, _u64p),
                                            ptr(v2, _u8p), ptr(m2, _u64p)))
        return v1[:n], m1[:n], v2[:n], m2[:n]

    def match_pairs_device(self, d_seqs, d_offs1, d_lens1, d_offs2, d_lens2, n, get_stranded, counters,
                           d_valid1=0, d_mat1=0, d_valid2=0, d_mat2=0, stream=0):
        self._check(load().fame_match_pairs_device(self.h, d_seqs, d_offs1, d_lens1, d_offs2, d_lens2, int(n),
                                                   int(get_stranded), ptr(counters, _u64p), d_valid1 or None,
                                                   d_mat1 or None, d_valid2 or None, d_mat2 or None,
                                                   stream or None))

    def set_form(self, form):
        """"plain" (one kernel per stage) or "batched" (one graph launch per batch)"""
        self._check(load().fame_set_form(self.h, {"plain": 0, "batched": 1}[form]))

    def set_pipeline(self, sub_batch_reads):
        self._check(load().fame_set_pipeline(self.h, int(sub_batch_reads)))

    def form_info(self):
        t = np.zeros(9, np.uint64)
        self._check(load().fame_form_info(self.h, ptr(t, _u64p)))
        return dict(zip(("form", "graph_launches", "fallbacks", "graphs", "classes", "verify_blocks",
                         "pipelined_calls", "sub_batch", "graphs_held"), (int(x) for x in t)))

    def decide_strand(self):
        f = C.c_uint64()
        r = C.c_uint64()
        m = C.c_int()
        self._check(load().fame_decide_strand(self.h, C.byref(f), C.byref(r), C.byref(m)))
        return bool(m.value), int(f.value), int(r.value)

    def meth(self):
        out = np.zeros((self.info["ncpg"], 4), np.uint16)
        if self.info["ncpg"]:
            self._check(load().fame_get_meth(self.h, ptr(out, _u16p), self.info["ncpg"]))
        return out

    def reset(self):
        self._check(load().fame_reset(self.h))

    def print_methylation_levels(self, basename):
        self._check(load().fame_print_methylation_levels(self.h, basename.encode()))

    def timings(self):
        t = np.zeros(7, np.float64)
        self._check(load().fame_last_timings(self.h, ptr(t, _dbl)))
        return dict(zip(("prep", "seed", "verify", "select", "meth", "total", "dp"), t.tolist()))

    def stats(self):
        t = np.zeros(13, np.uint64)
        self._check(load().fame_last_stats(self.h, ptr(t, _u64p)))
        return dict(zip(("kmers", "entries", "seeded", "jobs", "window_bytes", "dp", "read_bytes",
                         "tab_bytes", "bucket_entries", "dp_rows", "word_letters", "random_slots",
                         "ordered_slots"), (int(x) for x in t)))

    def close(self):
        if getattr(self, "h", None):
            load().fame_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop over a Context (fame_serve_*): submit chunks of reads as
    they arrive, wait for each one's results.  The chunk arrays passed to
    submit are kept alive here until the chunk completes."""

    def __init__(self, ctx, hbm_budget=0, max_batch_reads=0):
        self.ctx = ctx
        h = C.c_void_p()
        ctx._check(load().fame_serve_open(ctx.h, int(hbm_budget), int(max_batch_reads), C.byref(h)))
        self.h = h.value
        self._live = {}

    def submit(self, seqs, offs, lens):
        n = int(len(lens))
        seqs = np.ascontiguousarray(seqs, np.uint8)
        if seqs.size == 0:
            seqs = np.zeros(1, np.uint8)
        offs = np.ascontiguousarray(offs, np.uint64)
        lens = np.ascontiguousarray(lens, np.uint32)
        valid = np.zeros(max(n, 1), np.uint8)
        mat = np.zeros(max(n, 1), np.uint64)
        t = C.c_uint64()
        rc = load().fame_serve_submit(self.h, seqs.ctypes.data, offs.ctypes.data, lens.ctypes.data, n,
                                      valid.ctypes.data, mat.ctypes.data, C.byref(t))
        self.ctx._check(rc)
        self._live[t.value] = (seqs, offs, lens, valid, mat, n)
        return t.value

    def wait(self, ticket):
        """(valid, mat, counters[3], times_ms {queued, gpu, latency, batch_reads})"""
        cnt = np.zeros(3, np.uint64)
        tm = np.zeros(4, np.float64)
        rc = load().fame_serve_wait(self.h, ticket, ptr(cnt, _u64p), ptr(tm, _dbl))
        keep = self._live.pop(ticket)
        self.ctx._check(rc)
        n = keep[5]
        return keep[3][:n], keep[4][:n], cnt, dict(zip(("queued", "gpu", "latency", "batch_reads"), tm.tolist()))

    def stats(self):
        out = np.zeros(4, np.uint64)
        busy = C.c_double()
        load().fame_serve_stats(self.h, ptr(out, _u64p), C.byref(busy))
        return dict(zip(("batches", "requests", "reads", "max_batch_reads"), (int(x) for x in out)),
                    busy_ms=busy.value)

    def close(self):
        if getattr(self, "h", None):
            load().fame_serve_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FastqReader:
    """Native FASTQ parser (fame_fastq_*): next(k) -> (seqs, offs, lens) of up
    to k records as numpy copies, empty arrays at the end of the file."""

    def __init__(self, path, gz=False):
        h = C.c_void_p()
        rc = load().fame_fastq_open(path.encode(), int(gz), C.byref(h))
        if rc != FAME_OK:
            raise FameError(f"{ERRORS.get(rc, rc)}: cannot open {path}")
        self.h = h.value

    def next(self, max_reads):
        n = C.c_uint32()
        s, o, l = _u8p(), _u64p(), _u32p()
        rc = load().fame_fastq_next(self.h, int(max_reads), C.byref(n), C.byref(s), C.byref(o), C.byref(l))
        if rc != FAME_OK:
            raise FameError(f"{ERRORS.get(rc, rc)}: FASTQ parse failed")
        k = n.value
        if k == 0:
            return np.zeros(0, np.uint8), np.zeros(0, np.uint64), np.zeros(0, np.uint32)
        offs = np.ctypeslib.as_array(o, shape=(k,)).copy()
        lens = np.ctypeslib.as_array(l, shape=(k,)).copy()
        total = int(offs[-1] + lens[-1])
        seqs = np.ctypeslib.as_array(s, shape=(max(total, 1),))[:total].copy()
        return seqs, offs, lens

    def close(self):
        if getattr(self, "h", None):
            load().fame_fastq_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ScMatrix:
    """Single-cell count matrix (fame_sc_*) written from a Context's counts."""

    def __init__(self, ctx, path):
        h = C.c_void_p()
        ctx._check(load().fame_sc_open(ctx.h, path.encode(), C.byref(h)))
        self.ctx, self.h = ctx, h.value

    def write_cell(self, sc_id):
        self.ctx._check(load().fame_sc_write_cell(self.h, sc_id.encode()))

    def close(self):
        if self.h:
            rc = load().fame_sc_close(self.h)
            self.h = None
            if rc != FAME_OK:
                raise FameError(f"{ERRORS.get(rc, rc)}: single-cell output")
