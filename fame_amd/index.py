"""Index construction into FAME's --load_index format (libfameindex.so).

build_index(chroms, path, ...)      from in-memory chromosome byte arrays
build_index_fasta(fasta, path, ...) from a FASTA file, parsed like FAME's
                                    readReference (no --human_opt)
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfameindex.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m fame_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.fame_build_index.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_char_p), C.c_uint32, C.c_uint32, C.c_uint64,
                                       C.c_uint64, C.c_int, C.c_char_p, C.c_char_p, C.c_size_t]
        L.fame_build_index_fasta.argtypes = [C.c_char_p, C.c_uint32, C.c_uint64, C.c_uint64, C.c_int,
                                             C.c_char_p, C.c_char_p, C.c_size_t]
        _lib = L
    return _lib


def build_index(chroms, path, names=None, readlen=100, htab=1 << 30, kmercutoff=1500,
                noloss=False):
    n = len(chroms)
    names = names or [f"chr{i + 1}" for i in range(n)]
    arrs = [np.ascontiguousarray(c, dtype=np.uint8) for c in chroms]
    ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in arrs])
    lens = (C.c_uint64 * n)(*[a.size for a in arrs])
    nm = (C.c_char_p * n)(*[s.encode() for s in names])
    err = C.create_string_buffer(256)
    rc = _load().fame_build_index(ptrs, lens, nm, n, readlen, htab, kmercutoff, int(noloss),
                                  path.encode(), err, 256)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return path


def build_index_fasta(fasta, path, readlen=100, htab=1 << 30, kmercutoff=1500, noloss=False):
    err = C.create_string_buffer(256)
    rc = _load().fame_build_index_fasta(fasta.encode(), readlen, htab, kmercutoff, int(noloss),
                                        path.encode(), err, 256)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return path
