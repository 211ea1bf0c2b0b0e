"""Writes tests/golden/units/*.npz: input/output vectors of FAME's building
blocks, produced by FAME's own code (the reference harness's `units` mode,
built from the reference sources by oracle/Makefile):

    ntps64     ntHash::NTPS64, direct and rolling (spaced_nthash/nthash.hpp:480-498)
    shiftand   ShiftAnd<E>::querySeq / queryRevSeq (ShiftAnd.h:130-255), E = 0, 1, 2, 6
    dp         LevenshtDP<uint16_t, band> fill + backtrack (LevenshtDP.h:207-544) with
               CompiFwd / CompiRev (ReadQueue.h:295-338) at band 8, and the equality /
               complement comparators at bands 0-2
    hopscotch  iteration order of tsl::hopscotch_map<uint32_t, uint16_t, MetaHash>

    python tests/golden/make_units.py

Inputs are seeded random cases; each .npz holds the packed input stream and
the reference's packed output stream, which tests/test_oracle_units.py
decodes.
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
from make_golden import CASES, ref_binary  # noqa: E402

ALPHA = b"ACGT"
COMP = bytes.maketrans(b"ACGTN", b"TGCAN")


def rand_seq(rng, n, alpha=ALPHA):
    return bytes(np.frombuffer(alpha, np.uint8)[rng.integers(0, len(alpha), n)])


def mutate(rng, s, nsub, nindel, bisulfite=False):
    s = bytearray(s)
    if bisulfite:
        for i in range(len(s)):
            if s[i] == ord("C") and rng.random() < 0.9:
                s[i] = ord("T")
    for _ in range(nsub):
        if s:
            s[rng.integers(0, len(s))] = ALPHA[rng.integers(0, 4)]
    for _ in range(nindel):
        if rng.random() < 0.5 and len(s) > 1:
            del s[rng.integers(0, len(s))]
        else:
            s.insert(int(rng.integers(0, len(s) + 1)), ALPHA[rng.integers(0, 4)])
    return bytes(s)


def ntps64_cases(rng):
    seqs = [rand_seq(rng, int(n)) for n in rng.integers(32, 300, 200)]
    seqs += [b"A" * 40, b"C" * 33, b"ACGT" * 16, b"T" * 32]
    buf = struct.pack("<I", len(seqs)) + b"".join(struct.pack("<I", len(s)) + s for s in seqs)
    return buf


def shiftand_cases(rng):
    out = []
    lens = [20, 31, 32, 50, 62, 63, 64, 65, 100, 126, 127, 128, 129, 150, 191, 192, 193, 255, 300]
    for _ in range(500):
        L = int(rng.choice(lens))
        E = int(rng.choice([0, 1, 2, 6, 6, 6]))
        rev = int(rng.integers(0, 2))
        text = bytearray(rand_seq(rng, int(rng.integers(L + 10, 2100))))
        # plant a (mutated, bisulfite-converted) copy of the pattern 1-3 times
        pat = rand_seq(rng, L)
        for _ in range(int(rng.integers(1, 4))):
            copy = mutate(rng, pat, int(rng.integers(0, 5)), int(rng.integers(0, 3)))
            if rev:
                copy = copy.translate(COMP)[::-1]
            at = int(rng.integers(0, max(1, len(text) - len(copy))))
            text[at:at + len(copy)] = copy
        # the pattern sees the converted read: C -> T mostly
        pat_bs = mutate(rng, pat, 0, 0, bisulfite=rng.random() < 0.5)
        if rng.random() < 0.15:
            for _ in range(int(rng.integers(1, 4))):
                text[int(rng.integers(0, len(text)))] = ord("N")
        out.append((E, rev, pat_bs, bytes(text)))
    buf = struct.pack("<I", len(out))
    for E, rev, p, t in out:
        buf += struct.pack("<BBI", E, rev, len(p)) + p + struct.pack("<I", len(t)) + t
    return buf


def dp_cases(rng):
    out = []
    for _ in range(600):
        small = rng.random() < 0.3
        band = int(rng.choice([0, 1, 2])) if small else 8
        L = int(rng.integers(max(band, 2), 12)) if small else int(rng.integers(30, 256))
        rev = int(rng.integers(0, 2))
        cmp = (3 if rev else 2) if small else rev  # CompiRev goes with the reverse fill
        row = rand_seq(rng, L)
        seg = mutate(rng, row, int(rng.integers(0, 4)), int(rng.integers(0, 4)))
        if not small:
            # the reference holds C where the bisulfite read holds T
            seg = bytes(b if b != ord("T") or rng.random() < 0.3 else ord("C") for b in seg)
        if rev:
            seg = seg.translate(COMP)[::-1]
        if rng.random() < 0.1:
            seg = rand_seq(rng, len(seg))
        pad = band + 4
        ref = rand_seq(rng, pad + L + 2 * band) + seg + rand_seq(rng, pad)
        end = pad + L + 2 * band + len(seg) - 1  # colPat = ref + end, the segment's last letter
        out.append((band, rev, cmp, row, ref, end))
    buf = struct.pack("<I", len(out))
    for band, rev, cmp, row, ref, end in out:
        buf += struct.pack("<BBBI", band, rev, cmp, len(row)) + row + struct.pack("<I", len(ref)) + ref
        buf += struct.pack("<Q", end)
    return buf


def hopscotch_cases(rng):
    out = []
    for bc in (0, 1, 2, 3, 4, 16, 100, 128, 1024):
        for n in (1, 5, 40, 300, 3000):
            out.append((bc, rng.integers(0, 4 * n + 8, n).astype(np.uint32)))
    # windows of a read's hits: runs of neighbouring ids with repeats
    for _ in range(20):
        base = int(rng.integers(0, 1 << 20))
        out.append((0, (base + rng.integers(0, 200, 2000)).astype(np.uint32)))
    # ids that share low bits: full neighbourhoods, growth, the overflow list
    for step in (64, 1 << 12, 1 << 20):
        out.append((0, (np.arange(200, dtype=np.uint64) * step % (1 << 32)).astype(np.uint32)))
        out.append((128, (np.arange(90, dtype=np.uint64) * step % (1 << 32)).astype(np.uint32)))
    buf = struct.pack("<I", len(out))
    for bc, keys in out:
        buf += struct.pack("<QI", bc, keys.size) + keys.astype("<u4").tobytes()
    return buf


def main():
    binp = ref_binary(CASES["se_basic"])
    out = os.path.join(HERE, "units")
    os.makedirs(out, exist_ok=True)
    rng = np.random.default_rng(2024)
    gens = {"ntps64": ntps64_cases, "shiftand": shiftand_cases, "dp": dp_cases, "hopscotch": hopscotch_cases}
    with tempfile.TemporaryDirectory() as td:
        for kind, gen in gens.items():
            inp = gen(rng)
            with open(os.path.join(td, "in.bin"), "wb") as f:
                f.write(inp)
            subprocess.run([binp, "units", kind, "in.bin", "out.bin"], cwd=td, check=True)
            res = open(os.path.join(td, "out.bin"), "rb").read()
            np.savez_compressed(os.path.join(out, kind + ".npz"), input=np.frombuffer(inp, np.uint8),
                                output=np.frombuffer(res, np.uint8))
            print(kind, len(inp), "->", len(res), "bytes")


if __name__ == "__main__":
    main()
