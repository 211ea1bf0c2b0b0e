"""Seeded synthetic WGBS workloads: reference genomes and bisulfite reads.

Genome model: i.i.d. bases with GC 41 % (A = T = 0.295, C = G = 0.205); CpG
dinucleotides are depleted to about 1 % of positions by turning the G of a
CpG into A.  Written as FASTA with ``>chrN`` headers and 80 columns, the form
FAME's reader parses (RefReader_istr.cpp:25-160).

Read model (directional library): a read is sampled uniformly from the
original top strand (OT) or the original bottom strand (OB, the reverse
complement).  Each genomic CpG carries a methylation level drawn from
Beta(0.5, 0.5); a read C in CpG context stays C with that probability and
becomes T otherwise; every other C becomes T with probability 0.995.  Then
sequencing noise: substitutions at ``sub_rate`` and 1-bp insertions or
deletions at ``indel_rate`` per base; qualities uniform in Phred 30-40.
The bisulfite toss follows the reference's simulator
(Synth/SynthDS.cpp:333-346) in spirit; the generator itself is our own.

Everything is numpy and vectorised so that a 300 Mbp genome and 10 M reads
are produced in tens of seconds.
"""

from __future__ import annotations

import os

import numpy as np

_BASES = np.frombuffer(b"ACGT", dtype=np.uint8)
_COMP = np.zeros(256, dtype=np.uint8)
for _a, _b in zip(b"ACGTN", b"TGCAN"):
    _COMP[_a] = _b


def make_genome(chrom_lengths, seed=1, cpg_target=0.01):
    """Return a list of uint8 arrays (one per chromosome) over b'ACGT'."""
    rng = np.random.default_rng(seed)
    p = np.array([0.295, 0.205, 0.205, 0.295])
    chroms = []
    for n in chrom_lengths:
        seq = _BASES[rng.choice(4, size=int(n), p=p)]
        is_cpg = (seq[:-1] == ord("C")) & (seq[1:] == ord("G"))
        natural = max(float(is_cpg.mean()), 1e-12)
        keep = min(1.0, cpg_target / natural)
        drop = is_cpg & (rng.random(is_cpg.size) >= keep)
        g_idx = np.nonzero(drop)[0] + 1
        seq[g_idx] = ord("A")
        chroms.append(seq)
    return chroms


def add_repeat_families(chroms, frac=0.2, unit=300, divergence=0.10, copies_per_family=200, seed=5):
    """Overwrite a fraction `frac` of the genome with copies of `unit`-bp
    repeat families (in place).  Family consensi are segments of the genome
    itself (same base and CpG model); every copy carries substitutions at
    `divergence` of its positions relative to its consensus.  Copies occupy
    whole `unit`-bp slots, so they never overlap.  Returns the number of
    copies placed."""
    rng = np.random.default_rng(seed)
    slots = [c.size // unit for c in chroms]
    total_slots = sum(slots)
    ncopy = int(frac * total_slots)
    nfam = max(1, ncopy // copies_per_family)
    # consensi from random genome positions
    cons = np.empty((nfam, unit), np.uint8)
    for f in range(nfam):
        c = chroms[int(rng.integers(0, len(chroms)))]
        p = int(rng.integers(0, c.size - unit))
        cons[f] = c[p:p + unit]
    chosen = rng.choice(total_slots, size=ncopy, replace=False)
    fam = rng.integers(0, nfam, size=ncopy)
    first = np.zeros(len(chroms) + 1, np.int64)
    first[1:] = np.cumsum(slots)
    placed = 0
    for ci, c in enumerate(chroms):
        sel = (chosen >= first[ci]) & (chosen < first[ci + 1])
        idx = chosen[sel] - first[ci]
        if idx.size == 0:
            continue
        cp = cons[fam[sel]].copy()
        mut = rng.random(cp.shape) < divergence
        cp[mut] = _BASES[(_CODE[cp[mut]] + rng.integers(1, 4, size=int(mut.sum()))) % 4]
        view = c[:slots[ci] * unit].reshape(slots[ci], unit)
        view[idx] = cp
        placed += idx.size
    return placed


def add_near_copies(chroms, segments=3000, seg_len=(1500, 4000), copies=(5, 50), divergence=0.01, seed=6):
    """Copy `segments` genome segments 5-50 times each (in place) with
    substitutions at `divergence`, so that reads from them have many
    candidate windows.  Returns (chrom, start, end) arrays of every copy and
    of the source segments, for sampling reads there (make_reads regions)."""
    rng = np.random.default_rng(seed)
    rc, rs, re_ = [], [], []
    for _ in range(segments):
        L = int(rng.integers(seg_len[0], seg_len[1] + 1))
        src = int(rng.integers(0, len(chroms)))
        p = int(rng.integers(0, chroms[src].size - L))
        seg = chroms[src][p:p + L].copy()
        rc.append(src)
        rs.append(p)
        re_.append(p + L)
        for _k in range(int(rng.integers(copies[0], copies[1] + 1))):
            dst = int(rng.integers(0, len(chroms)))
            d = int(rng.integers(0, chroms[dst].size - L))
            cp = seg.copy()
            mut = rng.random(L) < divergence
            cp[mut] = _BASES[(_CODE[cp[mut]] + rng.integers(1, 4, size=int(mut.sum()))) % 4]
            chroms[dst][d:d + L] = cp
            rc.append(dst)
            rs.append(d)
            re_.append(d + L)
    return np.array(rc, np.int64), np.array(rs, np.int64), np.array(re_, np.int64)


def write_fasta(path, chroms, names=None, width=80):
    with open(path, "wb") as f:
        for i, seq in enumerate(chroms):
            name = names[i] if names else f"chr{i + 1}"
            f.write(b">" + name.encode() + b"\n")
            b = seq.tobytes()
            for s in range(0, len(b), width):
                f.write(b[s:s + width] + b"\n")


# Phred 30-40 quality lines (qualities do not enter the matching path)
_QUAL = (np.random.default_rng(5).integers(30, 41, size=(4096, 256)) + 33).astype(np.uint8)


def revcomp(a):
    return _COMP[a[..., ::-1]]


class ReadSet:
    """Reads of one length class held as a dense (n, L) uint8 matrix."""

    def __init__(self, seqs, lens, quals=None, truth=None):
        self.seqs = seqs          # (n, Lmax) uint8, rows padded with 0 past lens
        self.lens = lens          # (n,) uint32
        self.quals = quals
        self.truth = truth        # dict of per-read generator truth

    def __len__(self):
        return int(self.lens.shape[0])

    def flat(self):
        """Concatenated bytes plus u64 offsets / u32 lengths (the C-ABI input form)."""
        n = len(self)
        lens = self.lens.astype(np.uint64)
        offs = np.zeros(n, dtype=np.uint64)
        if n:
            offs[1:] = np.cumsum(lens)[:-1]
        if n and np.all(self.lens == self.seqs.shape[1]):
            buf = np.ascontiguousarray(self.seqs).reshape(-1)
        else:
            mask = np.arange(self.seqs.shape[1])[None, :] < self.lens[:, None]
            buf = self.seqs[mask]
        return np.ascontiguousarray(buf, dtype=np.uint8), offs, self.lens.astype(np.uint32)

    def write_fastq(self, path, start=0, stop=None, name_prefix="r"):
        stop = len(self) if stop is None else stop
        with open(path, "wb") as f:
            for i in range(start, stop):
                L = int(self.lens[i])
                s = self.seqs[i, :L].tobytes()
                q = (self.quals[i, :L] if self.quals is not None
                     else _QUAL[i % _QUAL.shape[0], :L]).tobytes()
                f.write(b"@" + f"{name_prefix}{i}".encode() + b"\n" + s + b"\n+\n" + q + b"\n")

    def subset(self, idx):
        t = None if self.truth is None else {k: v[idx] for k, v in self.truth.items()}
        return ReadSet(self.seqs[idx], self.lens[idx],
                       None if self.quals is None else self.quals[idx], t)


def meth_levels(chroms, seed=7):
    """Per-position methylation level arrays (only CpG C positions are used)."""
    rng = np.random.default_rng(seed)
    out = []
    for seq in chroms:
        lv = np.zeros(seq.size, dtype=np.float32)
        c = np.nonzero((seq[:-1] == ord("C")) & (seq[1:] == ord("G")))[0]
        lv[c] = rng.beta(0.5, 0.5, size=c.size).astype(np.float32)
        out.append(lv)
    return out


_CODE = np.full(256, 0, dtype=np.uint8)
for _i, _b in enumerate(b"ACGT"):
    _CODE[_b] = _i


def make_reads(chroms, n, read_len, seed=3, levels=None, sub_rate=0.005, indel_rate=0.0005,
               len_range=None, ob_fraction=0.5, batch=1 << 18, regions=None,
               non_cpg_conv=0.995, threads=None, origins=None, lengths=None):
    """Sample `n` bisulfite reads.

    read_len   nominal (maximum) length; with `len_range=(lo, hi)` each read's
               length is uniform in [lo, hi] (hi <= read_len).
    regions    optional (chrom_idx, start, end) arrays restricting the origin
               of every read (used for repeat-heavy mixes).
    origins    optional (chrom_idx, pos, ob) arrays fixing every read's origin:
               an OT read covers [pos, pos + L), an OB read is the reverse
               complement of [pos + read_len + 8 - L, pos + read_len + 8);
               lengths then gives each read's L.
    Batches are generated on a thread pool, each from its own seeded stream,
    so the output depends only on the arguments.
    """
    from numpy.lib.stride_tricks import sliding_window_view

    if levels is None:
        levels = meth_levels(chroms, seed=seed + 11)
    lens_c = np.array([c.size for c in chroms], dtype=np.int64)
    coff = np.zeros(len(chroms), np.int64)
    coff[1:] = np.cumsum(lens_c)[:-1]
    G = np.concatenate(chroms)
    LV = np.concatenate(levels)
    pad = read_len + 8
    Gw = sliding_window_view(G, pad + 2)
    Lw = sliding_window_view(LV, pad + 2)
    seqs = np.zeros((n, read_len), dtype=np.uint8)
    lens = np.zeros(n, dtype=np.uint32)
    t_chrom = np.zeros(n, dtype=np.int32)
    t_pos = np.zeros(n, dtype=np.int64)
    t_ob = np.zeros(n, dtype=bool)
    def one(s):
        rng = np.random.default_rng([seed, s])
        m = min(batch, n - s)
        if origins is not None:
            ci = np.asarray(origins[0][s:s + m], np.int64)
            pos = np.asarray(origins[1][s:s + m], np.int64)
        elif regions is None:
            ci = rng.choice(lens_c.size, size=m, p=lens_c / lens_c.sum())
            pos = (rng.random(m) * (lens_c[ci] - pad - 2)).astype(np.int64) + 1
        else:
            rc, rs, re_ = regions
            k = rng.integers(0, rc.size, size=m)
            ci = rc[k]
            span = np.maximum(re_[k] - rs[k] - pad - 2, 1)
            pos = rs[k] + (rng.random(m) * span).astype(np.int64) + 1
        ob = rng.random(m) < ob_fraction
        if origins is not None:
            ob = np.asarray(origins[2][s:s + m], bool)
        if lengths is not None:
            L = np.asarray(lengths[s:s + m], np.int64)
        elif len_range is None:
            L = np.full(m, read_len, dtype=np.int64)
        else:
            L = rng.integers(len_range[0], len_range[1] + 1, size=m)
        # genome span [pos-1, pos+pad] of every read (one base of context on
        # both sides for CpG calls at the read ends)
        g0 = coff[ci] + pos - 1
        win = Gw[g0]
        lvl = Lw[g0]
        # OT: bases win[1 : 1+pad], next base win[2 : 2+pad].  OB: the reverse
        # complement of the same span; a bottom-strand CpG C is the complement
        # of the G of a top-strand CpG whose C sits one position to the left,
        # so its level is read there.
        read = np.where(ob[:, None], revcomp(win[:, 1:1 + pad]), win[:, 1:1 + pad])
        nxt = np.where(ob[:, None], revcomp(win[:, 0:pad]), win[:, 2:2 + pad])
        lv = np.where(ob[:, None], lvl[:, 0:pad][:, ::-1], lvl[:, 1:1 + pad])
        # bisulfite: CpG C kept with the CpG's level, other C -> T at non_cpg_conv
        cpos = np.flatnonzero(read == ord("C"))
        u = rng.random(cpos.size, dtype=np.float32)
        in_cpg = nxt.reshape(-1)[cpos] == ord("G")
        conv = np.where(in_cpg, u >= lv.reshape(-1)[cpos], u < non_cpg_conv)
        flat = read.reshape(-1)
        flat[cpos[conv]] = ord("T")
        # substitutions: Binomial count, uniform positions, a different base
        ns = rng.binomial(flat.size, sub_rate)
        sp = rng.integers(0, flat.size, size=ns)
        flat[sp] = _BASES[(_CODE[flat[sp]] + rng.integers(1, 4, size=ns)) % 4]
        # at most one 1-bp indel per read, applied by a gather
        has = rng.random(m) < 1.0 - (1.0 - indel_rate) ** pad
        p_ev = np.where(has, rng.integers(0, read_len, size=m), pad + 1)[:, None]
        is_del = (rng.random(m) < 0.5)[:, None]
        j = np.arange(read_len)[None, :]
        src = np.where(is_del, j + (j >= p_ev), j - (j > p_ev))
        out = np.take_along_axis(read, np.minimum(src, pad - 1), axis=1)
        rows = np.flatnonzero(has & ~is_del[:, 0])
        out[rows, p_ev[rows, 0]] = _BASES[rng.integers(0, 4, size=rows.size)]
        if len_range is not None or lengths is not None:
            out[j >= L[:, None]] = 0
        seqs[s:s + m] = out
        lens[s:s + m] = L
        t_chrom[s:s + m] = ci
        t_pos[s:s + m] = pos
        t_ob[s:s + m] = ob

    import concurrent.futures as cf
    with cf.ThreadPoolExecutor(threads or min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(one, range(0, n, batch)))
    quals = None
    return ReadSet(seqs, lens, quals, {"chrom": t_chrom, "pos": t_pos, "ob": t_ob})


def make_pairs(chroms, n, read_len, gap=(20, 400), seed=3, len_range=None, ob_fraction=0.5, **kw):
    """Sample `n` read pairs of a directional bisulfite library.

    A pair comes from a fragment of length L1 + g + L2 (g uniform in `gap`,
    mate lengths uniform in `len_range`, else read_len) on the converted
    original top (OT) or bottom (OB) strand: mate 1 is the fragment's first
    L1 letters, mate 2 the reverse complement of its last L2 letters.  Both
    mates are drawn by `make_reads` at fixed origins (same bisulfite, error
    and CpG level model; the two ends of a fragment do not overlap, so their
    conversions are independent).  Returns (ReadSet mate 1, ReadSet mate 2);
    truth holds the fragment's chromosome, top-strand start, strand and g.
    """
    rng = np.random.default_rng([seed, 17])
    pad = read_len + 8
    lens_c = np.array([c.size for c in chroms], dtype=np.int64)
    if len_range is None:
        l1 = np.full(n, read_len, np.int64)
        l2 = np.full(n, read_len, np.int64)
    else:
        l1 = rng.integers(len_range[0], len_range[1] + 1, size=n)
        l2 = rng.integers(len_range[0], len_range[1] + 1, size=n)
    g = rng.integers(gap[0], gap[1] + 1, size=n)
    F = l1 + g + l2
    ci = rng.choice(lens_c.size, size=n, p=lens_c / lens_c.sum())
    span = 2 * read_len + gap[1] + 2 * pad + 8
    p = (rng.random(n) * (lens_c[ci] - span)).astype(np.int64) + pad + 2
    ob = rng.random(n) < ob_fraction
    e = p + F
    # mate 1: OT -> [p, p + L1) top; OB -> revcomp of [e - L1, e)
    pos1 = np.where(ob, e - pad, p)
    # mate 2 = revcomp of X: OT X = converted top [e - L2, e); OB X = the OB
    # read of [p, p + L2)
    pos2 = np.where(ob, p + l2 - pad, e - l2)
    a = make_reads(chroms, n, read_len, seed=seed * 7919 + 1, origins=(ci, pos1, ob), lengths=l1, **kw)
    x = make_reads(chroms, n, read_len, seed=seed * 7919 + 2, origins=(ci, pos2, ob), lengths=l2, **kw)
    m2 = np.zeros_like(x.seqs)
    j = np.arange(read_len)[None, :]
    src = np.clip(l2[:, None] - 1 - j, 0, read_len - 1)
    m2[:] = _COMP[np.take_along_axis(x.seqs, src, axis=1)]
    m2[j >= l2[:, None]] = 0
    truth = {"chrom": ci.astype(np.int32), "pos": p, "ob": ob, "gap": g}
    return (ReadSet(a.seqs, l1.astype(np.uint32), None, truth),
            ReadSet(m2, l2.astype(np.uint32), None, dict(truth)))


def config_genome(name, scale=1.0, seed=1):
    """Chromosome lengths for the named workloads.

    small   : 2 chromosomes x 5 Mbp (the reference's CPU-runnable case)
    genome  : 24 chromosomes, 300 Mbp in total
    """
    if name == "small":
        return [int(5_000_000 * scale)] * 2
    if name == "genome":
        rng = np.random.default_rng(seed + 100)
        w = rng.uniform(0.5, 1.5, size=24)
        L = (w / w.sum() * 300_000_000 * scale).astype(np.int64)
        return [int(x) for x in L]
    raise ValueError(name)
