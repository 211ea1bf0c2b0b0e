"""FAME compile-time configurations the drop-in binary is built for: those of
the golden fixtures (each --load_index file only loads into a FAME built with
its own constants, RefGenome.cpp:239-280)."""

# HTAB, READLEN, CHROMNUM, MINPD, MAXPD, CHUNK, CUT
CONFIGS = [
    (65536, 100, 2, 50, 550, 300000, 1500),    # se_basic / se_long / pe_basic / sc_pe
    (32768, 100, 2, 50, 550, 700, 1500),       # se_ga_chunked / pe_ga_chunked / sc_se (700-read chunks)
    (16384, 100, 2, 50, 550, 300000, 24),      # se_filt / se_noloss / pe_filt / pe_noloss (KMERCUTOFF 24)
    (16384, 100, 3, 50, 550, 300000, 1500),    # se_repeats / se_unord / pe_repeats / pe_unord
    (32768, 150, 2, 50, 550, 300000, 1500),    # se_rl150
    (32768, 150, 2, 20, 400, 300000, 1500),    # pe_rl150
]


def name(htab, readlen, nchrom, minpd, maxpd, chunk, cut):
    return f"fame_hip_h{htab}_r{readlen}_n{nchrom}_p{minpd}-{maxpd}_k{chunk}_x{cut}"


def make_args(cfg):
    htab, readlen, nchrom, minpd, maxpd, chunk, cut = cfg
    return [f"HTAB={htab}", f"READLEN={readlen}", f"CHROMNUM={nchrom}", f"MINPD={minpd}", f"MAXPD={maxpd}",
            f"CHUNK={chunk}", f"CUT={cut}"]
