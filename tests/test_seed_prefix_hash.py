"""The seed kernel's prefix form of the spaced-seed hash (go_hash in
fame_device.hip) against FAME's own ntHash::NTPS64 outputs (vectors written by
the reference's code, tests/golden/units/ntps64.npz).

With Q[t] = rol(seed(c_t), -t) and P[j] = Q[0] ^ ... ^ Q[j-1], the hash of
the k-mer at k is rol(XOR of P[k + e] over the ends e of the runs of care
positions of SEEDBITS, 31 + k) -- the form the kernel evaluates with ten LDS
reads per k-mer.  A dropped run end, a rotation off by one or a wrong seed
for a letter each change most of these hashes."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))
from test_oracle_units import SEEDBITS, _load  # noqa: E402

M64 = (1 << 64) - 1
K = 32
# nthash.hpp:24-27 as the reduced table uses them: A, G, and T (C hashes as T)
SEEDS = {ord("A"): 0x3c8bfbb395c60474, ord("G"): 0x20323ed082572324, ord("T"): 0x295549f54be24456,
         ord("C"): 0x295549f54be24456}


def rol(x, r):
    r &= 63
    return ((x << r) | (x >> (64 - r))) & M64 if r else x


def run_ends(bits):
    ends, prev = [], False
    for i in range(K + 1):
        cur = i < K and ((bits >> (K - 1 - i)) & 1) == 1
        if cur != prev:
            ends.append(i)
        prev = cur
    return ends


def prefix_hashes(seq, bits):
    ends = run_ends(bits)
    P = [0]
    for t, c in enumerate(seq):
        P.append(P[-1] ^ rol(SEEDS.get(c, 0), (64 - (t & 63)) & 63))
    out = []
    for k in range(len(seq) - K + 1):
        x = 0
        for e in ends:
            x ^= P[k + e]
        out.append(rol(x, (K - 1 + k) & 63))
    return out


def test_run_ends_of_the_shipped_seed():
    # 0b11011110111111011111111111111101: runs [0,1] [3,6] [8,13] [15,29] [31]
    assert run_ends(SEEDBITS) == [0, 2, 3, 7, 8, 14, 15, 30, 31, 32]


def test_prefix_form_equals_fame_ntps64():
    i, o = _load("ntps64")
    checked = 0
    for _ in range(i.get("I")):
        seq = i.bytes(i.get("I"))
        nk = len(seq) - 31
        direct = list(o.get(f"{nk}Q")) if nk > 1 else [o.get("Q")]
        o.get(f"{nk}Q") if nk > 1 else o.get("Q")  # the rolling form (same values)
        if any(c not in SEEDS for c in seq):
            continue  # reads holding other letters are never seeded
        assert prefix_hashes(seq, SEEDBITS) == direct
        checked += nk
    assert checked > 10000


def test_a_planted_mistake_is_caught():
    i, o = _load("ntps64")
    i.get("I")
    seq = i.bytes(i.get("I"))
    nk = len(seq) - 31
    direct = list(o.get(f"{nk}Q"))
    ends = run_ends(SEEDBITS)
    # dropping one run end (a run read as open) must change the hashes
    bad = SEEDBITS & ~(1 << (K - 1 - ends[2]))
    assert prefix_hashes(seq, bad) != direct
