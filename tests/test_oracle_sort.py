"""Pins the oracle's std::sort emulation (used for the candidate order of the
paired path, ReadQueue.cpp:2179-2181 and 2259-2261) against the host's own
libstdc++ std::sort, on arrays with many equal counts: introsort is not
stable, so the order among ties is the algorithm's, and every branch
(insertion sort below 16 elements, median-of-three partitions, the heap sort
fallback after 2*log2(n) levels) must match."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import orc

# Same element type and comparator shape as ReadQueue.h:339-351: pairs of
# (metaId, (c0, c1, flag, flag)), ordered by one count, greater first.
SRC = r"""
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <tuple>
#include <vector>
typedef std::pair<uint32_t, std::tuple<uint8_t, uint8_t, bool, bool>> E;
struct C0 { bool operator()(E& a, E& b) const { return std::get<0>(a.second) > std::get<0>(b.second); } };
struct C1 { bool operator()(E& a, E& b) const { return std::get<1>(a.second) > std::get<1>(b.second); } };
// McIlroy, "A Killer Adversary for Quicksort" (1999): answers std::sort's
// comparisons lazily so that every partition is as uneven as possible; the
// frozen values are an input that drives introsort into its heap sort.
struct Adv {
    std::vector<int>* val; int* gas; int* cand; int* ns;
    bool operator()(int x, int y) const {
        std::vector<int>& v = *val;
        if (v[x] == *gas && v[y] == *gas) { if (x == *cand) v[x] = (*ns)++; else v[y] = (*ns)++; }
        if (v[x] == *gas) *cand = x; else if (v[y] == *gas) *cand = y;
        return v[x] < v[y];
    }
};
int main(int argc, char** argv) {
    if (argc > 2) {
        int n = atoi(argv[2]);
        std::vector<int> val(n, n - 1), ptr(n);
        int gas = n - 1, cand = -1, ns = 0;
        for (int i = 0; i < n; ++i) ptr[i] = i;
        std::sort(ptr.begin(), ptr.end(), Adv{&val, &gas, &cand, &ns});
        for (int i = 0; i < n; ++i) printf("%d ", val[i]);
        printf("\n");
        return 0;
    }
    unsigned n, sec;
    while (scanf("%u %u", &n, &sec) == 2) {
        std::vector<E> v(n);
        for (auto& e : v) {
            unsigned k, c0, c1;
            scanf("%u %u %u", &k, &c0, &c1);
            e = E(k, std::make_tuple((uint8_t)c0, (uint8_t)c1, false, false));
        }
        if (sec) std::sort(v.begin(), v.end(), C1()); else std::sort(v.begin(), v.end(), C0());
        for (auto& e : v) printf("%u ", e.first);
        printf("\n");
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def std_sort_bin(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("stdsort")
    src = os.path.join(str(d), "s.cpp")
    open(src, "w").write(SRC)
    exe = os.path.join(str(d), "s")
    subprocess.run(["g++", "-O2", "-o", exe, src], check=True)
    return exe


def test_std_sort_emulation_matches_libstdcxx(std_sort_bin):
    rng = np.random.default_rng(4)
    sizes = list(range(0, 40)) + [47, 64, 100, 128, 200, 255, 300, 513, 1000, 2048]
    cases = []
    for n in sizes:
        for spread in (1, 2, 3, 8, 60, 256):
            for sec in (0, 1):
                c0 = rng.integers(0, spread, size=n)
                c1 = rng.integers(0, spread, size=n)
                if spread == 60 and n > 16:
                    # sorted and reverse-sorted runs drive the partitions
                    # towards their worst case and the heap sort fallback
                    c0 = np.sort(c0)
                    c1 = np.sort(c1)[::-1]
                cases.append((n, sec, c0, c1))
    inp = []
    for n, sec, c0, c1 in cases:
        inp.append(f"{n} {sec}\n" + " ".join(f"{i} {a} {b}" for i, a, b in zip(range(n), c0, c1)))
    r = subprocess.run([std_sort_bin], input="\n".join(inp) + "\n", capture_output=True, text=True,
                       check=True)
    lines = r.stdout.split("\n")
    assert len(lines) >= len(cases)
    for (n, sec, c0, c1), ln in zip(cases, lines):
        want = [int(x) for x in ln.split()]
        vals = (c0.astype(np.uint32) | (c1.astype(np.uint32) << 8))
        keys, _ = orc.std_sort_counts(np.arange(n), vals, sec)
        assert keys.tolist() == want, (n, sec)


def test_std_sort_emulation_heapsort_path(std_sort_bin):
    """Inputs from McIlroy's adversary against libstdc++'s std::sort drive it
    past its depth limit (2*log2 n) into heap sort; the emulation must follow
    it there, also with ties folded in (counts halved)."""
    for n in (64, 100, 128, 200, 256):
        r = subprocess.run([std_sort_bin, "adv", str(n)], capture_output=True, text=True, check=True)
        val = np.array([int(x) for x in r.stdout.split()], np.int64)
        assert sorted(val.tolist()) == list(range(n))
        for counts in ((n - 1 - val), (n - 1 - val) // 2):
            for sec in (0, 1):
                c = counts.astype(np.uint32)
                c0, c1 = (c, 0 * c) if sec == 0 else (0 * c, c)
                inp = f"{n} {sec}\n" + " ".join(f"{i} {a} {b}" for i, (a, b) in enumerate(zip(c0, c1))) + "\n"
                r = subprocess.run([std_sort_bin], input=inp, capture_output=True, text=True, check=True)
                want = [int(x) for x in r.stdout.split()]
                keys, _ = orc.std_sort_counts(np.arange(n), c0 | (c1 << 8), sec)
                assert keys.tolist() == want, (n, sec)
