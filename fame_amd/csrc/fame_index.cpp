// Native index builder writing FAME's --load_index format (the --store_index
// path of main.cpp:318-351), multithreaded with OpenMP.
//
// It follows the reference construction step by step, so that the bucket
// contents and their order -- which decide matching results -- are the
// reference's:
//   readReference / readLine   RefReader_istr.cpp:25-160, RefReader_istr.h:48-113
//   generateWindows            RefGenome.cpp:532-568
//   estimateTablesizes         RefGenome.cpp:1787-1913 (k-mers per window:
//                              rev k-mers then fwd k-mers per N-free
//                              interval, every SKIPMOD-th after the first)
//   generateHashes             RefGenome.cpp:570-759 (buckets filled by
//                              decrementing tabIndex, i.e. descending
//                              insertion order inside a bucket)
//   filterHashTable            RefGenome.cpp:2267-2352 (KMERCUTOFF)
//   filterRedundancyInHashTable RefGenome.cpp:2191-2264
//   save / write_strands       RefGenome.cpp:83-228, 392-412 (tmask by
//                              getTMask RefGenome.h:84-101)
// Bytes the reference reads outside a chromosome (rev k-mers of the last,
// shorter window of a chromosome) read as 0 here.  Struct padding is written
// as 0, and filteredKmers / chrMap in ascending key order (the reference
// writes hash-container order; the loader only uses them as sets / maps).
//
// C entry points (used through ctypes by fame_amd/index.py):
//   fame_build_index(chroms, lens, names, n, readlen, htab, kmercutoff,
//                    noloss, out_path)       from in-memory sequences
//   fame_build_index_fasta(fasta, ..., out_path)  from a FASTA file

#include <omp.h>
#include <sys/mman.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

constexpr uint32_t KMERLEN = 32;
constexpr uint32_t WINLEN = 2048;
constexpr uint32_t SKIPMOD = 2;
constexpr uint32_t SEEDBITS = 0b11011110111111011111111111111101u; // CONST.h:58
constexpr uint32_t CPGDUMMY = 0xFFFFFFFFu;

struct CpG {
    uint8_t chrom;
    uint32_t pos;
};
struct Win {
    uint32_t startPos;
    uint8_t chrom;
    uint32_t startInd, endInd;
};

inline uint64_t rol64(uint64_t v, int s)
{
    s &= 63;
    return s ? (v << s) | (v >> (64 - s)) : v;
}
inline uint64_t red_seed(uint8_t c)
{
    switch (c) {
    case 'A': case 'a': case 4:
        return 0x3c8bfbb395c60474ULL;
    case 'C': case 'c': case 'T': case 't': case 1: case 7:
        return 0x295549f54be24456ULL;
    case 'G': case 'g': case 3:
        return 0x20323ed082572324ULL;
    default:
        return 0;
    }
}
inline uint64_t ntps64(const uint8_t* k)
{
    uint64_t s = 0;
    for (uint32_t i = 0; i < KMERLEN; ++i)
        if ((SEEDBITS >> (KMERLEN - 1 - i)) & 1) s ^= rol64(red_seed(k[i]), (int)(KMERLEN - 1 - i));
    return s;
}

struct Genome {
    std::vector<std::vector<uint8_t>> seq;
    std::vector<CpG> cpg, cpgStart;
    std::map<uint8_t, std::string> names;
};

// readLine RefReader_istr.h:48-113 for one chromosome's text
void add_chrom_text(Genome& g, uint8_t chrIndex1, const uint8_t* s, size_t n, std::vector<uint8_t>& seq,
                    bool& lastC, uint32_t readlen)
{
    const uint8_t ci = (uint8_t)(chrIndex1 - 1);
    for (size_t i = 0; i < n; ++i) {
        switch (s[i]) {
        case 'g':
        case 'G':
            if (lastC) {
                if (seq.size() >= readlen - 2)
                    g.cpg.push_back({ci, (uint32_t)seq.size() - readlen + 1});
                else
                    g.cpgStart.push_back({ci, (uint32_t)seq.size() - 1});
                lastC = false;
            }
            seq.push_back('G');
            break;
        case 'c':
        case 'C':
            seq.push_back('C');
            lastC = true;
            break;
        case 'a':
        case 'A':
            seq.push_back('A');
            lastC = false;
            break;
        case 't':
        case 'T':
            seq.push_back('T');
            lastC = false;
            break;
        default:
            seq.push_back('N');
            lastC = false;
            break;
        }
    }
}

// readReference RefReader_istr.cpp:25-160 (without --human_opt)
bool read_fasta(const char* path, Genome& g, uint32_t readlen)
{
    std::ifstream ifs(path);
    if (!ifs) return false;
    std::string line;
    uint8_t chrIndex = 0;
    bool cont = false, lastC = false;
    uint32_t unID = 1;
    std::vector<uint8_t> seq;
    while (std::getline(ifs, line)) {
        if (!line.empty() && line[0] == '>' && line.size() > 1) {
            if (cont) {
                g.seq.emplace_back(std::move(seq));
                seq = std::vector<uint8_t>();
                lastC = false;
            }
            ++chrIndex;
            std::string hdr(line.begin() + 1, line.end());
            size_t p = hdr.find_first_of(" \t");
            std::string id(hdr, 0, (uint32_t)p);
            for (auto& kv : g.names)
                if (kv.second == id) {
                    id.append("_");
                    id.append(std::to_string(unID++));
                    break;
                }
            g.names.insert({(uint8_t)(chrIndex - 1), id});
            cont = true;
            continue;
        }
        if (cont) add_chrom_text(g, chrIndex, (const uint8_t*)line.data(), line.size(), seq, lastC, readlen);
    }
    if (cont) g.seq.emplace_back(std::move(seq));
    return true;
}

// generateWindows RefGenome.cpp:532-568.  cpgTable[cpgTabInd] one past the
// end reads as {chrom 0, pos 0} (the reference reads past its vector there).
std::vector<Win> make_windows(const Genome& g, uint32_t readlen)
{
    std::vector<Win> w;
    uint32_t ind = 0;
    const uint32_t ncpg = (uint32_t)g.cpg.size();
    auto at = [&](uint32_t i) { return i < ncpg ? g.cpg[i] : CpG{0, 0}; };
    for (uint32_t c = 0; c < g.seq.size(); ++c) {
        const uint8_t chr = (uint8_t)c;
        for (uint32_t ws = 0; ws < g.seq[c].size(); ws = ws + WINLEN - readlen) {
            const uint32_t we = ws + WINLEN;
            uint32_t is = CPGDUMMY, ie = CPGDUMMY;
            while (ind < ncpg && g.cpg[ind].chrom == chr && g.cpg[ind].pos <= we) {
                if (is == CPGDUMMY) is = ind;
                ie = ind;
                ++ind;
            }
            while (ind > 0 && at(ind).chrom == chr && at(ind).pos >= we - readlen) --ind;
            if (ind > 0) ++ind;
            w.push_back({ws, chr, is, ie});
        }
    }
    return w;
}

// the reduced forward / reverse window sequences and their N-free intervals
// (RefGenome.cpp:586-656, including the write of 'N' into redSeq at the
// mirrored position)
struct WinSeq {
    uint8_t red[WINLEN];
    uint8_t rev[WINLEN];
    uint32_t iv[WINLEN][2];
    uint32_t niv;
};

void window_seq(const Genome& g, const Win& w, WinSeq& s)
{
    memset(s.red, 0, WINLEN);
    memset(s.rev, 0, WINLEN);
    s.niv = 0;
    const auto& seq = g.seq[w.chrom];
    bool lastN = true;
    uint32_t nStart = 0, j = 0;
    const size_t end = std::min<size_t>((size_t)w.startPos + WINLEN, seq.size());
    for (uint32_t i = w.startPos; i < end; ++i, ++j) {
        const uint32_t rp = WINLEN - 1 - j;
        switch (seq[i]) {
        case 'N':
            if (!lastN) {
                s.iv[s.niv][0] = nStart;
                s.iv[s.niv][1] = j - 1;
                ++s.niv;
            }
            lastN = true;
            s.red[j] = 'N';
            s.red[rp] = 'N';
            break;
        case 'C':
            if (lastN) {
                nStart = j;
                lastN = false;
            }
            s.red[j] = 'T';
            s.rev[rp] = 'G';
            break;
        case 'G':
            if (lastN) {
                nStart = j;
                lastN = false;
            }
            s.red[j] = 'G';
            s.rev[rp] = 'T';
            break;
        case 'T':
            if (lastN) {
                nStart = j;
                lastN = false;
            }
            s.red[j] = 'T';
            s.rev[rp] = 'A';
            break;
        case 'A':
            if (lastN) {
                nStart = j;
                lastN = false;
            }
            s.red[j] = 'A';
            s.rev[rp] = 'T';
            break;
        default:
            break;
        }
    }
    if (!lastN) {
        s.iv[s.niv][0] = nStart;
        s.iv[s.niv][1] = j - 1;
        ++s.niv;
    }
}

// visits the k-mers of one window in insertion order: f(key_hash, offset, fwd)
template <typename F>
void window_kmers(const WinSeq& s, uint32_t readlen, F&& f)
{
    for (uint32_t v = 0; v < s.niv; ++v) {
        const uint32_t a = s.iv[v][0], b = s.iv[v][1];
        const uint32_t dist = b - a + 1;
        if (dist < readlen) continue;
        const uint8_t* sr = s.rev + (WINLEN - 1 - b);
        uint64_t kpr = WINLEN - a - KMERLEN;
        f(ntps64(sr), kpr, false);
        for (uint32_t i = 0; i < dist - KMERLEN; ++i) {
            --kpr;
            if (!(i % SKIPMOD)) f(ntps64(sr + i + 1), kpr, false);
        }
        const uint8_t* sf = s.red + a;
        uint64_t kp = a;
        f(ntps64(sf), kp, true);
        for (uint32_t i = 0; i < dist - KMERLEN; ++i) {
            ++kp;
            if (!(i % SKIPMOD)) f(ntps64(sf + i + 1), kp, true);
        }
    }
}

inline uint8_t chr_at(const Genome& g, uint8_t chrom, uint64_t p)
{
    const auto& s = g.seq[chrom];
    return p < s.size() ? s[p] : 0;
}

// reproduceKmerSeq RefGenome.cpp:1953-2006
uint64_t kmer_code(const Genome& g, const Win& w, uint64_t off, bool fwd)
{
    const uint64_t base = (uint64_t)w.startPos + off;
    uint64_t k = 0;
    if (fwd) {
        for (uint32_t i = 0; i < KMERLEN; ++i) {
            k <<= 2;
            if ((SEEDBITS >> (KMERLEN - 1 - i)) & 1) {
                const uint8_t c = chr_at(g, w.chrom, base + i);
                if (c == 'C' || c == 'T') k += 3;
                else if (c == 'G') k += 2;
            }
        }
    } else {
        for (uint32_t i = KMERLEN; i > 0; --i) {
            k <<= 2;
            if ((SEEDBITS >> (i - 1)) & 1) { // SEED[KMERLEN - i]
                const uint8_t c = chr_at(g, w.chrom, base + i - 1);
                if (c == 'G' || c == 'A') k += 3;
                else if (c == 'C') k += 2;
            }
        }
    }
    return k;
}

// reproduceTMask RefGenome.cpp:2013-2045
uint32_t repro_tmask(const Genome& g, const Win& w, uint64_t off, bool fwd)
{
    const uint64_t base = (uint64_t)w.startPos + off;
    uint32_t m = 0;
    if (fwd) {
        for (uint32_t i = 0; i < KMERLEN; ++i) m = (m << 1) | (chr_at(g, w.chrom, base + i) == 'T');
    } else {
        for (uint32_t i = KMERLEN; i > 0; --i) m = (m << 1) | (chr_at(g, w.chrom, base + i - 1) == 'A');
    }
    return m;
}

// getTMask RefGenome.h:84-101 as used by save(): the strand is read from
// strandTable[metaId] and the reverse loop covers positions 31..1
uint32_t save_tmask(const Genome& g, const Win& w, uint64_t off, bool strand_of_meta)
{
    const uint64_t base = (uint64_t)w.startPos + off;
    uint32_t m = 0;
    if (strand_of_meta) {
        for (uint32_t i = 0; i < KMERLEN; ++i) m = (m << 1) | (chr_at(g, w.chrom, base + i) == 'T');
    } else {
        for (uint32_t i = KMERLEN - 1; i > 0; --i) m = (m << 1) | (chr_at(g, w.chrom, base + i) == 'A');
    }
    return m;
}

// large uninitialised array on transparent huge pages (first touch happens
// in the parallel loops that fill it)
template <typename T>
struct BigArray {
    T* p = nullptr;
    size_t n = 0;
    explicit BigArray(size_t count) : n(count)
    {
        size_t bytes = ((count * sizeof(T)) + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
        p = static_cast<T*>(aligned_alloc(2u << 20, bytes ? bytes : (2u << 20)));
        if (!p) throw std::bad_alloc();
        madvise(p, bytes, MADV_HUGEPAGE);
    }
    ~BigArray() { free(p); }
    BigArray(const BigArray&) = delete;
    T& operator[](size_t i) { return p[i]; }
    const T& operator[](size_t i) const { return p[i]; }
    T* data() { return p; }
};

// exclusive prefix sum out[i] = sum(v(j), j < i) for i <= n, in parallel blocks
template <typename V, typename O>
void par_prefix(uint64_t n, V&& v, O& out)
{
    const int T = omp_get_max_threads();
    std::vector<uint64_t> part(T + 1, 0);
    const uint64_t blk = (n + T - 1) / T;
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        const uint64_t a = std::min<uint64_t>(n, blk * t), b = std::min<uint64_t>(n, a + blk);
        uint64_t s = 0;
        for (uint64_t i = a; i < b; ++i) s += v(i);
        part[t + 1] = s;
#pragma omp barrier
#pragma omp single
        for (int k = 0; k < T; ++k) part[k + 1] += part[k];
        s = part[t];
        for (uint64_t i = a; i < b; ++i) {
            out[i] = s;
            s += v(i);
        }
    }
    out[n] = part[T];
}

struct Entry {
    uint32_t meta; // metaId (isStart is always 0)
    uint16_t off;
    uint8_t fwd;
};

int build(Genome& g, uint32_t readlen, uint64_t htab, uint64_t kmercutoff, bool noloss, const char* out,
          char* err, size_t errlen)
{
    const bool verbose = getenv("FAME_INDEX_VERBOSE") != nullptr;
    double t0 = omp_get_wtime();
    auto phase = [&](const char* what) {
        if (verbose) fprintf(stderr, "[index] %-12s %.2fs\n", what, omp_get_wtime() - t0);
    };
    if (readlen <= KMERLEN || htab == 0) {
        snprintf(err, errlen, "bad parameters");
        return -4;
    }
    const std::vector<Win> wins = make_windows(g, readlen);
    const uint64_t nwin = wins.size();
    // pass 1: entries per window
    std::vector<uint64_t> wbase(nwin + 1, 0);
#pragma omp parallel
    {
        WinSeq* s = new WinSeq;
#pragma omp for schedule(dynamic, 256)
        for (uint64_t w = 0; w < nwin; ++w) {
            window_seq(g, wins[w], *s);
            uint64_t n = 0;
            window_kmers(*s, readlen, [&](uint64_t, uint64_t, bool) { ++n; });
            wbase[w + 1] = n;
        }
        delete s;
    }
    for (uint64_t w = 0; w < nwin; ++w) wbase[w + 1] += wbase[w];
    const uint64_t nent = wbase[nwin];
    phase("windows+count");
    // pass 2: bucket and payload of every entry, indexed by insertion order g
    std::vector<uint32_t> ekey(nent);
    std::vector<Entry> eval(nent);
    const bool hp2 = (htab & (htab - 1)) == 0;
#pragma omp parallel
    {
        WinSeq* s = new WinSeq;
#pragma omp for schedule(dynamic, 256)
        for (uint64_t w = 0; w < nwin; ++w) {
            window_seq(g, wins[w], *s);
            uint64_t i = wbase[w];
            window_kmers(*s, readlen, [&](uint64_t h, uint64_t off, bool fwd) {
                ekey[i] = (uint32_t)(hp2 ? (h & (htab - 1)) : (h % htab));
                eval[i] = {(uint32_t)w, (uint16_t)off, (uint8_t)fwd};
                ++i;
            });
        }
        delete s;
    }
    phase("hash");
    if (htab > 0xFFFFFFFFull) {
        snprintf(err, errlen, "HTABSIZE above 2^32 is not supported by this builder");
        return -4;
    }
    // counting sort; inside a bucket descending insertion order (the
    // reference decrements tabIndex while inserting in ascending order)
    BigArray<uint64_t> tab(htab + 1);
    {
        BigArray<std::atomic<uint32_t>>* cnt = new BigArray<std::atomic<uint32_t>>(htab);
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < htab; ++i) (*cnt)[i].store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < nent; ++i) (*cnt)[ekey[i]].fetch_add(1, std::memory_order_relaxed);
        par_prefix(htab, [&](uint64_t i) { return (uint64_t)(*cnt)[i].load(std::memory_order_relaxed); }, tab);
        // fill slots in any order, then order each bucket by descending g
        std::vector<uint64_t> slot_g(nent);
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < htab; ++i) (*cnt)[i].store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < nent; ++i) {
            const uint32_t b = ekey[i];
            slot_g[tab[b] + (*cnt)[b].fetch_add(1, std::memory_order_relaxed)] = i;
        }
        delete cnt;
#pragma omp parallel for schedule(dynamic, 65536)
        for (uint64_t b = 0; b < htab; ++b)
            if (tab[b + 1] - tab[b] > 1)
                std::sort(slot_g.begin() + (ptrdiff_t)tab[b], slot_g.begin() + (ptrdiff_t)tab[b + 1],
                          std::greater<uint64_t>());
        std::vector<uint32_t>().swap(ekey);
        // payload in table order
        std::vector<Entry> tbl(nent);
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < nent; ++i) tbl[i] = eval[slot_g[i]];
        eval.swap(tbl);
    }
    phase("sort");
    std::vector<Entry>& kt = eval;
    // filterHashTable: buckets with >= KMERCUTOFF entries lose the k-mers
    // occurring >= KMERCUTOFF times in them
    std::set<uint64_t> filtered;
    std::vector<uint8_t> keep(nent, 1);
    if (!noloss) {
        std::vector<uint64_t> big;
#pragma omp parallel
        {
            std::vector<uint64_t> mine;
#pragma omp for schedule(static) nowait
            for (uint64_t b = 0; b < htab; ++b)
                if (tab[b + 1] - tab[b] >= kmercutoff) mine.push_back(b);
#pragma omp critical
            big.insert(big.end(), mine.begin(), mine.end());
        }
        std::sort(big.begin(), big.end());
#pragma omp parallel
        {
            std::set<uint64_t> mine;
#pragma omp for schedule(dynamic, 1)
            for (size_t t = 0; t < big.size(); ++t) {
                const uint64_t b = big[t];
                std::unordered_map<uint64_t, uint64_t> cnt;
                for (uint64_t j = tab[b]; j < tab[b + 1]; ++j)
                    ++cnt[kmer_code(g, wins[kt[j].meta], kt[j].off, kt[j].fwd)];
                for (uint64_t j = tab[b]; j < tab[b + 1]; ++j) {
                    const uint64_t k = kmer_code(g, wins[kt[j].meta], kt[j].off, kt[j].fwd);
                    if (cnt[k] >= kmercutoff) {
                        keep[j] = 0;
                        mine.insert(k);
                    }
                }
            }
#pragma omp critical
            filtered.insert(mine.begin(), mine.end());
        }
    }
    phase("cutoff");
    // filterRedundancyInHashTable: inside a run of entries with the same
    // (window, strand), keep the first entry of every distinct tmask
#pragma omp parallel for schedule(dynamic, 65536)
    for (uint64_t b = 0; b < htab; ++b) {
        const uint64_t lo = tab[b], hi = tab[b + 1];
        if (hi - lo < 2) continue;
        uint64_t metaID = ~0ull;
        bool isFwd = false;
        std::vector<uint32_t> masks;
        for (uint64_t j = lo; j < hi; ++j) {
            if (!keep[j]) continue;
            const Entry& e = kt[j];
            const uint32_t tm = repro_tmask(g, wins[e.meta], e.off, e.fwd);
            if (e.meta != metaID || (bool)e.fwd != isFwd) {
                masks.assign(1, tm);
                metaID = e.meta;
                isFwd = e.fwd;
            } else if (std::find(masks.begin(), masks.end(), tm) == masks.end()) {
                masks.push_back(tm);
            } else {
                keep[j] = 0;
            }
        }
    }
    phase("redundancy");
    // compaction
    BigArray<uint64_t> ntab(htab + 1);
    par_prefix(htab, [&](uint64_t b) {
        uint64_t k = 0;
        for (uint64_t j = tab[b]; j < tab[b + 1]; ++j) k += keep[j];
        return k;
    }, ntab);
    const uint64_t nk = ntab[htab];
    std::vector<Entry> fk(nk);
#pragma omp parallel for schedule(dynamic, 65536)
    for (uint64_t b = 0; b < htab; ++b) {
        uint64_t o = ntab[b];
        for (uint64_t j = tab[b]; j < tab[b + 1]; ++j)
            if (keep[j]) fk[o++] = kt[j];
    }
    std::vector<Entry>().swap(kt);
    std::vector<uint8_t>().swap(keep);
    phase("compact");
    // save RefGenome.cpp:83-228
    FILE* f = fopen(out, "wb");
    if (!f) {
        snprintf(err, errlen, "cannot write %s", out);
        return -1;
    }
    std::vector<char> iobuf(64 << 20);
    setvbuf(f, iobuf.data(), _IOFBF, iobuf.size());
    auto put = [&](const void* p, size_t n) { fwrite(p, 1, n, f); };
    const uint32_t RL = readlen, WL = WINLEN, KL = KMERLEN, SB = SEEDBITS;
    put(&htab, 8);
    put(&RL, 4);
    put(&WL, 4);
    put(&KL, 4);
    put(&kmercutoff, 8);
    put(&SB, 4);
    auto put_cpgs = [&](const std::vector<CpG>& v) {
        uint64_t n = v.size();
        put(&n, 8);
        std::vector<uint8_t> raw(8 * n, 0);
        for (uint64_t i = 0; i < n; ++i) {
            raw[8 * i] = v[i].chrom;
            memcpy(&raw[8 * i + 4], &v[i].pos, 4);
        }
        put(raw.data(), raw.size());
    };
    put_cpgs(g.cpg);
    put_cpgs(g.cpgStart);
    uint64_t nchrom = g.seq.size();
    put(&nchrom, 8);
    for (auto& s : g.seq) {
        uint64_t n = s.size();
        put(&n, 8);
        put(s.data(), n);
    }
    uint64_t tablen = htab + 1;
    put(&tablen, 8);
    put(ntab.data(), 8 * tablen);
    put(&nk, 8);
    {
        // strand of entry index metaId, as getTMask reads it
        auto strand_at = [&](uint64_t i) { return i < nk ? (bool)fk[i].fwd : false; };
        std::vector<uint32_t> rec(2 * std::min<uint64_t>(nk, 1 << 22));
        for (uint64_t first = 0; first < nk; first += (1 << 22)) {
            const uint64_t n = std::min<uint64_t>(nk - first, 1 << 22);
#pragma omp parallel for schedule(static)
            for (uint64_t i = 0; i < n; ++i) {
                const Entry& e = fk[first + i];
                rec[2 * i] = e.meta;
                rec[2 * i + 1] = save_tmask(g, wins[e.meta], e.off, strand_at(e.meta));
            }
            put(rec.data(), 8 * n);
        }
    }
    uint64_t zero = 0;
    put(&zero, 8); // metaCpGs
    put(&zero, 8); // metaStartCpGs
    put(&nwin, 8);
    {
        std::vector<uint32_t> raw(4 * nwin);
        for (uint64_t i = 0; i < nwin; ++i) {
            raw[4 * i] = wins[i].startPos;
            raw[4 * i + 1] = wins[i].chrom;
            raw[4 * i + 2] = wins[i].startInd;
            raw[4 * i + 3] = wins[i].endInd;
        }
        put(raw.data(), 16 * nwin);
    }
    uint64_t nf = filtered.size();
    put(&nf, 8);
    for (uint64_t k : filtered) put(&k, 8);
    uint64_t nmap = g.names.size();
    put(&nmap, 8);
    for (auto& kv : g.names) {
        uint8_t id = kv.first;
        uint64_t sl = kv.second.size();
        put(&id, 1);
        put(&sl, 8);
        put(kv.second.data(), sl);
    }
    const bool ok = !ferror(f);
    fclose(f);
    phase("write");
    // <index>_strands, write_strands RefGenome.cpp:392-412
    std::string sp = std::string(out) + "_strands";
    FILE* s = fopen(sp.c_str(), "wb");
    if (!s || !ok) {
        if (s) fclose(s);
        snprintf(err, errlen, "write error");
        return -1;
    }
    fwrite(&nk, 8, 1, s);
    std::vector<uint8_t> bits((nk + 7) / 8, 0);
    for (uint64_t i = 0; i < nk; ++i)
        if (fk[i].fwd) bits[i >> 3] |= (uint8_t)(1u << (i & 7));
    fwrite(bits.data(), 1, bits.size(), s);
    const bool ok2 = !ferror(s);
    fclose(s);
    if (!ok2) {
        snprintf(err, errlen, "write error");
        return -1;
    }
    return 0;
}

} // namespace

extern "C" {

// chroms: n pointers to chromosome byte strings (any characters; normalised
// like readLine), names: n NUL-terminated names (chrMap entries)
int fame_build_index(const uint8_t* const* chroms, const uint64_t* lens, const char* const* names, uint32_t n,
                     uint32_t readlen, uint64_t htab, uint64_t kmercutoff, int noloss, const char* out, char* err,
                     size_t errlen)
{
    if (n > 255) {
        snprintf(err, errlen, "more than 255 chromosomes");
        return -4;
    }
    Genome g;
    for (uint32_t c = 0; c < n; ++c) {
        std::vector<uint8_t> seq;
        seq.reserve(lens[c]);
        bool lastC = false;
        add_chrom_text(g, (uint8_t)(c + 1), chroms[c], lens[c], seq, lastC, readlen);
        g.seq.emplace_back(std::move(seq));
        g.names.insert({(uint8_t)c, names[c]});
    }
    return build(g, readlen, htab, kmercutoff, noloss != 0, out, err, errlen);
}

int fame_build_index_fasta(const char* fasta, uint32_t readlen, uint64_t htab, uint64_t kmercutoff, int noloss,
                           const char* out, char* err, size_t errlen)
{
    Genome g;
    if (!read_fasta(fasta, g, readlen)) {
        snprintf(err, errlen, "cannot read %s", fasta);
        return -1;
    }
    return build(g, readlen, htab, kmercutoff, noloss != 0, out, err, errlen);
}

} // extern "C"
