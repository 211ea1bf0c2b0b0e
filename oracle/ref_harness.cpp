// Reference-bench harness: links FAME's own translation units (compiled
// unmodified from their source tree, with a configured CONST.h) and replaces
// only main.cpp, so that the reference CPU path can be timed with
// steady_clock and its per-read results dumped for parity checks.
//
// This file is test infrastructure: it is the checker and the CPU baseline,
// never the product path.
//
// Modes
//   build <genome.fa> <index> [--no_loss] [--human_opt]
//       mirrors main.cpp:318-351 (readReference -> RefGenome ctor -> save)
//   match <index> <reads.fq> <out_prefix> [--unord_reads] [--max_reads N]
//       mirrors queryRoutine main.cpp:363-398; writes
//         <out_prefix>_cpg.tsv      via ReadQueue::printMethylationLevels
//         <out_prefix>_reads.bin    "FAMERES1", u64 n, u8 valid[n], u64 mat[n]
//         <out_prefix>_stats.json   counters, strand tallies, per-chunk timings
//   match_pe <index> <r1.fq> <r2.fq> <out_prefix> [--unord_reads] [--max_reads N]
//       mirrors queryRoutinePaired main.cpp:400-441; writes the same TSV and
//       stats, and <out_prefix>_pairs.bin: "FAMEPE01", u64 n, u8 valid1[n],
//       u8 valid2[n], u64 mat1[n], u64 mat2[n] (mat as left in Read::mat)
//   parse <index> <reads.fq[.gz]> <out> [--gz]
//       ReadQueue::parseChunk / parseChunkGZ (ReadQueue.cpp:218-373) until it
//       reports the end; writes <out>: "FAMEFQ01", u64 n, u32 len[n], then the
//       n sequences back to back, and per parse call (u32 procReads, u8 more)
//       after them, preceded by a u64 count of calls
//   sc <index> <meta> <out> [--unord_reads] [--gz] [--pe]
//       the single-cell driver queryRoutineSC / queryRoutineSCPaired
//       (main.cpp:443-499) on FAME's single-cell ReadQueue (ReadQueue.cpp:
//       181-216): per line "<id> <fastq> [<fastq2>]" of <meta>,
//       matchSCBatch / matchSCBatchPaired; <out> is the count matrix
//   units ntps64|shiftand|dp|hopscotch <in.bin> <out.bin>
//       runs one of FAME's building blocks on the cases of <in.bin> (formats
//       in do_units below): ntHash::NTPS64 direct and rolling
//       (nthash.hpp:480-498), ShiftAnd<E>::querySeq / queryRevSeq, the
//       LevenshtDP fill + backtrack used by computeMethLvl, and the
//       iteration order of tsl::hopscotch_map<uint32_t, uint16_t, MetaHash>
//
// ReadQueue keeps readBuffer / strand tallies private; they are reached by
// compiling this translation unit only with `private` defined as `public`.

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <list>
#include <numeric>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#include <omp.h>

#define private public
#include "ReadQueue.h"
#include "RefReader_istr.h"
#undef private

using clk = std::chrono::steady_clock;

static double secs(clk::time_point a, clk::time_point b)
{
    return std::chrono::duration<double>(b - a).count();
}

struct ChunkTime {
    unsigned n;
    double s;
    bool stranded;
};

static void dump_reads(const std::string& path, const std::vector<uint8_t>& valid,
                       const std::vector<uint64_t>& mat)
{
    std::ofstream of(path, std::ios::binary | std::ios::trunc);
    const char magic[8] = {'F', 'A', 'M', 'E', 'R', 'E', 'S', '1'};
    of.write(magic, 8);
    uint64_t n = valid.size();
    of.write(reinterpret_cast<const char*>(&n), 8);
    of.write(reinterpret_cast<const char*>(valid.data()), n);
    of.write(reinterpret_cast<const char*>(mat.data()), 8 * n);
}

static void collect(std::vector<Read>& buf, unsigned n, std::vector<uint8_t>& valid,
                    std::vector<uint64_t>& mat)
{
    for (unsigned i = 0; i < n; ++i) {
        const Read& r = buf[i];
        valid.push_back(r.isInvalid ? 0 : 1);
        mat.push_back(r.isInvalid ? 0 : r.mat);
    }
}

static void write_stats(const std::string& path, const char* mode, uint64_t nreads,
                        uint64_t succ, uint64_t nonu, uint64_t unsucc, uint64_t succPaired,
                        uint64_t tooShort, const ReadQueue& q, double load_s,
                        const std::vector<ChunkTime>& chunks)
{
    double total = 0;
    for (auto& c : chunks) total += c.s;
    std::ofstream of(path, std::ios::trunc);
    of << "{\"mode\": \"" << mode << "\", \"reads\": " << nreads << ", \"succ\": " << succ
       << ", \"nonunique\": " << nonu << ", \"unsucc\": " << unsucc
       << ", \"succ_paired\": " << succPaired << ", \"too_short\": " << tooShort
       << ", \"r1_fwd\": " << q.r1FwdMatches << ", \"r1_rev\": " << q.r1RevMatches
       << ", \"match_r1_fwd\": " << (q.matchR1Fwd ? "true" : "false")
       << ", \"corenum\": " << CORENUM << ", \"htabsize\": " << MyConst::HTABSIZE
       << ", \"readlen\": " << MyConst::READLEN << ", \"chunksize\": " << MyConst::CHUNKSIZE
       << ", \"load_s\": " << load_s << ", \"match_s\": " << total << ", \"chunks\": [";
    for (size_t i = 0; i < chunks.size(); ++i) {
        of << (i ? ", " : "") << "{\"n\": " << chunks[i].n << ", \"s\": " << chunks[i].s
           << ", \"stranded\": " << (chunks[i].stranded ? "true" : "false") << "}";
    }
    of << "]}\n";
}

static int do_build(int argc, char** argv)
{
    if (argc < 4) return 2;
    bool noloss = false, human = false;
    for (int i = 4; i < argc; ++i) {
        if (!strcmp(argv[i], "--no_loss")) noloss = true;
        if (!strcmp(argv[i], "--human_opt")) human = true;
    }
    std::vector<struct CpG> cpgTab;
    std::vector<struct CpG> cpgStartTab;
    std::vector<std::vector<char>> genSeq;
    std::unordered_map<uint8_t, std::string> chrMap;
    readReference(argv[2], cpgTab, cpgStartTab, genSeq, chrMap, human);
    RefGenome ref(std::move(cpgTab), std::move(cpgStartTab), genSeq, noloss, chrMap);
    ref.save(argv[3]);
    return 0;
}

// parseChunk, truncated so that at most `budget` reads are taken from the
// file in total (budget 0 = no limit); used for bounded CPU samples.
static bool parse_bounded(ReadQueue& q, unsigned& n, uint64_t& seen, uint64_t budget)
{
    bool more = q.parseChunk(n);
    if (budget && seen + n >= budget) {
        n = static_cast<unsigned>(budget - seen);
        more = false;
    }
    seen += n;
    return more;
}

static int do_match(int argc, char** argv)
{
    if (argc < 5) return 2;
    bool both = false;
    uint64_t budget = 0;
    for (int i = 5; i < argc; ++i) {
        if (!strcmp(argv[i], "--unord_reads")) both = true;
        if (!strcmp(argv[i], "--max_reads") && i + 1 < argc) budget = std::stoull(argv[++i]);
    }
    auto t0 = clk::now();
    RefGenome ref{std::string(argv[2])};
    double load_s = secs(t0, clk::now());
    ReadQueue q(argv[3], ref, false, both);
    std::vector<uint8_t> valid;
    std::vector<uint64_t> mat;
    std::vector<ChunkTime> chunks;
    uint64_t succ = 0, nonu = 0, unsucc = 0, seen = 0;
    unsigned n = 0;
    auto run = [&](bool stranded) {
        auto a = clk::now();
        q.matchReads(n, succ, nonu, unsucc, stranded);
        chunks.push_back({n, secs(a, clk::now()), stranded});
        collect(q.readBuffer, n, valid, mat);
    };
    // queryRoutine, main.cpp:374-391: the first chunk is matched in both
    // orientations and decides the strand, full chunks follow, then the
    // remainder of the last parse (possibly empty) is matched.
    if (!both) {
        parse_bounded(q, n, seen, budget);
        run(true);
        q.decideStrand();
    }
    while (parse_bounded(q, n, seen, budget)) run(false);
    run(false);
    std::string out = argv[4];
    q.printMethylationLevels(out);
    dump_reads(out + "_reads.bin", valid, mat);
    write_stats(out + "_stats.json", "se", valid.size(), succ, nonu, unsucc, 0, 0, q, load_s,
                chunks);
    std::cout << "HARNESS reads=" << valid.size() << " succ=" << succ << " nonunique=" << nonu
              << " unsucc=" << unsucc << "\n";
    return 0;
}

static int do_match_pe(int argc, char** argv)
{
    if (argc < 6) return 2;
    bool both = false;
    uint64_t budget = 0;
    for (int i = 6; i < argc; ++i) {
        if (!strcmp(argv[i], "--unord_reads")) both = true;
        if (!strcmp(argv[i], "--max_reads") && i + 1 < argc) budget = std::stoull(argv[++i]);
    }
    auto t0 = clk::now();
    RefGenome ref{std::string(argv[2])};
    double load_s = secs(t0, clk::now());
    ReadQueue q(argv[3], argv[4], ref, false, both);
    std::vector<uint8_t> v1, v2;
    std::vector<uint64_t> m1, m2;
    std::vector<ChunkTime> chunks;
    uint64_t succ = 0, nonu = 0, unsucc = 0, succP = 0, tooShort = 0, seen = 0;
    unsigned n = 0;
    auto run = [&](bool stranded) {
        // matchPairedReads writes mat only for matched pairs (ReadQueue.cpp:1593-1594)
        // and parseChunk leaves it unset, so clear it first: unmatched pairs then
        // read 0, as in the oracle and the HIP path.
        for (unsigned i = 0; i < n; ++i) q.readBuffer[i].mat = q.readBuffer2[i].mat = 0;
        auto a = clk::now();
        q.matchPairedReads(n, succ, nonu, unsucc, succP, tooShort, stranded);
        chunks.push_back({n, secs(a, clk::now()), stranded});
        for (unsigned i = 0; i < n; ++i) {
            v1.push_back(q.readBuffer[i].isInvalid ? 0 : 1);
            v2.push_back(q.readBuffer2[i].isInvalid ? 0 : 1);
            m1.push_back(q.readBuffer[i].mat);
            m2.push_back(q.readBuffer2[i].mat);
        }
    };
    // queryRoutinePaired, main.cpp:413-431
    if (!both) {
        parse_bounded(q, n, seen, budget);
        run(true);
        q.decideStrand();
    }
    while (parse_bounded(q, n, seen, budget)) run(false);
    run(false);
    std::string out = argv[5];
    q.printMethylationLevels(out);
    {
        std::ofstream of(out + "_pairs.bin", std::ios::binary | std::ios::trunc);
        const char magic[8] = {'F', 'A', 'M', 'E', 'P', 'E', '0', '1'};
        of.write(magic, 8);
        uint64_t nn = v1.size();
        of.write(reinterpret_cast<const char*>(&nn), 8);
        of.write(reinterpret_cast<const char*>(v1.data()), nn);
        of.write(reinterpret_cast<const char*>(v2.data()), nn);
        of.write(reinterpret_cast<const char*>(m1.data()), 8 * nn);
        of.write(reinterpret_cast<const char*>(m2.data()), 8 * nn);
    }
    write_stats(out + "_stats.json", "pe", v1.size(), succ, nonu, unsucc, succP, tooShort, q, load_s,
                chunks);
    std::cout << "HARNESS pairs=" << v1.size() << " succ=" << succ << " nonunique=" << nonu
              << " unsucc=" << unsucc << " paired=" << succP << " tooshort=" << tooShort << "\n";
    return 0;
}

static int do_parse(int argc, char** argv)
{
    if (argc < 5) return 2;
    bool gz = false;
    for (int i = 5; i < argc; ++i)
        if (!strcmp(argv[i], "--gz")) gz = true;
    RefGenome ref{std::string(argv[2])};
    ReadQueue q(argv[3], ref, gz, false);
    std::vector<uint32_t> lens;
    std::string seqs;
    std::vector<std::pair<uint32_t, uint8_t>> calls;
    for (;;) {
        unsigned n = 0;
        const bool more = gz ? q.parseChunkGZ(n) : q.parseChunk(n);
        calls.push_back({n, more ? 1 : 0});
        for (unsigned i = 0; i < n; ++i) {
            lens.push_back(static_cast<uint32_t>(q.readBuffer[i].seq.size()));
            seqs += q.readBuffer[i].seq;
        }
        if (!more) break;
    }
    std::ofstream of(argv[4], std::ios::binary | std::ios::trunc);
    const char magic[8] = {'F', 'A', 'M', 'E', 'F', 'Q', '0', '1'};
    of.write(magic, 8);
    uint64_t n = lens.size(), nc = calls.size();
    of.write(reinterpret_cast<const char*>(&n), 8);
    of.write(reinterpret_cast<const char*>(lens.data()), 4 * n);
    of.write(seqs.data(), static_cast<std::streamsize>(seqs.size()));
    of.write(reinterpret_cast<const char*>(&nc), 8);
    for (auto& c : calls) {
        of.write(reinterpret_cast<const char*>(&c.first), 4);
        of.write(reinterpret_cast<const char*>(&c.second), 1);
    }
    std::cout << "HARNESS parsed=" << n << " calls=" << nc << "\n";
    return 0;
}

static int do_sc(int argc, char** argv)
{
    if (argc < 5) return 2;
    bool both = false, gz = false, pe = false;
    for (int i = 5; i < argc; ++i) {
        if (!strcmp(argv[i], "--unord_reads")) both = true;
        if (!strcmp(argv[i], "--gz")) gz = true;
        if (!strcmp(argv[i], "--pe")) pe = true;
    }
    RefGenome ref{std::string(argv[2])};
    ReadQueue q(argv[4], ref, gz, both, pe);
    std::ifstream meta(argv[3]);
    std::string line;
    uint64_t cells = 0;
    while (std::getline(meta, line)) {
        // fields separated by single spaces, as main.cpp:451-457 splits them
        std::vector<std::string> f;
        size_t old = 0, pos;
        while ((pos = line.find_first_of(' ', old)) != std::string::npos) {
            f.push_back(line.substr(old, pos - old));
            old = pos + 1;
        }
        f.push_back(line.substr(old));
        if (f.size() < (pe ? 3u : 2u)) continue;
        if (pe) q.matchSCBatchPaired(f[1].c_str(), f[2].c_str(), f[0], gz);
        else q.matchSCBatch(f[1].c_str(), f[0], gz);
        ++cells;
    }
    std::cerr << "HARNESS cells=" << cells << "\n";
    return 0;
}

// ---- units ---------------------------------------------------------------
// Input and output are little-endian binary streams; every case is
// self-describing.

struct In {
    std::ifstream f;
    template <typename T> T get()
    {
        T v{};
        f.read(reinterpret_cast<char*>(&v), sizeof v);
        return v;
    }
    std::string bytes(size_t n)
    {
        std::string s(n, '\0');
        f.read(&s[0], static_cast<std::streamsize>(n));
        return s;
    }
};

struct Out {
    std::ofstream f;
    template <typename T> void put(T v) { f.write(reinterpret_cast<const char*>(&v), sizeof v); }
};

// ShiftAnd<E>: case = u8 E, u8 rev, u32 plen, pattern, u32 tlen, text.
// rev scans text[tlen-1] down to text[0] (queryRevSeq with `end` one before
// text[0], which a leading pad letter keeps inside the vector); out = u32 n,
// then n x (u64 position, u8 errors)
template <size_t E>
static void sa_case(std::string& pat, const std::string& text, bool rev, std::array<uint8_t, 16>& lmap, Out& o)
{
    ShiftAnd<E> sa(pat, lmap);
    std::vector<char> t;
    t.push_back('A');
    t.insert(t.end(), text.begin(), text.end());
    std::vector<uint64_t> pos;
    std::vector<uint8_t> err;
    if (rev) sa.queryRevSeq(t.end() - 1, t.begin(), pos, err);
    else sa.querySeq(t.begin() + 1, t.end(), pos, err);
    o.put<uint32_t>(static_cast<uint32_t>(pos.size()));
    for (size_t i = 0; i < pos.size(); ++i) {
        o.put<uint64_t>(pos[i]);
        o.put<uint8_t>(err[i]);
    }
}

// LevenshtDP<uint16_t, B>: case = u8 band, u8 rev, u8 cmp (0 CompiFwd,
// 1 CompiRev, 2 equality, 3 complement), u32 L, row, u32 rlen, ref,
// u64 ref_end (colPat = ref + ref_end); out = u16 getEditDist(), u32 n,
// n x u8 ERROR_T
struct CmpEq {
    uint8_t operator()(const char a, const char b) { return a == b ? 0 : 1; }
};
struct CmpWc {
    uint8_t operator()(const char a, const char b)
    {
        switch (a) {
        case 'A': return b == 'T' ? 0 : 1;
        case 'C': return b == 'G' ? 0 : 1;
        case 'G': return b == 'C' ? 0 : 1;
        case 'T': return b == 'A' ? 0 : 1;
        }
        return 1;
    }
};

template <size_t B, typename C>
static void dp_run(const std::string& row, const char* col, bool rev, Out& o)
{
    C comp;
    LevenshtDP<uint16_t, B> lev(row, col);
    std::vector<ERROR_T> al;
    if (rev) {
        lev.template runDPFillRev<C>(comp);
        lev.template backtrackDPRev<C>(comp, al);
    } else {
        lev.template runDPFill<C>(comp);
        lev.template backtrackDP<C>(comp, al);
    }
    o.put<uint16_t>(lev.getEditDist());
    o.put<uint32_t>(static_cast<uint32_t>(al.size()));
    for (auto e : al) o.put<uint8_t>(static_cast<uint8_t>(e));
}

template <size_t B>
static void dp_case(const std::string& row, const char* col, bool rev, int cmp, Out& o)
{
    switch (cmp) {
    case 0: dp_run<B, ReadQueue::CompiFwd>(row, col, rev, o); break;
    case 1: dp_run<B, ReadQueue::CompiRev>(row, col, rev, o); break;
    case 2: dp_run<B, CmpEq>(row, col, rev, o); break;
    default: dp_run<B, CmpWc>(row, col, rev, o); break;
    }
}

static int do_units(int argc, char** argv)
{
    if (argc < 5) return 2;
    const std::string kind = argv[2];
    In in;
    in.f.open(argv[3], std::ios::binary);
    Out o;
    o.f.open(argv[4], std::ios::binary | std::ios::trunc);
    if (!in.f || !o.f) return 1;
    const uint32_t ncase = in.get<uint32_t>();
    if (kind == "ntps64") {
        // case = u32 len, bases; out = (len-31) direct hashes, then the same
        // k-mers by the rolling update as getSeedRefs runs it (ReadQueue.cpp:2363-2404)
        for (uint32_t c = 0; c < ncase; ++c) {
            const uint32_t len = in.get<uint32_t>();
            std::string seq = in.bytes(len);
            const unsigned K = MyConst::KMERLEN;
            uint64_t h = 0;
            for (uint32_t k = 0; k + K <= len; ++k) o.put<uint64_t>(ntHash::NTPS64(seq.data() + k, MyConst::SEED, K, h));
            uint64_t fh = 0;
            o.put<uint64_t>(ntHash::NTPS64(seq.data(), MyConst::SEED, K, fh));
            for (uint32_t k = 0; k + K < len; ++k)
                o.put<uint64_t>(ntHash::NTPS64(seq.data() + k + 1, MyConst::SEED, seq[k], seq[k + K], K, fh));
        }
    } else if (kind == "shiftand") {
        std::array<uint8_t, 16> lmap{};
        lmap['A' % 16] = 0;
        lmap['C' % 16] = 1;
        lmap['G' % 16] = 2;
        lmap['T' % 16] = 3;
        for (uint32_t c = 0; c < ncase; ++c) {
            const uint8_t E = in.get<uint8_t>();
            const bool rev = in.get<uint8_t>() != 0;
            std::string pat = in.bytes(in.get<uint32_t>());
            const std::string text = in.bytes(in.get<uint32_t>());
            switch (E) {
            case 0: sa_case<0>(pat, text, rev, lmap, o); break;
            case 1: sa_case<1>(pat, text, rev, lmap, o); break;
            case 2: sa_case<2>(pat, text, rev, lmap, o); break;
            case 3: sa_case<3>(pat, text, rev, lmap, o); break;
            case 4: sa_case<4>(pat, text, rev, lmap, o); break;
            case 5: sa_case<5>(pat, text, rev, lmap, o); break;
            case 6: sa_case<6>(pat, text, rev, lmap, o); break;
            default: return 1;
            }
        }
    } else if (kind == "dp") {
        for (uint32_t c = 0; c < ncase; ++c) {
            const uint8_t band = in.get<uint8_t>();
            const bool rev = in.get<uint8_t>() != 0;
            const int cmp = in.get<uint8_t>();
            const std::string row = in.bytes(in.get<uint32_t>());
            const std::string ref = in.bytes(in.get<uint32_t>());
            const uint64_t end = in.get<uint64_t>();
            const char* col = ref.data() + end;
            switch (band) {
            case 0: dp_case<0>(row, col, rev, cmp, o); break;
            case 1: dp_case<1>(row, col, rev, cmp, o); break;
            case 2: dp_case<2>(row, col, rev, cmp, o); break;
            case 8: dp_case<8>(row, col, rev, cmp, o); break;
            default: return 1;
            }
        }
    } else if (kind == "hopscotch") {
        // case = u64 bucket_count, u32 n, n x u32 keys (repeats bump the
        // count, as ++map[metaId] in getSeedRefs); out = u32 m, m x (u32 key,
        // u32 count) in iteration order, u64 bucket_count() at the end
        for (uint32_t c = 0; c < ncase; ++c) {
            const uint64_t bc = in.get<uint64_t>();
            const uint32_t n = in.get<uint32_t>();
            tsl::hopscotch_map<uint32_t, uint16_t, ReadQueue::MetaHash> m(bc);
            for (uint32_t i = 0; i < n; ++i) ++m[in.get<uint32_t>()];
            o.put<uint32_t>(static_cast<uint32_t>(m.size()));
            for (auto it = m.begin(); it != m.end(); ++it) {
                o.put<uint32_t>(it->first);
                o.put<uint32_t>(it->second);
            }
            o.put<uint64_t>(m.bucket_count());
        }
    } else {
        return 2;
    }
    return in.f ? 0 : 1;
}

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::cerr << "usage: build|match ...\n";
        return 2;
    }
    MyConst::sanityChecks();
    std::string mode = argv[1];
    if (mode == "build") return do_build(argc, argv);
    if (mode == "match") return do_match(argc, argv);
    if (mode == "match_pe") return do_match_pe(argc, argv);
    if (mode == "parse") return do_parse(argc, argv);
    if (mode == "units") return do_units(argc, argv);
    if (mode == "sc") return do_sc(argc, argv);
    std::cerr << "unknown mode\n";
    return 2;
}
