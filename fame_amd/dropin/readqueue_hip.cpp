// RefGenome / ReadQueue of FAME's own main.cpp over libfamehip.so and
// libfameindex.so (the declarations are in the RefGenome.h / ReadQueue.h of
// this folder, which the build puts ahead of FAME's own headers).

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "../../include/fame_hip.h"
#include "../../include/fame_index.h"
#include "ReadQueue.h"
#include "RefGenome.h"

// ---------------------------------------------------------------------------
// RefGenome
// ---------------------------------------------------------------------------

RefGenome::RefGenome(std::vector<struct CpG>&& cpgTab, std::vector<struct CpG>&& cpgStartTab,
                     std::vector<std::vector<char> >& genSeq, const bool noLoss,
                     std::unordered_map<uint8_t, std::string>& chromMap)
    : seqs(genSeq), noloss(noLoss)
{
    (void)cpgTab;
    (void)cpgStartTab;
    names.resize(seqs.size());
    for (size_t i = 0; i < seqs.size(); ++i) {
        auto it = chromMap.find((uint8_t)i);
        names[i] = it == chromMap.end() ? std::string() : it->second;
    }
}

RefGenome::RefGenome(std::string filePath) : path(std::move(filePath))
{
    // the header checks of RefGenome::load (RefGenome.cpp:239-280)
    std::cout << "Start reading index file " << path << "\n";
    std::ifstream ifs(path, std::ifstream::binary);
    uint64_t htab = 0, cutoff = 0;
    uint32_t readlen = 0, winlen = 0, kmerlen = 0, seedbits = 0;
    ifs.read(reinterpret_cast<char*>(&htab), 8);
    ifs.read(reinterpret_cast<char*>(&readlen), 4);
    ifs.read(reinterpret_cast<char*>(&winlen), 4);
    ifs.read(reinterpret_cast<char*>(&kmerlen), 4);
    ifs.read(reinterpret_cast<char*>(&cutoff), 8);
    ifs.read(reinterpret_cast<char*>(&seedbits), 4);
    if (!ifs) {
        std::cerr << "Cannot read index file " << path << "\n\n";
        exit(1);
    }
    const char* bad = nullptr;
    if (htab != MyConst::HTABSIZE)
        bad = "Hash table size in source code and index file are different!";
    else if (readlen != MyConst::READLEN)
        bad = "Read length used in source code and index file are different!";
    else if (winlen != MyConst::WINLEN)
        bad = "Meta CpG length used in source code and index file are different!";
    else if (kmerlen != MyConst::KMERLEN)
        bad = "k-mer length used in source code and index file are different!";
    else if (cutoff != MyConst::KMERCUTOFF)
        bad = "k-mer cutoff used in source code and index file are different!";
    else if (seedbits != MyConst::SEEDBITS)
        bad = "Different seed used in index and methylation prediction!";
    if (bad) {
        std::cerr << bad << "\n\n";
        exit(1);
    }
}

void RefGenome::save(const std::string& filepath)
{
    std::vector<const uint8_t*> ptrs(seqs.size());
    std::vector<uint64_t> lens(seqs.size());
    std::vector<const char*> nm(seqs.size());
    for (size_t i = 0; i < seqs.size(); ++i) {
        ptrs[i] = reinterpret_cast<const uint8_t*>(seqs[i].data());
        lens[i] = seqs[i].size();
        nm[i] = names[i].c_str();
    }
    char err[256] = {0};
    std::cout << "Writing index to " << filepath << "\n";
    if (fame_build_index(ptrs.data(), lens.data(), nm.data(), (uint32_t)seqs.size(), MyConst::READLEN,
                         MyConst::HTABSIZE, MyConst::KMERCUTOFF, noloss ? 1 : 0, filepath.c_str(), err,
                         sizeof err) != 0) {
        std::cerr << "Index construction failed: " << err << "\n";
        exit(1);
    }
    path = filepath;
}

// ---------------------------------------------------------------------------
// ReadQueue
// ---------------------------------------------------------------------------

void ReadQueue::check(int rc) const
{
    if (rc != FAME_OK) {
        std::cerr << "ERROR: " << fame_last_error(ctx) << "\n";
        exit(1);
    }
}

void ReadQueue::open(const RefGenome& ref, bool paired)
{
    isPaired = paired;
    const int rc = fame_open(ref.indexPath().c_str(), bothStrands ? 1 : 0, &ctx);
    check(rc);
    if (paired) check(fame_set_pair_dist(ctx, MyConst::MINPDIST, MyConst::MAXPDIST));
}

ReadQueue::ReadQueue(const char* filePath, RefGenome& ref, const bool isGZ, const bool bsFlag) : bothStrands(bsFlag)
{
    open(ref, false);
    if (fame_fastq_open(filePath, isGZ ? 1 : 0, &fq) != FAME_OK) {
        std::cerr << "ERROR: Opening file `" << filePath << "' failed.\n";
        exit(1);
    }
}

ReadQueue::ReadQueue(const char* filePath, const char* filePath2, RefGenome& reference, const bool isGZ,
                     const bool bsFlag)
    : bothStrands(bsFlag)
{
    open(reference, true);
    if (fame_fastq_open(filePath, isGZ ? 1 : 0, &fq) != FAME_OK ||
        fame_fastq_open(filePath2, isGZ ? 1 : 0, &fq2) != FAME_OK) {
        std::cerr << "ERROR: Opening read files failed.\n";
        exit(1);
    }
}

ReadQueue::ReadQueue(const char* scOutputPath, RefGenome& reference, const bool isGZ, const bool bsFlag,
                     const bool isP)
    : bothStrands(bsFlag)
{
    (void)isGZ;
    open(reference, isP);
    check(fame_sc_open(ctx, scOutputPath, &sc));
}

ReadQueue::~ReadQueue()
{
    if (fq) fame_fastq_close(fq);
    if (fq2) fame_fastq_close(fq2);
    if (sc) fame_sc_close(sc);
    if (ctx) fame_close(ctx);
}

// parseChunk / parseChunkGZ (ReadQueue.cpp:218-373): up to CHUNKSIZE records;
// true while a full chunk was read.  The parser was opened for the file's
// compression, so both members read the same way.
bool ReadQueue::parse(unsigned int& procReads)
{
    uint32_t n = 0;
    if (fame_fastq_next(fq, MyConst::CHUNKSIZE, &n, &seqs, &offs, &lens) != FAME_OK) {
        std::cerr << "ERROR: reading the read file failed.\n";
        exit(1);
    }
    if (isPaired) {
        uint32_t n2 = 0;
        if (fame_fastq_next(fq2, MyConst::CHUNKSIZE, &n2, &seqs2, &offs2, &lens2) != FAME_OK) {
            std::cerr << "ERROR: reading the second read file failed.\n";
            exit(1);
        }
        if (n2 != n) { // ReadQueue.cpp:280-286
            std::cerr << "Not the same number of reads available in the paired read files. Terminating...\n\n";
            exit(1);
        }
    }
    procReads = n;
    return n >= MyConst::CHUNKSIZE;
}

bool ReadQueue::parseChunk(unsigned int& procReads) { return parse(procReads); }
bool ReadQueue::parseChunkGZ(unsigned int& procReads) { return parse(procReads); }

// decideStrand ReadQueue.cpp:376-400, with its console report
void ReadQueue::decideStrand()
{
    uint64_t f = 0, r = 0;
    int fwd = 1;
    check(fame_decide_strand(ctx, &f, &r, &fwd));
    std::cout << "\nFwd matches: " << f;
    std::cout << "\nRev matches: " << r;
    const double odds = (double)(f + 1) / (double)(r + 1);
    std::cout << "\nOdds of matching to fwd/rev strand: " << odds << "\n\n";
    if (odds > 0.005 && odds < 200)
        std::cout << "Warning! Many of the reads are mapped in different orientation. Stranding might harm the "
                     "prediction performance.\n\n";
    std::cout << "Deciding conversion status of read 1.\n";
    std::cout << (fwd ? "\tMatching read 1 as C->T converted.\n\n" : "\tMatching read 1 as G->A converted.\n\n");
}

bool ReadQueue::matchReads(const unsigned int& procReads, uint64_t& succMatch, uint64_t& nonUniqueMatch,
                           uint64_t& unSuccMatch, const bool getStranded)
{
    uint64_t c[3] = {0, 0, 0};
    check(fame_match_reads(ctx, seqs, offs, lens, procReads, getStranded ? 1 : 0, c, nullptr, nullptr));
    succMatch += c[0];
    nonUniqueMatch += c[1];
    unSuccMatch += c[2];
    return true;
}

bool ReadQueue::matchPairedReads(const unsigned int& procReads, uint64_t& succMatch, uint64_t& nonUniqueMatch,
                                 uint64_t& unSuccMatch, uint64_t& succPairedMatch, uint64_t& tooShortCountMatch,
                                 const bool getStranded)
{
    uint64_t c[5] = {0, 0, 0, 0, 0};
    check(fame_match_pairs(ctx, seqs, offs, lens, seqs2, offs2, lens2, procReads, getStranded ? 1 : 0, c, nullptr,
                           nullptr, nullptr, nullptr));
    succMatch += c[0];
    nonUniqueMatch += c[1];
    unSuccMatch += c[2];
    succPairedMatch += c[3];
    tooShortCountMatch += c[4];
    return true;
}

// matchSCBatch ReadQueue.cpp:1660-1731: one cell from zero counts through the
// chunk loop of queryRoutine, then its rows of the matrix
bool ReadQueue::matchSCBatch(const char* scFile, const std::string scId, const bool isGZ)
{
    if (fame_fastq_open(scFile, isGZ ? 1 : 0, &fq) != FAME_OK) {
        std::cerr << "ERROR: Opening file `" << std::string(scFile) << "' failed.\n";
        return EXIT_FAILURE;
    }
    check(fame_reset(ctx));
    unsigned int n = 0;
    uint64_t succ = 0, nonUnique = 0, unSucc = 0;
    if (!bothStrands) {
        parse(n);
        matchReads(n, succ, nonUnique, unSucc, true);
        decideStrand();
    }
    while (parse(n)) matchReads(n, succ, nonUnique, unSucc, false);
    matchReads(n, succ, nonUnique, unSucc, false);
    std::cout << "Finished " << std::string(scFile) << "\n\tOverall successfully matched: " << succ
              << "\n\tUnsuccessfully matched: " << unSucc << "\n\tNonunique matches: " << nonUnique << "\n";
    fame_fastq_close(fq);
    fq = nullptr;
    check(fame_sc_write_cell(sc, scId.c_str()));
    return true;
}

// matchSCBatchPaired ReadQueue.cpp:1732-1854
bool ReadQueue::matchSCBatchPaired(const char* scFile1, const char* scFile2, const std::string scId,
                                   const bool isGZ)
{
    if (fame_fastq_open(scFile1, isGZ ? 1 : 0, &fq) != FAME_OK ||
        fame_fastq_open(scFile2, isGZ ? 1 : 0, &fq2) != FAME_OK) {
        std::cerr << "ERROR: Opening files of cell " << scId << " failed.\n";
        return EXIT_FAILURE;
    }
    check(fame_reset(ctx));
    unsigned int n = 0;
    uint64_t succ = 0, nonUnique = 0, unSucc = 0, succPaired = 0, tooShort = 0;
    if (!bothStrands) {
        parse(n);
        matchPairedReads(n, succ, nonUnique, unSucc, succPaired, tooShort, true);
        decideStrand();
    }
    while (parse(n)) matchPairedReads(n, succ, nonUnique, unSucc, succPaired, tooShort, false);
    matchPairedReads(n, succ, nonUnique, unSucc, succPaired, tooShort, false);
    std::cout << "Finished cell " << scId << "\n\tOverall successfully matched: " << succ
              << "\n\tSuccessfully matched pairs: " << succPaired << "\n";
    fame_fastq_close(fq);
    fame_fastq_close(fq2);
    fq = fq2 = nullptr;
    check(fame_sc_write_cell(sc, scId.c_str()));
    return true;
}

// printMethylationLevels ReadQueue.cpp:1860-1893: <filename>_cpg.tsv
void ReadQueue::printMethylationLevels(std::string& filename)
{
    check(fame_print_methylation_levels(ctx, filename.c_str()));
}
