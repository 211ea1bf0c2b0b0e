// ReadQueue for FAME's own main.cpp linked against libfamehip.so: the public
// members main.cpp calls (FischerJo/FAME ReadQueue.h:44-129), each forwarding
// to the C ABI of include/fame_hip.h.  FASTQ chunks are parsed by
// fame_fastq_* straight into pinned buffers (parseChunk / parseChunkGZ,
// ReadQueue.cpp:218-373); matching, strand decision, CpG counts and the
// report run on the GPU.  Fatal errors print the library's message and
// exit(1), FAME's own convention.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <string>

#include "CONST.h"
#include "RefGenome.h"

struct fame_ctx;
struct fame_fastq;
struct fame_sc;

class ReadQueue
{
  public:
    ReadQueue() = delete;
    // single end (ReadQueue.cpp:25-78)
    ReadQueue(const char* filePath, RefGenome& ref, const bool isGZ, const bool bsFlag);
    // paired end (ReadQueue.cpp:79-180)
    ReadQueue(const char* filePath, const char* filePath2, RefGenome& reference, const bool isGZ,
              const bool bsFlag);
    // single cell (ReadQueue.cpp:181-216)
    ReadQueue(const char* scOutputPath, RefGenome& reference, const bool isGZ, const bool bsFlag,
              const bool isP);
    ~ReadQueue();
    ReadQueue(const ReadQueue&) = delete;
    ReadQueue& operator=(const ReadQueue&) = delete;

    bool parseChunk(unsigned int& procReads);
    bool parseChunkGZ(unsigned int& procReads);
    void decideStrand();
    bool matchReads(const unsigned int& procReads, uint64_t& succMatch, uint64_t& nonUniqueMatch,
                    uint64_t& unSuccMatch, const bool getStranded);
    bool matchPairedReads(const unsigned int& procReads, uint64_t& succMatch, uint64_t& nonUniqueMatch,
                          uint64_t& unSuccMatch, uint64_t& succPairedMatch, uint64_t& tooShortCountMatch,
                          const bool getStranded);
    bool matchSCBatch(const char* scFile, const std::string scId, const bool isGZ);
    bool matchSCBatchPaired(const char* scFile1, const char* scFile2, const std::string scId, const bool isGZ);
    void printMethylationLevels(std::string& filename);

  private:
    void open(const RefGenome& ref, bool paired);
    void check(int rc) const;
    bool parse(unsigned int& procReads);

    fame_ctx* ctx = nullptr;
    fame_fastq* fq = nullptr;
    fame_fastq* fq2 = nullptr;
    fame_sc* sc = nullptr;
    bool bothStrands = false;
    bool isPaired = false;
    // the chunk last parsed (buffers owned by the parsers)
    const uint8_t* seqs = nullptr;
    const uint64_t* offs = nullptr;
    const uint32_t* lens = nullptr;
    const uint8_t* seqs2 = nullptr;
    const uint64_t* offs2 = nullptr;
    const uint32_t* lens2 = nullptr;
};
