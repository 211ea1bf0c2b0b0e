// RefGenome for FAME's own main.cpp linked against libfamehip.so: the two
// constructors and save() that main.cpp uses (FischerJo/FAME
// RefGenome.h:40-105).  The index itself never lives in host memory here:
// RefGenome(path) checks the file's header against the compiled constants as
// RefGenome::load does (RefGenome.cpp:239-280) and keeps the path, which the
// ReadQueue hands to fame_open (HBM-resident index); the build constructor
// keeps the sequences and save() writes the index through libfameindex.so.
#pragma once

// the standard headers FAME's RefGenome.h brings in, which main.cpp relies on
#include <cstdint>
#include <fstream>
#include <istream>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "CONST.h"
#include "structs.h"

class RefGenome
{
  public:
    RefGenome() = delete;

    // --store_index path (RefGenome.cpp:26-74); cpgTab / cpgStartTab are
    // derived again from the sequences by the builder
    RefGenome(std::vector<struct CpG>&& cpgTab, std::vector<struct CpG>&& cpgStartTab,
              std::vector<std::vector<char> >& genSeq, const bool noloss,
              std::unordered_map<uint8_t, std::string>& chromMap);

    // --load_index path (RefGenome.cpp:77-80)
    RefGenome(std::string filePath);

    ~RefGenome() = default;

    // RefGenome::save (RefGenome.cpp:83-228), <filepath> and <filepath>_strands
    void save(const std::string& filepath);

    // the index file a ReadQueue opens on the device
    const std::string& indexPath() const { return path; }

  private:
    std::string path;
    std::vector<std::vector<char> > seqs;
    std::vector<std::string> names;
    bool noloss = false;
};
