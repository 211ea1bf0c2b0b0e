/* fame_index.h -- C ABI of the index builder (libfameindex.so).
 *
 * Writes FAME's --load_index file and its <index>_strands companion, the
 * output of FAME's --store_index path: the RefGenome construction constructor
 * (FischerJo/FAME RefGenome.cpp:26-74, with generateWindows :532,
 * generateHashes :570, filterHashTable :2267, filterRedundancyInHashTable
 * :2191) followed by RefGenome::save (RefGenome.cpp:83-228).  HTABSIZE,
 * READLEN and KMERCUTOFF are arguments instead of compile-time constants.
 * Returns 0 on success, a negative code with a message in err otherwise.
 */
#ifndef FAME_INDEX_H
#define FAME_INDEX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* chroms: n chromosome byte strings as readReference leaves them
 * (RefReader_istr.cpp:25-160), names[i] the chrMap entry of chromosome i */
int fame_build_index(const uint8_t* const* chroms, const uint64_t* lens, const char* const* names, uint32_t n,
                     uint32_t readlen, uint64_t htab, uint64_t kmercutoff, int noloss, const char* out, char* err,
                     size_t errlen);

/* the same from a FASTA file, parsed like readReference without --human_opt */
int fame_build_index_fasta(const char* fasta, uint32_t readlen, uint64_t htab, uint64_t kmercutoff, int noloss,
                           const char* out, char* err, size_t errlen);

#ifdef __cplusplus
}
#endif
#endif
