// FASTQ ingestion for the matching path: ReadQueue::parseChunk /
// parseChunkGZ (ReadQueue.cpp:218-373) as a multi-threaded scan into pinned
// host buffers (reads back to back, u64 offsets, u32 lengths), the form the
// C ABI and the serving loop take and copy to the device directly.
//
// Record semantics follow the reference's std::getline loop: the input is a
// sequence of lines ('\n' ends a line and is dropped; text after the last
// '\n' is one more line); a record starts at every fourth line, whatever it
// holds, and its sequence is the next line (empty if the input ends first).

#include <hip/hip_runtime.h>
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fame_hip.h"

namespace {

constexpr size_t kBlock = 64ull << 20; // bytes read per refill

// output buffer: pinned (hipHostMalloc) so that copies to the device run as
// DMA; plain host memory where no device is present (the parser itself needs
// none)
struct Pinned {
    void* p = nullptr;
    size_t bytes = 0;
    bool pinned = false;
    ~Pinned() { release(); }
    void release()
    {
        if (p) {
            if (pinned) (void)hipHostFree(p);
            else std::free(p);
        }
        p = nullptr;
        bytes = 0;
    }
    bool reserve(size_t want)
    {
        if (want <= bytes && p) return true;
        release();
        want = std::max<size_t>(want, 1);
        if (hipHostMalloc(&p, want, 0) == hipSuccess) {
            pinned = true;
        } else {
            pinned = false;
            p = std::malloc(want);
            if (!p) return false;
        }
        bytes = want;
        return true;
    }
};

} // namespace

struct fame_fastq {
    FILE* f = nullptr;
    gzFile gz = nullptr;
    bool eof = false;
    bool ioerr = false;
    std::vector<char> buf; // unconsumed input
    size_t len = 0;        // bytes of buf in use
    unsigned threads = 1;
    Pinned seqs, offs, lens;
    std::string err;
};

namespace {

size_t refill(fame_fastq* q)
{
    if (q->eof) return 0;
    if (q->buf.size() < q->len + kBlock) q->buf.resize(q->len + kBlock);
    size_t got = 0;
    if (q->gz) {
        const int r = gzread(q->gz, q->buf.data() + q->len, (unsigned)kBlock);
        if (r < 0) q->ioerr = true;
        got = r > 0 ? (size_t)r : 0;
    } else {
        got = fread(q->buf.data() + q->len, 1, kBlock, q->f);
        if (got == 0 && ferror(q->f)) q->ioerr = true;
    }
    if (got == 0) q->eof = true;
    q->len += got;
    return got;
}

// appends the newline positions in [from, to) to out, found by `threads`
// scanners in parallel
void newlines(const char* b, size_t from, size_t to, unsigned threads, std::vector<size_t>& out)
{
    if (to <= from) return;
    const size_t len = to - from;
    const unsigned T = (unsigned)std::max<size_t>(1, std::min<size_t>(threads, len / (1 << 20) + 1));
    std::vector<std::vector<size_t>> part(T);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t) {
        th.emplace_back([&, t]() {
            const char* p = b + from + len * t / T;
            const char* e = b + from + len * (t + 1) / T;
            while (p < e) {
                const char* x = (const char*)memchr(p, '\n', (size_t)(e - p));
                if (!x) break;
                part[t].push_back((size_t)(x - b));
                p = x + 1;
            }
        });
    }
    for (auto& x : th) x.join();
    for (auto& v : part) out.insert(out.end(), v.begin(), v.end());
}

} // namespace

extern "C" {

int fame_fastq_open(const char* path, int gz, fame_fastq** out)
{
    if (!path || !out) return FAME_ERR_ARG;
    *out = nullptr;
    fame_fastq* q = new fame_fastq();
    if (gz) q->gz = gzopen(path, "rb");
    else q->f = fopen(path, "rb");
    if (!q->gz && !q->f) {
        delete q;
        return FAME_ERR_IO;
    }
    if (q->gz) gzbuffer(q->gz, 1 << 20);
    q->threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    *out = q;
    return FAME_OK;
}

int fame_fastq_next(fame_fastq* q, uint32_t max_reads, uint32_t* n, const uint8_t** seqs, const uint64_t** offs,
                    const uint32_t** lens)
{
    if (!q || !n || !seqs || !offs || !lens || !max_reads) return FAME_ERR_ARG;
    *n = 0;
    // enough input for max_reads records, or all of it
    std::vector<size_t> nl;
    size_t scanned = 0;
    for (;;) {
        newlines(q->buf.data(), scanned, q->len, q->threads, nl);
        scanned = q->len;
        if (nl.size() >= 4ull * max_reads || q->eof) break;
        refill(q);
    }
    if (q->ioerr) return FAME_ERR_IO;
    // lines: ends at the newlines, plus a last unterminated line at EOF
    size_t nlines = nl.size();
    const bool tail = q->eof && q->len > (nl.empty() ? 0 : nl.back() + 1);
    if (tail) nlines += 1;
    auto line_end = [&](size_t i) { return i < nl.size() ? nl[i] : q->len; };
    auto line_start = [&](size_t i) { return i == 0 ? (size_t)0 : nl[i - 1] + 1; };
    // whole records only, unless the input has ended
    size_t nrec = q->eof ? (nlines + 3) / 4 : nlines / 4;
    nrec = std::min<size_t>(nrec, max_reads);
    if (!nrec) return FAME_OK;
    if (!q->offs.reserve(8 * nrec + 8) || !q->lens.reserve(4 * nrec + 4)) return FAME_ERR_CAPACITY;
    uint64_t* o = (uint64_t*)q->offs.p;
    uint32_t* l = (uint32_t*)q->lens.p;
    uint64_t total = 0;
    for (size_t r = 0; r < nrec; ++r) {
        const size_t li = 4 * r + 1;
        const size_t len = li < nlines ? line_end(li) - line_start(li) : 0;
        if (len > 0xFFFFFFFFull) return FAME_ERR_ARG;
        o[r] = total;
        l[r] = (uint32_t)len;
        total += len;
    }
    if (!q->seqs.reserve(total + 64)) return FAME_ERR_CAPACITY;
    uint8_t* s = (uint8_t*)q->seqs.p;
    {
        const unsigned T = (unsigned)std::max<size_t>(1, std::min<size_t>(q->threads, nrec / 4096 + 1));
        std::vector<std::thread> th;
        for (unsigned t = 0; t < T; ++t) {
            th.emplace_back([&, t]() {
                for (size_t r = nrec * t / T; r < nrec * (t + 1) / T; ++r) {
                    const size_t li = 4 * r + 1;
                    if (li < nlines && l[r]) memcpy(s + o[r], q->buf.data() + line_start(li), l[r]);
                }
            });
        }
        for (auto& x : th) x.join();
    }
    // drop the consumed records' bytes
    const size_t used_lines = std::min(4 * nrec, nlines);
    const size_t cut = used_lines == 0 ? 0 : (used_lines <= nl.size() ? nl[used_lines - 1] + 1 : q->len);
    memmove(q->buf.data(), q->buf.data() + cut, q->len - cut);
    q->len -= cut;
    *n = (uint32_t)nrec;
    *seqs = s;
    *offs = o;
    *lens = l;
    return FAME_OK;
}

int fame_fastq_close(fame_fastq* q)
{
    if (!q) return FAME_OK;
    if (q->gz) gzclose(q->gz);
    if (q->f) fclose(q->f);
    delete q;
    return FAME_OK;
}

} // extern "C"
