// Host side of libfamehip.so: index loading into HBM, per-batch
// orchestration of the kernels, and the C ABI declared in include/fame_hip.h.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fame_hip.h"
#include "fame_internal.h"

using namespace fame;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // grows (never shrinks); contents are not preserved
    hipError_t reserve(size_t want)
    {
        if (want <= bytes && p) return hipSuccess;
        release();
        size_t b = want ? want : 1;
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    template <typename T>
    T* as() const
    {
        return static_cast<T*>(p);
    }
};

struct HipError {
    hipError_t e;
    const char* what;
};

#define HIPC(call)                                                                                           \
    do {                                                                                                     \
        hipError_t _e = (call);                                                                              \
        if (_e != hipSuccess) throw HipError{_e, #call};                                                     \
    } while (0)

// pinned host memory that grows on demand (staging of small host arrays, so
// that their copies to the device are asynchronous)
struct HostStage {
    void* p = nullptr;
    size_t cap = 0;
    HostStage() = default;
    HostStage(const HostStage&) = delete;
    HostStage& operator=(const HostStage&) = delete;
    ~HostStage()
    {
        if (p) (void)hipHostFree(p);
    }
    template <typename T>
    T* get(size_t count)
    {
        const size_t bytes = count * sizeof(T) + 64;
        if (bytes > cap) {
            if (p) (void)hipHostFree(p);
            p = nullptr;
            cap = 0;
            HIPC(hipHostMalloc(&p, bytes, 0));
            cap = bytes;
        }
        return (T*)p;
    }
};

struct FameError {
    int code;
    std::string msg;
};

} // namespace

struct GraphPlan;

struct fame_ctx {
    std::string err;
    // a serving worker: its own batch buffers, stream and counters over the
    // parent's index and CpG accumulators
    fame_ctx* parent = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr; // second lane of the batched form
    bool both = false;
    bool match_r1_fwd = true;
    // header
    uint64_t htab = 0, kmercutoff = 0, ncpg = 0, nchrom = 0, nkmer = 0, nwin = 0, nfilt = 0;
    uint32_t readlen = 0, winlen = 0, kmerlen = 0, seedbits = 0;
    // host copies needed for the report
    std::vector<uint32_t> cpg_pos;
    std::vector<uint8_t> cpg_chrom;
    std::vector<std::string> chr_name; // 256 entries, "" when unmapped
    // device index
    DevBuf d_tab, d_slot, d_kent, d_kpe, d_filt32, d_win, d_seq, d_coff, d_clen, d_cpg_pos, d_cpg_chrom, d_pk,
        d_pkoff, d_wtext, d_gkey, d_gslot, d_kpos;
    uint64_t seq_total = 0;                // letters of all chromosomes
    std::vector<uint64_t> h_pkoff, h_clen; // per chromosome, for d_wtext
    DevBuf d_meth, d_counters, d_flags, d_stats;
    uint64_t stats[FAME_NSTATS] = {0};
    DevIndex ix{};
    // batch buffers
    DevBuf b_seqs, b_offs, b_lens, b_rev, b_pre, b_codes, b_cand_off, b_cand_nf, b_cand_nr, b_pool_n;
    DevBuf b_job_meta[FAME_NPOOL], b_job_ro[FAME_NPOOL], b_job_nm[FAME_NPOOL], b_job_match[FAME_NPOOL],
        b_job_ovf_off[FAME_NPOOL];
    DevBuf b_ovf_n, b_ovf_pool, b_ovf_job, b_ovf_buf, b_ovf_total;
    DevBuf b_heavy_n, b_heavy_ro, b_heavy_e, b_heavy_bc, b_heavy_scratch, b_mid_ovl;
    DevBuf b_status, b_use, b_mat, b_valid, b_dp, b_dec, b_trace, b_direct, b_dplist, b_methn;
    uint32_t pool_cap[FAME_NPOOL] = {0, 0, 0, 0};
    uint32_t match_cap = 4;
    // paired-end: MINPDIST / MAXPDIST (CONST.h:41-42 defaults) and buffers
    uint32_t minpd = 50, maxpd = 550;
    DevBuf d_pe_counters;
    DevBuf b2_pool_n, b2_job_meta[FAME_NPOOL], b2_job_ro[FAME_NPOOL], b2_job_nm[FAME_NPOOL],
        b2_job_match[FAME_NPOOL], b2_job_ovf_off[FAME_NPOOL];
    DevBuf b2_ovf_n, b2_ovf_pool, b2_ovf_job, b2_ovf_buf, b2_ovf_total;
    DevBuf b_acc1[FAME_NPOOL], b_acc2[FAME_NPOOL];
    DevBuf b_pe_off, b_pe_nf, b_pe_nr, b_pe_m1, b_pe_mapn, b_pe_map, b_pe_srt, b_pe_big, b_pe_list, b_pe_wlist;
    uint32_t pool2_cap[FAME_NPOOL] = {0, 0, 0, 0};
    uint32_t pe_map_cap = 0;
    uint32_t meth_threads = 65536; // lanes of the DP kernel: as many as are resident at once
    double timings[FAME_TIMING_FIELDS] = {0};
    uint64_t last_heavy = 0, last_ovf = 0;
    // batched form: plans per capacity class, their graphs, and counts
    int form = 0; // 0 plain, 1 batched
    std::vector<std::unique_ptr<GraphPlan>> plans;
    uint64_t graph_calls = 0, graph_fallbacks = 0, graph_captures = 0;
    uint32_t pipe_batch = 3u << 20; // largest sub-batch of a batched call split over two lanes
    DevBuf b_pvalid, b_pmat, b_pstatus;  // outputs of a pipelined call
    // reads over FAME_DEV_MAX_READ letters: counters (long_n, longest, DP
    // count), list, rows, DP list and DP scratch
    DevBuf b_long, b_long_list, b_long_codes, b_dp_long, b_ldp, b_ldec, b_ltrace;
    HostStage h_offs;                    // host-buffer calls: read offsets rebased to the span copied
    // batched form, paired: the second lane's context and the lanes' mate
    // offsets / lengths of their sub-batch
    fame_ctx* pair_worker = nullptr;
    DevBuf b_lane_offs[2], b_lane_lens[2];
    uint64_t pipe_calls = 0;
    double big_ms = 0; // paired seed runs with global scratch, last call
};

namespace {

void set_err(fame_ctx* c, const std::string& m)
{
    if (c) c->err = m;
}

// the CpG accumulators the context's batches add to
uint32_t* meth_buf(fame_ctx* c) { return (c->parent ? c->parent : c)->d_meth.as<uint32_t>(); }

// ---------------------------------------------------------------------------
// index loading (RefGenome::load RefGenome.cpp:230-387, read_strands :414-433)
// ---------------------------------------------------------------------------

struct File {
    FILE* f = nullptr;
    explicit File(const std::string& p) { f = fopen(p.c_str(), "rb"); }
    ~File()
    {
        if (f) fclose(f);
    }
    void read(void* p, size_t n)
    {
        if (n && fread(p, 1, n, f) != n) throw FameError{FAME_ERR_IO, "truncated index file"};
    }
    template <typename T>
    T get()
    {
        T v;
        read(&v, sizeof v);
        return v;
    }
    void skip(uint64_t n)
    {
        if (n && fseeko(f, (off_t)n, SEEK_CUR) != 0) throw FameError{FAME_ERR_IO, "truncated index file"};
    }
};

constexpr size_t kStage = 256ull << 20; // host staging chunk for large tables

void load_index(fame_ctx* c, const std::string& path)
{
    File f(path);
    if (!f.f) throw FameError{FAME_ERR_IO, "cannot open index " + path};
    c->htab = f.get<uint64_t>();
    c->readlen = f.get<uint32_t>();
    c->winlen = f.get<uint32_t>();
    c->kmerlen = f.get<uint32_t>();
    c->kmercutoff = f.get<uint64_t>();
    c->seedbits = f.get<uint32_t>();
    if (c->kmerlen != 32) throw FameError{FAME_ERR_FORMAT, "index k-mer length is not 32"};
    if (c->htab == 0) throw FameError{FAME_ERR_FORMAT, "index hash table size is 0"};
    if (c->seedbits != FAME_SEEDBITS)
        throw FameError{FAME_ERR_FORMAT, "index built with another spaced seed (SEEDBITS)"};
    if (c->readlen <= 32 || c->readlen > FAME_MAX_MATE)
        throw FameError{FAME_ERR_FORMAT, "index READLEN outside (32, 255]"};
    hipStream_t st = c->stream;
    // CpG table: {u8 chrom; 3 pad; u32 pos}
    c->ncpg = f.get<uint64_t>();
    {
        std::vector<uint8_t> raw(c->ncpg * 8);
        f.read(raw.data(), raw.size());
        c->cpg_pos.resize(c->ncpg);
        c->cpg_chrom.resize(c->ncpg);
        for (uint64_t i = 0; i < c->ncpg; ++i) {
            c->cpg_chrom[i] = raw[8 * i];
            memcpy(&c->cpg_pos[i], &raw[8 * i + 4], 4);
        }
        HIPC(c->d_cpg_pos.reserve(4 * c->ncpg + 4));
        HIPC(c->d_cpg_chrom.reserve(c->ncpg + 1));
        HIPC(hipMemcpy(c->d_cpg_pos.p, c->cpg_pos.data(), 4 * c->ncpg, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(c->d_cpg_chrom.p, c->cpg_chrom.data(), c->ncpg, hipMemcpyHostToDevice));
    }
    f.skip(8 * f.get<uint64_t>()); // cpgStartTable (only used by dead code)
    // chromosomes
    c->nchrom = f.get<uint64_t>();
    if (c->nchrom > 256) throw FameError{FAME_ERR_FORMAT, "more than 256 chromosomes"};
    std::vector<uint64_t> coff(c->nchrom), clen(c->nchrom);
    {
        long pos0 = ftello(f.f);
        uint64_t total = 0;
        for (uint64_t i = 0; i < c->nchrom; ++i) {
            clen[i] = f.get<uint64_t>();
            coff[i] = total;
            total += clen[i];
            f.skip(clen[i]);
        }
        c->seq_total = total;
        HIPC(c->d_seq.reserve(total + 64));
        HIPC(hipMemset(c->d_seq.p, 0, total + 64));
        fseeko(f.f, pos0, SEEK_SET);
        std::vector<uint8_t> buf;
        for (uint64_t i = 0; i < c->nchrom; ++i) {
            f.get<uint64_t>();
            uint64_t left = clen[i], o = coff[i];
            while (left) {
                size_t n = (size_t)std::min<uint64_t>(left, kStage);
                buf.resize(n);
                f.read(buf.data(), n);
                HIPC(hipMemcpy(c->d_seq.as<uint8_t>() + o, buf.data(), n, hipMemcpyHostToDevice));
                o += n;
                left -= n;
            }
        }
        HIPC(c->d_coff.reserve(8 * c->nchrom + 8));
        HIPC(c->d_clen.reserve(8 * c->nchrom + 8));
        HIPC(hipMemcpy(c->d_coff.p, coff.data(), 8 * c->nchrom, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(c->d_clen.p, clen.data(), 8 * c->nchrom, hipMemcpyHostToDevice));
        // the verify stage's packed copy: 16 letters per word, two words of
        // padding after each chromosome
        std::vector<uint64_t> pkoff(c->nchrom);
        uint64_t words = 0;
        for (uint64_t i = 0; i < c->nchrom; ++i) {
            pkoff[i] = words;
            words += (clen[i] + 15) / 16 + 2;
        }
        HIPC(c->d_pk.reserve(8 * words + 8));
        HIPC(hipMemset(c->d_pk.p, 0, 8 * words + 8));
        HIPC(c->d_pkoff.reserve(8 * c->nchrom + 8));
        HIPC(hipMemcpy(c->d_pkoff.p, pkoff.data(), 8 * c->nchrom, hipMemcpyHostToDevice));
        if (words >= (1ull << 32)) throw FameError{FAME_ERR_FORMAT, "genome of more than 2^36 letters"};
        c->h_pkoff = pkoff;
        c->h_clen.assign(clen.begin(), clen.end());
        for (uint64_t i = 0; i < c->nchrom; ++i)
            launch_pack_seq(c->d_seq.as<uint8_t>() + coff[i], clen[i], (clen[i] + 15) / 16,
                            c->d_pk.as<uint64_t>() + pkoff[i], st);
        HIPC(hipStreamSynchronize(st));
    }
    // tabIndex: u64[htab + 1], narrowed to u32 when every prefix fits
    const uint64_t tablen = f.get<uint64_t>();
    if (tablen != c->htab + 1) throw FameError{FAME_ERR_FORMAT, "tabIndex length != HTABSIZE + 1"};
    DevBuf tab64;
    HIPC(tab64.reserve(8 * tablen));
    {
        std::vector<uint8_t> buf;
        uint64_t left = 8 * tablen, o = 0;
        while (left) {
            size_t n = (size_t)std::min<uint64_t>(left, kStage);
            buf.resize(n);
            f.read(buf.data(), n);
            HIPC(hipMemcpy(tab64.as<uint8_t>() + o, buf.data(), n, hipMemcpyHostToDevice));
            o += n;
            left -= n;
        }
    }
    c->nkmer = f.get<uint64_t>();
    uint64_t last = 0;
    HIPC(hipMemcpy(&last, tab64.as<uint64_t>() + c->htab, 8, hipMemcpyDeviceToHost));
    if (last > c->nkmer) throw FameError{FAME_ERR_FORMAT, "tabIndex points past the k-mer table"};
    // FAME_TAB64=1 (tests only) keeps the 64-bit tabIndex, and with it the
    // kernels of indexes of 2^32 k-mers or more, on a smaller index
    const char* tab64_env = getenv("FAME_TAB64");
    const bool keep64 = tab64_env && tab64_env[0] == '1';
    if (c->nkmer < (1ull << 32) && !keep64) {
        HIPC(c->d_tab.reserve(4 * tablen));
        launch_narrow_tab(tab64.as<uint64_t>(), tablen, c->d_tab.as<uint32_t>(), st);
        HIPC(hipStreamSynchronize(st));
        tab64.release();
        c->ix.tab_is64 = 0;
    } else {
        std::swap(c->d_tab.p, tab64.p);
        std::swap(c->d_tab.bytes, tab64.bytes);
        c->ix.tab_is64 = 1;
    }
    // strands companion file, then the k-mer table packed with it
    DevBuf sbits;
    {
        File g(path + "_strands");
        if (!g.f) throw FameError{FAME_ERR_IO, "missing strand file " + path + "_strands"};
        uint64_t n = g.get<uint64_t>();
        if (n != c->nkmer) throw FameError{FAME_ERR_FORMAT, "strand file size differs from k-mer table"};
        std::vector<uint8_t> bits((n + 7) / 8);
        g.read(bits.data(), bits.size());
        HIPC(sbits.reserve(bits.size() + 1));
        HIPC(hipMemcpy(sbits.p, bits.data(), bits.size(), hipMemcpyHostToDevice));
    }
    HIPC(c->d_kent.reserve(4 * c->nkmer + 4));
    HIPC(c->d_kpe.reserve(8 * c->nkmer + 8));
    {
        DevBuf stage;
        HIPC(stage.reserve(kStage));
        std::vector<uint8_t> buf;
        const uint64_t per = kStage / 8;
        for (uint64_t first = 0; first < c->nkmer; first += per) {
            uint64_t n = std::min<uint64_t>(per, c->nkmer - first);
            buf.resize(8 * n);
            f.read(buf.data(), 8 * n);
            HIPC(hipMemcpy(stage.p, buf.data(), 8 * n, hipMemcpyHostToDevice));
            launch_pack_kmers(stage.as<uint32_t>(), sbits.as<uint8_t>(), first, n, c->d_kent.as<uint32_t>(), st);
            launch_pack_tmask(stage.as<uint32_t>(), c->d_kent.as<uint32_t>() + first, n,
                              c->d_kpe.as<uint64_t>() + first, st);
            HIPC(hipStreamSynchronize(st));
        }
    }
    // slot table of the single-end seed stage (u32 tabIndex only; the
    // environment variable FAME_NO_SLOT_TABLE=1 keeps the tabIndex path)
    const char* no_slot = getenv("FAME_NO_SLOT_TABLE");
    if (!c->ix.tab_is64 && !(no_slot && no_slot[0] == '1')) {
        HIPC(c->d_slot.reserve(8 * c->htab));
        DevBuf bad;
        HIPC(bad.reserve(4));
        HIPC(hipMemsetAsync(bad.p, 0, 4, st));
        launch_build_slots(c->d_tab.as<uint32_t>(), c->d_kent.as<uint32_t>(), c->htab, c->d_slot.as<uint64_t>(),
                           bad.as<uint32_t>(), st);
        uint32_t b = 0;
        HIPC(hipMemcpyAsync(&b, bad.p, 4, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        if (b) c->d_slot.release();
    }
    // the slot table in genome order (DevIndex::gkey): positions and keys
    // must fit 32 bits with ~0u to spare, and the three arrays must leave
    // room for the batches; FAME_NO_GORDER=1 keeps every lookup random
    const char* no_go = getenv("FAME_NO_GORDER");
    const uint64_t gn = 2 * c->seq_total;
    if (c->d_slot.p && !(no_go && no_go[0] == '1') && gn >= 64 && gn < 0xFFFFFFFFull &&
        c->htab <= 0xFFFFFFFFull) {
        const uint64_t need = 12 * gn + 4 * c->htab;
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr > need + (8ull << 30)) {
            HIPC(c->d_gkey.reserve(4 * gn));
            HIPC(c->d_gslot.reserve(8 * gn));
            HIPC(c->d_kpos.reserve(4 * c->htab));
            launch_gorder(c->d_seq.as<uint8_t>(), c->seq_total, c->htab, c->d_slot.as<uint64_t>(),
                          c->d_gkey.as<uint32_t>(), c->d_gslot.as<uint64_t>(), c->d_kpos.as<uint32_t>(), st);
            HIPC(hipStreamSynchronize(st));
        }
    }
    f.skip(8 * f.get<uint64_t>()); // metaCpGs (empty in current builds)
    f.skip(8 * f.get<uint64_t>()); // metaStartCpGs
    // windows {u32 startPos; u8 chrom; 3 pad; u32 startInd; u32 endInd}
    c->nwin = f.get<uint64_t>();
    if (c->nwin >= (1ull << 30)) throw FameError{FAME_ERR_FORMAT, "more than 2^30 windows"};
    {
        std::vector<uint32_t> raw(4 * c->nwin);
        f.read(raw.data(), 16 * c->nwin);
        for (uint64_t i = 0; i < c->nwin; ++i) {
            raw[4 * i + 1] &= 0xFFu;
            if (raw[4 * i + 1] >= c->nchrom) throw FameError{FAME_ERR_FORMAT, "window on unknown chromosome"};
        }
        HIPC(c->d_win.reserve(16 * c->nwin + 16));
        HIPC(hipMemcpy(c->d_win.p, raw.data(), 16 * c->nwin, hipMemcpyHostToDevice));
        std::vector<uint32_t> wt(2 * c->nwin);
        for (uint64_t i = 0; i < c->nwin; ++i) {
            const uint32_t ch = raw[4 * i + 1];
            wt[2 * i] = (uint32_t)c->h_pkoff[ch];
            wt[2 * i + 1] = (uint32_t)std::min<uint64_t>(c->h_clen[ch], 0xFFFFFFFFull);
        }
        HIPC(c->d_wtext.reserve(8 * c->nwin + 8));
        HIPC(hipMemcpy(c->d_wtext.p, wt.data(), 8 * c->nwin, hipMemcpyHostToDevice));
    }
    c->nfilt = f.get<uint64_t>();
    {
        // filteredKmers: the paired path looks up a u32 k-mer code, so only
        // the values below 2^32 can ever be found (ReadQueue.cpp:2527, 2643)
        std::vector<uint64_t> fk(c->nfilt);
        f.read(fk.data(), 8 * c->nfilt);
        std::vector<uint32_t> f32;
        for (uint64_t v : fk)
            if (v < (1ull << 32)) f32.push_back((uint32_t)v);
        std::sort(f32.begin(), f32.end());
        f32.erase(std::unique(f32.begin(), f32.end()), f32.end());
        c->ix.nfilt32 = (uint32_t)f32.size();
        HIPC(c->d_filt32.reserve(4 * f32.size() + 4));
        if (!f32.empty()) HIPC(hipMemcpy(c->d_filt32.p, f32.data(), 4 * f32.size(), hipMemcpyHostToDevice));
    }
    // chrMap
    c->chr_name.assign(256, std::string());
    uint64_t nmap = 0;
    if (fread(&nmap, 8, 1, f.f) == 1) {
        for (uint64_t i = 0; i < nmap; ++i) {
            uint8_t id = f.get<uint8_t>();
            uint64_t sl = f.get<uint64_t>();
            std::string s(sl, '\0');
            f.read(&s[0], sl);
            if (c->chr_name[id].empty()) c->chr_name[id] = s; // unordered_map::insert keeps the first
        }
    }
    // methylation accumulators and counters
    HIPC(c->d_meth.reserve(16 * c->ncpg + 16));
    HIPC(hipMemset(c->d_meth.p, 0, 16 * c->ncpg + 16));
    HIPC(c->d_counters.reserve(8 * 8));
    HIPC(hipMemset(c->d_counters.p, 0, 64));
    HIPC(c->d_flags.reserve(16));
    // device index view
    DevIndex& ix = c->ix;
    ix.tab = c->d_tab.p;
    ix.htab = c->htab;
    ix.hpow2 = (c->htab & (c->htab - 1)) == 0;
    ix.hmask = c->htab - 1;
    ix.kent = c->d_kent.as<uint32_t>();
    ix.slot = c->d_slot.p ? c->d_slot.as<uint64_t>() : nullptr;
    ix.gkey = c->d_gkey.p ? c->d_gkey.as<uint32_t>() : nullptr;
    ix.gslot = c->d_gkey.p ? c->d_gslot.as<uint64_t>() : nullptr;
    ix.kpos = c->d_gkey.p ? c->d_kpos.as<uint32_t>() : nullptr;
    ix.gn = c->d_gkey.p ? 2 * c->seq_total : 0;
    ix.kpe = c->d_kpe.as<uint64_t>();
    ix.filt32 = c->d_filt32.as<uint32_t>();
    ix.win = c->d_win.as<uint4>();
    ix.seq = c->d_seq.as<uint8_t>();
    ix.coff = c->d_coff.as<uint64_t>();
    ix.clen = c->d_clen.as<uint64_t>();
    ix.pk = c->d_pk.as<uint64_t>();
    ix.pkoff = c->d_pkoff.as<uint64_t>();
    ix.wtext = c->d_wtext.as<uint2>();
    ix.cpg_pos = c->d_cpg_pos.as<uint32_t>();
    ix.cpg_chrom = c->d_cpg_chrom.as<uint8_t>();
    ix.readlen = c->readlen;
    ix.winlen = c->winlen;
    ix.seedbits = c->seedbits;
}

// ---------------------------------------------------------------------------
// one batch
// ---------------------------------------------------------------------------

template <typename T>
T* grow(DevBuf& b, size_t count)
{
    HIPC(b.reserve(sizeof(T) * count));
    return b.as<T>();
}

struct Timer {
    hipEvent_t e[17];
    hipStream_t st;
    explicit Timer(hipStream_t s) : st(s)
    {
        for (auto& x : e) (void)hipEventCreate(&x);
    }
    ~Timer()
    {
        for (auto& x : e) (void)hipEventDestroy(x);
    }
    void mark(int i) { (void)hipEventRecord(e[i], st); }
    float ms(int a, int b)
    {
        float t = 0;
        (void)hipEventElapsedTime(&t, e[a], e[b]);
        return t;
    }
};

// FAME_VERIFY_QUEUE_LIMIT (tests only): verify's per-block run queue takes
// at most this many runs per round of 256 jobs, so that small inputs reach
// the in-place re-verification of jobs whose runs do not fit; 0 or unset:
// what the queue's LDS holds
uint32_t verify_queue_limit()
{
    const char* v = getenv("FAME_VERIFY_QUEUE_LIMIT");
    if (!v || !*v) return 0;
    return (uint32_t)strtoul(v, nullptr, 10);
}

// the heavy scratch (which the context already holds) can grow to `bytes`
// with 1 GiB of device memory left over
// FAME_NO_MID_TIER=1 (tests): every deferred pair on global scratch
static bool no_mid_tier()
{
    const char* e = getenv("FAME_NO_MID_TIER");
    return e && e[0] == '1';
}

bool scratch_fits(fame_ctx* c, size_t bytes)
{
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return false;
    return bytes + (1ull << 30) <= fr + c->b_heavy_scratch.bytes;
}

int run_batch(fame_ctx* c, const uint8_t* d_seqs, const uint64_t* d_offs, const uint32_t* d_lens, uint32_t n,
              uint64_t total_bytes, int get_stranded, uint64_t counters[3], hipStream_t st)
{
    auto t0 = std::chrono::steady_clock::now();
    Timer tm(st);
    BatchDev bd{};
    bd.seqs = d_seqs;
    bd.offs = d_offs;
    bd.lens = d_lens;
    bd.rev = grow<uint8_t>(c->b_rev, total_bytes + 64);
    bd.pre = grow<uint8_t>(c->b_pre, n + 1);
    const bool o0 = c->both || get_stranded || c->match_r1_fwd;
    const bool o1 = c->both || get_stranded || !c->match_r1_fwd;
    bd.ori_mask = (o0 ? 1u : 0u) | (o1 ? 2u : 0u);
    const uint32_t n_ro = 2 * n;
    bd.codes = grow<uint8_t>(c->b_codes, (size_t)n_ro * 256 + 256);
    bd.cand_off = grow<uint32_t>(c->b_cand_off, n_ro + 1);
    bd.cand_nf = grow<uint32_t>(c->b_cand_nf, n_ro + 1);
    bd.cand_nr = grow<uint32_t>(c->b_cand_nr, n_ro + 1);
    bd.pool_n = grow<uint32_t>(c->b_pool_n, FAME_NPOOL);
    bd.match_cap = c->match_cap;
    bd.verify_runs_cap = verify_queue_limit();
    // room for the slots persistent seed waves leave unused at chunk ends
    const uint32_t want_cap = std::max<uint32_t>(1u << 16, 4 * n) + 64u * FAME_SEED_WAVES;
    for (int p = 0; p < FAME_NPOOL; ++p)
        if (c->pool_cap[p] < want_cap) c->pool_cap[p] = want_cap;
    bd.ovf_n = grow<uint32_t>(c->b_ovf_n, 1);
    bd.ovf_list_cap = std::max<uint32_t>(1024, n / 4);
    bd.ovf_pool = grow<uint32_t>(c->b_ovf_pool, bd.ovf_list_cap);
    bd.ovf_job = grow<uint32_t>(c->b_ovf_job, bd.ovf_list_cap);
    bd.heavy_n = grow<uint32_t>(c->b_heavy_n, 1);
    bd.heavy_cap = std::max<uint32_t>(1024, n_ro);
    bd.heavy_ro = grow<uint32_t>(c->b_heavy_ro, bd.heavy_cap);
    bd.heavy_e = grow<uint32_t>(c->b_heavy_e, bd.heavy_cap);
    bd.heavy_bc = grow<uint32_t>(c->b_heavy_bc, bd.heavy_cap);
    bd.status = grow<uint8_t>(c->b_status, n + 1);
    bd.use = grow<uint8_t>(c->b_use, n + 1);
    bd.mat = grow<uint64_t>(c->b_mat, n + 1);
    bd.valid = grow<uint8_t>(c->b_valid, n + 1);
    bd.counters = c->d_counters.as<uint64_t>();
    bd.meth = meth_buf(c);
    bd.flags = c->d_flags.as<uint32_t>();
    bd.max_len = bd.flags + 1;
    bd.stats = grow<unsigned long long>(c->d_stats, FAME_NSTATS);
    HIPC(hipMemsetAsync(bd.stats, 0, 8 * FAME_NSTATS, st));
    bd.direct_list = grow<uint32_t>(c->b_direct, n + 1);
    bd.dp_list = grow<uint32_t>(c->b_dplist, n + 1);
    bd.direct_n = grow<uint32_t>(c->b_methn, 3);
    bd.dp_n = bd.direct_n + 1;
    bd.dp_long_n = bd.direct_n + 2;
    bd.dp_long_list = grow<uint32_t>(c->b_dp_long, n + 1);
    HIPC(hipMemsetAsync(bd.direct_n, 0, 12, st));
    bd.long_n = grow<uint32_t>(c->b_long, 4);
    bd.long_list = grow<uint32_t>(c->b_long_list, n + 1);
    HIPC(hipMemsetAsync(bd.long_n, 0, 8, st)); // long reads and the longest of them

    HIPC(hipMemsetAsync(bd.flags, 0, 8, st)); // flags and the longest read
    tm.mark(0);
    launch_prep(bd, n, c->readlen, st);
    tm.mark(1);
    // seed: retried with larger pools when a pool overflowed
    uint32_t pool_n[FAME_NPOOL];
    uint32_t heavy_n = 0, flags = 0;
    uint32_t long_info[2] = {0, 0}; // long reads, the longest of them
    for (int attempt = 0;; ++attempt) {
        for (int p = 0; p < FAME_NPOOL; ++p) {
            bd.pool_cap[p] = c->pool_cap[p];
            bd.job_meta[p] = grow<uint32_t>(c->b_job_meta[p], c->pool_cap[p]);
            bd.job_ro[p] = grow<uint32_t>(c->b_job_ro[p], c->pool_cap[p]);
        }
        HIPC(hipMemsetAsync(bd.pool_n, 0, 4 * FAME_NPOOL, st));
        HIPC(hipMemsetAsync(bd.heavy_n, 0, 4, st));
        tm.mark(2);
        launch_seed_light(c->ix, bd, n_ro, st);
        tm.mark(3);
        HIPC(hipMemcpyAsync(pool_n, bd.pool_n, 4 * FAME_NPOOL, hipMemcpyDeviceToHost, st));
        HIPC(hipMemcpyAsync(&heavy_n, bd.heavy_n, 4, hipMemcpyDeviceToHost, st));
        HIPC(hipMemcpyAsync(&flags, bd.flags, 4, hipMemcpyDeviceToHost, st));
        HIPC(hipMemcpyAsync(long_info, bd.long_n, 8, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        if (long_info[0]) {
            // long reads: their rows and bucket sums, then the heavy list
            // holds them too
            const uint32_t lmax = std::min<uint32_t>(long_info[1], FAME_LONG_MAX_READ);
            bd.long_stride = (lmax + 63) & ~63u;
            bd.long_codes = grow<uint8_t>(c->b_long_codes, 2ull * long_info[0] * bd.long_stride + 64);
            launch_long_measure(c->ix, bd, 2 * long_info[0], st);
            HIPC(hipMemcpyAsync(&heavy_n, bd.heavy_n, 4, hipMemcpyDeviceToHost, st));
            HIPC(hipMemcpyAsync(&flags, bd.flags, 4, hipMemcpyDeviceToHost, st));
            HIPC(hipStreamSynchronize(st));
        }
        if (flags & FLAG_TOO_LONG) {
            set_err(c, "a read longer than FAME_MAX_READ letters could match a window (WINLEN over "
                       "FAME_MAX_READ - 12)");
            return FAME_ERR_ARG;
        }
        if (flags & FLAG_HEAVY_FULL) {
            set_err(c, "heavy list overflow");
            return FAME_ERR_CAPACITY;
        }
        // heavy (read, orientation) pairs: global-memory scratch, in slices
        if (heavy_n) {
            std::vector<uint32_t> he(heavy_n), hb(heavy_n), hr(heavy_n);
            HIPC(hipMemcpy(he.data(), bd.heavy_e, 4 * heavy_n, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(hb.data(), bd.heavy_bc, 4 * heavy_n, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(hr.data(), bd.heavy_ro, 4 * heavy_n, hipMemcpyDeviceToHost));
            uint64_t emax_all = 1;
            for (uint32_t i = 0; i < heavy_n; ++i) emax_all = std::max<uint64_t>(emax_all, he[i]);
            if (emax_all > 0x3FFFFFFFull) {
                set_err(c, "a read has more than 2^30 index hits, or a bucket of more than 2^24 entries");
                return FAME_ERR_CAPACITY;
            }
            // size classes (caps 1024 * 4^k): each run's scratch slices are
            // sized for its own largest read, so one outlier does not shrink
            // the number of slices of the others
            std::vector<uint32_t> ord(heavy_n);
            for (uint32_t i = 0; i < heavy_n; ++i) ord[i] = i;
            std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return he[a] < he[b]; });
            // the LDS tier first: pairs of up to seed_mid_max_e() index hits
            // (reads of at most FAME_DEV_MAX_READ letters); those whose
            // windows do not fit it come back in its overflow list and join
            // the others in the global-scratch classes below
            const uint32_t mid_e = c->ix.tab_is64 || no_mid_tier() ? 0u : seed_mid_max_e();
            std::stable_partition(ord.begin(), ord.end(),
                                  [&](uint32_t i) { return he[i] <= mid_e && !(hr[i] >> 31); });
            uint32_t nmid = 0;
            while (nmid < heavy_n && he[ord[nmid]] <= mid_e && !(hr[ord[nmid]] >> 31)) ++nmid;
            {
                std::vector<uint32_t> r2(heavy_n);
                for (uint32_t i = 0; i < heavy_n; ++i) r2[i] = hr[ord[i]];
                HIPC(hipMemcpyAsync(bd.heavy_ro, r2.data(), 4 * heavy_n, hipMemcpyHostToDevice, st));
                HIPC(hipStreamSynchronize(st));
            }
            tm.mark(12);
            if (nmid) {
                uint32_t* ovl = grow<uint32_t>(c->b_mid_ovl, nmid + 1);
                HIPC(hipMemsetAsync(ovl, 0, 4, st));
                launch_seed_mid(c->ix, bd, 0, nmid, ovl, std::min<uint32_t>(nmid, seed_mid_blocks()), st);
                uint32_t novl = 0;
                HIPC(hipMemcpyAsync(&novl, ovl, 4, hipMemcpyDeviceToHost, st));
                HIPC(hipStreamSynchronize(st));
                std::vector<uint32_t> back(novl);
                if (novl) HIPC(hipMemcpy(back.data(), ovl + 1, 4ull * novl, hipMemcpyDeviceToHost));
                // the global classes: the overflowed pairs and those past the
                // tier, sorted by hits again
                std::vector<uint32_t> rest;
                rest.reserve(novl + heavy_n - nmid);
                for (uint32_t x : back) rest.push_back(ord[x]);
                for (uint32_t i = nmid; i < heavy_n; ++i) rest.push_back(ord[i]);
                std::sort(rest.begin(), rest.end(), [&](uint32_t a, uint32_t b) { return he[a] < he[b]; });
                ord.swap(rest);
                std::vector<uint32_t> r3(ord.size());
                for (size_t i = 0; i < ord.size(); ++i) r3[i] = hr[ord[i]];
                if (!r3.empty())
                    HIPC(hipMemcpyAsync(bd.heavy_ro, r3.data(), 4 * r3.size(), hipMemcpyHostToDevice, st));
                HIPC(hipStreamSynchronize(st));
            }
            const uint32_t nglob = (uint32_t)ord.size();
            auto rup = [](uint64_t v) {
                uint64_t p = 1;
                while (p < v) p <<= 1;
                return p;
            };
            // the classes and their scratch first (the buffer is sized once,
            // before any launch uses it)
            struct Cls {
                uint32_t c0, c1, per;
                HeavyScratch hs;
            };
            std::vector<Cls> classes;
            size_t need = 0;
            const uint32_t resident = seed_heavy_blocks(c->ix.tab_is64);
            for (uint32_t c0 = 0; c0 < nglob;) {
                uint64_t cap = 1024;
                while (cap < he[ord[c0]]) cap *= 4;
                uint32_t c1 = c0;
                uint64_t emax = 1, bcmax = 1;
                while (c1 < nglob && he[ord[c1]] <= cap) {
                    emax = std::max<uint64_t>(emax, he[ord[c1]]);
                    bcmax = std::max<uint64_t>(bcmax, hb[ord[c1]]);
                    ++c1;
                }
                HeavyScratch hs{};
                hs.cap_e = (uint32_t)emax;
                if (long_info[0])
                    hs.kmax = std::min<uint32_t>(long_info[1], FAME_LONG_MAX_READ); // >= its k-mers
                hs.hsize = (uint32_t)rup(2 * emax);
                uint64_t tbb = std::max(rup(bcmax), rup(2 * emax));
                hs.tb_cap = (uint32_t)(2 * tbb + 64);
                const size_t stride = heavy_scratch_stride(hs);
                hs.stride = stride;
                // one launch per class: as many blocks (slices) as are resident
                // at once, within the scratch budget, striding over the class
                const size_t budget = 2ull << 30;
                const uint32_t cnt = c1 - c0;
                const uint32_t per = (uint32_t)std::max<size_t>(
                    1, std::min<size_t>(std::min<size_t>(cnt, budget / stride), resident));
                need = std::max(need, stride * per);
                classes.push_back({c0, c1, per, hs});
                c0 = c1;
            }
            if (!scratch_fits(c, need)) {
                set_err(c, "a read's index hits need " + std::to_string(need >> 20) +
                               " MiB of scratch, more than the device has free");
                return FAME_ERR_CAPACITY;
            }
            if (need) HIPC(c->b_heavy_scratch.reserve(need));
            for (Cls& k : classes) {
                k.hs.base = c->b_heavy_scratch.as<uint8_t>();
                k.hs.first = k.c0;
                k.hs.count = k.c1 - k.c0;
                launch_seed_heavy(c->ix, bd, k.hs, k.per, st);
            }
            tm.mark(13);
            HIPC(hipMemcpyAsync(pool_n, bd.pool_n, 4 * FAME_NPOOL, hipMemcpyDeviceToHost, st));
            HIPC(hipMemcpyAsync(&flags, bd.flags, 4, hipMemcpyDeviceToHost, st));
            HIPC(hipStreamSynchronize(st));
            if (flags & FLAG_CAPACITY) {
                set_err(c, "a read's candidate map exceeded the device emulation limits");
                return FAME_ERR_CAPACITY;
            }
        }
        if (!(flags & FLAG_POOL_FULL)) break;
        if (attempt > 3) {
            set_err(c, "candidate pools keep overflowing");
            return FAME_ERR_CAPACITY;
        }
        for (int p = 0; p < FAME_NPOOL; ++p)
            c->pool_cap[p] = std::max<uint32_t>(c->pool_cap[p], pool_n[p] + (pool_n[p] >> 2) + 1024);
        HIPC(hipMemsetAsync(bd.flags, 0, 4, st));
    }
    c->last_heavy = heavy_n;
    // verify
    for (int p = 0; p < FAME_NPOOL; ++p) {
        const uint32_t cap = std::max<uint32_t>(pool_n[p], 1);
        bd.job_nm[p] = grow<uint32_t>(c->b_job_nm[p], cap);
        bd.job_match[p] = grow<uint32_t>(c->b_job_match[p], (size_t)cap * c->match_cap);
        bd.job_ovf_off[p] = grow<uint32_t>(c->b_job_ovf_off[p], cap);
    }
    bd.ovf_buf = grow<uint32_t>(c->b_ovf_buf, 1);
    HIPC(hipMemsetAsync(bd.ovf_n, 0, 4, st));
    tm.mark(4);
    for (int p = 0; p < FAME_NPOOL; ++p) launch_verify(c->ix, bd, p, pool_n[p], nullptr, st);
    tm.mark(5);
    uint32_t ovf_n = 0;
    HIPC(hipMemcpyAsync(&ovf_n, bd.ovf_n, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(&flags, bd.flags, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (flags & FLAG_OVF_FULL) {
        set_err(c, "match overflow list full");
        return FAME_ERR_CAPACITY;
    }
    c->last_ovf = ovf_n;
    if (ovf_n) {
        uint32_t* d_total = grow<uint32_t>(c->b_ovf_total, 2);
        launch_ovf_offsets(bd, ovf_n, d_total, st);
        uint32_t total = 0;
        HIPC(hipMemcpyAsync(&total, d_total, 4, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        bd.ovf_buf = grow<uint32_t>(c->b_ovf_buf, total + 1);
        tm.mark(14);
        for (int p = 0; p < FAME_NPOOL; ++p) launch_verify(c->ix, bd, p, ovf_n, bd.ovf_job, st);
        tm.mark(15);
    }
    tm.mark(6);
    launch_select(c->ix, bd, n, get_stranded, st);
    tm.mark(7);
    {
        const size_t cells = (size_t)FAME_DP_SROWS * (2 * (FAME_DPBAND + 1) + 1);
        uint16_t* dp = grow<uint16_t>(c->b_dp, cells * c->meth_threads);
        uint64_t* dec = grow<uint64_t>(c->b_dec, (size_t)(FAME_DEV_MAX_READ + 1) * c->meth_threads);
        uint64_t* tr = grow<uint64_t>(c->b_trace, (size_t)FAME_DP_TRACE_WORDS * c->meth_threads);
        tm.mark(8);
        launch_meth(c->ix, bd, n, dp, dec, tr, c->meth_threads, st, tm.e[16]);
        if (long_info[0]) {
            // the long reads' DP on scratch sized for the longest of them
            const uint32_t lmax = std::min<uint32_t>(long_info[1], FAME_LONG_MAX_READ);
            const uint32_t lanes = std::min<uint32_t>(8192, std::max<uint32_t>(256, (long_info[0] + 255) & ~255u));
            const uint32_t cap = 2 * lmax + 16;
            uint16_t* ldp = grow<uint16_t>(c->b_ldp, cells * lanes);
            uint64_t* ldec = grow<uint64_t>(c->b_ldec, (size_t)(lmax + 1) * lanes);
            uint64_t* ltr = grow<uint64_t>(c->b_ltrace, (size_t)(cap / 32 + 1) * lanes);
            launch_meth_dp_long(c->ix, bd, ldp, ldec, ltr, lanes, cap, st);
        }
        tm.mark(9);
    }
    uint64_t cnt[5];
    HIPC(hipMemcpyAsync(cnt, bd.counters, 40, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(c->stats, bd.stats, 8 * FAME_NSTATS, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(&flags, bd.flags, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (flags & FLAG_DP_BROKEN) {
        set_err(c, "a banded alignment's traceback left the band (internal error)");
        return FAME_ERR_CAPACITY;
    }
    // per-call counters: the device accumulates; report the increment
    counters[0] += cnt[0];
    counters[1] += cnt[1];
    counters[2] += cnt[2];
    HIPC(hipMemsetAsync(bd.counters, 0, 24, st));
    HIPC(hipStreamSynchronize(st));
    // kernel-only intervals (events bracket the launches, not the host waits)
    c->timings[0] = tm.ms(0, 1);
    c->timings[1] = tm.ms(2, 3) + (heavy_n ? tm.ms(12, 13) : 0.0);
    c->timings[2] = tm.ms(4, 5) + (c->last_ovf ? tm.ms(14, 15) : 0.0);
    c->timings[3] = tm.ms(6, 7);
    c->timings[4] = tm.ms(8, 9);
    c->timings[6] = n ? tm.ms(16, 9) : 0.0;
    c->timings[5] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return FAME_OK;
}

// ---------------------------------------------------------------------------
// batched form: the whole batch as one graph launch
// ---------------------------------------------------------------------------

} // namespace

// candidate persistent grids of the verify kernels, tried on the first
// batches of a plan (online: every trial batch is real work), then fixed
constexpr uint32_t kVerifyBlocks[] = {1024, 2048, 4096, 8192};
constexpr int kNCand = 4;

struct GraphKey {
    const void* s;
    const void* o;
    const void* l;
    int stranded;
    uint32_t ori;
    int cand;
    bool operator==(const GraphKey& k) const
    {
        return s == k.s && o == k.o && l == k.l && stranded == k.stranded && ori == k.ori && cand == k.cand;
    }
};

struct GraphPlan {
    uint32_t cap = 0; // reads per batch
    int lane = 0;     // 0: single calls; 1, 2: the two lanes of a pipelined call
    uint64_t rev_bytes = 0;
    DevBuf rev, pre, codes, cand_off, cand_nf, cand_nr, pool_n;
    DevBuf job_meta[FAME_NPOOL], job_ro[FAME_NPOOL], job_nm[FAME_NPOOL], job_match[FAME_NPOOL],
        job_ovf_off[FAME_NPOOL];
    DevBuf ovf_n, ovf_pool, ovf_job, ovf_buf, ovf_total, heavy_n, heavy_ro, heavy_e, heavy_bc, heavy_scratch,
        heavy_scratch2;
    DevBuf status, use, mat, valid, stats, direct, dplist, methn, flags, ndev, dp, dec, trace, counters;
    uint32_t pool_cap = 0, ovf_list_cap = 0, ovf_buf_cap = 0, heavy_cap = 0, heavy_blocks = 1024, heavy_blocks2 = 16,
             mid_blocks = 1024;
    HeavyScratch hs{}, hs2{};
    // captured graphs by input buffers and variant, at most kMaxGraphs: the
    // least recently used one is destroyed for a new one (callers that pass
    // new buffers every call would otherwise add one per call)
    static constexpr size_t kMaxGraphs = 64;
    std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs;
    std::vector<uint64_t> graph_used;
    uint64_t tick = 0;
    hipGraphExec_t find_graph(const GraphKey& k)
    {
        for (size_t i = 0; i < graphs.size(); ++i)
            if (graphs[i].first == k) {
                graph_used[i] = ++tick;
                return graphs[i].second;
            }
        return nullptr;
    }
    // the plan's launches have completed (its stream was synchronized)
    void add_graph(const GraphKey& k, hipGraphExec_t e)
    {
        if (graphs.size() >= kMaxGraphs) {
            size_t lru = 0;
            for (size_t i = 1; i < graphs.size(); ++i)
                if (graph_used[i] < graph_used[lru]) lru = i;
            (void)hipGraphExecDestroy(graphs[lru].second);
            graphs[lru] = {k, e};
            graph_used[lru] = ++tick;
            return;
        }
        graphs.push_back({k, e});
        graph_used.push_back(++tick);
    }
    // pinned: [0] n, [1] flags, [2..7] counters (3 x u64), [8..11] pool
    // counts; pipelined calls: [16 + k] n, [32 + k] flags and [64 + 4k ..]
    // pool counts of the lane's k-th sub-batch
    uint32_t* h_io = nullptr;
    // autotuner: ms per read of each candidate, the choice (-1 while trying)
    double cand_ms[kNCand] = {0, 0, 0, 0};
    int cand_reads[kNCand] = {0, 0, 0, 0};
    int chosen = -1;
    int next_trial = 0;
    void drop_graphs()
    {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
        graphs.clear();
        graph_used.clear();
    }
    ~GraphPlan()
    {
        drop_graphs();
        if (h_io) (void)hipHostFree(h_io);
    }
};

namespace {

// verify pools of a plan (the same capacity for each read-length class)
void plan_pools(fame_ctx* c, GraphPlan* pl, uint32_t cap)
{
    pl->pool_cap = cap;
    for (int p = 0; p < FAME_NPOOL; ++p) {
        HIPC(pl->job_meta[p].reserve(4ull * cap));
        HIPC(pl->job_ro[p].reserve(4ull * cap));
        HIPC(pl->job_nm[p].reserve(4ull * cap));
        HIPC(pl->job_match[p].reserve(4ull * cap * c->match_cap));
        HIPC(pl->job_ovf_off[p].reserve(4ull * cap));
    }
}

// after a batch overflowed a pool: pools for the demand it showed (the
// device counters keep counting past the capacity), graphs re-captured
void plan_grow_pools(fame_ctx* c, GraphPlan* pl, const uint32_t* demand)
{
    uint64_t need = 0;
    for (int p = 0; p < FAME_NPOOL; ++p) need = std::max<uint64_t>(need, demand[p]);
    need = 2 * need + 65536;
    if (need <= pl->pool_cap || need > 0xFFFFFFF0ull) return;
    pl->drop_graphs();
    plan_pools(c, pl, (uint32_t)need);
}

GraphPlan* plan_for(fame_ctx* c, uint32_t n, uint64_t total_bytes, int lane = 0)
{
    uint32_t cap = 1u << 16;
    while (cap < n) cap <<= 1;
    GraphPlan* pl = nullptr;
    for (auto& p : c->plans)
        if (p->cap == cap && p->lane == lane) pl = p.get();
    if (!pl) {
        c->plans.emplace_back(new GraphPlan());
        pl = c->plans.back().get();
        pl->cap = cap;
        pl->lane = lane;
        HIPC(pl->counters.reserve(64));
        const uint64_t n_ro = 2ull * cap;
        HIPC(pl->pre.reserve(cap + 1));
        HIPC(pl->codes.reserve(n_ro * 256 + 256));
        HIPC(pl->cand_off.reserve(4 * n_ro + 4));
        HIPC(pl->cand_nf.reserve(4 * n_ro + 4));
        HIPC(pl->cand_nr.reserve(4 * n_ro + 4));
        HIPC(pl->pool_n.reserve(4 * FAME_NPOOL));
        plan_pools(c, pl, std::max<uint32_t>(1u << 16, 4 * cap) + 64u * FAME_SEED_WAVES);
        pl->ovf_list_cap = std::max<uint32_t>(1024, cap / 4);
        pl->ovf_buf_cap = 1u << 24;
        HIPC(pl->ovf_n.reserve(8));
        HIPC(pl->ovf_pool.reserve(4ull * pl->ovf_list_cap));
        HIPC(pl->ovf_job.reserve(4ull * pl->ovf_list_cap));
        HIPC(pl->ovf_buf.reserve(4ull * pl->ovf_buf_cap));
        HIPC(pl->ovf_total.reserve(8));
        pl->heavy_cap = (uint32_t)n_ro;
        HIPC(pl->heavy_n.reserve(8));
        HIPC(pl->heavy_ro.reserve(4 * n_ro));
        HIPC(pl->heavy_e.reserve(4 * n_ro));
        HIPC(pl->heavy_bc.reserve(4 * n_ro));
        // heavy (read, orientation) pairs in graph, two tiers: up to 16384
        // index hits on 1024 scratch slices (or more, below), up to 2^20 on
        // 16; larger ones make the batch fall back to the plain path
        auto dims = [](HeavyScratch& hs, uint64_t emax) {
            hs.cap_e = (uint32_t)emax;
            hs.hsize = (uint32_t)(2 * emax);
            hs.tb_cap = (uint32_t)(2 * (2 * emax) + 64);
            hs.stride = heavy_scratch_stride(hs);
        };
        dims(pl->hs, 16384);
        // batches of a million reads or more: the first tier on as many
        // slices as are resident at once (1.25 MB each)
        if (cap >= (1u << 20)) pl->heavy_blocks = std::max<uint32_t>(1024, seed_heavy_blocks(c->ix.tab_is64));
        // the LDS tier takes the pairs of up to seed_mid_max_e() hits (and
        // hands those it cannot hold to the first global tier)
        const bool mid = !c->ix.tab_is64 && !no_mid_tier();
        pl->hs.e_lo = 0;
        pl->hs.e_hi = 16384;
        pl->hs.handed = mid ? 1u : 0u;
        if (mid) pl->mid_blocks = seed_mid_blocks();
        dims(pl->hs2, 1u << 20);
        pl->hs2.e_lo = 16385;
        pl->hs2.e_hi = 0xFFFFFFFFu;
        pl->hs2.handed = pl->hs.handed;
        HIPC(pl->heavy_scratch.reserve(pl->hs.stride * pl->heavy_blocks));
        HIPC(pl->heavy_scratch2.reserve(pl->hs2.stride * pl->heavy_blocks2));
        pl->hs.base = pl->heavy_scratch.as<uint8_t>();
        pl->hs2.base = pl->heavy_scratch2.as<uint8_t>();
        HIPC(pl->status.reserve(cap + 1));
        HIPC(pl->use.reserve(cap + 1));
        HIPC(pl->mat.reserve(8ull * cap + 8));
        HIPC(pl->valid.reserve(cap + 1));
        HIPC(pl->stats.reserve(8 * FAME_NSTATS));
        HIPC(pl->direct.reserve(4ull * cap + 4));
        HIPC(pl->dplist.reserve(4ull * cap + 4));
        HIPC(pl->methn.reserve(8));
        HIPC(pl->flags.reserve(16));
        HIPC(pl->ndev.reserve(8));
        const size_t cells = (size_t)FAME_DP_SROWS * (2 * (FAME_DPBAND + 1) + 1);
        HIPC(pl->dp.reserve(2 * cells * c->meth_threads));
        HIPC(pl->dec.reserve(8ull * (FAME_DEV_MAX_READ + 1) * c->meth_threads));
        HIPC(pl->trace.reserve(8ull * FAME_DP_TRACE_WORDS * c->meth_threads));
        HIPC(hipHostMalloc((void**)&pl->h_io, 1024, 0));
    }
    if (total_bytes + 64 > pl->rev_bytes) {
        // the graphs hold the old pointer
        pl->drop_graphs();
        HIPC(pl->rev.reserve(total_bytes + 64));
        pl->rev_bytes = total_bytes + 64;
    }
    return pl;
}

// FAME_POOL_CAP_LIMIT (tests only): verify pools of the fixed-capacity forms
// (batched graphs, the allocation-free entry point) hold at most this many
// slots, so that small inputs reach their overflow handling; 0 or unset: no
// limit
uint32_t pool_cap_limit(uint32_t cap)
{
    const char* v = getenv("FAME_POOL_CAP_LIMIT");
    if (!v || !*v) return cap;
    const unsigned long k = strtoul(v, nullptr, 10);
    return k && k < cap ? (uint32_t)k : cap;
}

BatchDev plan_batch(fame_ctx* c, GraphPlan* pl, const uint8_t* d_seqs, const uint64_t* d_offs,
                    const uint32_t* d_lens, uint32_t ori_mask)
{
    BatchDev bd{};
    bd.n_dev = pl->ndev.as<uint32_t>();
    bd.seqs = d_seqs;
    bd.offs = d_offs;
    bd.lens = d_lens;
    bd.rev = pl->rev.as<uint8_t>();
    bd.pre = pl->pre.as<uint8_t>();
    bd.codes = pl->codes.as<uint8_t>();
    bd.ori_mask = ori_mask;
    bd.cand_off = pl->cand_off.as<uint32_t>();
    bd.cand_nf = pl->cand_nf.as<uint32_t>();
    bd.cand_nr = pl->cand_nr.as<uint32_t>();
    bd.pool_n = pl->pool_n.as<uint32_t>();
    for (int p = 0; p < FAME_NPOOL; ++p) {
        bd.pool_cap[p] = pool_cap_limit(pl->pool_cap);
        bd.job_meta[p] = pl->job_meta[p].as<uint32_t>();
        bd.job_ro[p] = pl->job_ro[p].as<uint32_t>();
        bd.job_nm[p] = pl->job_nm[p].as<uint32_t>();
        bd.job_match[p] = pl->job_match[p].as<uint32_t>();
        bd.job_ovf_off[p] = pl->job_ovf_off[p].as<uint32_t>();
    }
    bd.match_cap = c->match_cap;
    bd.verify_runs_cap = verify_queue_limit();
    bd.ovf_n = pl->ovf_n.as<uint32_t>();
    bd.ovf_pool = pl->ovf_pool.as<uint32_t>();
    bd.ovf_job = pl->ovf_job.as<uint32_t>();
    bd.ovf_list_cap = pl->ovf_list_cap;
    bd.ovf_buf = pl->ovf_buf.as<uint32_t>();
    bd.ovf_buf_cap = pl->ovf_buf_cap;
    bd.heavy_n = pl->heavy_n.as<uint32_t>();
    bd.heavy_ro = pl->heavy_ro.as<uint32_t>();
    bd.heavy_e = pl->heavy_e.as<uint32_t>();
    bd.heavy_bc = pl->heavy_bc.as<uint32_t>();
    bd.heavy_cap = pl->heavy_cap;
    bd.status = pl->status.as<uint8_t>();
    bd.use = pl->use.as<uint8_t>();
    bd.mat = pl->mat.as<uint64_t>();
    bd.valid = pl->valid.as<uint8_t>();
    // single calls count into the context (strand tallies included); the
    // lanes of a pipelined call count into their own arrays
    bd.counters = pl->lane ? pl->counters.as<uint64_t>() : c->d_counters.as<uint64_t>();
    bd.meth = meth_buf(c);
    bd.direct_list = pl->direct.as<uint32_t>();
    bd.dp_list = pl->dplist.as<uint32_t>();
    bd.direct_n = pl->methn.as<uint32_t>();
    bd.dp_n = bd.direct_n + 1;
    bd.flags = pl->flags.as<uint32_t>();
    bd.max_len = bd.flags + 1;
    bd.stats = pl->stats.as<unsigned long long>();
    return bd;
}

// the batch pipeline of plan pl over the given inputs, captured into a graph
hipGraphExec_t capture_graph(fame_ctx* c, GraphPlan* pl, const GraphKey& key, const uint8_t* d_seqs,
                             const uint64_t* d_offs, const uint32_t* d_lens, uint32_t n, uint32_t ori,
                             int get_stranded, int cand, hipStream_t st)
{
    BatchDev bd = plan_batch(c, pl, d_seqs, d_offs, d_lens, ori);
    const uint32_t cap = pl->cap;
    // grids for the read lengths of the batch being captured: verify pools
    // and the seed variant its lengths need get full persistent grids, the
    // others small ones (complete whatever data later replays bring)
    uint32_t lr[2] = {0xFFFFFFFFu, 0};
    {
        uint32_t* d_lr = reinterpret_cast<uint32_t*>(pl->ovf_total.as<uint8_t>());
        HIPC(hipMemcpyAsync(d_lr, lr, 8, hipMemcpyHostToDevice, st));
        launch_len_range(d_lens, n, d_lr, st);
        HIPC(hipMemcpyAsync(lr, d_lr, 8, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
    }
    const uint32_t wmin = lr[0] ? (lr[0] + 63) / 64 : 1, wmax = lr[1] ? (lr[1] + 63) / 64 : 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    launch_batch_reset(bd, st);
    launch_prep(bd, cap, c->readlen, st);
    launch_seed_light(c->ix, bd, 2 * cap, st, std::max<uint32_t>(lr[1], 1));
    if (pl->hs.handed) launch_seed_mid(c->ix, bd, 0, FAME_JOBS_ON_DEVICE, nullptr, pl->mid_blocks, st);
    launch_seed_heavy_dev(c->ix, bd, pl->hs, pl->heavy_blocks, st);
    launch_seed_heavy_dev(c->ix, bd, pl->hs2, pl->heavy_blocks2, st);
    for (int p = 0; p < FAME_NPOOL; ++p) {
        const bool used = (uint32_t)p + 1 >= wmin && (uint32_t)p + 1 <= wmax;
        launch_verify(c->ix, bd, p, FAME_JOBS_ON_DEVICE, nullptr, st, used ? kVerifyBlocks[cand] : 16);
    }
    launch_ovf_offsets(bd, FAME_JOBS_ON_DEVICE, pl->ovf_total.as<uint32_t>(), st);
    for (int p = 0; p < FAME_NPOOL; ++p) launch_verify(c->ix, bd, p, FAME_JOBS_ON_DEVICE, bd.ovf_job, st, 64);
    launch_select(c->ix, bd, cap, get_stranded, st);
    launch_meth(c->ix, bd, cap, pl->dp.as<uint16_t>(), pl->dec.as<uint64_t>(), pl->trace.as<uint64_t>(),
                c->meth_threads, st);
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        HIPC(e);
    }
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPC(e);
    pl->add_graph(key, exec);
    c->graph_captures += 1;
    return exec;
}

// one batch through the plan's graph; 1 = the batch outgrew the graph's
// fixed capacities and must run on the plain path (nothing was changed)
int run_batch_graph(fame_ctx* c, const uint8_t* d_seqs, const uint64_t* d_offs, const uint32_t* d_lens, uint32_t n,
                    uint64_t total_bytes, int get_stranded, uint64_t counters[3], hipStream_t st,
                    const uint8_t** out_valid, const uint64_t** out_mat, const uint8_t** out_status)
{
    auto t0 = std::chrono::steady_clock::now();
    GraphPlan* pl = plan_for(c, n, total_bytes);
    const bool o0 = c->both || get_stranded || c->match_r1_fwd;
    const bool o1 = c->both || get_stranded || !c->match_r1_fwd;
    const uint32_t ori = (o0 ? 1u : 0u) | (o1 ? 2u : 0u);
    if (pl->chosen < 0)
        for (auto& q : c->plans)
            if (q->chosen >= 0) pl->chosen = q->chosen; // the grid suits the GPU, not the class
    const int cand = pl->chosen >= 0 ? pl->chosen : pl->next_trial;
    GraphKey key{d_seqs, d_offs, d_lens, get_stranded, ori, cand};
    hipGraphExec_t exec = pl->find_graph(key);
    if (!exec) exec = capture_graph(c, pl, key, d_seqs, d_offs, d_lens, n, ori, get_stranded, cand, st);
    Timer tm(st);
    pl->h_io[0] = n;
    HIPC(hipMemsetAsync(c->d_counters.p, 0, 24, st));
    HIPC(hipMemcpyAsync(pl->ndev.p, pl->h_io, 4, hipMemcpyHostToDevice, st));
    tm.mark(0);
    HIPC(hipGraphLaunch(exec, st));
    tm.mark(1);
    HIPC(hipMemcpyAsync(pl->h_io + 1, pl->flags.p, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(pl->h_io + 2, c->d_counters.p, 24, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(pl->h_io + 8, pl->pool_n.p, 4 * FAME_NPOOL, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(c->stats, pl->stats.p, 8 * FAME_NSTATS, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    c->graph_calls += 1;
    const uint32_t flags = pl->h_io[1];
    // the batch's counters stay on the device until read: clear them for the
    // next call of either form
    HIPC(hipMemsetAsync(c->d_counters.p, 0, 24, st));
    HIPC(hipStreamSynchronize(st));
    if (flags & FLAG_DP_BROKEN) {
        // raised after reads were counted: an error, never a re-run
        set_err(c, "a banded alignment's traceback left the band (internal error)");
        return FAME_ERR_CAPACITY;
    }
    if (flags & FLAG_ABORT) {
        c->graph_fallbacks += 1;
        if (flags & FLAG_POOL_FULL) plan_grow_pools(c, pl, pl->h_io + 8);
        return 1;
    }
    const double ms = tm.ms(0, 1);
    if (pl->chosen < 0 && n >= pl->cap / 4) {
        pl->cand_ms[cand] += ms;
        pl->cand_reads[cand] += (int)std::min<uint32_t>(n, 0x7FFFFFFF);
        pl->next_trial += 1;
        if (pl->next_trial >= kNCand) {
            int best = 0;
            for (int k = 1; k < kNCand; ++k)
                if (pl->cand_ms[k] / std::max(1, pl->cand_reads[k]) < pl->cand_ms[best] / std::max(1, pl->cand_reads[best]))
                    best = k;
            pl->chosen = best;
        }
    }
    uint64_t cnt[3];
    memcpy(cnt, pl->h_io + 2, 24);
    counters[0] += cnt[0];
    counters[1] += cnt[1];
    counters[2] += cnt[2];
    *out_valid = pl->valid.as<uint8_t>();
    *out_mat = pl->mat.as<uint64_t>();
    if (out_status) *out_status = pl->status.as<uint8_t>();
    for (double& x : c->timings) x = 0;
    c->timings[1] = ms; // the whole graph
    c->timings[5] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return FAME_OK;
}

// a large batched call: sub-batches of at most pipe_batch reads alternate
// between two lanes (own graph buffers, own stream), so that one lane's
// memory-bound seeding overlaps the other's compute-bound verification and
// DP.  The call is cut into an even number of equal sub-batches, so that the
// lanes end together (a last sub-batch of its own runs with no other lane's
// work beside it).  Every read is independent and the shared CpG counts are
// atomic sums, so the split does not change any result.  Sub-batches that
// outgrew the graphs are re-run on the plain path at the end.  The stranded
// first chunk (get_stranded: both orientations, strand tallies) is split the
// same way; the lanes' tallies are added to the context's.
int run_pipelined(fame_ctx* c, const uint8_t* d_seqs, const uint64_t* d_offs, const uint32_t* d_lens, uint32_t n,
                  uint64_t total_bytes, int get_stranded, uint64_t counters[3], const uint8_t** out_valid,
                  const uint64_t** out_mat, const uint8_t** out_status)
{
    auto t0 = std::chrono::steady_clock::now();
    const bool o0 = c->both || get_stranded || c->match_r1_fwd;
    const bool o1 = c->both || get_stranded || !c->match_r1_fwd;
    const uint32_t ori = (o0 ? 1u : 0u) | (o1 ? 2u : 0u);
    hipStream_t ls[2] = {c->stream, c->stream2};
    GraphPlan* pl[2] = {plan_for(c, c->pipe_batch, total_bytes, 1), plan_for(c, c->pipe_batch, total_bytes, 2)};
    int cand = 2;
    for (auto& q : c->plans)
        if (q->chosen >= 0) cand = q->chosen;
    uint8_t* ov = grow<uint8_t>(c->b_pvalid, n + 1);
    uint64_t* om = grow<uint64_t>(c->b_pmat, n + 1);
    uint8_t* os = grow<uint8_t>(c->b_pstatus, n + 1);
    uint32_t nb = (n + c->pipe_batch - 1) / c->pipe_batch;
    if (nb & 1) ++nb; // n > pipe_batch: at least two
    const uint32_t B = (uint32_t)(((uint64_t)n + nb - 1) / nb);
    nb = (n + B - 1) / B; // no empty sub-batch
    struct Sub {
        uint32_t s, n, lane, slot;
    };
    std::vector<Sub> subs;
    for (int l = 0; l < 2; ++l) HIPC(hipMemsetAsync(pl[l]->counters.p, 0, 64, ls[l]));
    // rounds of up to 16 sub-batches per lane (the pinned ring of each lane)
    for (uint32_t k0 = 0; k0 < nb; k0 += 32) {
        const uint32_t k1 = std::min(nb, k0 + 32);
        // graphs first: nothing is captured while a lane has work queued
        std::vector<hipGraphExec_t> ex(k1 - k0, nullptr);
        std::vector<Sub> round;
        for (uint32_t k = k0; k < k1; ++k) {
            const int l = (int)(k & 1);
            const uint32_t s0 = k * B;
            GraphKey key{d_seqs, d_offs + s0, d_lens + s0, get_stranded, ori, cand};
            ex[k - k0] = pl[l]->find_graph(key);
            if (!ex[k - k0]) {
                HIPC(hipStreamSynchronize(ls[l]));
                ex[k - k0] = capture_graph(c, pl[l], key, d_seqs, d_offs + s0, d_lens + s0, std::min(n - s0, B), ori,
                                           get_stranded, cand, ls[l]);
            }
        }
        for (uint32_t k = k0; k < k1; ++k) {
            const int l = (int)(k & 1);
            const uint32_t s0 = k * B, m = std::min(n - s0, B);
            const uint32_t slot = (k - k0) >> 1;
            GraphPlan* p = pl[l];
            p->h_io[16 + slot] = m;
            HIPC(hipMemcpyAsync(p->ndev.p, p->h_io + 16 + slot, 4, hipMemcpyHostToDevice, ls[l]));
            HIPC(hipGraphLaunch(ex[k - k0], ls[l]));
            HIPC(hipMemcpyAsync(p->h_io + 32 + slot, p->flags.p, 4, hipMemcpyDeviceToHost, ls[l]));
            HIPC(hipMemcpyAsync(p->h_io + 64 + 4 * slot, p->pool_n.p, 4 * FAME_NPOOL, hipMemcpyDeviceToHost, ls[l]));
            HIPC(hipMemcpyAsync(ov + s0, p->valid.p, m, hipMemcpyDeviceToDevice, ls[l]));
            HIPC(hipMemcpyAsync(om + s0, p->mat.p, 8ull * m, hipMemcpyDeviceToDevice, ls[l]));
            HIPC(hipMemcpyAsync(os + s0, p->status.p, m, hipMemcpyDeviceToDevice, ls[l]));
            round.push_back({s0, m, (uint32_t)l, slot});
        }
        HIPC(hipStreamSynchronize(ls[0]));
        HIPC(hipStreamSynchronize(ls[1]));
        // the round's flags, read before the ring is reused; aborted
        // sub-batches are carried to the end
        for (const Sub& x : round) {
            const uint32_t f = pl[x.lane]->h_io[32 + x.slot];
            if (f & FLAG_DP_BROKEN) {
                set_err(c, "a banded alignment's traceback left the band (internal error)");
                return FAME_ERR_CAPACITY;
            }
            if (!(f & FLAG_ABORT)) continue;
            subs.push_back(x);
            if (f & FLAG_POOL_FULL) plan_grow_pools(c, pl[x.lane], pl[x.lane]->h_io + 64 + 4 * x.slot);
        }
        round.clear();
    }
    c->graph_calls += nb;
    c->pipe_calls += 1;
    uint64_t cnt[2][8];
    HIPC(hipMemcpy(cnt[0], pl[0]->counters.p, 64, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(cnt[1], pl[1]->counters.p, 64, hipMemcpyDeviceToHost));
    for (int k = 0; k < 3; ++k) counters[k] += cnt[0][k] + cnt[1][k];
    if (get_stranded) { // the strand tallies, where fame_decide_strand reads them
        uint64_t t[2];
        HIPC(hipMemcpy(t, c->d_counters.as<uint64_t>() + 3, 16, hipMemcpyDeviceToHost));
        t[0] += cnt[0][3] + cnt[1][3];
        t[1] += cnt[0][4] + cnt[1][4];
        HIPC(hipMemcpy(c->d_counters.as<uint64_t>() + 3, t, 16, hipMemcpyHostToDevice));
    }
    // the sub-batches that outgrew the graphs, on the plain path
    for (Sub& x : subs) {
        c->graph_fallbacks += 1;
        int rc = run_batch(c, d_seqs, d_offs + x.s, d_lens + x.s, x.n, total_bytes, get_stranded, counters, c->stream);
        if (rc != FAME_OK) return rc;
        HIPC(hipMemcpyAsync(ov + x.s, c->b_valid.p, x.n, hipMemcpyDeviceToDevice, c->stream));
        HIPC(hipMemcpyAsync(om + x.s, c->b_mat.p, 8ull * x.n, hipMemcpyDeviceToDevice, c->stream));
        HIPC(hipMemcpyAsync(os + x.s, c->b_status.p, x.n, hipMemcpyDeviceToDevice, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
    }
    *out_valid = ov;
    *out_mat = om;
    if (out_status) *out_status = os;
    memset(c->stats, 0, sizeof c->stats);
    for (double& x : c->timings) x = 0;
    c->timings[5] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->timings[1] = c->timings[5];
    return FAME_OK;
}

// one single-end batch in the context's form; the per-read outputs stay in
// device buffers whose addresses are returned
int match_batch(fame_ctx* c, const uint8_t* d_seqs, const uint64_t* d_offs, const uint32_t* d_lens, uint32_t n,
                uint64_t total_bytes, int get_stranded, uint64_t counters[3], hipStream_t st,
                const uint8_t** out_valid, const uint64_t** out_mat, const uint8_t** out_status)
{
    // a large call on two lanes; the stranded first chunk (300 000 reads in
    // FAME's loop) also from 2^16 reads on, since nothing else can overlap it
    if (c->form == 1 && (n > c->pipe_batch || (get_stranded && n >= (1u << 16))))
        return run_pipelined(c, d_seqs, d_offs, d_lens, n, total_bytes, get_stranded, counters, out_valid, out_mat,
                             out_status);
    if (c->form == 1 && n) {
        const int rc = run_batch_graph(c, d_seqs, d_offs, d_lens, n, total_bytes, get_stranded, counters, st,
                                       out_valid, out_mat, out_status);
        if (rc != 1) return rc;
    }
    const int rc = run_batch(c, d_seqs, d_offs, d_lens, n, total_bytes, get_stranded, counters, st);
    *out_valid = c->b_valid.as<uint8_t>();
    *out_mat = c->b_mat.as<uint64_t>();
    if (out_status) *out_status = c->b_status.as<uint8_t>();
    return rc;
}

// verify jobs of one phase: the match arrays, then the jobs whose matches
// overflowed match_cap once more into an exactly sized buffer
int verify_phase(fame_ctx* c, BatchDev& bd, const uint32_t pool_n[FAME_NPOOL], DevBuf (&nm)[FAME_NPOOL],
                 DevBuf (&match)[FAME_NPOOL], DevBuf (&ovf_off)[FAME_NPOOL], DevBuf& ovf_buf, DevBuf& ovf_total,
                 hipStream_t st)
{
    for (int p = 0; p < FAME_NPOOL; ++p) {
        const uint32_t cap = std::max<uint32_t>(pool_n[p], 1);
        bd.job_nm[p] = grow<uint32_t>(nm[p], cap);
        bd.job_match[p] = grow<uint32_t>(match[p], (size_t)cap * c->match_cap);
        bd.job_ovf_off[p] = grow<uint32_t>(ovf_off[p], cap);
    }
    bd.ovf_buf = grow<uint32_t>(ovf_buf, 1);
    HIPC(hipMemsetAsync(bd.ovf_n, 0, 4, st));
    for (int p = 0; p < FAME_NPOOL; ++p) launch_verify(c->ix, bd, p, pool_n[p], nullptr, st);
    uint32_t ovf_n = 0, flags = 0;
    HIPC(hipMemcpyAsync(&ovf_n, bd.ovf_n, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(&flags, bd.flags, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (flags & FLAG_OVF_FULL) {
        set_err(c, "match overflow list full");
        return FAME_ERR_CAPACITY;
    }
    c->last_ovf += ovf_n;
    if (ovf_n) {
        uint32_t* d_total = grow<uint32_t>(ovf_total, 2);
        launch_ovf_offsets(bd, ovf_n, d_total, st);
        uint32_t total = 0;
        HIPC(hipMemcpyAsync(&total, d_total, 4, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        bd.ovf_buf = grow<uint32_t>(ovf_buf, total + 1);
        for (int p = 0; p < FAME_NPOOL; ++p) launch_verify(c->ix, bd, p, ovf_n, bd.ovf_job, st);
    }
    return FAME_OK;
}

// global scratch slice of the paired seed stage for reads of at most cap_e
// index hits (the order of k_pe_seed_big's carving)
void pe_scratch_dims(PairDev& pd, uint32_t cap_e)
{
    auto rup = [](uint64_t v) {
        uint64_t p = 1;
        while (p < v) p <<= 1;
        return p;
    };
    pd.cap_e = cap_e;
    pd.hsize = (uint32_t)rup(2ull * cap_e);
    // a strand map holds at most the first read's windows (<= E / QTHRESH)
    // and the windows the second read's entries create (<= E)
    pd.mcap = cap_e + cap_e / 5 + 64;
    pd.cap_t = pd.mcap;
    pd.tb_cap = (uint32_t)(2 * rup(2ull * pd.mcap) + 64);
    pd.mhsz = (uint32_t)rup(2ull * (2ull * pd.mcap + cap_e));
    size_t stride = 0;
    auto add = [&](size_t b) { stride += (b + 255) & ~(size_t)255; };
    add(8 * 256);
    add(4 * 256);
    add(4 * 256);
    add(4 * 256);
    add(256);
    add(4ull * cap_e);
    add(cap_e);
    add(4ull * cap_e);
    for (int i = 0; i < 3; ++i) add(4ull * pd.hsize);
    add(8ull * cap_e);
    add(4ull * cap_e);
    add(4ull * pd.tb_cap);
    add(4ull * pd.cap_t);
    add(4ull * pd.cap_t);
    add(8ull * pd.mcap);
    add(8ull * pd.mcap);
    for (int i = 0; i < 4; ++i) add(4ull * pd.mhsz);
    pd.stride = stride;
}

// ReadQueue::matchPairedReads over n pairs held as 2n device reads (mate 1 of
// pair i is read i, mate 2 read n + i)
int run_pairs(fame_ctx* c, const uint8_t* d_seqs, const uint64_t* d_offs, const uint32_t* d_lens, uint32_t n,
              uint64_t total_bytes, int get_stranded, uint64_t counters[5], hipStream_t st)
{
    auto t0 = std::chrono::steady_clock::now();
    Timer tm(st);
    const uint32_t nr = 2 * n;
    const uint32_t nitems = 2 * n;
    c->last_ovf = 0;
    c->last_heavy = 0;
    c->big_ms = 0;
    BatchDev bd{};
    bd.seqs = d_seqs;
    bd.offs = d_offs;
    bd.lens = d_lens;
    bd.rev = grow<uint8_t>(c->b_rev, total_bytes + 64);
    bd.pre = grow<uint8_t>(c->b_pre, nr + 1);
    const bool o0 = c->both || get_stranded || c->match_r1_fwd;
    const bool o1 = c->both || get_stranded || !c->match_r1_fwd;
    bd.ori_mask = (o0 ? 1u : 0u) | (o1 ? 2u : 0u);
    bd.codes = grow<uint8_t>(c->b_codes, (size_t)2 * nr * 256 + 256);
    bd.pool_n = grow<uint32_t>(c->b_pool_n, FAME_NPOOL);
    bd.match_cap = c->match_cap;
    bd.verify_runs_cap = verify_queue_limit();
    const uint32_t seed_grid = std::min<uint32_t>(std::max<uint32_t>(nitems, 1), 8192);
    const uint32_t want_cap = std::max<uint32_t>(1u << 16, 4 * n) + 64u * seed_grid;
    for (int p = 0; p < FAME_NPOOL; ++p) c->pool_cap[p] = std::max(c->pool_cap[p], want_cap);
    bd.ovf_n = grow<uint32_t>(c->b_ovf_n, 1);
    bd.ovf_list_cap = std::max<uint32_t>(1024, n / 4);
    bd.ovf_pool = grow<uint32_t>(c->b_ovf_pool, bd.ovf_list_cap);
    bd.ovf_job = grow<uint32_t>(c->b_ovf_job, bd.ovf_list_cap);
    bd.status = grow<uint8_t>(c->b_status, nr + 1);
    bd.use = grow<uint8_t>(c->b_use, nr + 1);
    bd.mat = grow<uint64_t>(c->b_mat, nr + 1);
    bd.valid = grow<uint8_t>(c->b_valid, nr + 1);
    bd.counters = c->d_counters.as<uint64_t>();
    bd.meth = meth_buf(c);
    bd.flags = c->d_flags.as<uint32_t>();
    bd.max_len = bd.flags + 1;
    bd.stats = grow<unsigned long long>(c->d_stats, FAME_NSTATS);
    HIPC(hipMemsetAsync(bd.stats, 0, 8 * FAME_NSTATS, st));
    bd.direct_list = grow<uint32_t>(c->b_direct, nr + 1);
    bd.dp_list = grow<uint32_t>(c->b_dplist, nr + 1);
    bd.direct_n = grow<uint32_t>(c->b_methn, 2);
    bd.dp_n = bd.direct_n + 1;
    HIPC(hipMemsetAsync(bd.direct_n, 0, 8, st));
    HIPC(hipMemsetAsync(bd.flags, 0, 8, st));

    PairDev pd{};
    pd.n = n;
    pd.combos = bd.ori_mask;
    pd.cw = c->maxpd / c->winlen + 1;
    pd.maxpd = c->maxpd;
    pd.map_off = grow<uint32_t>(c->b_pe_off, nitems + 1);
    pd.map_nf = grow<uint32_t>(c->b_pe_nf, nitems + 1);
    pd.map_nr = grow<uint32_t>(c->b_pe_nr, nitems + 1);
    pd.m1_n = grow<uint32_t>(c->b_pe_m1, nitems + 1);
    pd.map_n = grow<uint32_t>(c->b_pe_mapn, 8);
    pd.big_n = pd.map_n + 1;
    pd.next = pd.map_n + 2;
    pd.wlist_n = pd.map_n + 3;
    pd.wnext = pd.map_n + 4;
    pd.wlist = grow<uint32_t>(c->b_pe_wlist, nitems + 1);
    // FAME_PE_WALK_LDS / FAME_PE_WALK_BIG_LDS lower the walk kernels' LDS
    // limits (0: every item through k_pe_walk1_big, every such item's arrays
    // in the map pool), so that tests reach those paths on small inputs
    auto env_cap = [](const char* name, uint32_t dflt) {
        const char* v = getenv(name);
        if (!v || !*v) return dflt;
        return (uint32_t)std::min<unsigned long>(strtoul(v, nullptr, 10), dflt);
    };
    pd.wcap = env_cap("FAME_PE_WALK_LDS", FAME_PE_WCAP);
    pd.wbig = env_cap("FAME_PE_WALK_BIG_LDS", FAME_PE_WBIG);
    c->pe_map_cap = std::max<uint32_t>(c->pe_map_cap, std::max<uint32_t>(1u << 16, 8 * n) +
                                                          FAME_PE_MAP_CHUNK * seed_grid);
    HIPC(c->d_pe_counters.reserve(64));
    pd.counters = c->d_pe_counters.as<uint64_t>();
    HIPC(hipMemsetAsync(pd.counters, 0, 40, st));
    pd.big_cap = std::max<uint32_t>(1024, nitems);
    pd.big_item = grow<uint32_t>(c->b_pe_big, 2ull * pd.big_cap);
    pd.big_e = pd.big_item + pd.big_cap;
    pd.nslices = 0;
    const uint32_t minlen = (uint32_t)std::ceil((float)c->readlen * 0.75);

    tm.mark(0);
    launch_prep_pairs(bd, nr, n, minlen, st);
    tm.mark(1);
    // seed: retried with larger pools when a pool overflowed
    uint32_t pool_n[FAME_NPOOL];
    uint32_t flags = 0;
    for (int attempt = 0;; ++attempt) {
        for (int p = 0; p < FAME_NPOOL; ++p) {
            bd.pool_cap[p] = c->pool_cap[p];
            bd.job_meta[p] = grow<uint32_t>(c->b_job_meta[p], c->pool_cap[p]);
            bd.job_ro[p] = grow<uint32_t>(c->b_job_ro[p], c->pool_cap[p]);
        }
        pd.map_cap = c->pe_map_cap;
        pd.map = grow<PeItem>(c->b_pe_map, pd.map_cap);
        pd.srt = grow<PeItem>(c->b_pe_srt, pd.map_cap);
        pd.list = nullptr;
        pd.list_n = 0;
        HIPC(hipMemsetAsync(bd.pool_n, 0, 4 * FAME_NPOOL, st));
        HIPC(hipMemsetAsync(pd.map_n, 0, 12, st));
        tm.mark(2);
        launch_pe_seed(c->ix, bd, pd, seed_grid, st);
        tm.mark(3);
        uint32_t cnt2[2] = {0, 0};
        HIPC(hipMemcpyAsync(cnt2, pd.map_n, 8, hipMemcpyDeviceToHost, st));
        HIPC(hipMemcpyAsync(&flags, bd.flags, 4, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        if (flags & FLAG_TOO_LONG) {
            set_err(c, "a mate is longer than FAME_MAX_MATE");
            return FAME_ERR_ARG;
        }
        if (flags & FLAG_HEAVY_FULL) {
            set_err(c, "paired seed overflow list full");
            return FAME_ERR_CAPACITY;
        }
        // items beyond the LDS kernel: global-scratch runs, each sized for
        // the largest listed read, whose own overflows are listed again
        // with twice the capacity
        uint32_t big_n = cnt2[1];
        c->last_heavy = big_n;
        uint64_t cap_prev = 0;
        for (int round = 0; big_n && !(flags & FLAG_POOL_FULL); ++round) {
            if (round >= 24) { // 1024 * 2^23 hits: beyond the 2^30 bound
                flags |= FLAG_CAPACITY;
                break;
            }
            std::vector<uint32_t> be(big_n), bi(big_n);
            HIPC(hipMemcpy(be.data(), pd.big_e, 4 * big_n, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(bi.data(), pd.big_item, 4 * big_n, hipMemcpyDeviceToHost));
            uint64_t emax = 1;
            for (uint32_t x = 0; x < big_n; ++x) emax = std::max<uint64_t>(emax, be[x]);
            // the first run is sized for the bulk (<= 1024 hits, many
            // slices); what still overflows goes to larger runs
            emax = std::min<uint64_t>(emax, std::max<uint64_t>(1024, 2 * cap_prev));
            emax = std::max<uint64_t>(emax, std::max<uint64_t>(2 * cap_prev, 256));
            // a mate's hits are bounded by the device memory its scratch
            // slice needs (FAME's maps grow on the heap)
            if (emax > 0x3FFFFFFFull) {
                set_err(c, "a mate has more than 2^30 index hits, or a bucket of more than 2^24 entries");
                return FAME_ERR_CAPACITY;
            }
            cap_prev = emax;
            PairDev pb = pd;
            pe_scratch_dims(pb, (uint32_t)emax);
            if (!scratch_fits(c, pb.stride)) {
                set_err(c, "a mate's index hits need " + std::to_string(pb.stride >> 20) +
                               " MiB of scratch, more than the device has free");
                return FAME_ERR_CAPACITY;
            }
            const size_t budget = 2ull << 30;
            pb.nslices = (uint32_t)std::max<size_t>(1, std::min<size_t>(big_n, budget / pb.stride));
            pb.scratch = grow<uint8_t>(c->b_heavy_scratch, pb.stride * pb.nslices);
            // largest first (by the bit length of the hit count, a counting
            // sort), so that no long item starts last
            uint32_t nbin[33] = {0};
            for (uint32_t x = 0; x < big_n; ++x) ++nbin[32 - (be[x] ? __builtin_clz(be[x]) : 32)];
            uint32_t at[33];
            for (int b = 32, s = 0; b >= 0; --b) {
                at[b] = (uint32_t)s;
                s += (int)nbin[b];
            }
            std::vector<uint32_t> bl(big_n);
            for (uint32_t x = 0; x < big_n; ++x) bl[at[32 - (be[x] ? __builtin_clz(be[x]) : 32)]++] = bi[x];
            uint32_t* lst = grow<uint32_t>(c->b_pe_list, big_n + 1);
            HIPC(hipMemcpyAsync(lst, bl.data(), 4 * big_n, hipMemcpyHostToDevice, st));
            pb.list = lst;
            pb.list_n = big_n;
            HIPC(hipMemsetAsync(pd.big_n, 0, 4, st));
            HIPC(hipMemsetAsync(pd.next, 0, 4, st));
            tm.mark(12);
            launch_pe_seed(c->ix, bd, pb, pb.nslices, st);
            tm.mark(13);
            HIPC(hipMemcpyAsync(&flags, bd.flags, 4, hipMemcpyDeviceToHost, st));
            HIPC(hipMemcpyAsync(&big_n, pd.big_n, 4, hipMemcpyDeviceToHost, st));
            HIPC(hipStreamSynchronize(st));
            c->big_ms += tm.ms(12, 13);
        }
        if (flags & FLAG_CAPACITY) {
            set_err(c, "a pair's candidate maps exceeded the device emulation limits");
            return FAME_ERR_CAPACITY;
        }
        HIPC(hipMemcpyAsync(pool_n, bd.pool_n, 4 * FAME_NPOOL, hipMemcpyDeviceToHost, st));
        HIPC(hipMemcpyAsync(cnt2, pd.map_n, 8, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        if (!(flags & FLAG_POOL_FULL)) break;
        if (attempt > 3) {
            set_err(c, "paired candidate pools keep overflowing");
            return FAME_ERR_CAPACITY;
        }
        for (int p = 0; p < FAME_NPOOL; ++p)
            c->pool_cap[p] = std::max<uint32_t>(c->pool_cap[p], pool_n[p] + (pool_n[p] >> 2) + 1024);
        c->pe_map_cap = std::max<uint32_t>(c->pe_map_cap, cnt2[0] + (cnt2[0] >> 2) + 1024);
        HIPC(hipMemsetAsync(bd.flags, 0, 4, st));
    }
    // first reads over their windows
    tm.mark(4);
    int rc = verify_phase(c, bd, pool_n, c->b_job_nm, c->b_job_match, c->b_job_ovf_off, c->b_ovf_buf,
                          c->b_ovf_total, st);
    if (rc != FAME_OK) return rc;
    tm.mark(5);
    for (int p = 0; p < FAME_NPOOL; ++p) {
        pd.acc1[p] = grow<uint8_t>(c->b_acc1[p], std::max<uint32_t>(pool_n[p], 1));
        if (pool_n[p]) HIPC(hipMemsetAsync(pd.acc1[p], 0, pool_n[p], st));
    }
    // first walks, then the second reads' jobs (retried with larger pools)
    BatchDev b2 = bd;
    b2.pool_n = grow<uint32_t>(c->b2_pool_n, FAME_NPOOL);
    b2.ovf_n = grow<uint32_t>(c->b2_ovf_n, 1);
    b2.ovf_pool = grow<uint32_t>(c->b2_ovf_pool, bd.ovf_list_cap);
    b2.ovf_job = grow<uint32_t>(c->b2_ovf_job, bd.ovf_list_cap);
    uint32_t pool2_n[FAME_NPOOL];
    for (int p = 0; p < FAME_NPOOL; ++p) c->pool2_cap[p] = std::max(c->pool2_cap[p], std::max<uint32_t>(1u << 16, 2 * n));
    for (int attempt = 0;; ++attempt) {
        for (int p = 0; p < FAME_NPOOL; ++p) {
            b2.pool_cap[p] = c->pool2_cap[p];
            b2.job_meta[p] = grow<uint32_t>(c->b2_job_meta[p], c->pool2_cap[p]);
            b2.job_ro[p] = grow<uint32_t>(c->b2_job_ro[p], c->pool2_cap[p]);
        }
        HIPC(hipMemsetAsync(b2.pool_n, 0, 4 * FAME_NPOOL, st));
        HIPC(hipMemsetAsync(pd.wlist_n, 0, 8, st));
        tm.mark(6);
        launch_pe_walk1(c->ix, bd, b2, pd, st);
        tm.mark(7);
        HIPC(hipMemcpyAsync(pool2_n, b2.pool_n, 4 * FAME_NPOOL, hipMemcpyDeviceToHost, st));
        HIPC(hipMemcpyAsync(&flags, bd.flags, 4, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        if (!(flags & FLAG_POOL_FULL)) break;
        if (attempt > 3) {
            set_err(c, "second-read pools keep overflowing");
            return FAME_ERR_CAPACITY;
        }
        for (int p = 0; p < FAME_NPOOL; ++p)
            c->pool2_cap[p] = std::max<uint32_t>(c->pool2_cap[p], pool2_n[p] + (pool2_n[p] >> 2) + 1024);
        HIPC(hipMemsetAsync(bd.flags, 0, 4, st));
    }
    tm.mark(8);
    rc = verify_phase(c, b2, pool2_n, c->b2_job_nm, c->b2_job_match, c->b2_job_ovf_off, c->b2_ovf_buf,
                      c->b2_ovf_total, st);
    if (rc != FAME_OK) return rc;
    tm.mark(9);
    for (int p = 0; p < FAME_NPOOL; ++p) {
        pd.acc2[p] = grow<uint8_t>(c->b_acc2[p], std::max<uint32_t>(pool2_n[p], 1));
        if (pool2_n[p]) HIPC(hipMemsetAsync(pd.acc2[p], 0, pool2_n[p], st));
    }
    tm.mark(10);
    launch_pe_final(c->ix, bd, b2, pd, get_stranded, st);
    tm.mark(11);
    {
        const size_t cells = (size_t)FAME_DP_SROWS * (2 * (FAME_DPBAND + 1) + 1);
        uint16_t* dp = grow<uint16_t>(c->b_dp, cells * c->meth_threads);
        uint64_t* dec = grow<uint64_t>(c->b_dec, (size_t)(FAME_DEV_MAX_READ + 1) * c->meth_threads);
        uint64_t* tr = grow<uint64_t>(c->b_trace, (size_t)FAME_DP_TRACE_WORDS * c->meth_threads);
        tm.mark(14);
        launch_meth(c->ix, bd, nr, dp, dec, tr, c->meth_threads, st, tm.e[16]);
        tm.mark(15);
    }
    uint64_t cnt[5];
    HIPC(hipMemcpyAsync(cnt, pd.counters, 40, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(c->stats, bd.stats, 8 * FAME_NSTATS, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(&flags, bd.flags, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (flags & FLAG_DP_BROKEN) {
        set_err(c, "a banded alignment's traceback left the band (internal error)");
        return FAME_ERR_CAPACITY;
    }
    for (int k = 0; k < 5; ++k) counters[k] += cnt[k];
    c->timings[0] = tm.ms(0, 1);
    c->timings[1] = tm.ms(2, 3) + c->big_ms;
    c->timings[2] = tm.ms(4, 5) + tm.ms(8, 9);
    c->timings[3] = tm.ms(6, 7) + tm.ms(10, 11);
    c->timings[4] = tm.ms(14, 15);
    c->timings[6] = nr ? tm.ms(16, 15) : 0.0;
    c->timings[5] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return FAME_OK;
}

template <typename F>
int guarded(fame_ctx* c, F&& f)
{
    try {
        return f();
    } catch (const HipError& e) {
        set_err(c, std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.e));
        return FAME_ERR_HIP;
    } catch (const FameError& e) {
        set_err(c, e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_err(c, "host out of memory");
        return FAME_ERR_CAPACITY;
    }
}

} // namespace

// ---------------------------------------------------------------------------
// allocation-free single-end matching (fame_match_reads_noalloc)
// ---------------------------------------------------------------------------

// The per-batch buffers of the allocation-free entry point, carved from one
// file-scope device array.  Every GPU holds its own copy until the process
// ends.  Host code never passes this name to a kernel or a copy (on the host
// it stands for a host-side shadow, not device memory): it uses the address
// that hipGetSymbolAddress returned for the current device (na_base).
static __device__ __align__(256) uint8_t scratch[FAME_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kNaMaxDevices = 64;
constexpr uint32_t kNaSeedWaves = 1024;      // persistent seed waves (pool slack: 64 per wave and pool)
constexpr uint32_t kNaPoolCap = 4 * FAME_NOALLOC_MAX_READS + 64 * kNaSeedWaves;
constexpr uint32_t kNaOvfList = 8192;        // jobs with more than match_cap matches
constexpr uint32_t kNaOvfBuf = 1u << 20;     // their matches
constexpr uint32_t kNaHeavy1 = 16384, kNaHeavy1Slices = 8; // index hits of a deferred (read, orientation)
constexpr uint32_t kNaHeavy2 = 1u << 18;     // one slice for the rare larger ones
constexpr uint32_t kNaDpLanes = 8192;
constexpr uint32_t kNaVerifyBlocks = 512;

std::mutex g_na_mu;
uint8_t* g_na_base[kNaMaxDevices];
int g_na_state[kNaMaxDevices]; // 0 not looked up, 1 ready, -1 failed
hipError_t g_na_err[kNaMaxDevices];

// at fame_open, once per device, with that device current
void noalloc_setup(int dev)
{
    if (dev < 0 || dev >= kNaMaxDevices) return;
    std::lock_guard<std::mutex> g(g_na_mu);
    if (g_na_state[dev] != 0) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) {
        g_na_base[dev] = static_cast<uint8_t*>(p);
        g_na_state[dev] = 1;
    } else {
        g_na_err[dev] = e == hipSuccess ? hipErrorInvalidSymbol : e;
        g_na_state[dev] = -1;
    }
}

// byte offsets of the buffers in the array (256-byte aligned), for batches of
// up to FAME_NOALLOC_MAX_READS reads ending within FAME_NOALLOC_MAX_BYTES
struct NaLayout {
    size_t rev, pre, codes, cand_off, cand_nf, cand_nr, pool_n, job_meta[FAME_NPOOL], job_ro[FAME_NPOOL],
        job_nm[FAME_NPOOL], job_match[FAME_NPOOL], job_ovf_off[FAME_NPOOL], ovf_n, ovf_pool, ovf_job, ovf_buf,
        ovf_total, heavy_n, heavy_ro, heavy_e, heavy_bc, heavy1, heavy2, status, use, mat, valid, stats, direct,
        dplist, methn, flags, ndev, end_word, dp, dec, trace, end;
    HeavyScratch hs1, hs2;
};

NaLayout na_layout(uint32_t match_cap)
{
    NaLayout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += (bytes + 255) & ~(size_t)255;
        return at;
    };
    const size_t N = FAME_NOALLOC_MAX_READS, nro = 2 * N;
    L.rev = take(FAME_NOALLOC_MAX_BYTES + 64);
    L.pre = take(N + 1);
    L.codes = take(nro * 256 + 256);
    L.cand_off = take(4 * nro + 4);
    L.cand_nf = take(4 * nro + 4);
    L.cand_nr = take(4 * nro + 4);
    L.pool_n = take(4 * FAME_NPOOL);
    for (int p = 0; p < FAME_NPOOL; ++p) {
        L.job_meta[p] = take(4ull * kNaPoolCap);
        L.job_ro[p] = take(4ull * kNaPoolCap);
        L.job_nm[p] = take(4ull * kNaPoolCap);
        L.job_match[p] = take(4ull * kNaPoolCap * match_cap);
        L.job_ovf_off[p] = take(4ull * kNaPoolCap);
    }
    L.ovf_n = take(8);
    L.ovf_pool = take(4ull * kNaOvfList);
    L.ovf_job = take(4ull * kNaOvfList);
    L.ovf_buf = take(4ull * kNaOvfBuf);
    L.ovf_total = take(8);
    L.heavy_n = take(8);
    L.heavy_ro = take(4 * nro);
    L.heavy_e = take(4 * nro);
    L.heavy_bc = take(4 * nro);
    auto dims = [](HeavyScratch& hs, uint32_t emax, uint32_t lo, uint32_t hi) {
        hs.cap_e = emax;
        hs.hsize = 2 * emax;
        hs.tb_cap = 2 * (2 * emax) + 64;
        hs.stride = heavy_scratch_stride(hs);
        hs.e_lo = lo;
        hs.e_hi = hi;
    };
    dims(L.hs1, kNaHeavy1, 0, kNaHeavy1);
    // the second tier also takes the items beyond its capacity, which fail
    // there (FLAG_CAPACITY) and abort the batch before anything is counted
    dims(L.hs2, kNaHeavy2, kNaHeavy1 + 1, 0xFFFFFFFFu);
    L.heavy1 = take(L.hs1.stride * kNaHeavy1Slices);
    L.heavy2 = take(L.hs2.stride);
    L.status = take(N + 1);
    L.use = take(N + 1);
    L.mat = take(8 * N + 8);
    L.valid = take(N + 1);
    L.stats = take(8 * FAME_NSTATS);
    L.direct = take(4 * N + 4);
    L.dplist = take(4 * N + 4);
    L.methn = take(8);
    L.flags = take(16);
    L.ndev = take(8);
    L.end_word = take(8);
    const size_t cells = (size_t)FAME_DP_SROWS * (2 * (FAME_DPBAND + 1) + 1);
    L.dp = take(2 * cells * kNaDpLanes);
    L.dec = take(8ull * (FAME_DEV_MAX_READ + 1) * kNaDpLanes);
    L.trace = take(8ull * FAME_DP_TRACE_WORDS * kNaDpLanes);
    L.end = o;
    return L;
}

// One batch of at most FAME_NOALLOC_MAX_READS reads through the batched
// form's kernels (counts on the device, fixed capacities, launched directly).
// 1 = the batch outgrew a capacity; it was detected before anything was
// counted (select and the CpG stage skip an aborted batch).
int na_run(fame_ctx* c, uint8_t* base, const NaLayout& L, const uint8_t* d_seqs, const uint64_t* d_offs,
           const uint32_t* d_lens, uint32_t n, int get_stranded, uint64_t counters[3], uint8_t* d_valid,
           uint64_t* d_mat, hipStream_t st)
{
    auto at = [&](size_t off) { return base + off; };
    const bool o0 = c->both || get_stranded || c->match_r1_fwd;
    const bool o1 = c->both || get_stranded || !c->match_r1_fwd;
    BatchDev bd{};
    uint32_t* ndev = (uint32_t*)at(L.ndev);
    bd.n_dev = ndev;
    bd.seqs = d_seqs;
    bd.offs = d_offs;
    bd.lens = d_lens;
    bd.rev = at(L.rev);
    bd.pre = at(L.pre);
    bd.codes = at(L.codes);
    bd.ori_mask = (o0 ? 1u : 0u) | (o1 ? 2u : 0u);
    bd.cand_off = (uint32_t*)at(L.cand_off);
    bd.cand_nf = (uint32_t*)at(L.cand_nf);
    bd.cand_nr = (uint32_t*)at(L.cand_nr);
    bd.pool_n = (uint32_t*)at(L.pool_n);
    for (int p = 0; p < FAME_NPOOL; ++p) {
        bd.pool_cap[p] = pool_cap_limit(kNaPoolCap);
        bd.job_meta[p] = (uint32_t*)at(L.job_meta[p]);
        bd.job_ro[p] = (uint32_t*)at(L.job_ro[p]);
        bd.job_nm[p] = (uint32_t*)at(L.job_nm[p]);
        bd.job_match[p] = (uint32_t*)at(L.job_match[p]);
        bd.job_ovf_off[p] = (uint32_t*)at(L.job_ovf_off[p]);
    }
    bd.match_cap = c->match_cap;
    bd.verify_runs_cap = verify_queue_limit();
    bd.ovf_n = (uint32_t*)at(L.ovf_n);
    bd.ovf_pool = (uint32_t*)at(L.ovf_pool);
    bd.ovf_job = (uint32_t*)at(L.ovf_job);
    bd.ovf_list_cap = kNaOvfList;
    bd.ovf_buf = (uint32_t*)at(L.ovf_buf);
    bd.ovf_buf_cap = kNaOvfBuf;
    bd.heavy_n = (uint32_t*)at(L.heavy_n);
    bd.heavy_ro = (uint32_t*)at(L.heavy_ro);
    bd.heavy_e = (uint32_t*)at(L.heavy_e);
    bd.heavy_bc = (uint32_t*)at(L.heavy_bc);
    bd.heavy_cap = 2 * FAME_NOALLOC_MAX_READS;
    bd.status = at(L.status);
    bd.use = at(L.use);
    bd.mat = (uint64_t*)at(L.mat);
    bd.valid = at(L.valid);
    bd.counters = c->d_counters.as<uint64_t>();
    bd.meth = meth_buf(c);
    bd.direct_list = (uint32_t*)at(L.direct);
    bd.dp_list = (uint32_t*)at(L.dplist);
    bd.direct_n = (uint32_t*)at(L.methn);
    bd.dp_n = bd.direct_n + 1;
    bd.flags = (uint32_t*)at(L.flags);
    bd.max_len = bd.flags + 1;
    bd.stats = (unsigned long long*)at(L.stats);
    HeavyScratch h1 = L.hs1, h2 = L.hs2;
    h1.base = at(L.heavy1);
    h2.base = at(L.heavy2);
    // the array keeps the last call's contents: the counters, lists and flags
    // that must start at zero are cleared by kernels (k_batch_reset, and the
    // context's per-call counters below)
    launch_set_u32(ndev, n, st);
    launch_zero_u32((uint32_t*)c->d_counters.p, 6, st);
    launch_batch_reset(bd, st);
    launch_prep(bd, n, c->readlen, st);
    launch_seed_light(c->ix, bd, 2 * n, st, 0, kNaSeedWaves);
    launch_seed_heavy_dev(c->ix, bd, h1, kNaHeavy1Slices, st);
    launch_seed_heavy_dev(c->ix, bd, h2, 1, st);
    for (int p = 0; p < FAME_NPOOL; ++p) launch_verify(c->ix, bd, p, FAME_JOBS_ON_DEVICE, nullptr, st, kNaVerifyBlocks);
    launch_ovf_offsets(bd, FAME_JOBS_ON_DEVICE, (uint32_t*)at(L.ovf_total), st);
    for (int p = 0; p < FAME_NPOOL; ++p) launch_verify(c->ix, bd, p, FAME_JOBS_ON_DEVICE, bd.ovf_job, st, 64);
    launch_select(c->ix, bd, n, get_stranded, st);
    launch_meth(c->ix, bd, n, (uint16_t*)at(L.dp), (uint64_t*)at(L.dec), (uint64_t*)at(L.trace), kNaDpLanes, st);
    uint32_t flags = 0;
    uint64_t cnt[3] = {0, 0, 0};
    HIPC(hipMemcpyAsync(&flags, bd.flags, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(cnt, c->d_counters.p, 24, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    launch_zero_u32((uint32_t*)c->d_counters.p, 6, st);
    if (flags & FLAG_DP_BROKEN) {
        HIPC(hipStreamSynchronize(st));
        set_err(c, "a banded alignment's traceback left the band (internal error)");
        return FAME_ERR_CAPACITY;
    }
    if (flags & FLAG_TOO_LONG) {
        HIPC(hipStreamSynchronize(st));
        set_err(c, "reads over FAME_MAX_MATE letters need fame_match_reads or fame_match_reads_device");
        return FAME_ERR_ARG;
    }
    if (flags & FLAG_ABORT) {
        HIPC(hipStreamSynchronize(st));
        return 1;
    }
    counters[0] += cnt[0];
    counters[1] += cnt[1];
    counters[2] += cnt[2];
    if (d_valid) HIPC(hipMemcpyAsync(d_valid, bd.valid, n, hipMemcpyDeviceToDevice, st));
    if (d_mat) HIPC(hipMemcpyAsync(d_mat, bd.mat, 8ull * n, hipMemcpyDeviceToDevice, st));
    HIPC(hipStreamSynchronize(st));
    return FAME_OK;
}

// a batch that outgrew a capacity is matched again in halves (nothing of it
// was counted); a single read that still does not fit is an error
int na_match(fame_ctx* c, uint8_t* base, const NaLayout& L, const uint8_t* d_seqs, const uint64_t* d_offs,
             const uint32_t* d_lens, uint32_t first, uint32_t n, int get_stranded, uint64_t counters[3],
             uint8_t* d_valid, uint64_t* d_mat, hipStream_t st)
{
    const int rc = na_run(c, base, L, d_seqs, d_offs + first, d_lens + first, n, get_stranded, counters,
                          d_valid ? d_valid + first : nullptr, d_mat ? d_mat + first : nullptr, st);
    if (rc != 1) return rc;
    c->graph_fallbacks += 1;
    if (n == 1) {
        set_err(c, "read " + std::to_string(first) +
                       " exceeds the allocation-free limits (index hits or candidate windows); the reads before it "
                       "were matched and counted");
        return FAME_ERR_CAPACITY;
    }
    const uint32_t h = n / 2;
    const int r1 = na_match(c, base, L, d_seqs, d_offs, d_lens, first, h, get_stranded, counters, d_valid, d_mat, st);
    if (r1 != FAME_OK) return r1;
    return na_match(c, base, L, d_seqs, d_offs, d_lens, first + h, n - h, get_stranded, counters, d_valid, d_mat,
                    st);
}

} // namespace

// hooks of the serving layer (fame_serve.cpp)
namespace fame {

int serve_run(fame_ctx* c, const uint8_t* d_seqs, const uint64_t* d_offs, const uint32_t* d_lens, uint32_t n,
              uint64_t total_bytes, hipStream_t st, const uint8_t** d_status, const uint8_t** d_valid,
              const uint64_t** d_mat)
{
    return guarded(c, [&]() {
        uint64_t cnt[3] = {0, 0, 0};
        return match_batch(c, d_seqs, d_offs, d_lens, n, total_bytes, 0, cnt, st, d_valid, d_mat, d_status);
    });
}

hipStream_t serve_stream(fame_ctx* c) { return c->stream; }
int serve_device(fame_ctx* c) { return c->device; }
uint32_t serve_readlen(fame_ctx* c) { return c->readlen; }
uint32_t serve_winlen(fame_ctx* c) { return c->winlen; }

fame_ctx* serve_worker_open(fame_ctx* c)
{
    auto w = std::make_unique<fame_ctx>();
    w->parent = c;
    int rc = guarded(w.get(), [&]() {
        w->device = c->device;
        HIPC(hipSetDevice(c->device));
        HIPC(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
        HIPC(hipStreamCreateWithFlags(&w->stream2, hipStreamNonBlocking));
        w->both = c->both;
        w->match_r1_fwd = c->match_r1_fwd;
        w->htab = c->htab;
        w->kmercutoff = c->kmercutoff;
        w->ncpg = c->ncpg;
        w->nchrom = c->nchrom;
        w->nkmer = c->nkmer;
        w->nwin = c->nwin;
        w->nfilt = c->nfilt;
        w->readlen = c->readlen;
        w->winlen = c->winlen;
        w->kmerlen = c->kmerlen;
        w->seedbits = c->seedbits;
        w->ix = c->ix;
        w->match_cap = c->match_cap;
        w->minpd = c->minpd;
        w->maxpd = c->maxpd;
        w->meth_threads = c->meth_threads;
        w->form = c->form;
        w->pipe_batch = c->pipe_batch;
        HIPC(w->d_counters.reserve(8 * 8));
        HIPC(hipMemset(w->d_counters.p, 0, 64));
        HIPC(w->d_flags.reserve(16));
        return FAME_OK;
    });
    if (rc != FAME_OK) {
        fame_close(w.release());
        return nullptr;
    }
    return w.release();
}

void serve_worker_close(fame_ctx* w) { fame_close(w); }

// Batched form, paired: a call of more than pipe_batch / 2 pairs (not the
// stranded first chunk) is split into an even number of equal sub-batches of
// at most that many pairs that alternate between two lanes, the context and a worker context over the
// same index and CpG accumulators, each with its own buffers and stream and
// a host thread of its own: one lane's host waits between its stages (the
// paired pipeline sizes its stages from device counts) overlap the other's
// kernels.  Pairs are independent and the CpG counts are atomic sums, so the
// split changes no result.  Outputs: 2n mates (mate 1, then mate 2).
int run_pairs_pipelined(fame_ctx* c, const uint8_t* d_seqs, const uint64_t* d_o, const uint32_t* d_l, uint32_t n,
                        uint64_t total_bytes, uint64_t counters[5], const uint8_t** out_valid,
                        const uint64_t** out_mat)
{
    auto t0 = std::chrono::steady_clock::now();
    // an even number of equal sub-batches of at most pipe_batch / 2 pairs, so
    // that the lanes end together
    const uint32_t B0 = std::max<uint32_t>(c->pipe_batch / 2, 128);
    uint32_t nb = (n + B0 - 1) / B0;
    if (nb & 1) ++nb;
    const uint32_t B = (uint32_t)(((uint64_t)n + nb - 1) / nb);
    nb = (n + B - 1) / B;
    if (!c->pair_worker) c->pair_worker = serve_worker_open(c);
    fame_ctx* w = c->pair_worker;
    if (!w) { // no second lane: the plain pipeline
        const int rc = run_pairs(c, d_seqs, d_o, d_l, n, total_bytes, 0, counters, c->stream);
        *out_valid = c->b_valid.as<uint8_t>();
        *out_mat = c->b_mat.as<uint64_t>();
        return rc;
    }
    // the worker follows the context's strand decision and pair distances
    w->both = c->both;
    w->match_r1_fwd = c->match_r1_fwd;
    w->minpd = c->minpd;
    w->maxpd = c->maxpd;
    uint8_t* ov = grow<uint8_t>(c->b_pvalid, 2ull * n + 1);
    uint64_t* om = grow<uint64_t>(c->b_pmat, 2ull * n + 1);
    uint64_t* lo[2];
    uint32_t* ll[2];
    for (int l = 0; l < 2; ++l) {
        lo[l] = grow<uint64_t>(c->b_lane_offs[l], 2ull * B + 1);
        ll[l] = grow<uint32_t>(c->b_lane_lens[l], 2ull * B + 1);
    }
    uint64_t lcnt[2][5] = {{0}};
    int lrc[2] = {FAME_OK, FAME_OK};
    auto lane = [&](int l) {
        fame_ctx* x = l ? w : c;
        lrc[l] = guarded(x, [&]() {
            HIPC(hipSetDevice(c->device));
            hipStream_t st = x->stream;
            for (uint32_t k = (uint32_t)l; k < nb; k += 2) {
                const uint32_t s = k * B, m = std::min(B, n - s);
                // the sub-batch's mates as 2m reads: mate 1 of its pairs, then mate 2
                HIPC(hipMemcpyAsync(lo[l], d_o + s, 8ull * m, hipMemcpyDeviceToDevice, st));
                HIPC(hipMemcpyAsync(lo[l] + m, d_o + n + s, 8ull * m, hipMemcpyDeviceToDevice, st));
                HIPC(hipMemcpyAsync(ll[l], d_l + s, 4ull * m, hipMemcpyDeviceToDevice, st));
                HIPC(hipMemcpyAsync(ll[l] + m, d_l + n + s, 4ull * m, hipMemcpyDeviceToDevice, st));
                uint64_t cnt[5] = {0, 0, 0, 0, 0};
                const int rc = run_pairs(x, d_seqs, lo[l], ll[l], m, total_bytes, 0, cnt, st);
                if (rc != FAME_OK) return rc;
                HIPC(hipMemcpyAsync(ov + s, x->b_valid.p, m, hipMemcpyDeviceToDevice, st));
                HIPC(hipMemcpyAsync(ov + n + s, x->b_valid.as<uint8_t>() + m, m, hipMemcpyDeviceToDevice, st));
                HIPC(hipMemcpyAsync(om + s, x->b_mat.p, 8ull * m, hipMemcpyDeviceToDevice, st));
                HIPC(hipMemcpyAsync(om + n + s, x->b_mat.as<uint64_t>() + m, 8ull * m, hipMemcpyDeviceToDevice, st));
                HIPC(hipStreamSynchronize(st));
                for (int q = 0; q < 5; ++q) lcnt[l][q] += cnt[q];
            }
            return (int)FAME_OK;
        });
    };
    std::thread t1(lane, 1);
    lane(0);
    t1.join();
    if (lrc[0] != FAME_OK) return lrc[0];
    if (lrc[1] != FAME_OK) {
        set_err(c, "second lane: " + w->err);
        return lrc[1];
    }
    for (int q = 0; q < 5; ++q) counters[q] += lcnt[0][q] + lcnt[1][q];
    c->pipe_calls += 1;
    *out_valid = ov;
    *out_mat = om;
    memset(c->stats, 0, sizeof c->stats);
    for (double& x : c->timings) x = 0;
    c->timings[5] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->timings[1] = c->timings[5];
    return FAME_OK;
}

// one paired batch in the context's form; the 2n per-mate outputs stay in
// device buffers whose addresses are returned
int match_pairs_batch(fame_ctx* c, const uint8_t* d_seqs, const uint64_t* d_o, const uint32_t* d_l, uint32_t n,
                      uint64_t total_bytes, int get_stranded, uint64_t counters[5], hipStream_t st,
                      const uint8_t** out_valid, const uint64_t** out_mat)
{
    if (c->form == 1 && !get_stranded && n > std::max<uint32_t>(c->pipe_batch / 2, 128))
        return run_pairs_pipelined(c, d_seqs, d_o, d_l, n, total_bytes, counters, out_valid, out_mat);
    const int rc = run_pairs(c, d_seqs, d_o, d_l, n, total_bytes, get_stranded, counters, st);
    *out_valid = c->b_valid.as<uint8_t>();
    *out_mat = c->b_mat.as<uint64_t>();
    return rc;
}
void serve_set_err(fame_ctx* c, const std::string& m) { set_err(c, m); }

} // namespace fame

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int fame_open(const char* index_path, int both_strands, fame_ctx** out)
{
    if (!out || !index_path) return FAME_ERR_ARG;
    *out = nullptr;
    auto c = std::make_unique<fame_ctx>();
    c->both = both_strands != 0;
    int rc = guarded(c.get(), [&]() {
        HIPC(hipGetDevice(&c->device));
        HIPC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPC(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        c->meth_threads = meth_dp_lanes();
        noalloc_setup(c->device);
        load_index(c.get(), index_path);
        return FAME_OK;
    });
    if (rc != FAME_OK) {
        // keep the context so that fame_last_error can report
        *out = c.release();
        return rc;
    }
    *out = c.release();
    return FAME_OK;
}

int fame_close(fame_ctx* c)
{
    if (!c) return FAME_OK;
    if (c->pair_worker) fame_close(c->pair_worker);
    c->pair_worker = nullptr;
    c->plans.clear();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    delete c;
    return FAME_OK;
}

const char* fame_last_error(const fame_ctx* c)
{
    return c ? c->err.c_str() : "null context";
}

int fame_index_info(const fame_ctx* c, uint64_t out[FAME_INFO_FIELDS])
{
    if (!c || !out) return FAME_ERR_ARG;
    const uint64_t v[FAME_INFO_FIELDS] = {c->htab, c->readlen, c->winlen, c->kmerlen, c->kmercutoff, c->seedbits,
                                          c->ncpg, c->nchrom,  c->nkmer,  c->nwin,    c->nfilt};
    memcpy(out, v, sizeof v);
    return FAME_OK;
}

int fame_match_reads(fame_ctx* c, const uint8_t* seqs, const uint64_t* offs, const uint32_t* lens, uint32_t n,
                     int get_stranded, uint64_t counters[3], uint8_t* valid, uint64_t* mat)
{
    if (!c || !counters || (n && (!seqs || !offs || !lens))) return FAME_ERR_ARG;
    if (!c->d_tab.p) {
        set_err(c, "context has no index");
        return FAME_ERR_ARG;
    }
    return guarded(c, [&]() {
        // only the span of seqs the reads occupy is copied; the offsets are
        // rebased to it
        uint64_t lo = ~0ull, hi = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (lens[i] > FAME_MAX_READ && lens[i] <= c->winlen + 12) {
                set_err(c, "read " + std::to_string(i) + " is longer than FAME_MAX_READ and could match a window");
                return (int)FAME_ERR_ARG;
            }
            lo = std::min<uint64_t>(lo, offs[i]);
            hi = std::max<uint64_t>(hi, offs[i] + lens[i]);
        }
        if (!n) lo = 0;
        const uint64_t total = hi - lo;
        hipStream_t st = c->stream;
        uint8_t* d_s = grow<uint8_t>(c->b_seqs, total + 64);
        uint64_t* d_o = grow<uint64_t>(c->b_offs, n + 1);
        uint32_t* d_l = grow<uint32_t>(c->b_lens, n + 1);
        uint64_t* h_o = c->h_offs.get<uint64_t>(n + 1);
        for (uint32_t i = 0; i < n; ++i) h_o[i] = offs[i] - lo;
        if (total) HIPC(hipMemcpyAsync(d_s, seqs + lo, total, hipMemcpyHostToDevice, st));
        if (n) {
            HIPC(hipMemcpyAsync(d_o, h_o, 8ull * n, hipMemcpyHostToDevice, st));
            HIPC(hipMemcpyAsync(d_l, lens, 4ull * n, hipMemcpyHostToDevice, st));
        }
        const uint8_t* ov = nullptr;
        const uint64_t* om = nullptr;
        int rc = match_batch(c, d_s, d_o, d_l, n, total, get_stranded, counters, st, &ov, &om, nullptr);
        if (rc != FAME_OK) return rc;
        if (n && valid) HIPC(hipMemcpyAsync(valid, ov, n, hipMemcpyDeviceToHost, st));
        if (n && mat) HIPC(hipMemcpyAsync(mat, om, 8ull * n, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        return (int)FAME_OK;
    });
}

int fame_match_reads_device(fame_ctx* c, const uint8_t* d_seqs, const uint64_t* d_offs, const uint32_t* d_lens,
                            uint32_t n, int get_stranded, uint64_t counters[3], uint8_t* d_valid, uint64_t* d_mat,
                            void* stream)
{
    if (!c || !counters || (n && (!d_seqs || !d_offs || !d_lens))) return FAME_ERR_ARG;
    return guarded(c, [&]() {
        hipStream_t st = stream ? (hipStream_t)stream : c->stream;
        uint64_t total = 0;
        if (n) {
            // the reverse complements use the same offsets: size by the
            // furthest read end
            unsigned long long* d_end = grow<unsigned long long>(c->b_ovf_total, 1);
            HIPC(hipMemsetAsync(d_end, 0, 8, st));
            launch_max_end(d_offs, d_lens, n, d_end, st);
            unsigned long long e = 0;
            HIPC(hipMemcpyAsync(&e, d_end, 8, hipMemcpyDeviceToHost, st));
            HIPC(hipStreamSynchronize(st));
            total = e;
        }
        const uint8_t* ov = nullptr;
        const uint64_t* om = nullptr;
        int rc = match_batch(c, d_seqs, d_offs, d_lens, n, total, get_stranded, counters, st, &ov, &om, nullptr);
        if (rc != FAME_OK) return rc;
        if (n && d_valid) HIPC(hipMemcpyAsync(d_valid, ov, n, hipMemcpyDeviceToDevice, st));
        if (n && d_mat) HIPC(hipMemcpyAsync(d_mat, om, 8ull * n, hipMemcpyDeviceToDevice, st));
        HIPC(hipStreamSynchronize(st));
        return (int)FAME_OK;
    });
}

int fame_match_reads_noalloc(fame_ctx* c, const uint8_t* d_seqs, const uint64_t* d_offs, const uint32_t* d_lens,
                             uint32_t n, int get_stranded, uint64_t counters[3], uint8_t* d_valid, uint64_t* d_mat,
                             void* stream)
{
    if (!c || !counters || (n && (!d_seqs || !d_offs || !d_lens))) return FAME_ERR_ARG;
    if (!c->d_tab.p) {
        set_err(c, "context has no index");
        return FAME_ERR_ARG;
    }
    if (n > FAME_NOALLOC_MAX_READS) {
        set_err(c, "fame_match_reads_noalloc takes at most FAME_NOALLOC_MAX_READS (" +
                       std::to_string(FAME_NOALLOC_MAX_READS) + ") reads, got " + std::to_string(n));
        return FAME_ERR_ARG;
    }
    const int dev = c->device;
    if (dev < 0 || dev >= kNaMaxDevices || g_na_state[dev] != 1) {
        set_err(c, std::string("allocation-free scratch unavailable on this device: ") +
                       (dev >= 0 && dev < kNaMaxDevices && g_na_state[dev] < 0 ? hipGetErrorString(g_na_err[dev])
                                                                               : "not set up"));
        return FAME_ERR_HIP;
    }
    if (!n) return FAME_OK;
    return guarded(c, [&]() {
        hipStream_t st = stream ? (hipStream_t)stream : c->stream;
        int cur = -1;
        HIPC(hipGetDevice(&cur));
        if (cur != dev) {
            set_err(c, "the context's device is not the current device");
            return (int)FAME_ERR_ARG;
        }
        const NaLayout L = na_layout(c->match_cap);
        if (L.end > FAME_NOALLOC_SCRATCH_BYTES) {
            set_err(c, "allocation-free layout exceeds its scratch array");
            return (int)FAME_ERR_CAPACITY;
        }
        uint8_t* base = g_na_base[dev];
        // the reverse complements sit at the reads' own offsets: the furthest
        // read end must fit the array's copy buffer
        uint32_t* ew = (uint32_t*)(base + L.end_word);
        launch_zero_u32(ew, 2, st);
        launch_max_end(d_offs, d_lens, n, (unsigned long long*)ew, st);
        unsigned long long e = 0;
        HIPC(hipMemcpyAsync(&e, ew, 8, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        if (e > FAME_NOALLOC_MAX_BYTES) {
            set_err(c, "reads end " + std::to_string(e) + " bytes into d_seqs; fame_match_reads_noalloc takes at most "
                       "FAME_NOALLOC_MAX_BYTES (" + std::to_string(FAME_NOALLOC_MAX_BYTES) + ")");
            return (int)FAME_ERR_ARG;
        }
        return na_match(c, base, L, d_seqs, d_offs, d_lens, 0, n, get_stranded, counters, d_valid, d_mat, st);
    });
}

int fame_set_pair_dist(fame_ctx* c, uint32_t minpd, uint32_t maxpd)
{
    if (!c) return FAME_ERR_ARG;
    if (c->winlen == 0 || maxpd / c->winlen + 1 > 31 || minpd > maxpd) {
        set_err(c, "MAXPDIST must be below 30 windows and not below MINPDIST");
        return FAME_ERR_ARG;
    }
    c->minpd = minpd;
    c->maxpd = maxpd;
    return FAME_OK;
}

int fame_match_pairs(fame_ctx* c, const uint8_t* seqs1, const uint64_t* offs1, const uint32_t* lens1,
                     const uint8_t* seqs2, const uint64_t* offs2, const uint32_t* lens2, uint32_t n,
                     int get_stranded, uint64_t counters[5], uint8_t* valid1, uint64_t* mat1, uint8_t* valid2,
                     uint64_t* mat2)
{
    if (!c || !counters || (n && (!seqs1 || !offs1 || !lens1 || !seqs2 || !offs2 || !lens2))) return FAME_ERR_ARG;
    if (!c->d_tab.p) {
        set_err(c, "context has no index");
        return FAME_ERR_ARG;
    }
    if ((uint32_t)std::ceil((float)c->readlen * 0.75) < 32) {
        set_err(c, "paired matching needs ceil(0.75 READLEN) >= 32");
        return FAME_ERR_FORMAT;
    }
    if (n > 0x7FFFFFFFu) return FAME_ERR_ARG;
    return guarded(c, [&]() {
        // the spans of seqs1 and seqs2 the mates occupy, copied one after
        // the other; the offsets are rebased to them
        uint64_t lo1 = ~0ull, hi1 = 0, lo2 = ~0ull, hi2 = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (lens1[i] > FAME_MAX_MATE || lens2[i] > FAME_MAX_MATE) {
                set_err(c, "mate of pair " + std::to_string(i) + " is longer than FAME_MAX_MATE");
                return (int)FAME_ERR_ARG;
            }
            lo1 = std::min<uint64_t>(lo1, offs1[i]);
            hi1 = std::max<uint64_t>(hi1, offs1[i] + lens1[i]);
            lo2 = std::min<uint64_t>(lo2, offs2[i]);
            hi2 = std::max<uint64_t>(hi2, offs2[i] + lens2[i]);
        }
        if (!n) lo1 = lo2 = 0;
        const uint64_t t1 = hi1 - lo1, t2 = hi2 - lo2;
        hipStream_t st = c->stream;
        const uint64_t total = t1 + t2;
        uint8_t* d_s = grow<uint8_t>(c->b_seqs, total + 64);
        uint64_t* d_o = grow<uint64_t>(c->b_offs, 2ull * n + 1);
        uint32_t* d_l = grow<uint32_t>(c->b_lens, 2ull * n + 1);
        uint64_t* o = c->h_offs.get<uint64_t>(2ull * n + 1);
        for (uint32_t i = 0; i < n; ++i) {
            o[i] = offs1[i] - lo1;
            o[n + i] = offs2[i] - lo2 + t1;
        }
        if (t1) HIPC(hipMemcpyAsync(d_s, seqs1 + lo1, t1, hipMemcpyHostToDevice, st));
        if (t2) HIPC(hipMemcpyAsync(d_s + t1, seqs2 + lo2, t2, hipMemcpyHostToDevice, st));
        if (n) {
            HIPC(hipMemcpyAsync(d_o, o, 16ull * n, hipMemcpyHostToDevice, st));
            HIPC(hipMemcpyAsync(d_l, lens1, 4ull * n, hipMemcpyHostToDevice, st));
            HIPC(hipMemcpyAsync(d_l + n, lens2, 4ull * n, hipMemcpyHostToDevice, st));
        }
        const uint8_t* ov = nullptr;
        const uint64_t* om = nullptr;
        int rc = match_pairs_batch(c, d_s, d_o, d_l, n, total, get_stranded, counters, st, &ov, &om);
        if (rc != FAME_OK) return rc;
        if (n && valid1) HIPC(hipMemcpyAsync(valid1, ov, n, hipMemcpyDeviceToHost, st));
        if (n && valid2) HIPC(hipMemcpyAsync(valid2, ov + n, n, hipMemcpyDeviceToHost, st));
        if (n && mat1) HIPC(hipMemcpyAsync(mat1, om, 8ull * n, hipMemcpyDeviceToHost, st));
        if (n && mat2) HIPC(hipMemcpyAsync(mat2, om + n, 8ull * n, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        return (int)FAME_OK;
    });
}

int fame_match_pairs_device(fame_ctx* c, const uint8_t* d_seqs, const uint64_t* d_offs1, const uint32_t* d_lens1,
                            const uint64_t* d_offs2, const uint32_t* d_lens2, uint32_t n, int get_stranded,
                            uint64_t counters[5], uint8_t* d_valid1, uint64_t* d_mat1, uint8_t* d_valid2,
                            uint64_t* d_mat2, void* stream)
{
    if (!c || !counters || (n && (!d_seqs || !d_offs1 || !d_lens1 || !d_offs2 || !d_lens2))) return FAME_ERR_ARG;
    if ((uint32_t)std::ceil((float)c->readlen * 0.75) < 32) {
        set_err(c, "paired matching needs ceil(0.75 READLEN) >= 32");
        return FAME_ERR_FORMAT;
    }
    if (n > 0x7FFFFFFFu) return FAME_ERR_ARG;
    return guarded(c, [&]() {
        hipStream_t st = stream ? (hipStream_t)stream : c->stream;
        // the 2n mates as one read list: offsets and lengths of mate 1, then
        // of mate 2 (device-to-device copies); both index the same d_seqs
        uint64_t* d_o = grow<uint64_t>(c->b_offs, 2ull * n + 1);
        uint32_t* d_l = grow<uint32_t>(c->b_lens, 2ull * n + 1);
        uint64_t total = 0;
        if (n) {
            HIPC(hipMemcpyAsync(d_o, d_offs1, 8ull * n, hipMemcpyDeviceToDevice, st));
            HIPC(hipMemcpyAsync(d_o + n, d_offs2, 8ull * n, hipMemcpyDeviceToDevice, st));
            HIPC(hipMemcpyAsync(d_l, d_lens1, 4ull * n, hipMemcpyDeviceToDevice, st));
            HIPC(hipMemcpyAsync(d_l + n, d_lens2, 4ull * n, hipMemcpyDeviceToDevice, st));
            unsigned long long* d_end = grow<unsigned long long>(c->b_ovf_total, 1);
            HIPC(hipMemsetAsync(d_end, 0, 8, st));
            launch_max_end(d_o, d_l, 2 * n, d_end, st);
            unsigned long long e = 0;
            HIPC(hipMemcpyAsync(&e, d_end, 8, hipMemcpyDeviceToHost, st));
            HIPC(hipStreamSynchronize(st));
            total = e;
        }
        const uint8_t* ov = nullptr;
        const uint64_t* om = nullptr;
        int rc = match_pairs_batch(c, d_seqs, d_o, d_l, n, total, get_stranded, counters, st, &ov, &om);
        if (rc != FAME_OK) return rc;
        if (n && d_valid1) HIPC(hipMemcpyAsync(d_valid1, ov, n, hipMemcpyDeviceToDevice, st));
        if (n && d_valid2) HIPC(hipMemcpyAsync(d_valid2, ov + n, n, hipMemcpyDeviceToDevice, st));
        if (n && d_mat1) HIPC(hipMemcpyAsync(d_mat1, om, 8ull * n, hipMemcpyDeviceToDevice, st));
        if (n && d_mat2) HIPC(hipMemcpyAsync(d_mat2, om + n, 8ull * n, hipMemcpyDeviceToDevice, st));
        HIPC(hipStreamSynchronize(st));
        return (int)FAME_OK;
    });
}

int fame_set_form(fame_ctx* c, int form)
{
    if (!c || (form != FAME_FORM_PLAIN && form != FAME_FORM_BATCHED)) return FAME_ERR_ARG;
    c->form = form;
    return FAME_OK;
}

int fame_set_pipeline(fame_ctx* c, uint32_t sub_batch_reads)
{
    if (!c || sub_batch_reads < 256) return FAME_ERR_ARG;
    c->pipe_batch = sub_batch_reads;
    return FAME_OK;
}

int fame_form_info(const fame_ctx* c, uint64_t out[FAME_FORM_FIELDS])
{
    if (!c || !out) return FAME_ERR_ARG;
    out[0] = (uint64_t)c->form;
    out[1] = c->graph_calls;
    out[2] = c->graph_fallbacks;
    out[3] = c->graph_captures;
    out[4] = c->plans.size();
    out[5] = 0;
    out[6] = c->pipe_calls;
    out[7] = c->pipe_batch;
    out[8] = 0;
    for (auto& p : c->plans) out[8] += p->graphs.size();
    uint32_t big = 0;
    for (auto& p : c->plans)
        if (p->cap >= big && p->chosen >= 0) {
            big = p->cap;
            out[5] = kVerifyBlocks[p->chosen];
        }
    return FAME_OK;
}

int fame_decide_strand(fame_ctx* c, uint64_t* r1_fwd, uint64_t* r1_rev, int* match_r1_fwd)
{
    if (!c) return FAME_ERR_ARG;
    return guarded(c, [&]() {
        uint64_t t[2];
        HIPC(hipMemcpy(t, c->d_counters.as<uint64_t>() + 3, 16, hipMemcpyDeviceToHost));
        c->match_r1_fwd = t[0] > t[1];
        if (r1_fwd) *r1_fwd = t[0];
        if (r1_rev) *r1_rev = t[1];
        if (match_r1_fwd) *match_r1_fwd = c->match_r1_fwd ? 1 : 0;
        return (int)FAME_OK;
    });
}

int fame_get_meth(fame_ctx* c, uint16_t* out, uint64_t n_cpg)
{
    if (!c || !out || n_cpg != c->ncpg) return FAME_ERR_ARG;
    return guarded(c, [&]() {
        DevBuf tmp;
        HIPC(tmp.reserve(8 * c->ncpg + 8));
        launch_meth_u16(c->d_meth.as<uint32_t>(), 4 * c->ncpg, tmp.as<uint16_t>(), c->stream);
        HIPC(hipMemcpyAsync(out, tmp.p, 8 * c->ncpg, hipMemcpyDeviceToHost, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        return (int)FAME_OK;
    });
}

int fame_reset(fame_ctx* c)
{
    if (!c) return FAME_ERR_ARG;
    return guarded(c, [&]() {
        HIPC(hipMemset(c->d_meth.p, 0, 16 * c->ncpg + 16));
        HIPC(hipMemset(c->d_counters.p, 0, 64));
        c->match_r1_fwd = true;
        return (int)FAME_OK;
    });
}

int fame_print_methylation_levels(fame_ctx* c, const char* basename)
{
    if (!c || !basename) return FAME_ERR_ARG;
    std::vector<uint16_t> m(4 * c->ncpg + 4);
    int rc = fame_get_meth(c, m.data(), c->ncpg);
    if (rc != FAME_OK) return rc;
    std::string path = std::string(basename) + "_cpg.tsv";
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) {
        set_err(c, "cannot write " + path);
        return FAME_ERR_IO;
    }
    std::vector<char> buf(1 << 20);
    setvbuf(f, buf.data(), _IOFBF, buf.size());
    for (uint64_t i = 0; i < c->ncpg; ++i) {
        const std::string& nm = c->chr_name[c->cpg_chrom[i]];
        fprintf(f, "%s\t%u\t%u\t%u\t%u\t%u\n", nm.c_str(), (unsigned)(c->cpg_pos[i] + c->readlen - 2),
                (unsigned)m[4 * i], (unsigned)m[4 * i + 1], (unsigned)m[4 * i + 2], (unsigned)m[4 * i + 3]);
    }
    fclose(f);
    return FAME_OK;
}

// single-cell count matrix (ReadQueue's single-cell constructor
// ReadQueue.cpp:181-216 and printSCMethylationLevels :1894-1921)
struct fame_sc {
    fame_ctx* ctx;
    FILE* f;
    std::vector<char> buf;
};

int fame_sc_open(fame_ctx* c, const char* path, fame_sc** out)
{
    if (!c || !path || !out) return FAME_ERR_ARG;
    *out = nullptr;
    FILE* f = fopen(path, "wb");
    if (!f) {
        set_err(c, std::string("cannot write ") + path);
        return FAME_ERR_IO;
    }
    fame_sc* sc = new fame_sc{c, f, std::vector<char>(1 << 20)};
    setvbuf(f, sc->buf.data(), _IOFBF, sc->buf.size());
    // two header rows: the chromosome and the position of every CpG
    fputs("\nSC_ID\tCount_Type\t", f);
    for (uint64_t i = 0; i < c->ncpg; ++i) {
        fputs(c->chr_name[c->cpg_chrom[i]].c_str(), f);
        fputc('\t', f);
    }
    fputs("\nSC_ID\tCount_Type\t", f);
    for (uint64_t i = 0; i < c->ncpg; ++i) fprintf(f, "%u\t", (unsigned)(c->cpg_pos[i] + c->readlen - 2));
    *out = sc;
    return FAME_OK;
}

int fame_sc_write_cell(fame_sc* sc, const char* sc_id)
{
    if (!sc || !sc_id) return FAME_ERR_ARG;
    fame_ctx* c = sc->ctx;
    std::vector<uint16_t> m(4 * c->ncpg + 4);
    int rc = fame_get_meth(c, m.data(), c->ncpg);
    if (rc != FAME_OK) return rc;
    // the fourth row holds unmethRev under the label the reference prints
    static const char* const label[4] = {"methFwd", "unmethFwd", "methRev", "unmethFwd"};
    for (int k = 0; k < 4; ++k) {
        fprintf(sc->f, "\n%s\t%s\t", sc_id, label[k]);
        for (uint64_t i = 0; i < c->ncpg; ++i) fprintf(sc->f, "%u\t", (unsigned)m[4 * i + k]);
    }
    if (ferror(sc->f)) {
        set_err(c, "single-cell output write failed");
        return FAME_ERR_IO;
    }
    // the counts start again from zero for the next cell
    return guarded(c, [&]() {
        HIPC(hipMemset(c->d_meth.p, 0, 16 * c->ncpg + 16));
        return (int)FAME_OK;
    });
}

int fame_sc_close(fame_sc* sc)
{
    if (!sc) return FAME_OK;
    const int bad = fclose(sc->f) != 0;
    delete sc;
    return bad ? FAME_ERR_IO : FAME_OK;
}

int fame_last_stats(const fame_ctx* c, uint64_t out[FAME_STATS_FIELDS])
{
    if (!c || !out) return FAME_ERR_ARG;
    static_assert(FAME_STATS_FIELDS == FAME_NSTATS, "stats width");
    memcpy(out, c->stats, sizeof c->stats);
    out[7] = c->ix.slot ? 8 : (c->ix.tab_is64 ? 16 : 8); // bytes read per k-mer lookup
    return FAME_OK;
}

int fame_last_timings(const fame_ctx* c, double out[FAME_TIMING_FIELDS])
{
    if (!c || !out) return FAME_ERR_ARG;
    memcpy(out, c->timings, sizeof c->timings);
    return FAME_OK;
}

} // extern "C"
