/* CPU restatement of FAME's single-end read-matching path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY -- the checker for the HIP path, never the thing
 * measured or shipped.  Citations are FischerJo/FAME file:line.
 *
 * Integer types follow the reference exactly (u16 counts that wrap, u32
 * positions that wrap near chromosome starts, size_t / long in the DP
 * indexing), because the reference's results depend on them.
 *
 * Pinning (tests/test_oracle_golden.py, test_oracle_units.py,
 * test_oracle_sort.py):
 *   - whole path: per-read (valid, MATCH) and CpG TSV against the reference
 *     binary built from its own sources (oracle/_ref) on committed fixtures;
 *   - orc_ntps64 / orc_ntps64_rolling: ntHash::NTPS64 run by the reference
 *     harness (direct and rolling), and direct == rolling;
 *   - orc_shiftand_e: ShiftAnd<E> run by the harness (E = 0, 1, 2, 6, both
 *     directions, N letters), ShiftAnd_unittest.cpp known answers (+MISCOUNT),
 *     brute-force bisulfite approximate matching on small inputs;
 *   - orc_dp_align_e: LevenshtDP run by the harness (band 8 with CompiFwd /
 *     CompiRev, bands 0-2), LevenshtDP_unittest.cpp known answers, a textbook
 *     banded Levenshtein for the distance;
 *   - orc_hopscotch_order: tsl::hopscotch_map itself, run by the harness;
 *   - orc_std_sort_counts: the host's libstdc++ std::sort.
 */
#include "fame_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* CONST.h:73-107 (values that no index header carries) */
#define KMERLEN 32u
#define QTHRESH 5u
#define MISCOUNT 2u
#define ADDMIS 4u
#define NERR (MISCOUNT + ADDMIS) /* ShiftAnd<E>, ReadQueue.cpp:532 */
#define DPBAND (2u * MISCOUNT + ADDMIS) /* LevenshtDP<uint16_t,8>, ReadQueue.cpp:3640 */
#define CPGDUMMY 0xFFFFFFFFu
#define MAXW 257 /* patterns up to 16 447 characters (FAME has no limit; longer reads abort the oracle) */

struct orc_cpg {
    uint8_t chrom;
    uint32_t pos;
};
struct orc_win {
    uint32_t startPos;
    uint8_t chrom;
    uint32_t startInd;
    uint32_t endInd;
};
_Static_assert(sizeof(struct orc_cpg) == 8, "CpG record is 8 bytes in the index file");
_Static_assert(sizeof(struct orc_win) == 16, "metaWindow record is 16 bytes in the index file");

struct orc_index {
    uint64_t htab;
    uint32_t readlen, winlen, kmerlen;
    uint64_t kmercutoff;
    uint32_t seedbits;
    uint64_t ncpg, ncpgs;
    struct orc_cpg *cpg, *cpgs;
    uint64_t nchrom;
    uint64_t* clen;
    uint8_t** cseq;
    uint64_t tablen;
    uint64_t* tab;
    uint64_t nkmer;
    uint32_t* kmeta;
    uint32_t* ktmask;
    uint8_t* strand;
    uint64_t nwin;
    struct orc_win* win;
    uint64_t nfilt;
    uint64_t* filt;
    char* chrname[256];
};

struct orc_ctx {
    orc_index* ix;
    uint16_t* meth; /* 4 per CpG */
    uint64_t r1fwd, r1rev;
    int match_r1_fwd;
    int both;
    uint64_t ub;
    uint32_t minpd, maxpd;  /* MINPDIST / MAXPDIST (CONST.h:41-42) */
    uint64_t* filt_sorted;  /* filteredKmers, sorted (paired path) */
};

/* ------------------------------------------------------------------ */
/* index loading: RefGenome::load RefGenome.cpp:230-387                */
/* ------------------------------------------------------------------ */

static int rd(FILE* f, void* p, size_t n)
{
    return n == 0 || fread(p, 1, n, f) == n;
}

static void* xmalloc(size_t n)
{
    void* p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}

void orc_free_index(orc_index* ix)
{
    if (!ix) return;
    free(ix->cpg);
    free(ix->cpgs);
    if (ix->cseq)
        for (uint64_t i = 0; i < ix->nchrom; ++i) free(ix->cseq[i]);
    free(ix->cseq);
    free(ix->clen);
    free(ix->tab);
    free(ix->kmeta);
    free(ix->ktmask);
    free(ix->strand);
    free(ix->win);
    free(ix->filt);
    for (int i = 0; i < 256; ++i) free(ix->chrname[i]);
    free(ix);
}

orc_index* orc_load_index(const char* path, char* err, size_t errlen)
{
    FILE* f = fopen(path, "rb");
    if (!f) {
        snprintf(err, errlen, "cannot open index %s", path);
        return NULL;
    }
    orc_index* ix = calloc(1, sizeof *ix);
    int ok = rd(f, &ix->htab, 8) && rd(f, &ix->readlen, 4) && rd(f, &ix->winlen, 4) &&
             rd(f, &ix->kmerlen, 4) && rd(f, &ix->kmercutoff, 8) && rd(f, &ix->seedbits, 4);
    if (!ok || ix->kmerlen != KMERLEN || ix->htab == 0) {
        snprintf(err, errlen, "bad index header");
        goto fail;
    }
    ok = rd(f, &ix->ncpg, 8);
    ix->cpg = xmalloc(ix->ncpg * 8);
    ok = ok && rd(f, ix->cpg, ix->ncpg * 8) && rd(f, &ix->ncpgs, 8);
    ix->cpgs = xmalloc(ix->ncpgs * 8);
    ok = ok && rd(f, ix->cpgs, ix->ncpgs * 8) && rd(f, &ix->nchrom, 8);
    if (!ok || ix->nchrom > 256) goto trunc;
    ix->clen = calloc(ix->nchrom, 8);
    ix->cseq = calloc(ix->nchrom, sizeof(uint8_t*));
    for (uint64_t i = 0; i < ix->nchrom; ++i) {
        if (!rd(f, &ix->clen[i], 8)) goto trunc;
        ix->cseq[i] = xmalloc(ix->clen[i]);
        if (!rd(f, ix->cseq[i], ix->clen[i])) goto trunc;
    }
    if (!rd(f, &ix->tablen, 8) || ix->tablen != ix->htab + 1) goto trunc;
    ix->tab = xmalloc(ix->tablen * 8);
    if (!rd(f, ix->tab, ix->tablen * 8) || !rd(f, &ix->nkmer, 8)) goto trunc;
    ix->kmeta = xmalloc(ix->nkmer * 4);
    ix->ktmask = xmalloc(ix->nkmer * 4);
    {
        uint32_t* tmp = xmalloc(ix->nkmer * 8);
        if (!rd(f, tmp, ix->nkmer * 8)) {
            free(tmp);
            goto trunc;
        }
        for (uint64_t i = 0; i < ix->nkmer; ++i) {
            ix->kmeta[i] = tmp[2 * i];
            ix->ktmask[i] = tmp[2 * i + 1];
        }
        free(tmp);
    }
    {
        uint64_t n;
        if (!rd(f, &n, 8)) goto trunc;
        if (fseek(f, (long)(n * 8), SEEK_CUR)) goto trunc; /* metaCpGs (unused) */
        if (!rd(f, &n, 8)) goto trunc;
        if (fseek(f, (long)(n * 8), SEEK_CUR)) goto trunc; /* metaStartCpGs (unused) */
    }
    if (!rd(f, &ix->nwin, 8)) goto trunc;
    ix->win = xmalloc(ix->nwin * 16);
    if (!rd(f, ix->win, ix->nwin * 16) || !rd(f, &ix->nfilt, 8)) goto trunc;
    ix->filt = xmalloc(ix->nfilt * 8);
    if (!rd(f, ix->filt, ix->nfilt * 8)) goto trunc;
    {
        uint64_t n;
        if (rd(f, &n, 8)) {
            for (uint64_t i = 0; i < n; ++i) {
                uint8_t id;
                uint64_t sl;
                if (!rd(f, &id, 1) || !rd(f, &sl, 8)) goto trunc;
                char* s = xmalloc(sl + 1);
                if (!rd(f, s, sl)) {
                    free(s);
                    goto trunc;
                }
                s[sl] = 0;
                free(ix->chrname[id]);
                ix->chrname[id] = s;
            }
        }
    }
    fclose(f);
    /* <index>_strands, RefGenome.cpp:414-433 */
    {
        char* sp = xmalloc(strlen(path) + 16);
        sprintf(sp, "%s_strands", path);
        FILE* g = fopen(sp, "rb");
        free(sp);
        uint64_t n;
        if (!g || !rd(g, &n, 8) || n != ix->nkmer) {
            if (g) fclose(g);
            snprintf(err, errlen, "bad or missing strand file");
            orc_free_index(ix);
            return NULL;
        }
        uint8_t* bits = xmalloc((n + 7) / 8);
        int okb = rd(g, bits, (n + 7) / 8);
        fclose(g);
        if (!okb) {
            free(bits);
            snprintf(err, errlen, "truncated strand file");
            orc_free_index(ix);
            return NULL;
        }
        ix->strand = xmalloc(n ? n : 1);
        for (uint64_t i = 0; i < n; ++i) ix->strand[i] = (bits[i >> 3] >> (i & 7)) & 1;
        free(bits);
    }
    return ix;
trunc:
    snprintf(err, errlen, "truncated index file");
fail:
    fclose(f);
    orc_free_index(ix);
    return NULL;
}

void orc_index_info(const orc_index* ix, uint64_t out[11])
{
    out[0] = ix->htab;
    out[1] = ix->readlen;
    out[2] = ix->winlen;
    out[3] = ix->kmerlen;
    out[4] = ix->kmercutoff;
    out[5] = ix->seedbits;
    out[6] = ix->ncpg;
    out[7] = ix->nchrom;
    out[8] = ix->nkmer;
    out[9] = ix->nwin;
    out[10] = ix->nfilt;
}

/* ------------------------------------------------------------------ */
/* spaced-seed ntHash: nthash.hpp:24-27 seeds, :88-121 msTab, :480-498  */
/* ------------------------------------------------------------------ */

/* seedC (0x3193c18562a02b4c) never appears: the reduced table maps C to seedT */
static const uint64_t SEED_A = 0x3c8bfbb395c60474ULL;
static const uint64_t SEED_G = 0x20323ed082572324ULL;
static const uint64_t SEED_T = 0x295549f54be24456ULL;

static inline uint64_t rol64(uint64_t v, int s)
{
    s &= 63;
    return s ? (v << s) | (v >> (64 - s)) : v;
}

/* msTab row for a byte: the reduced (C->T) alphabet table of nthash.hpp:88-121 */
static inline uint64_t red_seed(uint8_t c)
{
    switch (c) {
    case 'A': case 'a': case 4:
        return SEED_A;
    case 'C': case 'c': case 'T': case 't': case 1: case 7:
        return SEED_T;
    case 'G': case 'g': case 3:
        return SEED_G;
    default:
        return 0;
    }
}

/* closed form of NTPS64: XOR over seed positions i of rol(msTab[c_i], 31-i)
 * (vecX[j] == rol(seedX, j) for the tables of nthash.hpp:30-76) */
uint64_t orc_ntps64(const uint8_t* kmer, uint32_t seedbits)
{
    uint64_t s = 0;
    for (unsigned i = 0; i < KMERLEN; ++i)
        if ((seedbits >> (KMERLEN - 1 - i)) & 1) s ^= rol64(red_seed(kmer[i]), (int)(KMERLEN - 1 - i));
    return s;
}

void orc_ntps64_rolling(const uint8_t* seq, uint32_t len, uint32_t seedbits, uint64_t* out)
{
    if (len < KMERLEN) return;
    /* NTPS64 first k-mer nthash.hpp:480-489 */
    uint64_t h = 0, s = 0;
    for (unsigned i = 0; i < KMERLEN; ++i) {
        uint64_t v = rol64(red_seed(seq[i]), (int)(KMERLEN - 1 - i));
        h ^= v;
        if ((seedbits >> (KMERLEN - 1 - i)) & 1) s ^= v;
    }
    out[0] = s;
    /* rolling update nthash.hpp:492-498 */
    for (uint32_t c = 0; c + KMERLEN < len; ++c) {
        h = rol64(h, 1) ^ rol64(red_seed(seq[c]), KMERLEN % 64) ^ red_seed(seq[c + KMERLEN]);
        s = h;
        const uint8_t* k = seq + c + 1;
        for (unsigned i = 0; i < KMERLEN; ++i)
            if (!((seedbits >> (KMERLEN - 1 - i)) & 1)) s ^= rol64(red_seed(k[i]), (int)(KMERLEN - 1 - i));
        out[c + 1] = s;
    }
}

/* ------------------------------------------------------------------ */
/* tsl::hopscotch_map<uint32_t, V, MetaHash> (ReadQueue.h:137-144,       */
/* hopscotch_map.h:76-79, hopscotch_hash.h) -- physical emulation        */
/* ------------------------------------------------------------------ */

#define NBH 62u /* NeighborhoodSize */
#define MAX_PROBES (12u * NBH)

typedef struct {
    uint64_t mask;  /* power_of_two_growth_policy::m_mask */
    uint64_t arr;   /* m_buckets.size() */
    uint64_t* info; /* bit 0 occupied, bit 1 overflow, bits 2.. neighbours */
    uint32_t* key;
    uint32_t* val; /* the value: a u16 count (single-end) or a packed
                      tuple<u8, u8, bool, bool> (paired: c0 | c1 << 8 | f2 << 16 | f3 << 17) */
    uint64_t nelem; /* m_nb_elements (includes the overflow list) */
    uint32_t* okey;
    uint32_t* oval;
    uint64_t nover, ocap;
    uint64_t max_thr, min_thr;
} hsmap;

static uint64_t rup_pow2(uint64_t v)
{
    if (v && !(v & (v - 1))) return v;
    if (!v) return 1;
    --v;
    for (unsigned i = 1; i < 64; i *= 2) v |= v >> i;
    return v + 1;
}

/* hopscotch_hash ctor hopscotch_hash.h:599-628 with max_load_factor(0.9) :1173-1177 */
static void hs_init(hsmap* m, uint64_t bucket_count)
{
    memset(m, 0, sizeof *m);
    if (bucket_count > 0) {
        bucket_count = rup_pow2(bucket_count);
        m->mask = bucket_count - 1;
        m->arr = bucket_count + NBH - 1;
        m->info = calloc(m->arr, 8);
        m->key = calloc(m->arr, 4);
        m->val = calloc(m->arr, 4);
    }
    const float ml = 0.9f;
    const float bc = (float)(m->arr ? m->arr - NBH + 1 : 0);
    m->max_thr = (uint64_t)(bc * ml);
    m->min_thr = (uint64_t)(bc * 0.1f);
}

static void hs_free(hsmap* m)
{
    free(m->info);
    free(m->key);
    free(m->val);
    free(m->okey);
    free(m->oval);
    memset(m, 0, sizeof *m);
}

static inline uint64_t hs_next_count(const hsmap* m) { return (m->mask + 1) * 2; }

static inline void hs_toggle(hsmap* m, uint64_t b, uint64_t i) { m->info[b] ^= 1ULL << (i + 2); }

static void hs_push_over(hsmap* m, uint32_t k, uint32_t v)
{
    if (m->nover == m->ocap) {
        m->ocap = m->ocap ? 2 * m->ocap : 8;
        m->okey = realloc(m->okey, m->ocap * 4);
        m->oval = realloc(m->oval, m->ocap * 4);
    }
    m->okey[m->nover] = k;
    m->oval[m->nover] = v;
    m->nover++;
}

/* find_value_impl hopscotch_hash.h:1589-1605 */
static uint32_t* hs_find(hsmap* m, uint32_t k)
{
    if (!m->arr) return NULL; /* static empty bucket: no neighbours, no overflow */
    uint64_t ib = k & m->mask;
    uint64_t nb = m->info[ib] >> 2, b = ib;
    while (nb) {
        if ((nb & 1) && m->key[b] == k) return &m->val[b];
        ++b;
        nb >>= 1;
    }
    if (m->info[ib] & 2)
        for (uint64_t i = 0; i < m->nover; ++i)
            if (m->okey[i] == k) return &m->oval[i];
    return NULL;
}

static void hs_insert_impl(hsmap* m, uint64_t ib, uint32_t k, uint32_t v);

/* rehash_impl hopscotch_hash.h:1244-1290 (+ rehash :1179-1182) */
static void hs_rehash(hsmap* m, uint64_t count)
{
    uint64_t need = (uint64_t)ceilf((float)m->nelem / 0.9f);
    if (need > count) count = need;
    hsmap n;
    hs_init(&n, count);
    if (m->nover) {
        n.okey = m->okey;
        n.oval = m->oval;
        n.nover = m->nover;
        n.ocap = m->ocap;
        m->okey = NULL;
        m->oval = NULL;
        m->nover = m->ocap = 0;
        n.nelem += n.nover;
        for (uint64_t i = 0; i < n.nover; ++i) n.info[n.okey[i] & n.mask] |= 2;
    }
    for (uint64_t b = 0; b < m->arr; ++b)
        if (m->info[b] & 1) hs_insert_impl(&n, m->key[b] & n.mask, m->key[b], m->val[b]);
    hs_free(m);
    *m = n;
}

/* swap_empty_bucket_closer hopscotch_hash.h:1545-1578 */
static int hs_swap_closer(hsmap* m, uint64_t* e)
{
    const uint64_t start = *e - NBH + 1;
    for (uint64_t tc = start; tc < *e; ++tc) {
        uint64_t nb = m->info[tc] >> 2, ts = tc;
        while (nb != 0 && ts < *e) {
            if (nb & 1) {
                m->key[*e] = m->key[ts];
                m->val[*e] = m->val[ts];
                m->info[*e] |= 1;
                m->info[ts] &= ~1ULL;
                hs_toggle(m, tc, *e - tc);
                hs_toggle(m, tc, ts - tc);
                *e = ts;
                return 1;
            }
            ++ts;
            nb >>= 1;
        }
    }
    return 0;
}

/* will_neighborhood_change_on_rehash hopscotch_hash.h:1466-1486 */
static int hs_will_change(const hsmap* m, uint64_t ib)
{
    const uint64_t mask2 = rup_pow2(hs_next_count(m)) - 1;
    for (uint64_t b = ib; b < m->arr && (b - ib) < NBH; ++b) {
        uint64_t h = m->key[b];
        if ((h & m->mask) != (h & mask2)) return 1;
    }
    return 0;
}

/* insert_impl hopscotch_hash.h:1426-1458 */
static void hs_insert_impl(hsmap* m, uint64_t ib, uint32_t k, uint32_t v)
{
    if (m->nelem - m->nover >= m->max_thr) {
        hs_rehash(m, hs_next_count(m));
        ib = k & m->mask;
    }
    /* find_empty_bucket :1491-1500 */
    uint64_t e = m->arr;
    {
        uint64_t lim = ib + MAX_PROBES < m->arr ? ib + MAX_PROBES : m->arr;
        for (uint64_t b = ib; b < lim; ++b)
            if (!(m->info[b] & 1)) {
                e = b;
                break;
            }
    }
    if (e < m->arr) {
        do {
            if (e - ib < NBH) {
                m->key[e] = k;
                m->val[e] = v;
                m->info[e] |= 1;
                hs_toggle(m, ib, e - ib);
                m->nelem++;
                return;
            }
        } while (hs_swap_closer(m, &e));
    }
    if (m->nelem < m->min_thr || !hs_will_change(m, ib)) {
        hs_push_over(m, k, v);
        m->info[ib] |= 2;
        m->nelem++;
        return;
    }
    hs_rehash(m, hs_next_count(m));
    hs_insert_impl(m, k & m->mask, k, v);
}

/* operator[]: the value of k, inserted value-initialised (0) when absent */
static uint32_t* hs_get(hsmap* m, uint32_t k)
{
    uint32_t* p = hs_find(m, k);
    if (p) return p;
    /* a map without buckets (static empty bucket): insert_impl sees 0 >= 0
     * and rehashes first */
    hs_insert_impl(m, m->arr ? (k & m->mask) : 0, k, 0);
    return hs_find(m, k);
}

/* operator[] followed by ++ on a u16 count */
static void hs_bump_or_insert(hsmap* m, uint32_t k)
{
    uint32_t* p = hs_get(m, k);
    *p = (uint16_t)(*p + 1);
}

/* iteration order: buckets in physical order, then the overflow list
 * (hopscotch_iterator hopscotch_hash.h:562-573, begin() :752-759) */
static int hs_items(const hsmap* m, uint32_t* keys, uint32_t* vals, int cap)
{
    int n = 0;
    for (uint64_t b = 0; b < m->arr; ++b)
        if (m->info[b] & 1) {
            if (n < cap) {
                keys[n] = m->key[b];
                if (vals) vals[n] = m->val[b];
            }
            ++n;
        }
    for (uint64_t i = 0; i < m->nover; ++i) {
        if (n < cap) {
            keys[n] = m->okey[i];
            if (vals) vals[n] = m->oval[i];
        }
        ++n;
    }
    return n;
}

int orc_hopscotch_order(const uint32_t* keys, uint32_t n, uint64_t bucket_count, uint32_t* out, uint32_t* vals,
                        uint64_t* final_bucket_count)
{
    hsmap m;
    hs_init(&m, bucket_count);
    for (uint32_t i = 0; i < n; ++i) hs_bump_or_insert(&m, keys[i]);
    int r = hs_items(&m, out, vals, (int)n);
    /* bucket_count() hopscotch_hash.h:1139-1150 */
    if (final_bucket_count) *final_bucket_count = m.arr ? m.arr - NBH + 1 : 0;
    hs_free(&m);
    return r;
}

/* ------------------------------------------------------------------ */
/* getSeedRefs ReadQueue.cpp:2329-2461                                  */
/* ------------------------------------------------------------------ */

static inline uint64_t hkey(const orc_index* ix, const uint8_t* k)
{
    return orc_ntps64(k, ix->seedbits) % ix->htab;
}

static void seed_refs(const orc_index* ix, const uint8_t* seq, uint32_t L, hsmap* fwd, hsmap* rev,
                      uint32_t* bc_out)
{
    const uint32_t nk = L - KMERLEN; /* rolling steps after the first k-mer */
    uint32_t bucketCount = 0;
    for (uint32_t c = 0; c <= nk; ++c) {
        uint64_t key = hkey(ix, seq + c);
        bucketCount += (uint32_t)(ix->tab[key + 1] - ix->tab[key]);
    }
    bucketCount = bucketCount / 5;
    if (bc_out) *bc_out = bucketCount;
    hs_init(fwd, bucketCount);
    hs_init(rev, bucketCount);
    const uint32_t maxQPos = L - KMERLEN + 1 - QTHRESH;
    for (uint32_t c = 0; c <= nk; ++c) {
        /* k-mer 0 always inserts; k-mer cIdx+1 only while cIdx < maxQPos */
        const int may_insert = (c == 0) || (c - 1 < maxQPos);
        const uint64_t key = hkey(ix, seq + c);
        uint64_t lastId = 0xffffffffffffffffULL;
        int wasFwd = 0;
        for (uint64_t i = ix->tab[key]; i < ix->tab[key + 1]; ++i) {
            const uint32_t metaId = ix->kmeta[i] & 0x7fffffffu;
            const int isFwd = ix->strand[i];
            if (metaId == lastId && isFwd == wasFwd) continue;
            lastId = metaId;
            wasFwd = isFwd;
            hsmap* m = isFwd ? fwd : rev;
            if (may_insert) {
                hs_bump_or_insert(m, metaId);
            } else {
                uint32_t* p = hs_find(m, metaId);
                if (p) *p = (uint16_t)(*p + 1);
            }
        }
    }
}

int orc_seed_maps(orc_ctx* c, const uint8_t* seq, uint32_t len, uint32_t* fwd_keys,
                  uint16_t* fwd_cnt, int* nfwd, uint32_t* rev_keys, uint16_t* rev_cnt, int* nrev,
                  int cap, uint32_t* bucket_count)
{
    if (len < KMERLEN) return -1;
    hsmap f, r;
    seed_refs(c->ix, seq, len, &f, &r, bucket_count);
    uint32_t* v = xmalloc((size_t)(cap > 0 ? cap : 1) * 4);
    *nfwd = hs_items(&f, fwd_keys, v, cap);
    for (int i = 0; i < *nfwd && i < cap; ++i) fwd_cnt[i] = (uint16_t)v[i];
    *nrev = hs_items(&r, rev_keys, v, cap);
    for (int i = 0; i < *nrev && i < cap; ++i) rev_cnt[i] = (uint16_t)v[i];
    free(v);
    hs_free(&f);
    hs_free(&r);
    return 0;
}

/* ------------------------------------------------------------------ */
/* ShiftAnd<6> ShiftAnd.h:121-432                                       */
/* ------------------------------------------------------------------ */

/* E is the error count of ShiftAnd<E>: NERR on the matching path; other
 * values serve the known-answer tests of ShiftAnd_unittest.cpp */
typedef struct {
    int W;
    unsigned E;
    uint64_t plen;
    uint64_t masks[4][MAXW];
    uint64_t acc[MAXW];
    uint64_t act[NERR + 1][MAXW];
} orc_sa;

/* lmap ReadQueue.cpp:73-76: A->0 C->1 G->2 T->3 */
static inline int lmap(uint8_t c)
{
    switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    default: return 3;
    }
}

/* loadBitmasks ShiftAnd.h:345-432 */
static void sa_load(orc_sa* s, const uint8_t* seq, uint32_t L, unsigned E)
{
    s->E = E;
    s->plen = L;
    s->W = (int)(L / 64 + 1);
    for (int i = 0; i < s->W; ++i) s->acc[i] = 0;
    s->acc[L / 64] = 1ULL << (L % 64);
    uint64_t mA = ~0ULL, mC = ~0ULL, mG = ~0ULL, mT = ~0ULL;
    for (uint64_t i = L; i > 0;) {
        mA <<= 1;
        mC <<= 1;
        mG <<= 1;
        mT <<= 1;
        switch (seq[i - 1]) {
        case 'A': mA |= 1; break;
        case 'C': mC |= 1; break;
        case 'G': mG |= 1; break;
        case 'T': mT |= 1; mC |= 1; break; /* bisulfite: pattern T also matches C */
        default: break;
        }
        if (i % 64 == 0) {
            s->masks[0][i / 64] = mA;
            s->masks[1][i / 64] = mC;
            s->masks[2][i / 64] = mG;
            s->masks[3][i / 64] = mT;
            mA = mC = mG = mT = ~0ULL;
        }
        --i;
    }
    s->masks[0][0] = (mA << 1) | 1;
    s->masks[1][0] = (mC << 1) | 1;
    s->masks[2][0] = (mG << 1) | 1;
    s->masks[3][0] = (mT << 1) | 1;
}

/* reset ShiftAnd.h:259-272 */
static inline void sa_reset(orc_sa* s)
{
    for (unsigned i = 0; i <= s->E; ++i) {
        for (int j = s->W - 1; j > 0; --j) s->act[i][j] = 0;
        s->act[i][0] = (1ULL << (i + 1)) - 1;
    }
}

/* queryLetter ShiftAnd.h:276-318 */
static inline void sa_letter(orc_sa* s, uint8_t c)
{
    const uint64_t* mask = s->masks[lmap(c)];
    const int W = s->W;
    for (unsigned i = s->E; i > 0; --i) {
        for (int j = W - 1; j > 0; --j)
            s->act[i][j] = ((s->act[i][j] << 1 | s->act[i][j - 1] >> 63) & mask[j]) | s->act[i - 1][j] |
                           (s->act[i - 1][j] << 1 | s->act[i - 1][j - 1] >> 63);
        s->act[i][0] = ((s->act[i][0] << 1 | 1) & mask[0]) | s->act[i - 1][0] | (s->act[i - 1][0] << 1);
    }
    for (int j = W - 1; j > 0; --j) s->act[0][j] = (s->act[0][j] << 1 | s->act[0][j - 1] >> 63) & mask[j];
    s->act[0][0] = (s->act[0][0] << 1 | 1) & mask[0];
    for (unsigned i = 1; i <= s->E; ++i) {
        for (int j = W - 1; j > 0; --j) s->act[i][j] |= s->act[i - 1][j] << 1 | s->act[i - 1][j - 1] >> 63;
        s->act[i][0] |= s->act[i - 1][0] << 1;
    }
}

/* isMatch ShiftAnd.h:323-341 */
static inline int sa_is_match(const orc_sa* s, uint8_t* err)
{
    for (unsigned i = 0; i <= s->E; ++i)
        for (int j = 0; j < s->W; ++j)
            if (s->act[i][j] & s->acc[j]) {
                *err = (uint8_t)i;
                return 1;
            }
    return 0;
}

typedef struct {
    uint64_t* pos;
    uint8_t* err;
    int n, cap;
} matchvec;

static void mv_push(matchvec* v, uint64_t p, uint8_t e)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 16;
        v->pos = realloc(v->pos, v->cap * 8);
        v->err = realloc(v->err, v->cap);
    }
    v->pos[v->n] = p;
    v->err[v->n] = e;
    v->n++;
}

/* querySeq ShiftAnd.h:130-182 over text[0..tlen); positions are
 * `it - start + MISCOUNT` */
static void sa_query_fwd(orc_sa* s, const uint8_t* text, uint64_t tlen, matchvec* out)
{
    sa_reset(s);
    int was = 0;
    uint8_t prevErrs = MISCOUNT + 1;
    uint64_t ncomp = 0;
    for (uint64_t p = 0; p < tlen; ++p) {
        const uint8_t c = text[p];
        if (c == 'N') {
            sa_reset(s);
            continue;
        }
        sa_letter(s, c);
        ++ncomp;
        if (ncomp < s->plen - s->E) continue;
        uint8_t e;
        if (sa_is_match(s, &e)) {
            if (was) {
                if (e <= prevErrs) {
                    out->pos[out->n - 1] = p + MISCOUNT;
                    out->err[out->n - 1] = e;
                    prevErrs = e;
                }
            } else {
                mv_push(out, p + MISCOUNT, e);
                was = 1;
                prevErrs = e;
            }
        } else {
            was = 0;
        }
    }
}

/* queryRevSeq ShiftAnd.h:186-255: scans text[tlen-1] down to text[0]
 * complementing; the iterator `end` sits one before text[0], so
 * `it - end` = p + 1 and the position is p + 1 + pLen - 1 + MISCOUNT. */
static void sa_query_rev(orc_sa* s, const uint8_t* text, uint64_t tlen, matchvec* out)
{
    sa_reset(s);
    int was = 0;
    uint8_t prevErrs = MISCOUNT + 1;
    uint64_t ncomp = 0;
    for (uint64_t q = tlen; q > 0; --q) {
        const uint64_t p = q - 1;
        uint8_t c;
        switch (text[p]) {
        case 'A': c = 'T'; break;
        case 'C': c = 'G'; break;
        case 'G': c = 'C'; break;
        case 'T': c = 'A'; break;
        case 'N':
            sa_reset(s);
            continue;
        default:
            fprintf(stderr, "[oracle ShiftAnd] unknown text letter\n");
            abort();
        }
        sa_letter(s, c);
        ++ncomp;
        if (ncomp < s->plen - s->E) continue;
        uint8_t e;
        if (sa_is_match(s, &e)) {
            const uint64_t pos = (p + 1) + s->plen - 1 + MISCOUNT;
            if (was) {
                if (e <= prevErrs) {
                    out->pos[out->n - 1] = pos;
                    out->err[out->n - 1] = e;
                    prevErrs = e;
                }
            } else {
                mv_push(out, pos, e);
                was = 1;
                prevErrs = e;
            }
        } else {
            was = 0;
        }
    }
}

int orc_shiftand_e(const uint8_t* pat, uint32_t plen, const uint8_t* text, uint32_t tlen, int rev,
                   unsigned E, uint64_t* pos, uint8_t* err, int cap, int* total)
{
    if (plen == 0 || plen / 64 + 1 > MAXW || E > NERR) return -1;
    orc_sa s;
    sa_load(&s, pat, plen, E);
    matchvec v = {0};
    if (rev)
        sa_query_rev(&s, text, tlen, &v);
    else
        sa_query_fwd(&s, text, tlen, &v);
    int n = v.n < cap ? v.n : cap;
    for (int i = 0; i < n; ++i) {
        pos[i] = v.pos[i];
        err[i] = v.err[i];
    }
    *total = v.n;
    free(v.pos);
    free(v.err);
    return n;
}

int orc_shiftand(const uint8_t* pat, uint32_t plen, const uint8_t* text, uint32_t tlen, int rev,
                 uint64_t* pos, uint8_t* err, int cap, int* total)
{
    return orc_shiftand_e(pat, plen, text, tlen, rev, NERR, pos, err, cap, total);
}

int orc_shiftand_state(const uint8_t* pat, uint32_t plen, unsigned E, uint64_t* masks, uint64_t* acc,
                       uint64_t* act)
{
    if (plen == 0 || plen / 64 + 1 > MAXW || E > NERR) return -1;
    orc_sa s;
    sa_load(&s, pat, plen, E);
    sa_reset(&s);
    for (int m = 0; m < 4; ++m)
        for (int j = 0; j < s.W; ++j) masks[m * s.W + j] = s.masks[m][j];
    for (int j = 0; j < s.W; ++j) acc[j] = s.acc[j];
    for (unsigned e = 0; e <= E; ++e)
        for (int j = 0; j < s.W; ++j) act[e * s.W + j] = s.act[e][j];
    return s.W;
}

/* ------------------------------------------------------------------ */
/* saQuerySeedSetRef ReadQueue.cpp:1928-2160                            */
/* ------------------------------------------------------------------ */

static inline uint64_t mk_match(uint16_t off, uint8_t err, uint64_t fwd, uint64_t start, uint64_t meta)
{
    /* MATCH::constructMatch structs.h:185-188 */
    return (uint64_t)off | ((uint64_t)err << 16) | (fwd << 24) | (start << 25) | (meta << 32);
}

static int sa_query_seed_set(const orc_index* ix, orc_sa* sa, hsmap* fwd, hsmap* rev, uint64_t* mat)
{
    uint8_t multi[NERR + 1] = {0};
    uint64_t uniq[NERR + 1] = {0};
    uint8_t prevChr = 0;
    uint64_t prevOff = 0xffffffffffffffffULL;
    const uint64_t span = (uint64_t)ix->winlen + MISCOUNT + ADDMIS - 1;
    matchvec v = {0};
    int ret = 0;
    for (int pass = 0; pass < 2; ++pass) {
        hsmap* m = pass == 0 ? fwd : rev;
        if (pass == 1) {
            prevChr = 0;
            prevOff = 0xffffffffffffffffULL;
        }
        int nk = hs_items(m, NULL, NULL, 0);
        uint32_t* keys = xmalloc((size_t)nk * 4 + 4);
        uint32_t* cnts = xmalloc((size_t)nk * 4 + 4);
        hs_items(m, keys, cnts, nk);
        for (int ki = 0; ki < nk; ++ki) {
            if (cnts[ki] < QTHRESH) continue; /* q-gram lemma :1950 / :2044 */
            const uint32_t id = keys[ki];
            const struct orc_win* w = &ix->win[id];
            const uint8_t* cs = ix->cseq[w->chrom];
            const uint64_t clen = ix->clen[w->chrom];
            v.n = 0;
            if (pass == 0) {
                /* text [startPos, startPos+span) clipped to the chromosome :1953-1960 */
                uint64_t end = (uint64_t)w->startPos + span;
                if (end > clen) end = clen;
                sa_query_fwd(sa, cs + w->startPos, end - w->startPos, &v);
            } else {
                /* endIt = begin + startPos - 1 (begin if startPos == 0);
                 * startIt = begin + startPos + span, clipped to end-1 :2048-2059 */
                uint64_t e0 = w->startPos > 0 ? (uint64_t)w->startPos - 1 : 0;
                uint64_t s0 = (uint64_t)w->startPos + span;
                if (s0 >= clen) s0 = clen - 1;
                /* scanned positions (e0, s0] */
                sa_query_rev(sa, cs + e0 + 1, s0 > e0 ? s0 - e0 : 0, &v);
            }
            int i = 0;
            if (v.n > 0 && v.pos[0] + w->startPos == prevOff && w->chrom == prevChr) ++i;
            for (; i < v.n; ++i) {
                const uint8_t e = v.err[i];
                if (multi[e]) {
                    const uint64_t m2 = uniq[e];
                    const int isStart = (m2 >> 25) & 1;
                    const int isFwd = (m2 >> 24) & 1;
                    const int same_strand = pass == 0 ? isFwd : !isFwd;
                    if (same_strand && !isStart &&
                        (uint64_t)ix->win[(uint32_t)(m2 >> 32)].startPos + (m2 & 0xffff) ==
                            (uint64_t)w->startPos + v.pos[i])
                        continue;
                    if (!e) {
                        ret = -1;
                        free(keys);
                        free(cnts);
                        goto done;
                    }
                    multi[e] = 2;
                } else {
                    uniq[e] = mk_match((uint16_t)v.pos[i], e, pass == 0 ? 1 : 0, 0, id);
                    multi[e] = 1;
                }
            }
            if (v.n > 0) {
                prevChr = w->chrom;
                prevOff = w->startPos + v.pos[v.n - 1];
            } else {
                prevChr = 0;
                prevOff = 0xffffffffffffffffULL;
            }
        }
        free(keys);
        free(cnts);
    }
    ret = 0;
    for (unsigned i = 0; i <= NERR; ++i) {
        if (multi[i] == 0) continue;
        *mat = uniq[i];
        ret = multi[i] > 1 ? -1 : 1;
        break;
    }
done:
    free(v.pos);
    free(v.err);
    return ret;
}

/* ------------------------------------------------------------------ */
/* LevenshtDP<uint16_t, 8> + BandedMatrix<uint16_t, 9>                  */
/* ------------------------------------------------------------------ */

typedef struct {
    uint16_t* m;
    long rows; /* L + 1 */
    long bmb;  /* BandedMatrix band: LevenshtDP band + 1 */
    int ub;
} bmat;

/* BandedMatrix::operator() BandedMatrix.h:67-75 (including its overlap of
 * cell (8,17) with (9,0) at band 8) */
static inline uint16_t* bm_at(bmat* b, long r, long c)
{
    const long BMB = b->bmb;
    long off = (BMB - r > 0 ? BMB - r : 0) + (c - r + BMB);
    long idx = r * (2 * BMB + 1) + off;
    if (idx < 0 || idx >= b->rows * (2 * BMB + 1)) {
        static uint16_t sink;
        b->ub = 1;
        sink = 0xffff;
        return &sink;
    }
    return &b->m[idx];
}

/* CompiFwd / CompiRev ReadQueue.h:295-338 */
static inline uint16_t cmp_fwd(uint8_t r, uint8_t g)
{
    if (r == 'T' && g == 'C') return 0;
    return r == g ? 0 : 1;
}
static inline uint16_t cmp_rev(uint8_t r, uint8_t g)
{
    switch (r) {
    case 'T': if (g == 'A' || g == 'G') return 0; break;
    case 'G': if (g == 'C') return 0; break;
    case 'A': if (g == 'T') return 0; break;
    case 'C': if (g == 'G') return 0; break;
    }
    return 1;
}

typedef struct {
    const uint8_t* ref;
    uint64_t len;
    int64_t end; /* colPat = ref + end */
    int ub;
} colpat;

/* *(colPat - j + 1) */
static inline uint8_t col_at(colpat* cp, long j)
{
    int64_t p = cp->end - j + 1;
    if (p < 0 || (uint64_t)p >= cp->len) {
        cp->ub = 1;
        return 0;
    }
    return cp->ref[p];
}

static inline int imin3(int a, int b, int c)
{
    int m = a < b ? a : b;
    return m < c ? m : c;
}

/* the comparators of LevenshtDP_unittest.cpp: equality, and Watson-Crick
 * complement (its reverse-query test) */
static inline uint16_t cmp_eq(uint8_t r, uint8_t g) { return r == g ? 0 : 1; }
static inline uint16_t cmp_wc(uint8_t r, uint8_t g)
{
    switch (r) {
    case 'A': return g == 'T' ? 0 : 1;
    case 'C': return g == 'G' ? 0 : 1;
    case 'G': return g == 'C' ? 0 : 1;
    case 'T': return g == 'A' ? 0 : 1;
    }
    return 1;
}

static inline int dp_cmp(int mode, uint8_t r, uint8_t g)
{
    switch (mode) {
    case ORC_CMP_FWD: return cmp_fwd(r, g);
    case ORC_CMP_REV: return cmp_rev(r, g);
    case ORC_CMP_EQ: return cmp_eq(r, g);
    default: return cmp_wc(r, g);
    }
}

int orc_dp_align_e(const uint8_t* row, uint32_t L, const uint8_t* ref, uint64_t ref_len, int64_t ref_end,
                   int rev, unsigned band_, int cmp, uint8_t* trace, int cap, int* ub, int* dist)
{
    const long band = (long)band_;
    const long BMB = band + 1;
    bmat b;
    b.rows = (long)L + 1;
    b.bmb = BMB;
    b.ub = 0;
    b.m = xmalloc((size_t)b.rows * (2 * BMB + 1) * 2);
    for (long i = 0; i < b.rows * (2 * BMB + 1); ++i) b.m[i] = 0xffff;
    colpat cp = {ref, ref_len, ref_end, 0};
    /* runDPFill / runDPFillRev LevenshtDP.h:207-323: border init */
    *bm_at(&b, 0, 0) = 0;
    for (long c = 1; c <= band; ++c) *bm_at(&b, 0, c) = (uint16_t)c;
    for (long r = 1; r <= band; ++r) *bm_at(&b, r, 0) = (uint16_t)r;
    for (long c = 0; c < (long)(L - band); ++c) *bm_at(&b, c + band + 1, c) = (uint16_t)(0xffff - c - band - 1);
    for (long r = 0; r < (long)L; ++r) *bm_at(&b, r, r + band + 1) = (uint16_t)(0xffff - r - band - 1);
    /* fill: LevRec / LevRecRev :165-176, computed in int, stored as u16 */
    for (long r = 1; r <= (long)L; ++r)
        for (long o = -band; o <= band; ++o) {
            if (r + o <= 0) continue;
            const long j = r + o;
            const uint8_t rc = rev ? row[r - 1] : row[L - r];
            const uint8_t gc = col_at(&cp, j);
            const int mm = dp_cmp(cmp, rc, gc);
            const int v = imin3((int)*bm_at(&b, r - 1, j) + 1, (int)*bm_at(&b, r, j - 1) + 1,
                                (int)*bm_at(&b, r - 1, j - 1) + mm);
            *bm_at(&b, r, j) = (uint16_t)v;
        }
    /* backtrackDP / backtrackDPRev :329-544 */
    uint16_t minimum = 0xffff;
    unsigned long col = 0; /* size_t */
    for (long o = -band; o <= band; ++o) {
        uint16_t v = *bm_at(&b, (long)L, (long)L + o);
        if (v < minimum) {
            col = (unsigned long)((long)L + o);
            minimum = v;
        }
    }
    int n = 0;
    for (long r = (long)L; r > 0;) {
        const uint8_t rc = rev ? row[r - 1] : row[L - r];
        const uint8_t gc = col_at(&cp, (long)col);
        const int mm = dp_cmp(cmp, rc, gc);
        const long c = (long)col;
        int t;
        if (*bm_at(&b, r - 1, c) < *bm_at(&b, r, c - 1)) {
            if ((int)*bm_at(&b, r - 1, c) + 1 < (int)*bm_at(&b, r - 1, c - 1) + mm) {
                t = 2; /* INSERTION */
                --r;
            } else {
                t = mm ? 1 : 0;
                --r;
                --col;
            }
        } else {
            if ((int)*bm_at(&b, r, c - 1) + 1 < (int)*bm_at(&b, r - 1, c - 1) + mm) {
                t = 3; /* DELETION */
                --col;
            } else {
                t = mm ? 1 : 0;
                --r;
                --col;
            }
        }
        if (n < cap) trace[n] = (uint8_t)t;
        ++n;
    }
    while (col > 0) {
        if (n < cap) trace[n] = 3;
        ++n;
        --col;
    }
    free(b.m);
    *ub = b.ub | cp.ub;
    if (dist) *dist = minimum;
    return n;
}

int orc_dp_align(const uint8_t* row, uint32_t L, const uint8_t* ref, uint64_t ref_len,
                 int64_t ref_end, int rev, uint8_t* trace, int cap, int* ub)
{
    return orc_dp_align_e(row, L, ref, ref_len, ref_end, rev, DPBAND, rev ? ORC_CMP_REV : ORC_CMP_FWD, trace,
                          cap, ub, NULL);
}

/* ------------------------------------------------------------------ */
/* computeMethLvl ReadQueue.cpp:3278-3947                               */
/* ------------------------------------------------------------------ */

static inline void bump(uint16_t* p) { __atomic_fetch_add(p, (uint16_t)1, __ATOMIC_RELAXED); }

static void compute_meth(orc_ctx* cx, uint64_t mat, uint8_t* seq, uint32_t L)
{
    const orc_index* ix = cx->ix;
    const int isFwd = (mat >> 24) & 1;
    const uint32_t metaID = (uint32_t)(mat >> 32);
    const uint16_t offset = (uint16_t)(mat & 0xffff);
    const uint8_t errNum = (uint8_t)((mat >> 16) & 0xff);
    const uint64_t READLEN = ix->readlen;
    const struct orc_win* m = &ix->win[metaID];
    if (m->startInd == CPGDUMMY) return;
    const uint32_t metaPos = m->startPos;
    /* minPos = metaPos + offset - (seq.size() - 1), size_t arithmetic stored in u32 */
    const uint32_t minPos = (uint32_t)((uint64_t)metaPos + offset - ((uint64_t)L - 1));
    const uint32_t maxPos = metaPos + offset;
    const uint32_t RL2 = (uint32_t)READLEN - 2;
    if (errNum == 0) {
        /* :3355-3484 */
        for (uint32_t id = m->startInd; id <= m->endInd; ++id) {
            const uint32_t pos = ix->cpg[id].pos;
            if (pos + RL2 < minPos) continue;
            if (isFwd) {
                if (pos + RL2 > maxPos) break;
            } else {
                if (pos + (uint32_t)READLEN - 1 > maxPos) break;
            }
            const uint32_t trueCpGPos = pos - 1 + (uint32_t)READLEN - 1;
            const uint32_t readCpGPos = (uint32_t)((uint64_t)L - (uint64_t)(uint32_t)(metaPos + offset - trueCpGPos) - 1);
            uint16_t* lv = &cx->meth[4 * (uint64_t)id];
            if (isFwd) {
                if ((uint64_t)(uint32_t)(readCpGPos + MISCOUNT) >= L) continue;
                const uint8_t ch = seq[(uint32_t)(readCpGPos + MISCOUNT)];
                if (ch == 'T')
                    bump(&lv[1]);
                else if (ch == 'C')
                    bump(&lv[0]);
            } else {
                if ((uint64_t)(uint32_t)(readCpGPos + 1 + MISCOUNT + 2) > L) continue;
                const uint8_t ch = seq[(uint64_t)L - readCpGPos - 1 - MISCOUNT - 2];
                if (ch == 'T')
                    bump(&lv[3]);
                else if (ch == 'C')
                    bump(&lv[2]);
            }
        }
        return;
    }
    /* DP branch :3625-3912 */
    const uint8_t chrom = ix->cpg[m->startInd].chrom;
    int32_t minIndex = (int32_t)m->startInd;
    int32_t maxIndex = (int32_t)m->endInd;
    for (int32_t id = (int32_t)m->startInd; (uint32_t)id <= m->endInd; ++id) {
        const uint32_t pos = ix->cpg[id].pos;
        if (pos + RL2 < minPos) {
            ++minIndex;
        } else if (pos + RL2 > maxPos) {
            maxIndex = id - 1;
            break;
        }
    }
    uint8_t trace_buf[2048];
    int ub = 0;
    const int tn = orc_dp_align(seq, L, ix->cseq[chrom], ix->clen[chrom], (int64_t)metaPos + offset,
                                !isFwd, trace_buf, (int)sizeof trace_buf, &ub);
    if (ub) __atomic_fetch_add(&cx->ub, 1, __ATOMIC_RELAXED);
    uint32_t refSeqPos = metaPos + offset;
    int32_t readSeqPos = (int32_t)L - 1;
    int32_t alignPos = tn - 1;
    if (!isFwd) {
        /* std::reverse(seq) in place :3796 */
        for (uint32_t i = 0, j = L - 1; i < j; ++i, --j) {
            uint8_t t = seq[i];
            seq[i] = seq[j];
            seq[j] = t;
        }
    }
    for (int32_t id = maxIndex; id >= minIndex; --id) {
        const uint32_t pos = ix->cpg[id].pos;
        while (alignPos >= 0 && pos + RL2 < refSeqPos) {
            switch (trace_buf[alignPos]) {
            case 0:
            case 1:
                --readSeqPos;
                --refSeqPos;
                break;
            case 3:
                --refSeqPos;
                break;
            case 2:
                --readSeqPos;
                break;
            }
            if (readSeqPos < 0) break;
            --alignPos;
        }
        if (readSeqPos < 0) break;
        uint16_t* lv = &cx->meth[4 * (uint64_t)id];
        if (isFwd) {
            const uint8_t ch = seq[readSeqPos];
            if (ch == 'C')
                bump(&lv[0]);
            else if (ch == 'T')
                bump(&lv[1]);
        } else {
            const uint8_t ch = seq[readSeqPos + 1];
            if (ch == 'C')
                bump(&lv[2]);
            else if (ch == 'T')
                bump(&lv[3]);
        }
    }
}

/* ------------------------------------------------------------------ */
/* matchReads ReadQueue.cpp:404-933                                     */
/* ------------------------------------------------------------------ */

orc_ctx* orc_ctx_new(orc_index* ix, int both_strands)
{
    orc_ctx* c = calloc(1, sizeof *c);
    c->ix = ix;
    c->meth = calloc(ix->ncpg * 4 + 4, 2);
    c->match_r1_fwd = 1;
    c->both = both_strands;
    c->minpd = 50;
    c->maxpd = 550;
    return c;
}

void orc_ctx_free(orc_ctx* c)
{
    if (!c) return;
    free(c->meth);
    free(c->filt_sorted);
    free(c);
}

/* one read; returns 0 matched, 1 nonunique, 2 unmatched, 3 invalid (no counter) */
static int match_one(orc_ctx* cx, const uint8_t* in, uint32_t L, int getStranded, uint64_t* mat_out)
{
    const orc_index* ix = cx->ix;
    if (L < ix->readlen - 20) return 3; /* :430-435 */
    if (L < KMERLEN) return 3;
    if (L / 64 + 1 > MAXW) {
        fprintf(stderr, "oracle: a read of %u letters is beyond the oracle's ShiftAnd (MAXW %d)\n", L, MAXW);
        abort();
    }
    /* std::string keeps a terminating '\0' at seq[size()], which the rev DP
     * walk reads (seq[readSeqPos+1] with readSeqPos = L-1, :3858) */
    uint8_t* seq = xmalloc(L + 1);
    uint8_t* revSeq = calloc(L + 1, 1); /* resize() zero-fills; unknown chars leave '\0' */
    memcpy(seq, in, L);
    seq[L] = 0;
    int nflag = 0;
    for (uint32_t p = 0, q = L - 1; p < L; ++p, --q) {
        switch (seq[p]) {
        case 'A': revSeq[q] = 'T'; break;
        case 'C': revSeq[q] = 'G'; break;
        case 'G': revSeq[q] = 'C'; break;
        case 'T': revSeq[q] = 'A'; break;
        case 'N': nflag = 1; break;
        default: break; /* reported on stderr by the reference, then kept */
        }
    }
    int res;
    if (nflag) {
        res = 3;
        goto out;
    }
    uint64_t matchFwd = 0, matchRev = 0;
    int sF = 0, sR = 0;
    hsmap f, r;
    orc_sa* sa = xmalloc(sizeof(orc_sa));
    if (cx->both || getStranded || cx->match_r1_fwd) {
        seed_refs(ix, seq, L, &f, &r, NULL);
        sa_load(sa, seq, L, NERR);
        sF = sa_query_seed_set(ix, sa, &f, &r, &matchFwd);
        hs_free(&f);
        hs_free(&r);
    }
    if (cx->both || getStranded || !cx->match_r1_fwd) {
        seed_refs(ix, revSeq, L, &f, &r, NULL);
        sa_load(sa, revSeq, L, NERR);
        sR = sa_query_seed_set(ix, sa, &f, &r, &matchRev);
        hs_free(&f);
        hs_free(&r);
    }
    free(sa);
    /* decision block :630-805 */
    const uint8_t eF = (uint8_t)((matchFwd >> 16) & 0xff), eR = (uint8_t)((matchRev >> 16) & 0xff);
    int useFwd = -1; /* 1 = matchFwd/r.seq, 0 = matchRev/revSeq */
    int tally = 0;   /* +1 fwd tally, -1 rev tally */
    if (sF == 1 && sR == 1) {
        if (eF < eR) {
            useFwd = 1;
            tally = 1;
        } else if (eF > eR) {
            useFwd = 0;
            tally = -1;
        } else {
            const uint32_t mF = (uint32_t)(matchFwd >> 32), mR = (uint32_t)(matchRev >> 32);
            const uint64_t oF = matchFwd & 0xffff, oR = matchRev & 0xffff;
            const int f1 = (matchFwd >> 24) & 1, f2 = (matchRev >> 24) & 1;
            const int s1 = (matchFwd >> 25) & 1, s2 = (matchRev >> 25) & 1;
            uint32_t p1, p2;
            if (s1 && s2) {
                p1 = (uint32_t)oF;
                p2 = (uint32_t)oR;
            } else {
                p1 = (uint32_t)(ix->win[mF].startPos + oF);
                p2 = (uint32_t)(ix->win[mR].startPos + oR);
            }
            if (s1 == s2 && f1 == f2 && p1 == p2) {
                useFwd = 1;
                tally = 1;
            } else {
                res = 1;
                goto out;
            }
        }
    } else if (sF == 1) {
        if (sR == -1 && !(eF < eR)) {
            res = 1;
            goto out;
        }
        useFwd = 1;
        tally = 1;
    } else if (sR == 1) {
        if (sF == -1 && !(eR < eF)) {
            res = 1;
            goto out;
        }
        useFwd = 0;
        tally = -1;
    } else {
        res = (sF == -1 || sR == -1) ? 1 : 2;
        goto out;
    }
    if (getStranded) {
        if (tally > 0)
            __atomic_fetch_add(&cx->r1fwd, 1, __ATOMIC_RELAXED);
        else
            __atomic_fetch_add(&cx->r1rev, 1, __ATOMIC_RELAXED);
    }
    *mat_out = useFwd ? matchFwd : matchRev;
    compute_meth(cx, *mat_out, useFwd ? seq : revSeq, L);
    res = 0;
out:
    free(seq);
    free(revSeq);
    return res;
}

int orc_match_reads(orc_ctx* c, const uint8_t* seqs, const uint64_t* offs, const uint32_t* lens,
                    uint32_t n, int get_stranded, uint64_t counters[3], uint8_t* valid,
                    uint64_t* mat, int nthreads)
{
    uint64_t s = 0, nu = 0, us = 0;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64) reduction(+ : s, nu, us)
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t m = 0;
        int r = match_one(c, seqs + offs[i], lens[i], get_stranded, &m);
        valid[i] = r == 0;
        mat[i] = r == 0 ? m : 0;
        s += r == 0;
        nu += r == 1;
        us += r == 2;
    }
    counters[0] += s;
    counters[1] += nu;
    counters[2] += us;
    return 0;
}

int orc_decide_strand(orc_ctx* c, uint64_t* fwd, uint64_t* rev)
{
    *fwd = c->r1fwd;
    *rev = c->r1rev;
    c->match_r1_fwd = c->r1fwd > c->r1rev;
    return c->match_r1_fwd;
}

void orc_get_meth(const orc_ctx* c, uint16_t* out)
{
    memcpy(out, c->meth, c->ix->ncpg * 8);
}

uint64_t orc_ub_reads(const orc_ctx* c) { return c->ub; }

int orc_write_tsv(const orc_ctx* c, const char* path)
{
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    const orc_index* ix = c->ix;
    for (uint64_t i = 0; i < ix->ncpg; ++i) {
        const char* nm = ix->chrname[ix->cpg[i].chrom];
        const uint16_t* lv = &c->meth[4 * i];
        fprintf(f, "%s\t%u\t%u\t%u\t%u\t%u\n", nm ? nm : "", (unsigned)(ix->cpg[i].pos + ix->readlen - 2),
                lv[0], lv[1], lv[2], lv[3]);
    }
    fclose(f);
    return 0;
}

/* ------------------------------------------------------------------ */
/* paired-end: matchPairedReads ReadQueue.cpp:935-1655                  */
/* ------------------------------------------------------------------ */

/* map values of the paired maps: tuple<u8 c0, u8 c1, bool f2, bool f3> */
#define P_C0(v) ((v) & 0xffu)
#define P_C1(v) (((v) >> 8) & 0xffu)
#define P_F2 (1u << 16)
#define P_F3 (1u << 17)
#define KMERDIST 10
static inline uint32_t p_inc0(uint32_t v) { return (v & ~0xffu) | ((v + 1) & 0xffu); }
static inline uint32_t p_inc1(uint32_t v) { return (v & ~0xff00u) | ((v + 0x100u) & 0xff00u); }

void orc_set_pdist(orc_ctx* c, uint32_t minpd, uint32_t maxpd)
{
    c->minpd = minpd;
    c->maxpd = maxpd;
}

static int cmp_u64(const void* a, const void* b)
{
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}

/* ref.filteredKmers.find(kSeq) with the u32 kSeq widened to u64 */
static int filt_has(const orc_ctx* cx, uint32_t kseq)
{
    const orc_index* ix = cx->ix;
    if (!ix->nfilt) return 0;
    uint64_t k = kseq;
    return bsearch(&k, cx->filt_sorted, ix->nfilt, 8, cmp_u64) != NULL;
}

/* the u32 kSeq (two bits per letter: C,T 3, G 2, others 0; the last 16
 * letters survive) and the C mask of a k-mer, as getSeedRefsFirstRead /
 * SecondRead build them (:2518-2550, :2606-2640) */
static inline uint32_t kseq_code(uint8_t c) { return (c == 'C' || c == 'T') ? 3u : (c == 'G' ? 2u : 0u); }

/* getSeedRefsFirstRead ReadQueue.cpp:2468-2794: counts into maps sized by
 * the bucket count, then the windows with c0 >= QTHRESH copied in that
 * map's iteration order into maps of 20 buckets (outF / outR) */
static void seed_first(const orc_ctx* cx, const uint8_t* seq, uint32_t L, hsmap* outF, hsmap* outR)
{
    const orc_index* ix = cx->ix;
    const uint32_t nk = L - KMERLEN;
    uint32_t bucketCount = 0;
    for (uint32_t c = 0; c <= nk; ++c) {
        const uint64_t key = hkey(ix, seq + c);
        bucketCount += (uint32_t)(ix->tab[key + 1] - ix->tab[key]);
    }
    bucketCount /= 5;
    hsmap f, r;
    hs_init(&f, bucketCount);
    hs_init(&r, bucketCount);
    uint32_t cMask = 0, kSeq = 0;
    for (uint32_t i = 0; i < KMERLEN; ++i) {
        cMask = (cMask << 1) | (seq[i] == 'C');
        kSeq = (kSeq << 2) + kseq_code(seq[i]);
    }
    const uint32_t maxQPos = L - KMERLEN + 1 - QTHRESH;
    for (uint32_t c = 0; c <= nk; ++c) {
        if (c > 0) {
            const uint8_t nx = seq[KMERLEN + c - 1];
            kSeq = (kSeq << 2) + kseq_code(nx);
            cMask = (cMask << 1) | (nx == 'C'); /* & KMERMASK = 0xffffffff */
            if (filt_has(cx, kSeq)) continue;   /* not checked for k-mer 0 */
        }
        const uint64_t key = hkey(ix, seq + c);
        uint64_t lastId = 0xffffffffffffffffULL;
        int wasFwd = 0, wasStart = 0;
        for (uint64_t i = ix->tab[key]; i < ix->tab[key + 1]; ++i) {
            if (cMask & ix->seedbits & ix->ktmask[i]) continue; /* read C where the reference has T */
            const uint32_t metaId = ix->kmeta[i] & 0x7fffffffu;
            const int isFwd = ix->strand[i];
            const int isStart = (ix->kmeta[i] >> 31) & 1;
            if (metaId == lastId && isFwd == wasFwd && isStart == wasStart) continue;
            lastId = metaId;
            wasFwd = isFwd;
            wasStart = isStart;
            hsmap* m = isFwd ? &f : &r;
            if (c == 0 || c - 1 < maxQPos) {
                uint32_t* p = hs_get(m, metaId);
                *p = p_inc0(*p);
            } else {
                uint32_t* p = hs_find(m, metaId);
                if (p) *p = p_inc0(*p);
            }
        }
    }
    for (int pass = 0; pass < 2; ++pass) {
        hsmap* src = pass ? &r : &f;
        hsmap* dst = pass ? outR : outF;
        hs_init(dst, 20);
        const int n = hs_items(src, NULL, NULL, 0);
        uint32_t* k = xmalloc((size_t)n * 4 + 4);
        uint32_t* v = xmalloc((size_t)n * 4 + 4);
        hs_items(src, k, v, n);
        for (int i = 0; i < n; ++i)
            if (P_C0(v[i]) >= QTHRESH) *hs_get(dst, k[i]) = P_C0(v[i]);
        free(k);
        free(v);
    }
    hs_free(&f);
    hs_free(&r);
}

/* getSeedRefsSecondRead ReadQueue.cpp:2801-3164; returns hasCpG */
static int seed_second(const orc_ctx* cx, const uint8_t* seq, uint32_t L, hsmap* F, hsmap* R)
{
    const orc_index* ix = cx->ix;
    const uint32_t nk = L - KMERLEN;
    const int cw = (int)((double)cx->maxpd / ix->winlen) + 1; /* contextWLen */
    uint32_t cMask = 0, kSeq = 0;
    for (uint32_t i = 0; i < KMERLEN; ++i) {
        cMask = (cMask << 1) | (seq[i] == 'C');
        kSeq = (kSeq << 2) + kseq_code(seq[i]);
    }
    const uint32_t maxQPos = L - KMERLEN + 1 - QTHRESH;
    for (uint32_t c = 0; c <= nk; ++c) {
        if (c > 0) {
            const uint8_t nx = seq[KMERLEN + c - 1];
            kSeq = (kSeq << 2) + kseq_code(nx);
            cMask = (cMask << 1) | (nx == 'C');
            if (filt_has(cx, kSeq)) continue;
        }
        const uint64_t key = hkey(ix, seq + c);
        uint64_t lastId = 0xffffffffffffffffULL;
        int wasFwd = 0;
        for (uint64_t i = ix->tab[key]; i < ix->tab[key + 1]; ++i) {
            if (cMask & ix->seedbits & ix->ktmask[i]) continue;
            const uint32_t metaId = ix->kmeta[i] & 0x7fffffffu;
            const int isFwd = ix->strand[i];
            if (metaId == lastId && isFwd == wasFwd) continue; /* isStart not compared here */
            lastId = metaId;
            wasFwd = isFwd;
            hsmap* m = isFwd ? F : R;
            if (c == 0 || c - 1 < maxQPos) {
                /* credit the window when it or a neighbour within contextWLen
                 * holds a first-read window (operator[] may insert it) */
                for (int o = -cw; o <= cw; ++o)
                    if (hs_find(m, metaId + (uint32_t)o)) {
                        uint32_t* p = hs_get(m, metaId);
                        *p = p_inc1(*p);
                        break;
                    }
            } else {
                uint32_t* p = hs_find(m, metaId);
                if (p) *p = p_inc1(*p);
            }
        }
    }
    int has = 0;
    for (int pass = 0; pass < 2; ++pass) {
        hsmap* m = pass ? R : F;
        const int n = hs_items(m, NULL, NULL, 0);
        uint32_t* k = xmalloc((size_t)n * 4 + 4);
        uint32_t* v = xmalloc((size_t)n * 4 + 4);
        hs_items(m, k, v, n);
        for (int i = 0; i < n; ++i) {
            if (P_C1(v[i]) < QTHRESH) continue;
            *hs_find(m, k[i]) |= P_F2;
            for (int u = 1; u <= cw; ++u) {
                uint32_t* p = hs_find(m, k[i] + (uint32_t)u);
                if (p) *p |= P_F2;
                p = hs_find(m, k[i] - (uint32_t)u);
                if (p) *p |= P_F2;
            }
        }
        free(k);
        free(v);
    }
    for (int pass = 0; pass < 2 && !has; ++pass) {
        hsmap* m = pass ? R : F;
        const int n = hs_items(m, NULL, NULL, 0);
        uint32_t* k = xmalloc((size_t)n * 4 + 4);
        uint32_t* v = xmalloc((size_t)n * 4 + 4);
        hs_items(m, k, v, n);
        for (int i = 0; i < n && !has; ++i)
            if ((v[i] & P_F2) && ix->win[k[i]].startInd != CPGDUMMY) has = 1;
        free(k);
        free(v);
    }
    return has;
}

/* std::sort of libstdc++ (GCC 11, bits/stl_algo.h / stl_heap.h): introsort
 * with median-of-three pivots, heap sort past depth 2*floor(log2 n), then
 * a final insertion sort; comparator "count greater" on c0 (cmpMetaFirst)
 * or c1 (cmpMetaSecond), ReadQueue.h:340-351.  The order of equal counts
 * follows the algorithm, so it is reproduced step by step. */
typedef struct {
    uint32_t k, v;
} pe_item;

static inline int pe_less(pe_item a, pe_item b, int sec)
{
    return sec ? P_C1(a.v) > P_C1(b.v) : P_C0(a.v) > P_C0(b.v);
}
static inline void pe_swap(pe_item* a, long i, long j)
{
    pe_item t = a[i];
    a[i] = a[j];
    a[j] = t;
}
static void pe_insertion(pe_item* a, long lo, long hi, int sec)
{
    if (lo == hi) return;
    for (long i = lo + 1; i < hi; ++i) {
        pe_item val = a[i];
        if (pe_less(val, a[lo], sec)) {
            memmove(&a[lo + 1], &a[lo], (size_t)(i - lo) * sizeof *a);
            a[lo] = val;
        } else {
            long j = i;
            while (pe_less(val, a[j - 1], sec)) {
                a[j] = a[j - 1];
                --j;
            }
            a[j] = val;
        }
    }
}
static void pe_unguarded_insertion(pe_item* a, long lo, long hi, int sec)
{
    for (long i = lo; i < hi; ++i) {
        pe_item val = a[i];
        long j = i;
        while (pe_less(val, a[j - 1], sec)) {
            a[j] = a[j - 1];
            --j;
        }
        a[j] = val;
    }
}
static void pe_push_heap(pe_item* a, long first, long hole, long top, pe_item val, int sec)
{
    long parent = (hole - 1) / 2;
    while (hole > top && pe_less(a[first + parent], val, sec)) {
        a[first + hole] = a[first + parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a[first + hole] = val;
}
static void pe_adjust_heap(pe_item* a, long first, long hole, long len, pe_item val, int sec)
{
    const long top = hole;
    long child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (pe_less(a[first + child], a[first + child - 1], sec)) child--;
        a[first + hole] = a[first + child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        a[first + hole] = a[first + child - 1];
        hole = child - 1;
    }
    pe_push_heap(a, first, hole, top, val, sec);
}
static void pe_make_heap(pe_item* a, long first, long last, int sec)
{
    const long len = last - first;
    if (len < 2) return;
    long parent = (len - 2) / 2;
    for (;;) {
        pe_adjust_heap(a, first, parent, len, a[first + parent], sec);
        if (parent == 0) return;
        parent--;
    }
}
static void pe_pop_heap(pe_item* a, long first, long last, long result, int sec)
{
    pe_item v = a[result];
    a[result] = a[first];
    pe_adjust_heap(a, first, 0, last - first, v, sec);
}
static void pe_heap_sort_range(pe_item* a, long first, long last, int sec)
{
    /* __partial_sort(first, last, last): heap_select (no element past the
     * heap) then sort_heap */
    pe_make_heap(a, first, last, sec);
    while (last - first > 1) {
        --last;
        pe_pop_heap(a, first, last, last, sec);
    }
}
static void pe_median_to_first(pe_item* a, long result, long x, long y, long z, int sec)
{
    if (pe_less(a[x], a[y], sec)) {
        if (pe_less(a[y], a[z], sec))
            pe_swap(a, result, y);
        else if (pe_less(a[x], a[z], sec))
            pe_swap(a, result, z);
        else
            pe_swap(a, result, x);
    } else if (pe_less(a[x], a[z], sec))
        pe_swap(a, result, x);
    else if (pe_less(a[y], a[z], sec))
        pe_swap(a, result, z);
    else
        pe_swap(a, result, y);
}
static long pe_partition(pe_item* a, long first, long last, long pivot, int sec)
{
    for (;;) {
        while (pe_less(a[first], a[pivot], sec)) ++first;
        --last;
        while (pe_less(a[pivot], a[last], sec)) --last;
        if (!(first < last)) return first;
        pe_swap(a, first, last);
        ++first;
    }
}
static void pe_introsort(pe_item* a, long first, long last, long depth, int sec)
{
    while (last - first > 16) {
        if (depth == 0) {
            pe_heap_sort_range(a, first, last, sec);
            return;
        }
        --depth;
        const long mid = first + (last - first) / 2;
        pe_median_to_first(a, first, first + 1, mid, last - 1, sec);
        const long cut = pe_partition(a, first + 1, last, first, sec);
        pe_introsort(a, cut, last, depth, sec);
        last = cut;
    }
}
static void pe_sort(pe_item* a, long n, int sec)
{
    if (n == 0) return;
    long lg = 0;
    while ((1L << (lg + 1)) <= n) ++lg;
    pe_introsort(a, 0, n, 2 * lg, sec);
    if (n > 16) {
        pe_insertion(a, 0, 16, sec);
        pe_unguarded_insertion(a, 16, n, sec);
    } else {
        pe_insertion(a, 0, n, sec);
    }
}

void orc_std_sort_counts(uint32_t* keys, uint32_t* vals, uint32_t n, int second)
{
    pe_item* a = xmalloc((size_t)n * sizeof *a + sizeof *a);
    for (uint32_t i = 0; i < n; ++i) {
        a[i].k = keys[i];
        a[i].v = vals[i];
    }
    pe_sort(a, n, second);
    for (uint32_t i = 0; i < n; ++i) {
        keys[i] = a[i].k;
        vals[i] = a[i].v;
    }
    free(a);
}

typedef struct {
    uint64_t* m;
    int n, cap;
} matlist;

static void ml_push(matlist* l, uint64_t m)
{
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 16;
        l->m = realloc(l->m, (size_t)l->cap * 8);
    }
    l->m[l->n++] = m;
}

/* matchFwdFirst / matchRevFirst / matchFwdSecond / matchRevSecond
 * ReadQueue.cpp:4538-4776: one candidate window; returns 0 to stop the
 * walk.  `second` selects the count (c1) and flag (f3) of the second read;
 * the first read also marks f3 on the window and its neighbours in `m`. */
static int pe_match_window(const orc_ctx* cx, orc_sa* sa, pe_item meta, int isFwd, int second, hsmap* m,
                           uint8_t* prevChr, uint64_t* prevOff, matlist* out, int32_t* bmCount, matchvec* v)
{
    const orc_index* ix = cx->ix;
    const int cw = (int)((double)cx->maxpd / ix->winlen) + 1;
    const int32_t cnt = (int32_t)(second ? P_C1(meta.v) : P_C0(meta.v));
    const int32_t lim = *bmCount - KMERDIST > (int32_t)QTHRESH ? *bmCount - KMERDIST : (int32_t)QTHRESH;
    if (cnt < lim) return 0;
    if (!(meta.v & (second ? P_F3 : P_F2))) return 1;
    const struct orc_win* w = &ix->win[meta.k];
    const uint8_t* cs = ix->cseq[w->chrom];
    const uint64_t clen = ix->clen[w->chrom];
    const uint64_t span = (uint64_t)ix->winlen + MISCOUNT + ADDMIS - 1;
    v->n = 0;
    if (isFwd) {
        uint64_t end = (uint64_t)w->startPos + span;
        if (end > clen) end = clen;
        sa_query_fwd(sa, cs + w->startPos, end - w->startPos, v);
    } else {
        /* matchRevFirst: endIt = begin + startPos - 1, or begin when startPos
         * is 0 (:4612-4616); matchRevSecond has no such guard (:4736), where
         * startPos 0 would put endIt one before the chromosome: the scan then
         * also covers position 0 */
        uint64_t s0 = (uint64_t)w->startPos + span;
        if (s0 >= clen) s0 = clen - 1;
        if (second && w->startPos == 0) {
            sa_query_rev(sa, cs, s0 + 1, v);
        } else {
            const uint64_t e0 = w->startPos > 0 ? (uint64_t)w->startPos - 1 : 0;
            sa_query_rev(sa, cs + e0 + 1, s0 > e0 ? s0 - e0 : 0, v);
        }
    }
    int i = 0;
    if (v->n > 0 && v->pos[0] + w->startPos == *prevOff && w->chrom == *prevChr) ++i;
    for (; i < v->n; ++i) ml_push(out, mk_match((uint16_t)v->pos[i], v->err[i], isFwd ? 1 : 0, 0, meta.k));
    if (v->n > 0) {
        if (cnt > *bmCount) *bmCount = cnt;
        if (!second) {
            *hs_find(m, meta.k) |= P_F3;
            for (int u = 1; u <= cw; ++u) {
                uint32_t* p = hs_find(m, meta.k + (uint32_t)u);
                if (p) *p |= P_F3;
                p = hs_find(m, meta.k - (uint32_t)u);
                if (p) *p |= P_F3;
            }
        }
        *prevChr = w->chrom;
        *prevOff = w->startPos + v->pos[v->n - 1];
    } else {
        *prevChr = 0;
        *prevOff = 0xffffffffffffffffULL;
    }
    return 1;
}

/* saQuerySeedSetRefFirst :2167-2243 / saQuerySeedSetRefSecond :2249-2323 */
static void pe_query(const orc_ctx* cx, orc_sa* sa, hsmap* F, hsmap* R, int second, matlist* out)
{
    pe_item* vec[2];
    int nv[2];
    for (int pass = 0; pass < 2; ++pass) {
        hsmap* m = pass ? R : F;
        const int n = hs_items(m, NULL, NULL, 0);
        uint32_t* k = xmalloc((size_t)n * 4 + 4);
        uint32_t* v = xmalloc((size_t)n * 4 + 4);
        hs_items(m, k, v, n);
        vec[pass] = xmalloc((size_t)n * sizeof(pe_item) + sizeof(pe_item));
        for (int i = 0; i < n; ++i) {
            vec[pass][i].k = k[i];
            vec[pass][i].v = v[i];
        }
        free(k);
        free(v);
        nv[pass] = n;
        pe_sort(vec[pass], n, second);
    }
    matchvec mv = {0};
    int32_t bmCount = 0;
    uint8_t prevChr = 0;
    uint64_t prevOff = 0xffffffffffffffffULL;
    /* the strand whose first entry has the larger c0 goes first (ties: rev);
     * both walks compare c0 even for the second read */
    int order[2] = {0, 1}, nord = 0;
    if (nv[0] > 0 && nv[1] > 0) {
        if (P_C0(vec[0][0].v) > P_C0(vec[1][0].v)) {
            order[0] = 0;
            order[1] = 1;
        } else {
            order[0] = 1;
            order[1] = 0;
        }
        nord = 2;
    } else if (nv[0] > 0) {
        order[0] = 0;
        nord = 1;
    } else if (nv[1] > 0) {
        order[0] = 1;
        nord = 1;
    }
    for (int o = 0; o < nord; ++o) {
        const int pass = order[o];
        if (o == 1) {
            prevChr = 0;
            prevOff = 0xffffffffffffffffULL;
        }
        for (int i = 0; i < nv[pass]; ++i)
            if (!pe_match_window(cx, sa, vec[pass][i], pass == 0, second, pass ? R : F, &prevChr, &prevOff, out,
                                 &bmCount, &mv))
                break;
    }
    free(mv.pos);
    free(mv.err);
    free(vec[0]);
    free(vec[1]);
}

/* extractPairedMatch ReadQueue.cpp:3255-3275 */
static int pe_extract(const orc_ctx* cx, uint64_t m1, uint64_t m2)
{
    const orc_index* ix = cx->ix;
    const struct orc_win* w1 = &ix->win[(uint32_t)(m1 >> 32)];
    const struct orc_win* w2 = &ix->win[(uint32_t)(m2 >> 32)];
    if (w1->chrom != w2->chrom) return -1;
    const uint32_t p1 = w1->startPos + (uint32_t)(m1 & 0xffff);
    const uint32_t p2 = w2->startPos + (uint32_t)(m2 & 0xffff);
    const uint32_t d = p2 > p1 ? p2 - p1 : p1 - p2;
    if (d <= cx->maxpd + ix->readlen) return (int)((m1 >> 16) & 0xff) + (int)((m2 >> 16) & 0xff);
    return -1;
}

/* the reverse complement of a mate; returns 0 when the mate holds an N or an
 * unknown letter (the paired path rejects both, :1016-1088) */
static int pe_revcomp(const uint8_t* s, uint32_t L, uint8_t* out)
{
    int ok = 1;
    for (uint32_t p = 0, q = L - 1; p < L; ++p, --q) {
        switch (s[p]) {
        case 'A': out[q] = 'T'; break;
        case 'C': out[q] = 'G'; break;
        case 'G': out[q] = 'C'; break;
        case 'T': out[q] = 'A'; break;
        default: ok = 0; break;
        }
    }
    return ok;
}

/* one pair: counters[0..4] += succ, nonunique, unsucc, succPaired, tooShort */
static void match_pair(orc_ctx* cx, const uint8_t* in1, uint32_t L1, const uint8_t* in2, uint32_t L2,
                       int getStranded, uint64_t cnt[5], uint8_t* v1, uint64_t* mat1, uint8_t* v2,
                       uint64_t* mat2)
{
    const orc_index* ix = cx->ix;
    if (L1 / 64 + 1 > MAXW || L2 / 64 + 1 > MAXW) {
        fprintf(stderr, "oracle: a mate of %u letters is beyond the oracle's ShiftAnd (MAXW %d)\n",
                L1 > L2 ? L1 : L2, MAXW);
        abort();
    }
    const uint32_t minlen = (uint32_t)ceilf((float)ix->readlen * 0.75f);
    int inv1 = L1 < minlen, inv2 = L2 < minlen;
    if (inv1 || inv2) cnt[4] += 1;
    *v1 = *v2 = 1;
    *mat1 = *mat2 = 0;
    uint8_t* s1 = calloc(L1 + 1, 1);
    uint8_t* s2 = calloc(L2 + 1, 1);
    uint8_t* rs1 = calloc(L1 + 1, 1);
    uint8_t* rs2 = calloc(L2 + 1, 1);
    memcpy(s1, in1, L1);
    memcpy(s2, in2, L2);
    if (!pe_revcomp(s1, L1, rs1)) inv1 = 1;
    if (!pe_revcomp(s2, L2, rs2)) inv2 = 1;
    if (inv1 || inv2) {
        cnt[4] += 2;
        *v1 = !inv1;
        *v2 = !inv2;
        goto out;
    }
    /* mates shorter than a k-mer never reach here: 0.75 READLEN > 32 */
    {
        matlist m1F = {0}, m2R = {0}, m1R = {0}, m2F = {0};
        orc_sa* sa = xmalloc(sizeof(orc_sa));
        hsmap F, R;
        const int c1 = cx->both || getStranded || cx->match_r1_fwd;
        const int c2 = cx->both || getStranded || !cx->match_r1_fwd;
        if (c1) {
            seed_first(cx, s1, L1, &F, &R);
            if (F.nelem || R.nelem) {
                if (seed_second(cx, rs2, L2, &F, &R)) {
                    sa_load(sa, s1, L1, NERR);
                    pe_query(cx, sa, &F, &R, 0, &m1F);
                    if (m1F.n) {
                        sa_load(sa, rs2, L2, NERR);
                        pe_query(cx, sa, &F, &R, 1, &m2R);
                    }
                }
            }
            hs_free(&F);
            hs_free(&R);
        }
        if (c2) {
            seed_first(cx, rs1, L1, &F, &R);
            if (F.nelem || R.nelem) {
                if (seed_second(cx, s2, L2, &F, &R)) {
                    sa_load(sa, rs1, L1, NERR);
                    pe_query(cx, sa, &F, &R, 0, &m1R);
                    if (m1R.n) {
                        sa_load(sa, s2, L2, NERR);
                        pe_query(cx, sa, &F, &R, 1, &m2F);
                    }
                }
            }
            hs_free(&F);
            hs_free(&R);
        }
        free(sa);
        if ((m1F.n == 0 && m1R.n == 0) || (m2F.n == 0 && m2R.n == 0)) {
            *v1 = *v2 = 0;
            cnt[2] += 2;
        } else {
            const int worst = 2 * (int)NERR + 1;
            int best = worst, nonUnique = 0, orig = 1;
            uint64_t b1 = 0, b2 = 0;
            for (int pass = 0; pass < 2; ++pass) {
                if (pass == 0 ? !c1 : !c2) continue;
                const matlist* A = pass == 0 ? &m1F : &m1R;
                const matlist* B = pass == 0 ? &m2R : &m2F;
                for (int a = 0; a < A->n; ++a)
                    for (int b = 0; b < B->n; ++b) {
                        const int e = pe_extract(cx, A->m[a], B->m[b]);
                        if (e < 0) continue;
                        if (e == best) {
                            nonUnique = 1;
                        } else if (e < best) {
                            best = e;
                            b1 = A->m[a];
                            b2 = B->m[b];
                            if (pass == 1) orig = 0;
                            nonUnique = 0;
                        }
                    }
            }
            if (best == worst) {
                *v1 = *v2 = 0;
                cnt[2] += 2;
            } else if (nonUnique) {
                *v1 = *v2 = 0;
                cnt[1] += 2;
            } else {
                *mat1 = b1;
                *mat2 = b2;
                if (orig) {
                    if (getStranded) __atomic_fetch_add(&cx->r1fwd, 1, __ATOMIC_RELAXED);
                    compute_meth(cx, b1, s1, L1);
                    compute_meth(cx, b2, rs2, L2);
                } else {
                    if (getStranded) __atomic_fetch_add(&cx->r1rev, 1, __ATOMIC_RELAXED);
                    compute_meth(cx, b1, rs1, L1);
                    compute_meth(cx, b2, s2, L2);
                }
                const int d1 = ix->win[(uint32_t)(b1 >> 32)].startInd == CPGDUMMY;
                const int d2 = ix->win[(uint32_t)(b2 >> 32)].startInd == CPGDUMMY;
                if (d1 && d2) {
                    cnt[4] += 2;
                } else if (d1 || d2) {
                    cnt[4] += 1;
                    cnt[0] += 1;
                } else {
                    cnt[3] += 1;
                    cnt[0] += 2;
                }
            }
        }
        free(m1F.m);
        free(m2R.m);
        free(m1R.m);
        free(m2F.m);
    }
out:
    free(s1);
    free(s2);
    free(rs1);
    free(rs2);
}

int orc_match_pairs(orc_ctx* c, const uint8_t* seqs1, const uint64_t* offs1, const uint32_t* lens1,
                    const uint8_t* seqs2, const uint64_t* offs2, const uint32_t* lens2, uint32_t n,
                    int get_stranded, uint64_t counters[5], uint8_t* valid1, uint64_t* mat1,
                    uint8_t* valid2, uint64_t* mat2, int nthreads)
{
    if (c->ix->nfilt && !c->filt_sorted) {
        c->filt_sorted = xmalloc(c->ix->nfilt * 8);
        memcpy(c->filt_sorted, c->ix->filt, c->ix->nfilt * 8);
        qsort(c->filt_sorted, c->ix->nfilt, 8, cmp_u64);
    }
    uint64_t a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 32) reduction(+ : a0, a1, a2, a3, a4)
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t cnt[5] = {0, 0, 0, 0, 0};
        match_pair(c, seqs1 + offs1[i], lens1[i], seqs2 + offs2[i], lens2[i], get_stranded, cnt, &valid1[i],
                   &mat1[i], &valid2[i], &mat2[i]);
        a0 += cnt[0];
        a1 += cnt[1];
        a2 += cnt[2];
        a3 += cnt[3];
        a4 += cnt[4];
    }
    counters[0] += a0;
    counters[1] += a1;
    counters[2] += a2;
    counters[3] += a3;
    counters[4] += a4;
    return 0;
}
