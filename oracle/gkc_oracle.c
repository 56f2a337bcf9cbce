/*
 * gkc_oracle.c -- CPU oracle for GKC's k-mer counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against (tests/, __graft_entry__.smoke(), and bench.py's CPU
 * baseline leg).  The product library never links, loads or calls it.
 *
 * It restates, stage by stage, the semantics of the reference path:
 *   encode        gkc_cuda.cu:25-93 (d_basemap / d_basemap_compl)
 *   HPC           gkc_cuda.cu:164-192 (GPU_HPCEncoding; the per-read index
 *                 uses the read id, fixing the latent tid/rid slip at :180)
 *   filter        gkc_cuda.cu:196-216 (mm_filter, new_filter2)
 *   minimizer     gkc_cuda.cu:234-292 (GPU_GenMinimizer, O(k) per window,
 *                 forward and reverse-complement p-mers)
 *   superkmers    gkc_cuda.cu:298-324 (GPU_GenSKM offset lists)
 *   partition     superkmers.hpp:167-186 (SKMStore::add_skms, mm % P)
 *   packing       superkmers.hpp:63-71, 105-128 (CRead 2-bit MSB-first,
 *                 get_2bit, get_2bit_kmer_V2)
 *   counting      kmer_counting.cu:33-44 (canonicalkmer), :95-113 (rolling
 *                 extraction), :134 (ascending sort), :148-191 (run-length,
 *                 clamp to MAX_KMER_CNT, intended [min,max] filter)
 * Reads shorter than k are skipped (the reference reads out of bounds for
 * them).  A byte other than A C G T (either case) splits its read, after
 * HPC: only windows of k such bases are k-mers (the reference codes such
 * bytes 255 in the minimizer and as T in CRead, gkc_cuda.cu:25-58,
 * superkmers.hpp:67; this build counts no k-mer that holds one, as KMC
 * does).  k <= 64 is supported with unsigned __int128 keys, p <= 32 with
 * 64-bit minimizer values; k <= 32 / p <= 16 is the reference's own range.
 * The per-window (minimizer, position, strand) output has no reference
 * implementation (only declared fields, gkc_cuda.cu:159-160); its rule is
 * the reference's tie-break order (later candidate wins on ties, forward
 * before reverse complement within a p-mer, gkc_cuda.cu:281-284).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* gkc_cuda.cu:25-58: A/a->0, C/c->1, G/g->2, T/t->3, else 255 */
static uint8_t BASEMAP[256];
/* gkc_cuda.cu:60-93: complement codes A->3, C->2, G->1, T->0 */
static uint8_t BASEMAP_C[256];
static int tables_ready = 0;
static void init_tables(void) {
    if (tables_ready) return;
    memset(BASEMAP, 255, 256); memset(BASEMAP_C, 255, 256);
    const char *u = "ACGT", *l = "acgt";
    for (int i = 0; i < 4; i++) {
        BASEMAP[(uint8_t)u[i]] = BASEMAP[(uint8_t)l[i]] = (uint8_t)i;
        BASEMAP_C[(uint8_t)u[i]] = BASEMAP_C[(uint8_t)l[i]] = (uint8_t)(3 - i);
    }
    tables_ready = 1;
}

uint8_t or_code(uint8_t c) { init_tables(); return BASEMAP[c]; }
uint8_t or_code_compl(uint8_t c) { init_tables(); return BASEMAP_C[c]; }

/* Filters. mode 0: mm_filter (accepts all, gkc_cuda.cu:196-209);
 * mode 1: new_filter2 (gkc_cuda.cu:214-216): reject p-mers whose first
 * three bases are AAA (0), ACA (4) or AGA (8). */
int or_filter(uint64_t v, int p, int mode) {
    if (mode == 0) return 1;
    uint64_t top = v >> (2 * (p - 3));
    return (top != 0) & (top != 4) & (top != 8);
}

/* Partition of a minimizer. mode 0: mm % P (superkmers.hpp:180);
 * mode 1: a 64-bit mix (murmur3 fmix64) of mm, then % P. */
uint32_t or_partition(uint64_t mm, uint32_t P, int mode) {
    if (mode == 1) {
        mm ^= mm >> 33; mm *= 0xff51afd7ed558ccdULL;
        mm ^= mm >> 33; mm *= 0xc4ceb9fe1a85ec53ULL;
        mm ^= mm >> 33;
    }
    return (uint32_t)(mm % P);
}

/* HPC, gkc_cuda.cu:178-190: collapse runs of identical bytes in place at the
 * read's own CSR start; orig_pos[start+h] = in-read index of the first byte
 * of run h; read_len[r] = number of runs.  offs are absolute, outputs are
 * batch-local (index 0 = offs[0]). */
void or_hpc(const uint8_t *reads, const uint64_t *offs, int64_t n,
            uint8_t *out, int32_t *read_len, int32_t *orig_pos) {
    const uint64_t base = offs[0];
    for (int64_t r = 0; r < n; r++) {
        uint64_t s = offs[r] - base, L = offs[r + 1] - offs[r];
        if (L == 0) { read_len[r] = 0; continue; }
        uint64_t h = 0;
        out[s] = reads[s]; orig_pos[s] = 0;
        for (uint64_t i = 1; i < L; i++) {
            if (reads[s + i] != reads[s + i - 1]) {
                h++;
                out[s + h] = reads[s + i];
                orig_pos[s + h] = (int32_t)i;
            }
        }
        read_len[r] = (int32_t)(h + 1);
    }
}

/* Window minimizer of window i of one read, gkc_cuda.cu:253-286: the minimum
 * of {4^p-1} and every forward / reverse-complement p-mer inside the k-mer
 * that passes the filter.  Also returns the position (offset in the window)
 * and strand of the candidate that the reference's update order selects
 * (ties go to the later candidate; 255/-1 when no candidate passes). */
static uint64_t window_mm(const uint8_t *s, int k, int p, int fmode,
                          uint8_t *pos, int8_t *strand) {
    const uint64_t M = (p == 32) ? ~0ULL : ((1ULL << (2 * p)) - 1);
    uint64_t best = M;
    int bpos = 255, bstr = -1;
    for (int j = 0; j + p <= k; j++) {
        uint64_t f = 0, rc = 0;
        for (int t = 0; t < p; t++) f = (f << 2) | BASEMAP[s[j + t]];
        for (int t = p - 1; t >= 0; t--) rc = (rc << 2) | BASEMAP_C[s[j + t]];
        if (or_filter(f, p, fmode) && f <= best) { best = f; bpos = j; bstr = 0; }
        if (or_filter(rc, p, fmode) && rc <= best) { best = rc; bpos = j; bstr = (rc == f) ? -1 : 1; }
    }
    if (pos) *pos = (uint8_t)bpos;
    if (strand) *strand = (int8_t)bstr;
    return best;
}

/* Length of the run of A/C/G/T bytes starting at s[0] (at most n). */
static int64_t acgt_run(const uint8_t *s, int64_t n) {
    int64_t i = 0;
    while (i < n && BASEMAP[s[i]] != 255) i++;
    return i;
}

/* Per-window minimizers for a batch (GPU_GenMinimizer).  lens[r] is the
 * (possibly HPC) length; reads of batch-local read r start at offs[r]-offs[0].
 * mm[start + i] for every window i of the read whose k bases are A/C/G/T;
 * the other entries are not written.  pos/strand optional. */
void or_window_minimizers(const uint8_t *reads, const uint64_t *offs, const int32_t *lens,
                          int64_t n, int k, int p, int fmode,
                          uint64_t *mm, uint8_t *pos, int8_t *strand) {
    init_tables();
    const uint64_t base = offs[0];
    for (int64_t r = 0; r < n; r++) {
        uint64_t s = offs[r] - base;
        int64_t L = lens ? lens[r] : (int64_t)(offs[r + 1] - offs[r]);
        for (int64_t a = 0; a < L;) {                   /* ACGT fragments [a, a + f) */
            const int64_t f = acgt_run(reads + s + a, L - a);
            for (int64_t i = a; i + k <= a + f; i++)
                mm[s + i] = window_mm(reads + s + i, k, p, fmode,
                                      pos ? pos + s + i : NULL, strand ? strand + s + i : NULL);
            a += f + 1;
        }
    }
}

/* Superkmer offset lists, GPU_GenSKM gkc_cuda.cu:308-322: at each read's
 * batch-local start: [0, i1, i2, ..., len-k+1, 0] where the i are the window
 * starts whose minimizer differs from the previous window's.  Reads shorter
 * than k get an empty list [0, 0]. */
void or_superkmer_offs(const uint64_t *offs, const int32_t *lens, int64_t n, int k,
                       const uint64_t *mm, int32_t *skm) {
    const uint64_t base = offs[0];
    for (int64_t r = 0; r < n; r++) {
        uint64_t s = offs[r] - base;
        int64_t L = lens ? lens[r] : (int64_t)(offs[r + 1] - offs[r]);
        if (L < k) { if (L >= 1) skm[s] = 0; if (L >= 2) skm[s + 1] = 0; continue; }
        int64_t c = 0;
        skm[s] = 0;
        for (int64_t i = 1; i + k <= L; i++)
            if (mm[s + i] != mm[s + i - 1]) skm[s + (++c)] = (int32_t)i;
        skm[s + c + 1] = (int32_t)(L - k + 1);
        if (c + 2 < L) skm[s + c + 2] = 0;
    }
}

/* ---------------- CRead packing and k-mer extraction ---------------- */

/* superkmers.hpp:63-71: 2-bit MSB-first, 4 bases per byte, tail shifted up. */
static void cread_pack(const uint8_t *raw, int len, uint8_t *out) {
    int nb = (len + 3) / 4;
    memset(out, 0, (size_t)nb);
    for (int i = 0; i < len; i++) out[i / 4] = (uint8_t)((out[i / 4] << 2) | (BASEMAP[raw[i]] & 3));
    if (len % 4) out[len / 4] = (uint8_t)(out[len / 4] << (2 * (4 - len % 4)));
}
/* superkmers.hpp:105-107 */
static inline unsigned cread_get2(const uint8_t *c, int idx) { return (c[idx / 4] >> (2 * (3 - idx % 4))) & 3; }
/* superkmers.hpp:115-128 (get_2bit_kmer_V2 for beg = 0), widened to 128 bits */
static u128 cread_first_kmer(const uint8_t *c, int k) {
    u128 res = 0; int i;
    for (i = 0; i <= k - 4; i += 4) res = (res << 8) | c[i / 4];
    for (i = (i >> 2) << 2; i < k; i++) res = (res << 2) | cread_get2(c, i);
    return res;
}

/* kmer_counting.cu:33-44: rc = 2-bit groups of ~x in reverse order over k
 * groups; canonical = min(x, rc). */
static inline uint64_t canon64(uint64_t x, int k) {
    uint64_t x1 = ~x, res = 0;
    for (int i = 0; i < k; i++) { res = (res << 2) | (x1 & 3); x1 >>= 2; }
    return res < x ? res : x;
}
static inline u128 canon128(u128 x, int k) {
    u128 x1 = ~x, res = 0;
    for (int i = 0; i < k; i++) { res = (res << 2) | (x1 & 3); x1 >>= 2; }
    return res < x ? res : x;
}

uint64_t or_canonical64(uint64_t x, int k) { return canon64(x, k); }

/* ---------------- sorting (ascending, kmer_counting.cu:134) ---------------- */

static void radix_sort64(uint64_t *a, uint64_t *tmp, size_t n, int bits) {
    const int D = 11, R = 1 << D;
    size_t *cnt = (size_t *)malloc(sizeof(size_t) * R);
    uint64_t *src = a, *dst = tmp;
    for (int sh = 0; sh < bits; sh += D) {
        memset(cnt, 0, sizeof(size_t) * R);
        for (size_t i = 0; i < n; i++) cnt[(src[i] >> sh) & (R - 1)]++;
        size_t s = 0;
        for (int d = 0; d < R; d++) { size_t c = cnt[d]; cnt[d] = s; s += c; }
        for (size_t i = 0; i < n; i++) dst[cnt[(src[i] >> sh) & (R - 1)]++] = src[i];
        uint64_t *t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, n * sizeof(uint64_t));
    free(cnt);
}
static void radix_sort128(u128 *a, u128 *tmp, size_t n, int bits) {
    const int D = 11, R = 1 << D;
    size_t *cnt = (size_t *)malloc(sizeof(size_t) * R);
    u128 *src = a, *dst = tmp;
    for (int sh = 0; sh < bits; sh += D) {
        memset(cnt, 0, sizeof(size_t) * R);
        for (size_t i = 0; i < n; i++) cnt[(unsigned)(src[i] >> sh) & (R - 1)]++;
        size_t s = 0;
        for (int d = 0; d < R; d++) { size_t c = cnt[d]; cnt[d] = s; s += c; }
        for (size_t i = 0; i < n; i++) dst[cnt[(unsigned)(src[i] >> sh) & (R - 1)]++] = src[i];
        u128 *t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, n * sizeof(u128));
    free(cnt);
}

/* ---------------- full pipeline ---------------- */

typedef struct { uint64_t off; uint32_t len; } skm_ref;      /* batch-local start, bases */
typedef struct { skm_ref *v; size_t n, cap; } skm_vec;
static void vec_push(skm_vec *v, skm_ref x) {
    if (v->n == v->cap) { v->cap = v->cap ? v->cap * 2 : 64; v->v = (skm_ref *)realloc(v->v, v->cap * sizeof(skm_ref)); }
    v->v[v->n++] = x;
}

typedef struct {
    const uint8_t *reads; const uint64_t *offs; const int32_t *lens;
    int64_t r0, r1; int k, p, fmode, pmode; uint32_t P;
    skm_vec *bins;         /* [P] thread-local bins */
} p1_job;

/* Phase 1 for a read range: minimizers per window, superkmer boundaries and
 * the partition scatter (add_skms).  Works read by read so the whole-batch
 * minimizer array is never materialised. */
static void *phase1_worker(void *arg) {
    p1_job *j = (p1_job *)arg;
    const uint64_t base = j->offs[0];
    int64_t maxL = 0;
    for (int64_t r = j->r0; r < j->r1; r++) {
        int64_t L = j->lens ? j->lens[r] : (int64_t)(j->offs[r + 1] - j->offs[r]);
        if (L > maxL) maxL = L;
    }
    uint64_t *mm = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(maxL + 1));
    for (int64_t r = j->r0; r < j->r1; r++) {
        uint64_t s = j->offs[r] - base;
        int64_t L = j->lens ? j->lens[r] : (int64_t)(j->offs[r + 1] - j->offs[r]);
        for (int64_t a = 0; a < L;) {                   /* each ACGT fragment as a read */
            const int64_t f = acgt_run(j->reads + s + a, L - a);
            const uint64_t fs = s + (uint64_t)a;
            a += f + 1;
            if (f < j->k) continue;
            int64_t W = f - j->k + 1;
            for (int64_t i = 0; i < W; i++) mm[i] = window_mm(j->reads + fs + i, j->k, j->p, j->fmode, NULL, NULL);
            int64_t beg = 0;
            for (int64_t i = 1; i <= W; i++) {
                if (i == W || mm[i] != mm[i - 1]) {
                    /* superkmer over windows [beg, i): bases [beg, i + k - 1) */
                    uint32_t part = or_partition(mm[beg], j->P, j->pmode);
                    skm_ref ref = {fs + (uint64_t)beg, (uint32_t)(i - beg + j->k - 1)};
                    vec_push(&j->bins[part], ref);
                    beg = i;
                }
            }
        }
    }
    free(mm);
    return NULL;
}

typedef struct {
    const uint8_t *reads; int k; uint32_t P; int nthreads;
    skm_vec *bins;                 /* [nthreads][P] */
    uint32_t min_cnt, max_cnt;
    /* outputs */
    uint64_t *key_lo, *key_hi; uint32_t *cnt; uint64_t *part_off;
    uint64_t *part_kmers, *part_distinct, *part_skms;
    uint64_t table_cap;
    /* per-partition scratch placement: output is written at part_off[p]
     * which is only known after all partitions are counted, so each
     * partition first writes into its own region of width part_kmers[p] */
    uint64_t *scratch_off;
    uint32_t next;                 /* work counter */
    pthread_mutex_t mu;
} p2_ctx;

static void count_partition(p2_ctx *c, uint32_t part) {
    const int k = c->k;
    size_t nk = 0, nskm = 0, maxlen = 0;
    for (int t = 0; t < c->nthreads; t++) {
        skm_vec *v = &c->bins[(size_t)t * c->P + part];
        nskm += v->n;
        for (size_t i = 0; i < v->n; i++) {
            nk += v->v[i].len - k + 1;
            if (v->v[i].len > maxlen) maxlen = v->v[i].len;
        }
    }
    c->part_kmers[part] = nk; c->part_skms[part] = nskm;
    uint64_t dst = c->scratch_off[part];
    if (nk == 0) { c->part_distinct[part] = 0; c->part_off[part + 1] = 0; return; }
    uint8_t *packed = (uint8_t *)malloc(maxlen / 4 + 2);
    size_t w = 0, distinct = 0, kept = 0;
    if (k <= 32) {
        const uint64_t mask = (k == 32) ? ~0ULL : ((1ULL << (2 * k)) - 1);
        uint64_t *keys = (uint64_t *)malloc(nk * 8), *tmp = (uint64_t *)malloc(nk * 8);
        for (int t = 0; t < c->nthreads; t++) {
            skm_vec *v = &c->bins[(size_t)t * c->P + part];
            for (size_t i = 0; i < v->n; i++) {
                int len = (int)v->v[i].len;
                cread_pack(c->reads + v->v[i].off, len, packed);
                uint64_t x = (uint64_t)cread_first_kmer(packed, k) & mask;
                keys[w++] = canon64(x, k);
                for (int jj = k; jj < len; jj++) {
                    x = ((x << 2) | cread_get2(packed, jj)) & mask;
                    keys[w++] = canon64(x, k);
                }
            }
        }
        radix_sort64(keys, tmp, nk, 2 * k);
        size_t i = 0;
        while (i < nk) {
            size_t e = i + 1;
            while (e < nk && keys[e] == keys[i]) e++;
            uint64_t run = e - i;
            uint32_t cc = run > 0xFFFFFFFFULL ? 0xFFFFFFFFu : (uint32_t)run;
            distinct++;
            if (cc >= c->min_cnt && (c->max_cnt == 0 || cc <= c->max_cnt)) {
                c->key_lo[dst + kept] = keys[i];
                if (c->key_hi) c->key_hi[dst + kept] = 0;
                c->cnt[dst + kept] = cc; kept++;
            }
            i = e;
        }
        free(keys); free(tmp);
    } else {
        const u128 mask = (k == 64) ? ~(u128)0 : (((u128)1 << (2 * k)) - 1);
        u128 *keys = (u128 *)malloc(nk * 16), *tmp = (u128 *)malloc(nk * 16);
        for (int t = 0; t < c->nthreads; t++) {
            skm_vec *v = &c->bins[(size_t)t * c->P + part];
            for (size_t i = 0; i < v->n; i++) {
                int len = (int)v->v[i].len;
                cread_pack(c->reads + v->v[i].off, len, packed);
                u128 x = cread_first_kmer(packed, k) & mask;
                keys[w++] = canon128(x, k);
                for (int jj = k; jj < len; jj++) {
                    x = ((x << 2) | cread_get2(packed, jj)) & mask;
                    keys[w++] = canon128(x, k);
                }
            }
        }
        radix_sort128(keys, tmp, nk, 2 * k);
        size_t i = 0;
        while (i < nk) {
            size_t e = i + 1;
            while (e < nk && keys[e] == keys[i]) e++;
            uint64_t run = e - i;
            uint32_t cc = run > 0xFFFFFFFFULL ? 0xFFFFFFFFu : (uint32_t)run;
            distinct++;
            if (cc >= c->min_cnt && (c->max_cnt == 0 || cc <= c->max_cnt)) {
                c->key_lo[dst + kept] = (uint64_t)keys[i];
                c->key_hi[dst + kept] = (uint64_t)(keys[i] >> 64);
                c->cnt[dst + kept] = cc; kept++;
            }
            i = e;
        }
        free(keys); free(tmp);
    }
    free(packed);
    c->part_distinct[part] = distinct;
    c->part_off[part + 1] = kept;     /* count; turned into offsets later */
}

static void *phase2_worker(void *arg) {
    p2_ctx *c = (p2_ctx *)arg;
    for (;;) {
        pthread_mutex_lock(&c->mu);
        uint32_t part = c->next++;
        pthread_mutex_unlock(&c->mu);
        if (part >= c->P) break;
        count_partition(c, part);
    }
    return NULL;
}

/*
 * End-to-end count.  reads/offs: a CSR batch (absolute offsets).  When hpc
 * is set the reads are HPC-compressed first (into an internal copy).
 * Outputs: per-partition tables concatenated in partition order
 * (key_lo/key_hi/cnt, key_hi may be NULL for k <= 32), part_off[P+1],
 * per-partition k-mer totals, distinct counts (before the count filter) and
 * superkmer counts.  table_cap must be >= the number of k-mer occurrences.
 * Returns the number of table entries, or -1 on bad arguments.
 * Timing hooks: t_phase[0..1] receive phase-1 / phase-2 wall seconds.
 */
#include <time.h>
static double now_s(void) { struct timespec ts; clock_gettime(CLOCK_MONOTONIC, &ts); return ts.tv_sec + ts.tv_nsec * 1e-9; }

int64_t or_count(const uint8_t *reads, const uint64_t *offs, int64_t n,
                 int k, int p, int hpc, uint32_t P, int fmode, int pmode,
                 uint32_t min_cnt, uint32_t max_cnt, int nthreads,
                 uint64_t *key_lo, uint64_t *key_hi, uint32_t *cnt, uint64_t table_cap,
                 uint64_t *part_off, uint64_t *part_kmers, uint64_t *part_distinct,
                 uint64_t *part_skms, double *t_phase) {
    init_tables();
    if (k < 1 || k > 64 || p < 3 || p > 32 || p > k || P == 0 || n < 0) return -1;
    if (k > 32 && !key_hi) return -1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    const uint64_t B = offs[n] - offs[0];
    const uint8_t *src = reads;
    uint8_t *hreads = NULL; int32_t *hlens = NULL, *hpos = NULL;
    double t0 = now_s();
    if (hpc) {
        hreads = (uint8_t *)malloc(B + 1); hlens = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
        hpos = (int32_t *)malloc(sizeof(int32_t) * (B + 1));
        or_hpc(reads, offs, n, hreads, hlens, hpos);
        src = hreads;
    }
    skm_vec *bins = (skm_vec *)calloc((size_t)nthreads * P, sizeof(skm_vec));
    pthread_t th[256]; p1_job jobs[256];
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = (p1_job){src, offs, hlens, n * t / nthreads, n * (t + 1) / nthreads,
                           k, p, fmode, pmode, P, bins + (size_t)t * P};
        pthread_create(&th[t], NULL, phase1_worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    double t1 = now_s();

    /* scratch placement: partition p writes at the prefix of k-mer totals */
    uint64_t *scratch_off = (uint64_t *)malloc(sizeof(uint64_t) * (P + 1));
    {
        uint64_t acc = 0;
        for (uint32_t q = 0; q < P; q++) {
            scratch_off[q] = acc;
            for (int t = 0; t < nthreads; t++) {
                skm_vec *v = &bins[(size_t)t * P + q];
                for (size_t i = 0; i < v->n; i++) acc += v->v[i].len - k + 1;
            }
        }
        scratch_off[P] = acc;
        if (acc > table_cap) {
            for (size_t i = 0; i < (size_t)nthreads * P; i++) free(bins[i].v);
            free(bins); free(scratch_off); free(hreads); free(hlens); free(hpos);
            return -1;
        }
    }
    p2_ctx c = {src, k, P, nthreads, bins, min_cnt, max_cnt, key_lo, key_hi, cnt, part_off,
                part_kmers, part_distinct, part_skms, table_cap, scratch_off, 0};
    pthread_mutex_init(&c.mu, NULL);
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, phase2_worker, &c);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&c.mu);
    /* compact the per-partition regions into consecutive tables */
    part_off[0] = 0;
    for (uint32_t q = 0; q < P; q++) {
        uint64_t kept = part_off[q + 1];
        uint64_t from = scratch_off[q], to = part_off[q];
        if (from != to && kept) {
            memmove(key_lo + to, key_lo + from, kept * 8);
            if (key_hi) memmove(key_hi + to, key_hi + from, kept * 8);
            memmove(cnt + to, cnt + from, kept * 4);
        }
        part_off[q + 1] = to + kept;
    }
    double t2 = now_s();
    if (t_phase) { t_phase[0] = t1 - t0; t_phase[1] = t2 - t1; }
    for (size_t i = 0; i < (size_t)nthreads * P; i++) free(bins[i].v);
    free(bins); free(scratch_off); free(hreads); free(hlens); free(hpos);
    return (int64_t)part_off[P];
}
