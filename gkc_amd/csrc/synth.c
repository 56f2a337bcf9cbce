/*
 * Seeded synthetic genomes and reads for the k-mer counting workloads.
 *
 * The reference ships no datasets and no generator (its inputs are external
 * PacBio files, Developer.md:34-37), so the workloads are fixed here once:
 *   - PRNG: splitmix64 seeding xoshiro256** (public-domain algorithms by
 *     Vigna / Blackman).
 *   - Genome: bases drawn i.i.d. with P(G)=P(C)=gc/2, P(A)=P(T)=(1-gc)/2,
 *     optionally alternating GC per fixed-size block; then segmental
 *     duplications (copy a segment to a random place, mutate at a given
 *     divergence) and tandem repeats (a random unit repeated in place) are
 *     overwritten until their target fractions are reached.
 *   - Reads: uniform start, optional log-normal length, reverse-complemented
 *     with probability rc_prob, i.i.d. per-base errors split among
 *     substitution / insertion / deletion.
 * Each read draws from its own xoshiro stream seeded by (seed, read index),
 * so generation is parallel and the output does not depend on the thread
 * count.  Output is a CSR: uppercase ACGT bytes plus n+1 u64 offsets, the
 * same layout the reference's PinnedCSR holds (gkc_cuda.hpp:33-46).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

typedef struct { uint64_t s[4]; } rng_t;

static inline uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t rng_next(rng_t *r) {
    uint64_t *s = r->s;
    uint64_t res = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return res;
}
static void rng_seed(rng_t *r, uint64_t seed, uint64_t stream) {
    uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ULL);
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}
/* uniform double in [0,1) */
static inline double rng_unif(rng_t *r) { return (rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }
/* uniform integer in [lo, hi] */
static inline uint64_t rng_range(rng_t *r, uint64_t lo, uint64_t hi) {
    return lo + (uint64_t)(rng_unif(r) * (double)(hi - lo + 1));
}
static inline double rng_normal(rng_t *r) {
    double u1 = rng_unif(r), u2 = rng_unif(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static const char ACGT[4] = {'A', 'C', 'G', 'T'};
static inline char draw_base(rng_t *r, double gc) {
    double u = rng_unif(r);
    if (u < gc) return (u < gc * 0.5) ? 'G' : 'C';
    return (u < gc + (1.0 - gc) * 0.5) ? 'A' : 'T';
}
static inline char other_base(rng_t *r, char b) {
    int code = (b == 'A') ? 0 : (b == 'C') ? 1 : (b == 'G') ? 2 : 3;
    int o = (int)(rng_next(r) % 3);
    return ACGT[(code + 1 + o) & 3];
}
static inline char compl_base(char b) {
    switch (b) { case 'A': return 'T'; case 'C': return 'G'; case 'G': return 'C'; default: return 'A'; }
}

typedef struct {
    uint64_t genome_len;
    double gc;            /* GC fraction of block 0 */
    double gc_alt;        /* GC fraction of odd blocks (if gc_block > 0) */
    uint64_t gc_block;    /* 0: uniform composition */
    double seg_frac;      /* fraction covered by segmental duplications */
    uint32_t seg_min, seg_max;
    double seg_div;       /* substitution rate applied to each copy */
    double tandem_frac;   /* fraction covered by tandem repeats */
    uint32_t unit_min, unit_max, copies_min, copies_max;
} gkc_genome_cfg;

typedef struct {
    uint64_t n_reads;     /* 0: ceil(coverage * G / len_mean) */
    double coverage;
    int len_mode;         /* 0: fixed len_mean, 1: log-normal(len_mean, len_sd) */
    double len_mean, len_sd;
    uint32_t len_min;     /* clamp (also clamps to G) */
    double rc_prob;
    double err_rate;      /* per-base error probability */
    double sub_w, ins_w, del_w;  /* error-type weights (normalised) */
} gkc_reads_cfg;

/* Fill genome[0..G). Returns 0 on success. */
int gkc_synth_genome(const gkc_genome_cfg *cfg, uint64_t seed, char *genome) {
    const uint64_t G = cfg->genome_len;
    if (G == 0) return 0;
    rng_t r; rng_seed(&r, seed, 0x6E6F6D65ULL);
    for (uint64_t i = 0; i < G; i++) {
        double gc = cfg->gc;
        if (cfg->gc_block && ((i / cfg->gc_block) & 1)) gc = cfg->gc_alt;
        genome[i] = draw_base(&r, gc);
    }
    /* segmental duplications */
    if (cfg->seg_frac > 0 && cfg->seg_max > 0 && cfg->seg_max < G) {
        uint64_t target = (uint64_t)(cfg->seg_frac * (double)G), done = 0;
        while (done < target) {
            uint64_t len = rng_range(&r, cfg->seg_min, cfg->seg_max);
            uint64_t src = rng_range(&r, 0, G - len);
            uint64_t dst = rng_range(&r, 0, G - len);
            memmove(genome + dst, genome + src, len);
            for (uint64_t j = 0; j < len; j++)
                if (rng_unif(&r) < cfg->seg_div) genome[dst + j] = other_base(&r, genome[dst + j]);
            done += len;
        }
    }
    /* tandem repeats (unit repeated in place, overwriting) */
    if (cfg->tandem_frac > 0 && cfg->unit_max > 0) {
        uint64_t target = (uint64_t)(cfg->tandem_frac * (double)G), done = 0;
        char unit[256];
        while (done < target) {
            uint64_t ulen = rng_range(&r, cfg->unit_min, cfg->unit_max);
            uint64_t copies = rng_range(&r, cfg->copies_min, cfg->copies_max);
            uint64_t len = ulen * copies;
            if (len >= G) len = G - 1;
            uint64_t pos = rng_range(&r, 0, G - len);
            double gc = cfg->gc;
            if (cfg->gc_block && ((pos / cfg->gc_block) & 1)) gc = cfg->gc_alt;
            for (uint64_t j = 0; j < ulen; j++) unit[j] = draw_base(&r, gc);
            for (uint64_t j = 0; j < len; j++) genome[pos + j] = unit[j % ulen];
            done += len;
        }
    }
    return 0;
}

static uint64_t read_count(const gkc_reads_cfg *cfg, uint64_t G) {
    if (cfg->n_reads) return cfg->n_reads;
    return (uint64_t)ceil(cfg->coverage * (double)G / cfg->len_mean);
}

/* Target length of read i (before indels). */
static uint64_t target_len(const gkc_reads_cfg *cfg, uint64_t G, uint64_t seed, uint64_t i) {
    if (cfg->len_mode == 0) {
        uint64_t L = (uint64_t)cfg->len_mean;
        return L > G ? G : L;
    }
    rng_t r; rng_seed(&r, seed ^ 0x4C454E47ULL, i);
    double m = cfg->len_mean, sd = cfg->len_sd;
    double s2 = log(1.0 + (sd * sd) / (m * m));
    double mu = log(m) - 0.5 * s2;
    double x = exp(mu + sqrt(s2) * rng_normal(&r));
    uint64_t L = (uint64_t)x;
    if (L < cfg->len_min) L = cfg->len_min;
    if (L > G) L = G;
    return L;
}

/* Upper bound of output bytes for read i: indels can add at most one base per
 * genome base consumed, so 2*L bounds every read. */
typedef struct {
    const gkc_reads_cfg *cfg; const char *genome; uint64_t G, seed;
    uint64_t lo, hi; const uint64_t *bound_offs; char *scratch; uint64_t *lens;
} job_t;

static uint64_t gen_one(const gkc_reads_cfg *cfg, const char *genome, uint64_t G,
                        uint64_t seed, uint64_t i, char *out) {
    uint64_t L = target_len(cfg, G, seed, i);
    rng_t r; rng_seed(&r, seed, i + 1);
    uint64_t start = rng_range(&r, 0, G - L);
    double wsum = cfg->sub_w + cfg->ins_w + cfg->del_w;
    double ps = wsum > 0 ? cfg->sub_w / wsum : 1.0, pi = wsum > 0 ? cfg->ins_w / wsum : 0.0;
    uint64_t o = 0, g = start;
    while (o < L && g < G) {
        char b = genome[g];
        if (cfg->err_rate > 0 && rng_unif(&r) < cfg->err_rate) {
            double u = rng_unif(&r);
            if (u < ps) { out[o++] = other_base(&r, b); g++; }
            else if (u < ps + pi) { out[o++] = ACGT[rng_next(&r) & 3]; if (o < L) out[o++] = b; g++; }
            else { g++; }            /* deletion */
        } else { out[o++] = b; g++; }
    }
    if (rng_unif(&r) < cfg->rc_prob) {
        for (uint64_t a = 0; a < o / 2; a++) {
            char x = compl_base(out[a]), y = compl_base(out[o - 1 - a]);
            out[a] = y; out[o - 1 - a] = x;
        }
        if (o & 1) out[o / 2] = compl_base(out[o / 2]);
    }
    return o;
}

static void *worker(void *arg) {
    job_t *j = (job_t *)arg;
    for (uint64_t i = j->lo; i < j->hi; i++)
        j->lens[i] = gen_one(j->cfg, j->genome, j->G, j->seed, i, j->scratch + j->bound_offs[i]);
    return NULL;
}

/* Number of reads for a configuration. */
uint64_t gkc_synth_read_count(const gkc_reads_cfg *cfg, uint64_t genome_len) {
    return read_count(cfg, genome_len);
}

/* Upper bound of total read bytes (for buffer sizing). */
uint64_t gkc_synth_bytes_bound(const gkc_reads_cfg *cfg, uint64_t genome_len, uint64_t seed) {
    uint64_t n = read_count(cfg, genome_len), tot = 0;
    for (uint64_t i = 0; i < n; i++) tot += target_len(cfg, genome_len, seed, i);
    return tot;
}

/* Generate reads into a CSR.  `reads` must hold gkc_synth_bytes_bound() bytes,
 * `offs` n+1 entries.  Returns the total number of bases written. */
uint64_t gkc_synth_reads(const gkc_reads_cfg *cfg, const char *genome, uint64_t G,
                         uint64_t seed, int n_threads, char *reads, uint64_t *offs) {
    uint64_t n = read_count(cfg, G);
    uint64_t *bound = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    uint64_t *lens = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    bound[0] = 0;
    for (uint64_t i = 0; i < n; i++) bound[i + 1] = bound[i] + target_len(cfg, G, seed, i);
    /* generate in place at the upper-bound offsets, then compact (reads only
     * shrink relative to their bound: output length <= target length) */
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 64) n_threads = 64;
    pthread_t th[64]; job_t jobs[64];
    for (int t = 0; t < n_threads; t++) {
        jobs[t] = (job_t){cfg, genome, G, seed, n * t / n_threads, n * (t + 1) / n_threads, bound, reads, lens};
        pthread_create(&th[t], NULL, worker, &jobs[t]);
    }
    for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
    offs[0] = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (offs[i] != bound[i]) memmove(reads + offs[i], reads + bound[i], lens[i]);
        offs[i + 1] = offs[i] + lens[i];
    }
    uint64_t tot = offs[n];
    free(bound); free(lens);
    return tot;
}

/* Write single-line FASTA records ">r<index>\n<seq>\n" (the record layout the
 * reference's loader parses, read_loader.hpp:175-176). Returns 0 on success. */
int gkc_write_fasta(const char *path, const char *reads, const uint64_t *offs, uint64_t n) {
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    for (uint64_t i = 0; i < n; i++) {
        fprintf(f, ">r%llu\n", (unsigned long long)i);
        fwrite(reads + offs[i], 1, offs[i + 1] - offs[i], f);
        fputc('\n', f);
    }
    return fclose(f);
}

/* Parse a single-line FASTA file into a CSR (sequence lines only). Two-call
 * protocol: with reads==NULL it returns the number of records and sets
 * *total_bases; otherwise it fills reads/offs. Returns -1 on I/O error. */
int64_t gkc_read_fasta(const char *path, char *reads, uint64_t *offs, uint64_t *total_bases) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    char *line = NULL; size_t cap = 0; ssize_t len;
    int64_t n = 0; uint64_t tot = 0;
    if (offs) offs[0] = 0;
    while ((len = getline(&line, &cap, f)) > 0) {
        if (line[0] == '>') continue;
        while (len > 0 && (line[len - 1] == '\n' || line[len - 1] == '\r')) len--;
        if (reads) memcpy(reads + tot, line, (size_t)len);
        tot += (uint64_t)len;
        n++;
        if (offs) offs[n] = tot;
    }
    free(line);
    fclose(f);
    if (total_bases) *total_bases = tot;
    return n;
}
