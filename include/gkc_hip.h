/*
 * gkc_hip.h -- C ABI of the MI355X (gfx950) k-mer counting library.
 *
 * Plain C types only (pointers, sizes, PODs); no HIP or torch types appear
 * in any signature.  Every function returns 0 (or a non-negative count) on
 * success and a negative GKC_E* code on failure; gkc_last_error() gives the
 * message of the calling thread's last failure.
 *
 * The entry points replace the two calls of the reference's counting path:
 *   GenSuperkmerGPU   reference gkc_cuda.hpp:58-61 (impl. gkc_cuda.cu:376-541)
 *   kmc_counting_GPU  reference kmer_counting.hpp:9-12 (impl. kmer_counting.cu:59-199)
 * The C++ drop-in wrappers with the reference's exact signatures are in
 * include/gkc/gkc_cuda.hpp and include/gkc/kmer_counting.hpp.
 */
#ifndef GKC_HIP_H
#define GKC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GKC_OK = 0,
    GKC_EINVAL = -1,     /* bad argument */
    GKC_EHIP = -2,       /* HIP runtime error */
    GKC_ENOMEM = -3,     /* device allocation failed */
    GKC_ESTATE = -4,     /* call out of sequence */
    GKC_ECAP = -5,       /* output buffer too small */
    GKC_EIO = -6         /* file could not be created or written */
};

/* Flags of gkc_partition / gkc_count. */
#define GKC_DEVICE_INPUT   1u   /* reads/offs are device pointers (already in HBM) */
#define GKC_BATCHED        2u   /* persistent kernels + device work queue + hipGraph replay */
#define GKC_PRESIZED       4u   /* gkc_count: size tables from gkc_set_table_caps / the learned
                                   predictor instead of a count pass (overflow falls back) */
#define GKC_SYNC_ONLY      8u   /* internal: no host sync beyond what the mode needs */
#define GKC_LEARNED       16u   /* gkc_count with GKC_PRESIZED: capacities from the predictor set by
                                   gkc_set_predictor, computed on the device before the sort */
#define GKC_EMIT_WINDOWS  32u   /* gkc_partition: also write per-window (minimizer, position, strand) */
#define GKC_WITH_POS      64u   /* gkc_partition: keep each superkmer's read and offset in the bins
                                   (every batch between resets alike) */

/* Counting parameters (the reference's PAR fields K_kmer, P_minimizer, HPC,
 * SKM_partitions, kmer_min_freq, kmer_max_freq; utilities.hpp:104-123). */
typedef struct gkc_params {
    int32_t k;             /* 1..64; k <= 32 uses 64-bit keys, else 128-bit (hi, lo) */
    int32_t p;             /* minimizer length, 3..min(k, 32) */
    int32_t hpc;           /* homopolymer compression of reads */
    uint32_t n_parts;      /* superkmer partitions, 1..1024 */
    int32_t filter_mode;   /* 1: new_filter2 (reference default), 0: accept every p-mer */
    int32_t part_mode;     /* 0: minimizer % P (reference), 1: hashed minimizer % P */
    uint32_t min_count;    /* keep (k-mer, count) with min_count <= count ... */
    uint32_t max_count;    /* ... <= max_count; 0 = unbounded */
} gkc_params;

typedef struct gkc_ctx gkc_ctx;
typedef struct gkc_mlp gkc_mlp;     /* the learned predictor (see below) */

/* Per-stage device times of the last calls (milliseconds, HIP events). */
typedef struct gkc_stats {
    double ms_partition;       /* last gkc_partition (all its kernels) */
    double ms_count;           /* last gkc_count */
    double ms_extract, ms_sort, ms_runlength;   /* inside the last gkc_count */
    double ms_sort_pass_avg;   /* average duration of one onesweep pass launch */
    uint64_t n_kmers;          /* k-mer occurrences in the bins */
    uint64_t n_superkmers;
    uint64_t n_words;          /* 2-bit words in the bins */
    uint64_t n_distinct;       /* distinct k-mers (before the count filter) */
    uint64_t n_entries;        /* table entries kept */
    uint32_t n_pass;           /* radix passes run per key */
    uint32_t fallbacks;        /* presize overflows that fell back to count-then-allocate */
    uint32_t host_syncs;       /* host synchronisations inside the last gkc_count */
    uint32_t partition_fallbacks;  /* presized batched scatters redone by count-then-allocate since gkc_reset */
    uint32_t heavy_ranges;     /* binned tiles too large for LDS, sorted through global memory (last gkc_count) */
    uint32_t binned;           /* 1 when the last gkc_count used the binned kernels */
} gkc_stats;

/* Kernel variants (picked by gkc_amd/autotune.py, keyed on k, read-length
 * distribution and partition size). */
typedef struct gkc_tuning {
    int32_t sort_items;          /* keys per thread of a onesweep tile: 8,12,16,20,24 (u64, 4 waves),
                                    8,10,12 (u64, 8 waves) / 4,8,12 (u128) */
    int32_t sort_blocks_per_cu;  /* persistent workgroups per CU in the sort passes (0: occupancy limit) */
    int32_t single_list;         /* 1: batched tiles in one global order instead of one work list per XCD */
    int32_t count_mode;          /* batched form: 0 binned (onesweep passes on the top key bits, then
                                    per-window hash count of whole bins; default for k >= 8), 1 full LSD sort */
    int32_t bin_passes;          /* binned form: 8-bit onesweep passes before the count, 1..3 (0: automatic) */
    int32_t count_window;        /* binned form: keys per count window (the bins starting in it), a multiple
                                    of 64 in 128..1024 (0: 256); the same table for every value */
    int32_t sort_waves;          /* waves per onesweep workgroup, 4 or 8 (0: 4; 8 for 64-bit keys only, with
                                    sort_items 8, 10 or 12) */
    int32_t digit_bits;          /* binned form: digit width of the onesweep passes, 6..8 (0: 8), so bins are
                                    bin_passes x digit_bits key bits; 6 and 7 with sort_items 16 or 20 (u64),
                                    8 or 12 (u128) and 4 waves.  The full LSD sort keeps 8 */
} gkc_tuning;

const char *gkc_last_error(void);
int gkc_version(void);

/* Create a context on `device` (one per host thread; calls on one context
 * are serialised by the library). NULL on failure (see gkc_last_error). */
gkc_ctx *gkc_create(int device, const gkc_params *prm);
void gkc_destroy(gkc_ctx *ctx);
/* The context's HIP stream (hipStream_t) as an opaque pointer. */
void *gkc_stream(gkc_ctx *ctx);
int gkc_synchronize(gkc_ctx *ctx);

/* ---- drop-in #1: GenSuperkmerGPU (gkc_cuda.hpp:58-61) ----
 * For one sub-batch of n reads (host CSR: reads at reads[0..offs[n]-offs[0]),
 * absolute offsets offs[0..n]) fill the host arrays the reference hands to its
 * callback in T_h_data (types.h:32-46), indexed by batch-local base position:
 *   minimizers[b]   u32 window minimizer (valid at [0, len-k] of each read)
 *   skm_offs[b]     i32 superkmer offset lists [0, i1, .., len-k+1, 0] per read
 *   reads_out[b]    the (HPC-compressed, when hpc) bases           (may be NULL)
 *   read_len[r]     HPC read lengths                 (HPC only, may be NULL)
 *   hpc_orig_pos[b] original in-read index of each HPC base (HPC only, may be NULL)
 * Requires p <= 16 (the reference's u32 minimizer). */
int gkc_superkmer_arrays(gkc_ctx *ctx, const char *reads, const uint64_t *offs, int64_t n,
                         uint32_t *minimizers, int32_t *skm_offs, char *reads_out,
                         int32_t *read_len, int32_t *hpc_orig_pos);

/* ---- on-device read ingest (replaces the host ReadLoader's record parsing,
 * read_loader.hpp:44-196: _load_fastq / _load_fasta) ----
 * FASTQ text (4-line records: '@' header, sequence, '+', quality) or
 * single-line FASTA text ('>' header, one sequence line) in device memory to
 * the read CSR gkc_partition takes, on the device: the reads' bases back to
 * back in `bases` and n + 1 absolute offsets in `offs` (offs[0] = 0).  The read
 * is the sequence line without a trailing '\r'; the last line may lack its
 * newline.  format: 0 from the first byte, 1 FASTA, 2 FASTQ.  Capacities:
 * bases n_bytes bytes, offs n_bytes / 2 + 2 entries.  Runs on `stream` (a
 * hipStream_t, NULL for the default stream); the library keeps its scratch
 * for the device between calls (grown as inputs need, held until the process
 * ends) and serialises calls on one device.  Returns when done: the
 * number of reads, or a negative GKC_E* code (GKC_EINVAL for text that is not
 * whole records of the format). */
int64_t gkc_parse_reads(const char *text, uint64_t n_bytes, int format, char *bases, uint64_t *offs, void *stream);

/* ---- gkc_parse_reads with no device allocation ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (set-up of stream-captured work, tight serving loops, hosts whose
 * allocator forbids hipMalloc): the same arguments, output, errors and
 * serialisation per device as gkc_parse_reads, but the call allocates
 * nothing.  Its scratch is a fixed array of GKC_PARSE_NOALLOC_SCRATCH_BYTES
 * in the library's device image, one per GPU, held until the process ends;
 * every call on a GPU reuses that GPU's array (calls on one device are
 * serialised by the library: overlapping calls would race on it).  A text
 * needs 8 * (blocks + records) bytes of it and a little more, blocks =
 * n_bytes / 4096: about 16 MB of FASTQ with 10 kb reads up to about 1 GB
 * of FASTQ with 150-base reads (3.9 M records) fit; a larger text is
 * rejected with GKC_ECAP before any of its records is written. */
#define GKC_PARSE_NOALLOC_SCRATCH_BYTES (32ull << 20)
int64_t gkc_parse_reads_noalloc(const char *text, uint64_t n_bytes, int format, char *bases, uint64_t *offs,
                                void *stream);

/* ---- device-resident path (what GenSuperkmerGPU + SKMStore::add_skms do,
 * superkmers.hpp:167-186, without leaving the GPU) ----
 * Encode, minimizer, superkmer and scatter one batch of reads into the
 * context's partition bins (appending across calls).  With
 * GKC_DEVICE_INPUT the pointers are device memory. */
int gkc_partition(gkc_ctx *ctx, const char *reads, const uint64_t *offs, int64_t n, uint32_t flags);
/* Per-window (minimizer, position, strand) records of one batch (host
 * arrays, batch-local base index, HPC coordinates when hpc): the reference
 * declares these fields but never fills them (gkc_cuda.cu:159-160,
 * CountTask::StoreMinimizerPos gkc_cuda.hpp:50).  pos = offset in the window
 * of the chosen p-mer (ties: the later start, forward before reverse
 * complement), 255 when no p-mer passes the filter; strand 0 forward,
 * 1 reverse complement, -1 when both strands tie or none passes.  p <= 16. */
int gkc_window_minimizers(gkc_ctx *ctx, const char *reads, const uint64_t *offs, int64_t n,
                          uint32_t *mm, uint8_t *pos, int8_t *strand);
/* Device pointers to the records written by the last gkc_partition with
 * GKC_EMIT_WINDOWS (valid until the next such call). */
int gkc_windows_device(gkc_ctx *ctx, uint32_t **mm, uint8_t **pos, int8_t **strand, uint64_t *n);
/* ---- position-carrying variant (the reference's CountTask::SKMPartWithPos,
 * gkc_cuda.hpp:50, declared for "count a portion of k-mers and store their
 * positions", README.md:15-18, never implemented) ----
 * After gkc_partition(..., GKC_WITH_POS): partition `part`'s superkmers as
 * (read, offset of the first base in the read, k-mers), reads numbered
 * across batches from 0 at the last gkc_reset; offsets are HPC coordinates
 * when hpc.  Returns the number of superkmers (GKC_ECAP when above cap; with
 * the three outputs null, only the number). */
int64_t gkc_superkmer_positions(gkc_ctx *ctx, int32_t part, uint32_t *read, uint32_t *offset, uint32_t *n_kmers,
                                uint64_t cap);
/* After gkc_count on GKC_WITH_POS bins (k <= 32): every occurrence of the
 * partition's sampled table k-mers -- those whose canonical value v has
 * murmur3-fmix64(v) % sample_mod == 0 (1: all) -- grouped by k-mer in key
 * order: keys[e], counts[e], positions[pos_off[e] .. pos_off[e + 1]) holding
 * read << 32 | (k-mer offset in the read) << 1 | strand (1 when the reverse
 * complement is the canonical form), ascending within each k-mer.  Sets
 * *n_entries and *n_positions; GKC_ECAP when a capacity is short (call with
 * 0 capacities to size the buffers).  Superkmers are listed and k-mers found
 * on the device; each k-mer's few positions are ordered on the host. */
int gkc_kmer_positions(gkc_ctx *ctx, int32_t part, uint32_t sample_mod, uint64_t *keys, uint32_t *counts,
                       uint64_t *pos_off, uint64_t *positions, uint64_t entry_cap, uint64_t pos_cap,
                       uint64_t *n_entries, uint64_t *n_positions);
/* Clear the bins and the table. */
int gkc_reset(gkc_ctx *ctx);

/* ---- drop-in #2: kmc_counting_GPU (kmer_counting.hpp:9-12) ----
 * Count every partition in the bins: extract canonical k-mers, onesweep
 * radix sort, run-length count with the [min_count, max_count] filter.
 * The table stays in device memory, partition by partition, keys ascending. */
int gkc_count(gkc_ctx *ctx, uint32_t flags);
/* Table layout after gkc_count: part_off[P+1] entry offsets, part_distinct[P]
 * distinct k-mers before the filter, part_kmers[P] k-mer occurrences. */
int gkc_table_info(gkc_ctx *ctx, uint64_t *part_off, uint64_t *part_distinct, uint64_t *part_kmers);
/* Copy the table (or partition `part`, or all when part < 0) to host memory:
 * key_lo[i] (k <= 32: the key), key_hi[i] (k > 32 only, else may be NULL), cnt[i]. */
int64_t gkc_table_copy(gkc_ctx *ctx, int32_t part, uint64_t *key_lo, uint64_t *key_hi, uint32_t *cnt,
                       uint64_t cap);
/* Physical start of each partition's entries in the device table (a presized
 * count leaves unused capacity between partitions; part_off is logical). */
int gkc_table_starts(gkc_ctx *ctx, uint64_t *starts);
/* ---- table file (no counterpart in the reference: kmc_counting_GPU computes
 * its table and drops it, kmer_counting.cu:187-190) ----
 * The last gkc_count's table to a new file at `path` (truncated if it exists),
 * streamed from the device through two pinned staging buffers: the copy of
 * one chunk runs while the previous one is written.  Little-endian layout:
 *   64-byte header: char magic[8] = "GKCTAB1\0"; u32 version = 1, k, p, hpc,
 *     n_parts, key_words (1: k <= 32, 2: k > 32), min_count, max_count;
 *     u64 n_entries, n_kmers (k-mer occurrences counted), reserved = 0
 *   u64 part_off[n_parts + 1], part_distinct[n_parts], part_kmers[n_parts]
 *   u64 keys[n_entries][key_words] (key_words 2: low then high 64 bits),
 *     partition by partition, ascending within each
 *   u32 counts[n_entries]
 * Returns the bytes written, or a negative GKC_E* code (GKC_EIO: the file). */
int64_t gkc_table_write(gkc_ctx *ctx, const char *path);
/* Device pointers of the table (valid until the next gkc_count/gkc_reset). */
int gkc_table_device(gkc_ctx *ctx, void **keys, uint32_t **counts, uint64_t *n_entries);

/* Count one partition given as packed superkmers in host memory (the
 * SKMStore of the drop-in): n_skm superkmers, lens[i] bases each, bases 2-bit
 * MSB-first 16 per u32 word, each superkmer starting on a word boundary in
 * `words`.  Writes up to `cap` entries, sets *n_entries / *n_distinct. */
int gkc_count_packed(gkc_ctx *ctx, const uint32_t *words, uint64_t n_words, const uint32_t *lens,
                     uint64_t n_skm, uint64_t *key_lo, uint64_t *key_hi, uint32_t *cnt, uint64_t cap,
                     uint64_t *n_entries, uint64_t *n_distinct);

int gkc_get_stats(gkc_ctx *ctx, gkc_stats *out);

/* Pinned host memory for read batches (the reference's PinnedCSR uses
 * cudaHostAlloc, gkc_cuda.cu:121-139).  NULL on failure. */
void *gkc_host_alloc(uint64_t bytes);
void gkc_host_free(void *p);
/* Make `device` current for the calling thread and wait for it to be idle
 * (the reference's GPUReset, gkc_cuda.cu:367-373, minus the device reset). */
int gkc_device_select(int device);

/* ---- measurement (no reference counterpart) ----
 * Copy `bytes` (a multiple of 16; both pointers 16-byte aligned, device
 * memory) on `stream` (hipStream_t, NULL: default) with the library's
 * streaming-copy kernel: the practical read + write ceiling the benchmark
 * reports its kernels against, and a kernel of known traffic for HBM byte
 * counters.  Asynchronous; GKC_EINVAL for misaligned or odd sizes. */
int gkc_stream_copy(void *dst, const void *src, uint64_t bytes, void *stream);
/* Copy n_keys 64-bit keys with the write pattern of one onesweep pass and
 * none of its work: segments of seg_keys keys (a multiple of 5120; n_keys a
 * multiple of seg_keys), tiles of 5120 keys read in order, each tile written
 * as `runs` runs (runs divides 5120) into the segment's `runs` bucket regions,
 * segments spread over 8 lists, `workgroups` (a multiple of 8) workgroups.
 * The ceiling of that access pattern, for the benchmark; asynchronous on
 * `stream`; GKC_EINVAL for shapes outside these rules. */
int gkc_scatter_copy(void *dst, const void *src, uint64_t n_keys, uint64_t seg_keys, uint32_t runs,
                     uint32_t workgroups, void *stream);
/* The vendor library's radix sort (rocPRIM radix_sort_keys) of n 64-bit keys
 * on bits [begin_bit, end_bit), device buffers, timed with HIP events on
 * `stream` into *ms (synchronous; allocates its temporary storage): a
 * baseline for the onesweep passes, no reference counterpart. */
int gkc_vendor_sort_u64(const void *keys_in, void *keys_out, uint64_t n, int begin_bit, int end_bit, void *stream,
                        float *ms);

/* Per-partition 4-mer sketch taken during gkc_partition's scatter: float
 * [P][256] holding log1p(count) of the leading 4-mers (first base most
 * significant) of the partition's sampled superkmers, as stored in the bins
 * (runs of windows with one minimizer, also cut where a 2048-window tile of
 * the batch starts).  A superkmer is sampled when the 32-bit product of its
 * first min(k, 16) bases' 2-bit codes (first base highest) and 0x9E3779B1
 * has its top three bits zero: about one in eight, decided by its bases
 * alone.  k < 4 takes no sketch. */
int gkc_sketch(gkc_ctx *ctx, float *out);
int gkc_set_tuning(gkc_ctx *ctx, const gkc_tuning *t);
int gkc_get_tuning(gkc_ctx *ctx, gkc_tuning *t);
/* Table capacities used by gkc_count(GKC_PRESIZED): caps[P] entries.  The
 * binned count (count_mode 0) lays the table out densely from the exact
 * per-partition counts it measures on the device, so only the sum of caps
 * must cover the table; the full LSD count (count_mode 1) gives partition q
 * exactly caps[q] entries (unused capacity stays between partitions).  A
 * shortfall is caught on the device and the emission is redone with exact
 * sizes. */
int gkc_set_table_caps(gkc_ctx *ctx, const uint32_t *caps);
/* Learned presizing on the device: gkc_count(GKC_BATCHED | GKC_PRESIZED |
 * GKC_LEARNED) computes each partition's capacity from the sketch with `mlp`
 * (BF16 matrix-core inference) on the context's stream before sorting:
 * features (log1p(sketch) - x_mean) / x_std, output 1 de-normalised as
 * y * t_std[1] + t_mean[1] = log(distinct / total), capacity
 * min(ceil(total * exp(min(that, 0)) * margin) + 16, total).  Only their sum
 * is read back, while the sort runs.  The MLP must outlive its use; NULL
 * detaches.  At most 1024 partitions. */
int gkc_set_predictor(gkc_ctx *ctx, gkc_mlp *mlp, const float *x_mean, const float *x_std, const float *t_mean,
                      const float *t_std, float margin);
/* The capacities (caps[P]) the last GKC_LEARNED count computed on the device. */
int gkc_learned_caps(gkc_ctx *ctx, uint32_t *caps);
/* Per-partition k-mer totals of the bins (known exactly after gkc_partition). */
int gkc_part_kmers(gkc_ctx *ctx, uint64_t *out);

/* ---- learned partition pre-sizing (no reference counterpart: the
 * reference sizes nothing ahead, kmer_counting.cu:95-191 grows host vectors)
 * MLP 256 -> 128 -> 128 -> 2 with ReLU, FP32 master weights; flat parameter
 * order W1[256][128], b1[128], W2[128][128], b2[128], W3[128][2], b3[2]. */
uint64_t gkc_mlp_n_params(void);
const char *gkc_mlp_last_error(void);
gkc_mlp *gkc_mlp_create(int device, uint64_t seed);
void gkc_mlp_destroy(gkc_mlp *m);
int gkc_mlp_get_params(gkc_mlp *m, float *out);
int gkc_mlp_set_params(gkc_mlp *m, const float *in);
/* FP32 forward of n rows x[n][256] -> y[n][2] (HIP kernels). */
int gkc_mlp_forward(gkc_mlp *m, const float *x, int64_t n, float *y);
/* BF16 matrix-core (MFMA) inference with FP32 accumulation. */
int gkc_mlp_forward_bf16(gkc_mlp *m, const float *x, int64_t n, float *y);
/* Mean-squared-error loss against t[n][2] and its gradient (flat, n_params). */
int gkc_mlp_loss_grad(gkc_mlp *m, const float *x, const float *t, int64_t n, float *loss, float *grad);
/* Adam training (HIP forward + backward kernels); losses[epochs] per epoch. */
int gkc_mlp_train(gkc_mlp *m, const float *x, const float *t, int64_t n, int epochs, int batch, float lr,
                  float weight_decay, uint64_t seed, float *losses);
/* C[32][N] = A[32][K] * B[K][N] on the matrix cores (bf16 in, fp32 out), the
 * tile routine of the inference kernel; K <= 256, K % 16 == 0. */
int gkc_bf16_matmul32(int device, const float *A, const float *B, int K, int N, float *C);

#ifdef __cplusplus
}
#endif
#endif
