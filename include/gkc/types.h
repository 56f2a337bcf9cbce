/*
 * Scalar types and host records of the counting path, laid out as the
 * reference declares them (reference types.h:10-53) so code written against
 * the reference compiles unchanged against the drop-in.
 */
#ifndef GKC_DROPIN_TYPES_H
#define GKC_DROPIN_TYPES_H

#include <cstddef>
#include <cstdint>

#ifndef _in_
#define _in_
#endif
#ifndef _out_
#define _out_
#endif

/* The drop-in's record types live in an inline namespace: code written
 * against these headers names them unqualified as before, but their mangled
 * names differ from the reference's, so an object compiled against the
 * reference's own headers (whose SKMStore and T_h_data have other layouts)
 * fails to link against libgkc_dropin.so instead of corrupting memory. */
#define GKC_DROPIN_BEGIN inline namespace gkc_mi355x_v1 {
#define GKC_DROPIN_END }

typedef size_t T_CSR_cap;          /* CSR offsets */
typedef int T_CSR_count;
typedef unsigned short T_skm_len;  /* the reference's superkmer length type */
typedef int T_read_len;
typedef int T_read_cnt;
typedef unsigned int T_minimizer;  /* minimizer values, p <= 16 */
const T_minimizer T_MM_MAX = (T_minimizer)0xFFFFFFFFu;
typedef unsigned char T_kvalue;

GKC_DROPIN_BEGIN
struct ReadPtr {
    const char *read;
    T_read_len len;
};

GKC_DROPIN_END

typedef unsigned long long T_kmer;       /* k <= 32 */
typedef unsigned long long T_2bitkmer;

GKC_DROPIN_BEGIN

/* Host arrays handed to GenSuperkmerGPU's callback once per sub-batch. */
struct T_h_data {
    T_read_cnt reads_cnt;
    char *reads;                  /* the sub-batch's bases (HPC-compressed when HPC) */
    T_CSR_cap *reads_offs;        /* absolute offsets, reads_cnt + 1 */
    T_read_len *read_len;         /* HPC read lengths (HPC only) */
    T_read_len *hpc_orig_pos;     /* original in-read index per HPC base (HPC only) */
    T_minimizer *minimizers;      /* per batch-local base: window minimizer */
    T_read_len *superkmer_offs;   /* per read: [0, i1, .., len-k+1, 0] at the read's batch-local start */
    /* CountTask::StoreMinimizerPos only (the reference keeps these fields on
     * the device, gkc_cuda.cu:159-160, and never fills them); null otherwise.
     * Per batch-local base: the window minimizer's offset in the window (255
     * none) and strand (0 forward, 1 reverse complement, -1 tie or none). */
    T_kvalue *mm_pos;
    char *mm_strand;
};

GKC_DROPIN_END

typedef unsigned int T_kmer_cnt;
const T_kmer_cnt MAX_KMER_CNT = (T_kmer_cnt)0xFFFFFFFFu;
GKC_DROPIN_BEGIN
struct T_kmc {
    T_kmer kmer;
    T_kmer_cnt cnt;
};
/* k in (32, 64]: the reference has no 128-bit record; this one holds the
 * canonical k-mer as hi:lo. */
struct T_kmc128 {
    unsigned long long hi, lo;
    T_kmer_cnt cnt;
};
GKC_DROPIN_END

#endif
