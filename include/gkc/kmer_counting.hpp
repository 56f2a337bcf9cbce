/*
 * Drop-in for the reference's partition-counting entry point
 * (reference kmer_counting.hpp:9-12).  Counts one partition's superkmers on
 * the GPU (extraction, onesweep radix sort, run-length count) and returns the
 * number of distinct canonical k-mers; fills kmc_result_curthread with the
 * (k-mer, count) entries whose count lies in [kmer_min_freq, kmer_max_freq],
 * keys ascending (kmer_max_freq == 0xFFFF is taken as "no upper bound").
 * The store is consumed and deleted, as the reference does
 * (kmer_counting.cu:197).
 */
#ifndef GKC_DROPIN_KMER_COUNTING_HPP
#define GKC_DROPIN_KMER_COUNTING_HPP

#include <vector>

#include "superkmers.hpp"
#include "types.h"

size_t kmc_counting_GPU(T_kvalue k, SKMStore &skms_store, int gpuid, unsigned short kmer_min_freq,
                        unsigned short kmer_max_freq, std::vector<T_kmc> &kmc_result_curthread);

/* k in (32, 64]: the same count with 128-bit records. */
size_t kmc_counting_GPU_128(T_kvalue k, SKMStore &skms_store, int gpuid, unsigned short kmer_min_freq,
                            unsigned short kmer_max_freq, std::vector<T_kmc128> &kmc_result_curthread);

#endif
