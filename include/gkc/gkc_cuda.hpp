/*
 * Drop-in for the reference's superkmer-partitioning entry point
 * (reference gkc_cuda.hpp:25-61).  Same names, signatures and argument
 * meaning; the work runs in the gfx950 HIP library through include/gkc_hip.h.
 * Errors print the library's message and exit(1), as the reference does
 * (its CUDA_CHECK, gkc_cuda.hpp:4-9).
 */
#ifndef GKC_DROPIN_GKC_CUDA_HPP
#define GKC_DROPIN_GKC_CUDA_HPP

#include <atomic>
#include <functional>
#include <string>
#include <vector>

#include "types.h"

GKC_DROPIN_BEGIN
struct CUDAParams {
    int NUM_BLOCKS_PER_GRID, NUM_THREADS_PER_BLOCK;   /* reads per sub-batch = their product */
    int n_streams;                                    /* accepted; the library pipelines on its own stream */
    int device_id;
};

/* Reads in a pinned host CSR (one-time use). */
class PinnedCSR {
public:
    char *reads_CSR;
    T_CSR_cap *reads_offs;
    T_read_cnt n_reads;
    T_CSR_cap size_capacity;
    PinnedCSR(std::vector<std::string> &reads);
    PinnedCSR(std::vector<ReadPtr> &reads, bool keep_original = true);
    ~PinnedCSR();
    char *get_reads_CSR() { return reads_CSR; }
    T_CSR_cap *get_reads_offs() { return reads_offs; }
    T_read_cnt get_n_reads() { return n_reads; }
    T_CSR_cap size() { return size_capacity; }
};

enum CountTask { SKMPartition, SKMPartWithPos, StoreMinimizerPos };
GKC_DROPIN_END

/* Bases per partition of one sub-batch (reference gkc_cuda.cu:350-365). */
void CalcSKMPartSize_instream(T_read_cnt reads_cnt, T_read_len *superkmer_offs, T_CSR_cap *reads_offs,
                              T_minimizer *minimizers, int n_partitions, int k, std::atomic<size_t> part_sizes[]);

/* Select the device and wait for it.  Unlike the reference (cudaDeviceReset)
 * it does not tear down the device's contexts, so it is safe next to other
 * users of the GPU. */
void GPUReset(int did);

/* Minimizers and superkmer offset lists of every read, handed to
 * process_func once per sub-batch of NUM_BLOCKS_PER_GRID*NUM_THREADS_PER_BLOCK
 * reads (reference gkc_cuda.cu:376-541).  The callback's arrays are valid
 * only during the call; with HPC the pinned reads are overwritten by the
 * compressed bases, as in the reference. */
void GenSuperkmerGPU(PinnedCSR &pinned_reads, int K_kmer, int P_minimizer, bool HPC, CUDAParams gpars,
                     CountTask task, int SKM_partitions, std::function<void(T_h_data)> process_func);

#endif
