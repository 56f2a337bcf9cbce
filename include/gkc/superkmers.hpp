/*
 * Per-partition superkmer store of the drop-in (the reference's SKMStore,
 * superkmers.hpp:138-187).  Superkmers are kept 2-bit packed, 16 bases per
 * 32-bit word MSB first, each starting on a word boundary -- the layout the
 * HIP counter reads directly.  Thread-safe appends; kmc_counting_GPU consumes
 * and deletes the store, as the reference's does.
 */
#ifndef GKC_DROPIN_SUPERKMERS_HPP
#define GKC_DROPIN_SUPERKMERS_HPP

#include <atomic>
#include <cstdint>
#include <mutex>
#include <vector>

#include "types.h"

GKC_DROPIN_BEGIN
class SKMStore {
public:
    std::atomic<size_t> tot_size{0};        /* total bases added */
    SKMStore() {}
    ~SKMStore() {}
    /* Append one superkmer of `len` ASCII bases (a superkmer holding bytes
     * other than A/C/G/T is stored as its runs of A/C/G/T bytes). */
    void add_skm(const char *beg, uint32_t len);
    /* Number of superkmers stored. */
    size_t size() const;
    /* Partition scatter of one sub-batch (reference superkmers.hpp:167-186):
     * superkmer j of read r spans bases [s_{j-1}, s_j + k - 1) of the read and
     * goes to partition minimizers[start] % skm_partitions.size(). */
    static void add_skms(std::vector<SKMStore *> &skm_partitions, T_kvalue k, T_read_cnt reads_cnt,
                         const char *csr_reads, T_CSR_cap *reads_offs, T_read_len *superkmer_offs,
                         T_minimizer *minimizers);
    /* Packed contents (for the counter). */
    void take(std::vector<uint32_t> &words, std::vector<uint32_t> &lens);

private:
    void add_run(const char *beg, uint32_t len);
    mutable std::mutex mu_;
    std::vector<uint32_t> words_;
    std::vector<uint32_t> lens_;
};
GKC_DROPIN_END

#endif
