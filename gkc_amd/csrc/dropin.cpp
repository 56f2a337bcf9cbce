// C++ drop-in wrappers with the reference's signatures (include/gkc/*.hpp)
// over the C ABI of the HIP library (include/gkc_hip.h).  Host code only.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "gkc/gkc_cuda.hpp"
#include "gkc/kmer_counting.hpp"
#include "gkc/superkmers.hpp"
#include "gkc_hip.h"

namespace {

[[noreturn]] void die(const char* what) {
    std::fprintf(stderr, "GKC error calling \"%s\": %s\n", what, gkc_last_error());
    std::exit(1);
}

struct CtxDeleter {
    void operator()(gkc_ctx* c) const { gkc_destroy(c); }
};
using CtxPtr = std::unique_ptr<gkc_ctx, CtxDeleter>;

// One context per (thread, device, parameters): the reference calls both
// entry points from host thread pools (read_loader.hpp:494-497,
// experiment.cpp:191-201); contexts are not shared between threads.
gkc_ctx* thread_ctx(int device, int k, int p, bool hpc, unsigned min_cnt, unsigned max_cnt) {
    thread_local std::map<std::tuple<int, int, int, bool, unsigned, unsigned>, CtxPtr> cache;
    auto key = std::make_tuple(device, k, p, hpc, min_cnt, max_cnt);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second.get();
    gkc_params prm{};
    prm.k = k; prm.p = p; prm.hpc = hpc ? 1 : 0; prm.n_parts = 1; prm.filter_mode = 1; prm.part_mode = 0;
    prm.min_count = min_cnt; prm.max_count = max_cnt;
    gkc_ctx* c = gkc_create(device, &prm);
    if (!c) die("gkc_create");
    cache.emplace(key, CtxPtr(c));
    return c;
}

inline bool is_acgt(unsigned char ch) {
    switch (ch | 0x20) {
        case 'a': case 'c': case 'g': case 't': return true;
        default: return false;
    }
}

// The reads of one sub-batch cut at every byte other than A/C/G/T: the runs
// of at least k such bytes, back to back, with offsets from 0 (a run shorter
// than k holds no k-mer, before or after HPC).
void split_reads(const char* reads, const T_CSR_cap* offs, T_read_cnt n, int k, std::vector<char>& out,
                 std::vector<T_CSR_cap>& out_offs) {
    out.clear();
    out_offs.assign(1, 0);
    for (T_read_cnt r = 0; r < n; r++) {
        const char* s = reads + (offs[r] - offs[0]);
        const T_CSR_cap L = offs[r + 1] - offs[r];
        for (T_CSR_cap a = 0; a < L;) {
            T_CSR_cap e = a;
            while (e < L && is_acgt((unsigned char)s[e])) e++;
            if (e - a >= (T_CSR_cap)k) {
                out.insert(out.end(), s + a, s + e);
                out_offs.push_back(out.size());
            }
            a = e + 1;
        }
    }
}

}  // namespace

// ---------------------------------------------------------------- PinnedCSR

PinnedCSR::PinnedCSR(std::vector<std::string>& reads) {
    n_reads = (T_read_cnt)reads.size();
    size_capacity = 0;
    for (auto& r : reads) size_capacity += r.size();
    reads_offs = static_cast<T_CSR_cap*>(gkc_host_alloc((n_reads + 1) * sizeof(T_CSR_cap)));
    reads_CSR = static_cast<char*>(gkc_host_alloc(size_capacity + 1));
    if (!reads_offs || !reads_CSR) die("gkc_host_alloc");
    T_CSR_cap o = 0;
    reads_offs[0] = 0;
    for (size_t i = 0; i < reads.size(); i++) {
        std::memcpy(reads_CSR + o, reads[i].data(), reads[i].size());
        o += reads[i].size();
        reads_offs[i + 1] = o;
    }
}

PinnedCSR::PinnedCSR(std::vector<ReadPtr>& reads, bool) {
    n_reads = (T_read_cnt)reads.size();
    size_capacity = 0;
    for (auto& r : reads) size_capacity += (T_CSR_cap)r.len;
    reads_offs = static_cast<T_CSR_cap*>(gkc_host_alloc((n_reads + 1) * sizeof(T_CSR_cap)));
    reads_CSR = static_cast<char*>(gkc_host_alloc(size_capacity + 1));
    if (!reads_offs || !reads_CSR) die("gkc_host_alloc");
    T_CSR_cap o = 0;
    reads_offs[0] = 0;
    for (size_t i = 0; i < reads.size(); i++) {
        std::memcpy(reads_CSR + o, reads[i].read, (size_t)reads[i].len);
        o += (T_CSR_cap)reads[i].len;
        reads_offs[i + 1] = o;
    }
}

PinnedCSR::~PinnedCSR() {
    gkc_host_free(reads_offs);
    gkc_host_free(reads_CSR);
}

// ---------------------------------------------------------------- partitioning

void CalcSKMPartSize_instream(T_read_cnt reads_cnt, T_read_len* superkmer_offs, T_CSR_cap* reads_offs,
                              T_minimizer* minimizers, int n_partitions, int k, std::atomic<size_t> part_sizes[]) {
    const T_CSR_cap base = reads_offs[0];
    for (T_read_cnt r = 0; r < reads_cnt; r++) {
        T_CSR_cap idx = reads_offs[r] - base + 1;
        while (superkmer_offs[idx] != 0) {
            const T_read_len len = superkmer_offs[idx] - superkmer_offs[idx - 1] + k - 1;
            part_sizes[minimizers[superkmer_offs[idx - 1] + reads_offs[r] - base] % n_partitions] += len;
            idx++;
        }
    }
}

void GPUReset(int did) {
    if (gkc_device_select(did) != 0) die("GPUReset");
}

void GenSuperkmerGPU(PinnedCSR& pinned_reads, int K_kmer, int P_minimizer, bool HPC, CUDAParams gpars,
                     CountTask task, int SKM_partitions, std::function<void(T_h_data)> process_func) {
    // SKMPartition and SKMPartWithPos hand on the same arrays (a superkmer's
    // read and offset are its place in them); StoreMinimizerPos adds each
    // window's minimizer position and strand
    (void)SKM_partitions;
    gkc_ctx* c = thread_ctx(gpars.device_id, K_kmer, P_minimizer, HPC, 1, 0);
    const T_read_cnt per = std::max(1, gpars.NUM_BLOCKS_PER_GRID * gpars.NUM_THREADS_PER_BLOCK);
    for (T_read_cnt cur = 0; cur < pinned_reads.n_reads; cur += per) {
        const T_read_cnt end = std::min(cur + per, pinned_reads.n_reads);
        T_CSR_cap* offs = pinned_reads.reads_offs + cur;
        T_CSR_cap bytes = offs[end - cur] - offs[0];
        char* reads = pinned_reads.reads_CSR + offs[0];
        T_read_cnt cnt = end - cur;
        // reads holding bytes other than A/C/G/T are handed on cut at them (KMC's
        // rule; the reference codes such bytes 255 in its minimizers and as T in
        // CRead, gkc_cuda.cu:25-58, superkmers.hpp:67): the callback then sees the
        // fragments of >= K_kmer bases as its reads, in a buffer this call owns
        std::vector<char> frag;
        std::vector<T_CSR_cap> frag_offs;
        bool clean = true;
        for (T_CSR_cap i = 0; i < bytes && clean; i++) clean = is_acgt((unsigned char)reads[i]);
        if (!clean) {
            split_reads(reads, offs, cnt, K_kmer, frag, frag_offs);
            frag.push_back(0);
            reads = frag.data();
            offs = frag_offs.data();
            cnt = (T_read_cnt)(frag_offs.size() - 1);
            bytes = frag_offs.back();
        }
        T_h_data hd{};
        hd.reads_cnt = cnt;
        hd.reads = reads;                 // overwritten in place by the (HPC) bases
        hd.reads_offs = offs;
        hd.minimizers = new T_minimizer[bytes + 1];
        hd.superkmer_offs = new T_read_len[bytes + 1];
        if (HPC) {
            hd.read_len = new T_read_len[hd.reads_cnt];
            hd.hpc_orig_pos = new T_read_len[bytes + 1];
        }
        if (task == StoreMinimizerPos) {  // before the HPC bases overwrite the reads
            hd.mm_pos = new T_kvalue[bytes + 1];
            hd.mm_strand = new char[bytes + 1];
            if (gkc_window_minimizers(c, reads, reinterpret_cast<const uint64_t*>(offs), hd.reads_cnt, hd.minimizers,
                                      hd.mm_pos, reinterpret_cast<int8_t*>(hd.mm_strand)) != 0)
                die("GenSuperkmerGPU");
        }
        if (gkc_superkmer_arrays(c, reads, reinterpret_cast<const uint64_t*>(offs), hd.reads_cnt, hd.minimizers,
                                 hd.superkmer_offs, reads, hd.read_len, hd.hpc_orig_pos) != 0)
            die("GenSuperkmerGPU");
        process_func(hd);
        delete[] hd.minimizers;
        delete[] hd.superkmer_offs;
        delete[] hd.read_len;
        delete[] hd.hpc_orig_pos;
        delete[] hd.mm_pos;
        delete[] hd.mm_strand;
    }
}

// ---------------------------------------------------------------- SKMStore

void SKMStore::add_skm(const char* beg, uint32_t len) {
    // a superkmer holding bytes other than A/C/G/T is stored as its runs of
    // A/C/G/T bytes (no k-mer holds such a byte)
    for (uint32_t a = 0; a < len;) {
        uint32_t e = a;
        while (e < len && is_acgt((unsigned char)beg[e])) e++;
        if (e > a) add_run(beg + a, e - a);
        a = e + 1;
    }
}

void SKMStore::add_run(const char* beg, uint32_t len) {
    const uint32_t nw = (len + 15) / 16;
    uint32_t tmp[64];
    std::vector<uint32_t> big;
    uint32_t* w = tmp;
    if (nw > 64) { big.resize(nw); w = big.data(); }
    for (uint32_t i = 0; i < nw; i++) {
        uint32_t x = 0;
        for (uint32_t t = 0; t < 16; t++) {
            const uint32_t b = i * 16 + t;
            uint32_t code = 0;
            if (b < len) {
                const unsigned char ch = (unsigned char)beg[b];
                code = ((ch >> 2) ^ (ch >> 1)) & 3u;       // A/a 0, C/c 1, G/g 2, T/t 3
            }
            x |= code << (30 - 2 * t);
        }
        w[i] = x;
    }
    std::lock_guard<std::mutex> lk(mu_);
    words_.insert(words_.end(), w, w + nw);
    lens_.push_back(len);
    tot_size += len;
}

size_t SKMStore::size() const {
    std::lock_guard<std::mutex> lk(mu_);
    return lens_.size();
}

void SKMStore::take(std::vector<uint32_t>& words, std::vector<uint32_t>& lens) {
    std::lock_guard<std::mutex> lk(mu_);
    words.swap(words_);
    lens.swap(lens_);
    words_.clear();
    lens_.clear();
}

void SKMStore::add_skms(std::vector<SKMStore*>& skm_partitions, T_kvalue k, T_read_cnt reads_cnt, const char* csr_reads,
                        T_CSR_cap* reads_offs, T_read_len* superkmer_offs, T_minimizer* minimizers) {
    const size_t P = skm_partitions.size();
    const T_CSR_cap base = reads_offs[0];
    for (T_read_cnt r = 0; r < reads_cnt; r++) {
        T_CSR_cap idx = reads_offs[r] - base + 1;
        while (superkmer_offs[idx] != 0) {
            const uint32_t len = (uint32_t)(superkmer_offs[idx] - superkmer_offs[idx - 1] + k - 1);
            const T_CSR_cap start = superkmer_offs[idx - 1] + reads_offs[r] - base;
            skm_partitions[minimizers[start] % P]->add_skm(csr_reads + start, len);
            idx++;
        }
    }
}

// ---------------------------------------------------------------- counting

template <typename Rec, typename Emit>
static size_t count_store(T_kvalue k, SKMStore& store, int gpuid, unsigned short mn, unsigned short mx, Emit&& emit) {
    std::vector<uint32_t> words, lens;
    store.take(words, lens);
    delete &store;
    uint64_t nk = 0;
    for (uint32_t L : lens) if (L >= k) nk += L - k + 1;
    if (nk == 0) return 0;
    const unsigned maxc = (mx == 0xFFFF) ? 0u : (unsigned)mx;
    gkc_ctx* c = thread_ctx(gpuid, k, std::min<int>(k, 3), false, mn ? mn : 1, maxc);
    std::vector<uint64_t> lo(nk), hi(k > 32 ? nk : 0);
    std::vector<uint32_t> cnt(nk);
    uint64_t ne = 0, nd = 0;
    if (gkc_count_packed(c, words.data(), words.size(), lens.data(), lens.size(), lo.data(),
                         k > 32 ? hi.data() : nullptr, cnt.data(), nk, &ne, &nd) != 0)
        die("kmc_counting_GPU");
    for (uint64_t i = 0; i < ne; i++) emit(lo[i], k > 32 ? hi[i] : 0ull, cnt[i]);
    return (size_t)nd;
}

size_t kmc_counting_GPU(T_kvalue k, SKMStore& skms_store, int gpuid, unsigned short kmer_min_freq,
                        unsigned short kmer_max_freq, std::vector<T_kmc>& kmc_result_curthread) {
    if (k > 32) { std::fprintf(stderr, "kmc_counting_GPU: k > 32 needs kmc_counting_GPU_128\n"); std::exit(1); }
    return count_store<T_kmc>(k, skms_store, gpuid, kmer_min_freq, kmer_max_freq,
                              [&](uint64_t lo, uint64_t, uint32_t c) { kmc_result_curthread.push_back(T_kmc{lo, c}); });
}

size_t kmc_counting_GPU_128(T_kvalue k, SKMStore& skms_store, int gpuid, unsigned short kmer_min_freq,
                            unsigned short kmer_max_freq, std::vector<T_kmc128>& kmc_result_curthread) {
    return count_store<T_kmc128>(k, skms_store, gpuid, kmer_min_freq, kmer_max_freq,
                                 [&](uint64_t lo, uint64_t hi, uint32_t c) { kmc_result_curthread.push_back(T_kmc128{hi, lo, c}); });
}
