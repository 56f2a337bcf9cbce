// gkc_count: a two-phase k-mer counter written only against the drop-in
// entry points (include/gkc/*.hpp), with the phase structure of the
// reference's driver (experiment.cpp:76-89, 156-219):
//   phase 1  batches of reads, sorted by length, pinned, GenSuperkmerGPU with
//            SKMStore::add_skms as callback, on a pool of host threads
//   phase 2  kmc_counting_GPU per partition on a pool of host threads
// Usage: gkc_count -read reads.fa [-k 31] [-p 11] [-skm 64] [-t 4] [-hpc]
//                  [-kmin 1] [-kmax 65535] [-batch 2000] [-out table.bin]
// The table file: u32 P, u32 k, then per partition u64 n_entries, u64
// n_distinct, and n_entries records (u64 hi, u64 lo, u32 count).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "gkc/gkc_cuda.hpp"
#include "gkc/kmer_counting.hpp"
#include "gkc/superkmers.hpp"

static bool by_len(const ReadPtr& a, const ReadPtr& b) { return a.len < b.len; }

int main(int argc, char** argv) {
    int K = 31, P = 11, SKM = 64, T = 4, batch = 2000;
    unsigned kmin = 1, kmax = 65535;
    bool hpc = false;
    std::string path, out;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> const char* { if (i + 1 >= argc) { std::cerr << "missing value for " << a << "\n"; std::exit(2); } return argv[++i]; };
        if (a == "-k") K = std::atoi(next());
        else if (a == "-p") P = std::atoi(next());
        else if (a == "-skm") SKM = std::atoi(next());
        else if (a == "-t") T = std::atoi(next());
        else if (a == "-hpc") hpc = true;
        else if (a == "-kmin") kmin = (unsigned)std::atoi(next());
        else if (a == "-kmax") kmax = (unsigned)std::atoi(next());
        else if (a == "-batch") batch = std::atoi(next());
        else if (a == "-read") path = next();
        else if (a == "-out") out = next();
        else { std::cerr << "unknown option " << a << "\n"; return 2; }
    }
    if (path.empty()) { std::cerr << "need -read file.fa\n"; return 2; }
    // single-line FASTA: every non-header line is one read
    std::ifstream f(path);
    if (!f) { std::cerr << "cannot open " << path << "\n"; return 1; }
    std::vector<std::string> lines;
    std::string line;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (!line.empty() && line[0] == '>') continue;
        lines.push_back(line);
    }
    std::vector<ReadPtr> all;
    all.reserve(lines.size());
    for (auto& s : lines) all.push_back(ReadPtr{s.data(), (T_read_len)s.size()});

    CUDAParams gpars{8, 256, 6, 0};
    GPUReset(gpars.device_id);
    std::vector<SKMStore*> stores;
    for (int i = 0; i < SKM; i++) stores.push_back(new SKMStore());

    auto t1 = std::chrono::steady_clock::now();
    std::atomic<size_t> next_batch{0};
    const size_t nb = (all.size() + batch - 1) / batch;
    auto phase1 = [&]() {
        for (;;) {
            size_t b = next_batch++;
            if (b >= nb) break;
            std::vector<ReadPtr> reads(all.begin() + b * batch, all.begin() + std::min(all.size(), (b + 1) * batch));
            std::sort(reads.begin(), reads.end(), by_len);
            PinnedCSR pinned(reads);
            GenSuperkmerGPU(pinned, K, P, hpc, gpars, SKMPartition, SKM, [&](T_h_data hd) {
                SKMStore::add_skms(stores, (T_kvalue)K, hd.reads_cnt, hd.reads, hd.reads_offs, hd.superkmer_offs,
                                   hd.minimizers);
            });
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++) th.emplace_back(phase1);
    for (auto& x : th) x.join();
    th.clear();
    auto t2 = std::chrono::steady_clock::now();

    std::vector<std::vector<T_kmc>> res64(SKM);
    std::vector<std::vector<T_kmc128>> res128(SKM);
    std::vector<size_t> distinct(SKM, 0);
    std::atomic<int> next_part{0};
    auto phase2 = [&]() {
        for (;;) {
            int q = next_part++;
            if (q >= SKM) break;
            if (K <= 32)
                distinct[q] = kmc_counting_GPU((T_kvalue)K, *stores[q], gpars.device_id, (unsigned short)kmin,
                                               (unsigned short)kmax, res64[q]);
            else
                distinct[q] = kmc_counting_GPU_128((T_kvalue)K, *stores[q], gpars.device_id, (unsigned short)kmin,
                                                   (unsigned short)kmax, res128[q]);
        }
    };
    for (int t = 0; t < T; t++) th.emplace_back(phase2);
    for (auto& x : th) x.join();
    auto t3 = std::chrono::steady_clock::now();
    size_t tot = 0;
    for (auto d : distinct) tot += d;
    std::cerr << "Phase 1 Time: " << std::chrono::duration<double>(t2 - t1).count() << " sec\n";
    std::cerr << "Phase 2 Time: " << std::chrono::duration<double>(t3 - t2).count() << " sec\n";
    std::cerr << "Total number of distinct kmers: " << tot << "\n";
    if (!out.empty()) {
        FILE* o = std::fopen(out.c_str(), "wb");
        if (!o) { std::cerr << "cannot write " << out << "\n"; return 1; }
        uint32_t hdr[2] = {(uint32_t)SKM, (uint32_t)K};
        std::fwrite(hdr, 4, 2, o);
        for (int q = 0; q < SKM; q++) {
            uint64_t n = K <= 32 ? res64[q].size() : res128[q].size(), d = distinct[q];
            std::fwrite(&n, 8, 1, o);
            std::fwrite(&d, 8, 1, o);
            for (uint64_t i = 0; i < n; i++) {
                uint64_t hi = K <= 32 ? 0 : res128[q][i].hi, lo = K <= 32 ? res64[q][i].kmer : res128[q][i].lo;
                uint32_t c = K <= 32 ? res64[q][i].cnt : res128[q][i].cnt;
                std::fwrite(&hi, 8, 1, o);
                std::fwrite(&lo, 8, 1, o);
                std::fwrite(&c, 4, 1, o);
            }
        }
        std::fclose(o);
    }
    return 0;
}
