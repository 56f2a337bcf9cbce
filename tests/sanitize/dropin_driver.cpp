// Host sanitizer driver for the C++ drop-in's host code (dropin.cpp,
// included whole so its internal helpers are reachable): split_reads on
// reads holding other bytes, CalcSKMPartSize_instream and SKMStore::add_skms
// from four threads into shared stores over the oracle's superkmer arrays;
// the bases stored per partition must equal the sizes computed.  Built with
// -fsanitize=address,undefined and with -fsanitize=thread.
#include <thread>
#include "dropin.cpp"
extern "C" {
void or_window_minimizers(const uint8_t *reads, const uint64_t *offs, const int32_t *lens, int64_t n, int k, int p,
                          int fmode, uint64_t *mm, uint8_t *pos, int8_t *strand);
void or_superkmer_offs(const uint64_t *offs, const int32_t *lens, int64_t n, int k, const uint64_t *mm, int32_t *skm);
}
static uint64_t st = 88172645463325252ull;
static uint64_t rnd() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; }
int main() {
    const int n = 600, k = 31, p = 11, P = 16;
    std::vector<T_CSR_cap> offs(n + 1, 0);
    for (int r = 0; r < n; r++) offs[r + 1] = offs[r] + (r % 40 == 0 ? rnd() % 25 : 80 + rnd() % 300);
    const size_t B = offs[n];
    std::vector<char> reads(B + 1, 0), dirty;
    for (size_t i = 0; i < B; i++) reads[i] = "ACGT"[rnd() % 4];
    dirty = reads;
    for (size_t i = 0; i < B; i += 1 + rnd() % 200) dirty[i] = "NnRY-."[rnd() % 6];
    // reads cut at the other bytes
    std::vector<char> frag; std::vector<T_CSR_cap> fo;
    split_reads(dirty.data(), offs.data(), n, k, frag, fo);
    size_t fk = 0;
    for (size_t r = 0; r + 1 < fo.size(); r++) fk += fo[r + 1] - fo[r] - k + 1;
    // superkmer arrays of the clean reads (oracle), then the host partition sizes and concurrent add_skms
    std::vector<uint64_t> mm(B + 1, ~0ull);
    std::vector<uint8_t> mp(B + 1); std::vector<int8_t> ms(B + 1);
    std::vector<int32_t> skm(B + 2, -7);
    or_window_minimizers((const uint8_t*)reads.data(), (const uint64_t*)offs.data(), nullptr, n, k, p, 1, mm.data(),
                         mp.data(), ms.data());
    or_superkmer_offs((const uint64_t*)offs.data(), nullptr, n, k, mm.data(), skm.data());
    std::vector<T_minimizer> m32(B + 1);
    for (size_t i = 0; i <= B; i++) m32[i] = (T_minimizer)mm[i];
    std::vector<std::atomic<size_t>> sizes(P);
    for (auto& s : sizes) s = 0;
    CalcSKMPartSize_instream(n, skm.data(), offs.data(), m32.data(), P, k, sizes.data());
    std::vector<SKMStore> stores(P);
    std::vector<SKMStore*> sp;
    for (auto& s : stores) sp.push_back(&s);
    const int T = 4;
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++)
        th.emplace_back([&, t] {
            const int a = n * t / T, b = n * (t + 1) / T;
            std::vector<T_CSR_cap> o(offs.begin() + a, offs.begin() + b + 1);
            SKMStore::add_skms(sp, (T_kvalue)k, b - a, reads.data() + o[0], o.data(), skm.data() + o[0], m32.data() + o[0]);
        });
    for (auto& x : th) x.join();
    size_t tot = 0, want = 0;
    for (int q = 0; q < P; q++) {
        std::vector<uint32_t> w, l;
        stores[q].take(w, l);
        size_t bases = 0;
        for (uint32_t x : l) bases += x;
        if (bases != sizes[q]) { std::printf("partition %d: %zu bases stored, %zu sized\n", q, bases, (size_t)sizes[q]); return 3; }
        tot += bases; want += sizes[q];
    }
    std::printf("fragments %zu (k-mers %zu), stored bases %zu\n", fo.size() - 1, fk, tot);
    return tot == want ? 0 : 4;
}
