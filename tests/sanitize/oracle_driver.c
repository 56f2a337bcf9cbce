/* Host sanitizer driver for the CPU oracle (oracle/gkc_oracle.c): ragged
 * reads (empty, shorter than k, up to 460 bases) holding N runs, lowercase and
 * IUPAC bytes; HPC, per-window records and superkmer lists for three (k, p)
 * including k > 32; multi-threaded counting with and without thresholds.
 * Built with -fsanitize=address,undefined and with -fsanitize=thread (the
 * latter runs the first configuration only: argv[1] = 1). */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
void or_hpc(const uint8_t *reads, const uint64_t *offs, int64_t n, uint8_t *out, int32_t *rl, int32_t *pos);
void or_window_minimizers(const uint8_t *reads, const uint64_t *offs, const int32_t *lens, int64_t n, int k, int p,
                          int fmode, uint64_t *mm, uint8_t *pos, int8_t *strand);
void or_superkmer_offs(const uint64_t *offs, const int32_t *lens, int64_t n, int k, const uint64_t *mm, int32_t *skm);
int64_t or_count(const uint8_t *reads, const uint64_t *offs, int64_t n, int k, int p, int hpc, uint32_t P, int fmode,
                 int pmode, uint32_t minc, uint32_t maxc, int threads, uint64_t *lo, uint64_t *hi, uint32_t *cnt,
                 uint64_t cap, uint64_t *part_off, uint64_t *pk, uint64_t *pd, uint64_t *ps, double *t);
static uint64_t st = 88172645463325252ull;
static uint64_t rnd(void) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; }
int main(int argc, char **argv) {
    const int n_cfg = argc > 1 ? atoi(argv[1]) : 3;     /* (k, p) configurations run */
    const int64_t n = 600;
    uint64_t *offs = malloc((n + 1) * 8);
    offs[0] = 0;
    for (int64_t r = 0; r < n; r++) offs[r + 1] = offs[r] + (r % 50 == 0 ? rnd() % 20 : 60 + rnd() % 400);
    const uint64_t B = offs[n];
    uint8_t *reads = malloc(B + 1);
    for (uint64_t i = 0; i < B; i++) {
        uint64_t x = rnd() % 1000;
        reads[i] = x < 5 ? "NnRY-"[x] : "ACGTacgt"[x % 8];
        if (x < 2 && i + 30 < B) { memset(reads + i, 'N', 1 + rnd() % 30); i += 29; }
    }
    uint8_t *h = malloc(B + 1); int32_t *rl = malloc(n * 4), *pos = malloc((B + 1) * 4);
    or_hpc(reads, offs, n, h, rl, pos);
    uint64_t *mm = malloc((B + 1) * 8); uint8_t *mp = malloc(B + 1); int8_t *ms = malloc(B + 1);
    int32_t *skm = malloc((B + 1) * 4);
    const int ks[3][2] = {{31, 11}, {63, 19}, {21, 7}};
    for (int c = 0; c < n_cfg; c++) {
        for (int hp = 0; hp < 2; hp++) {
            memset(mm, 0xFF, (B + 1) * 8);
            or_window_minimizers(hp ? h : reads, offs, hp ? rl : NULL, n, ks[c][0], ks[c][1], 1, mm, mp, ms);
            or_superkmer_offs(offs, hp ? rl : NULL, n, ks[c][0], mm, skm);
        }
        uint64_t cap = B, *lo = malloc(cap * 8), *hi = malloc(cap * 8), po[65], pk[64], pd[64], ps[64];
        uint32_t *cnt = malloc(cap * 4); double t[2];
        for (int hp = 0; hp < 2; hp++) {
            int64_t r = or_count(reads, offs, n, ks[c][0], ks[c][1], hp, 64, 1, hp, c == 2 ? 2 : 1, c == 2 ? 40 : 0, 4, lo,
                                 ks[c][0] > 32 ? hi : NULL, cnt, cap, po, pk, pd, ps, t);
            if (r < 0) return 2;
            printf("k=%d hpc=%d entries=%lld\n", ks[c][0], hp, (long long)r);
        }
        free(lo); free(hi); free(cnt);
    }
    free(offs); free(reads); free(h); free(rl); free(pos); free(mm); free(mp); free(ms); free(skm);
    return 0;
}
