/* Checks the exactness claim of the DV kernel's fused transforms (mmrecode_amd/csrc/dv.hip:
 * dv_row_pass / dv_column_pass with kDvRowA, kDvCol88, kDvCol248 and kDvTieWindow) against the
 * oracle's reference-order inverse_88 / inverse_248 (oracle_dv_inverse_88, oracle_dv_inverse_248):
 * on every block whose fused values (divided by 4) all lie farther than the window from a
 * half-integer, every sample must equal the oracle's. Test infrastructure
 * (tests/test_idct_fast_path.py); the cosine tables are the oracle's.
 *
 * usage: dv_fast_check <liboracle.so> <blocks> <seed> <88|248> [window]
 *        (prints: blocks flagged_blocks mismatches)
 * stdin (optional, after the count line): extra blocks, 64 integers each (natural order).
 *
 *        dv_fast_check <liboracle.so> search <88|248> <blocks> <seed> <want>
 * searches sparse seeded level blocks (the device's input: levels by scan position and the block's
 * info byte) whose dequantised coefficients the fused form, without the window, rounds differently
 * from the reference; prints up to <want> of them, one per line: info, then position:level pairs. */
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static double kWindow = 0x1p-25;
static double ROWA[64], COL88[64], COL248[64];

static uint64_t rng_state;
static uint64_t next(void) {   /* splitmix64 */
    uint64_t z = (rng_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

/* The kernel's tables (dv_upload_tables): one rounded product of the reference's factors each. */
static void tables(const double* cos16, const double* cos8) {
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) {
            volatile double cj = j == 0 ? M_SQRT1_2 : 1.0;
            ROWA[i * 8 + j] = cj * cos16[i * 8 + j];
            COL88[i * 8 + j] = cj * cos16[i * 8 + j];
        }
    for (int y = 0; y < 4; ++y)
        for (int f = 0; f < 2; ++f)
            for (int v = 0; v < 4; ++v) {
                volatile double cv = v == 0 ? M_SQRT1_2 : 1.0;
                double e = cv * cos8[y * 4 + v];
                COL248[(2 * y + f) * 8 + 2 * v] = e;
                COL248[(2 * y + f) * 8 + 2 * v + 1] = f ? -e : e;
            }
}

/* The kernel's fused path for one block: row pass (lane v), column pass (lane x) over all 8 rows,
 * values / 4. Returns 1 when a value lies within the window of a half-integer. */
static int fast(const int32_t c[64], int m248, double out[64]) {
    const double* B = m248 ? COL248 : COL88;
    double tmp[64];
    for (int v = 0; v < 8; ++v)
        for (int x = 0; x < 8; ++x) {
            double s = ROWA[x * 8] * (double)c[v * 8];
            for (int u = 1; u < 8; ++u) s = fma(ROWA[x * 8 + u], (double)c[v * 8 + u], s);
            tmp[v * 8 + x] = s;
        }
    int near = 0;
    for (int x = 0; x < 8; ++x)
        for (int y = 0; y < 8; ++y) {
            double s = B[y * 8] * tmp[x];
            for (int v = 1; v < 8; ++v) s = fma(B[y * 8 + v], tmp[v * 8 + x], s);
            s = s * 0.25;
            out[y * 8 + x] = s;
            double a = fabs(s), f = a - floor(a);
            near |= fabs(f - 0.5) <= kWindow;
        }
    return near;
}

static uint8_t round_clamp(double v) {
    double r = round(v);
    return (uint8_t)(r < 0 ? 0 : r > 255 ? 255 : r);
}

static int search(void (*ref)(const int32_t*, uint8_t*), void (*deq)(const int16_t*, uint8_t, int32_t*), int m248,
                  long n, int want) {
    const double window = kWindow;
    int found = 0;
    for (long b = 0; b < n && found < want; ++b) {
        int16_t lv[64] = {0};
        lv[0] = (int16_t)((int)(next() % 512) - 256);
        const int k = (int)(next() % 4);
        for (int j = 0; j < k; ++j) lv[1 + next() % 63] = (int16_t)((int)(next() % 9) - 4);
        const uint8_t info = (uint8_t)(m248 | (next() % 4) << 1 | (next() % 16) << 3);
        int32_t c[64];
        double f[64];
        uint8_t out[64];
        deq(lv, info, c);
        kWindow = -1.0;
        fast(c, m248, f);
        kWindow = window;
        ref(c, out);
        int bad = 0;
        for (int i = 0; i < 64; ++i) bad += round_clamp(f[i]) != out[i];
        if (!bad) continue;
        printf("%d", info);
        for (int i = 0; i < 64; ++i)
            if (lv[i]) printf(" %d:%d", i, lv[i]);
        printf("\n");
        ++found;
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 5) return 2;
    void* lib = dlopen(argv[1], RTLD_NOW);
    if (!lib) { fprintf(stderr, "%s\n", dlerror()); return 2; }
    const int searching = argv[2][0] == 's';
    const int m248 = argv[searching ? 3 : 4][0] == '2';
    if (!searching && argc > 5) kWindow = atof(argv[5]);
    void (*ref88)(const int32_t*, uint8_t*) = (void (*)(const int32_t*, uint8_t*))dlsym(lib, "oracle_dv_inverse_88");
    void (*ref248)(const int32_t*, uint8_t*) = (void (*)(const int32_t*, uint8_t*))dlsym(lib, "oracle_dv_inverse_248");
    void (*deq)(const int16_t*, uint8_t, int32_t*) = (void (*)(const int16_t*, uint8_t, int32_t*))dlsym(lib, "oracle_dv_dequantize");
    const double* (*c16)(void) = (const double* (*)(void))dlsym(lib, "oracle_dv_cos16_table");
    const double* (*c8)(void) = (const double* (*)(void))dlsym(lib, "oracle_dv_cos8_table");
    if (!ref88 || !ref248 || !deq || !c16 || !c8) return 2;
    tables(c16(), c8());
    if (searching) {
        if (argc < 7) return 2;
        rng_state = strtoull(argv[5], 0, 10);
        return search(m248 ? ref248 : ref88, deq, m248, atol(argv[4]), atoi(argv[6]));
    }
    long n = atol(argv[2]);
    rng_state = strtoull(argv[3], 0, 10);
    long flagged = 0, bad = 0, count = 0;
    int32_t c[64];
    double f[64];
    uint8_t want[64];
    for (long b = 0;; ++b) {
        if (b < n) {
            const uint64_t kind = next() % 8;
            for (int i = 0; i < 64; ++i) c[i] = 0;
            if (kind == 0) {            /* dense, up to the dequantiser's range (|c| < 2^17) */
                for (int i = 0; i < 64; ++i) c[i] = (int32_t)(next() % 262143) - 131071;
            } else if (kind == 1) {     /* the largest sum |c|: every AC at the range end, DC too */
                for (int i = 1; i < 64; ++i) c[i] = (next() & 1) ? 131071 : -131072;
                c[0] = (next() & 1) ? 131068 + 1024 : -131072 + 1024;
            } else if (kind < 5) {      /* the dequantiser on seeded levels, as the benchmark draws them */
                int16_t lv[64];
                lv[0] = (int16_t)((int)(next() % 512) - 256);
                for (int i = 1; i < 64; ++i) {
                    int mag = 0;
                    if (next() % 5 == 0) { mag = 1; while (next() % 3 != 0) ++mag; }
                    lv[i] = (int16_t)((next() & 1) ? mag : -mag);
                }
                const uint8_t info = (uint8_t)(m248 | (next() % 4) << 1 | (next() % 16) << 3);
                deq(lv, info, c);
            } else {                    /* sparse, small: where exact ties live */
                const int k = 1 + (int)(next() % 6);
                for (int j = 0; j < k; ++j) c[next() % 64] = ((int32_t)(next() % 25) - 12) * (1 << (next() % 4));
                c[0] = (int32_t)(next() % 2048);
            }
        } else {
            int ok = 1;
            for (int i = 0; i < 64 && ok; ++i) ok = scanf("%d", &c[i]) == 1;
            if (!ok) break;
        }
        ++count;
        const int near = fast(c, m248, f);
        if (m248) ref248(c, want);
        else ref88(c, want);
        if (near) { ++flagged; continue; }
        for (int i = 0; i < 64; ++i)
            if (round_clamp(f[i]) != want[i]) ++bad;
    }
    printf("%ld %ld %ld\n", count, flagged, bad);
    return 0;
}
