/* Checks the exactness claim of the device's fast IDCT (mmrecode_amd/csrc/device_common.h,
 * idct8x8_row_pass_fast / idct8x8_column_pass_fast / idct_near_half, used by mpeg2_kernel and
 * mjpeg_blocks_kernel) against the oracle's reference-order IDCTs (oracle_mpeg2_inverse_dct,
 * oracle_mjpeg_inverse_dct): on every block, each sample whose fused evaluation (MJPEG: plus 128)
 * lies farther than the window from a half-integer must round to the oracle's sample. MJPEG blocks
 * with a row whose sum of |c| passes kRowAbsLimit (or a block sum above 8 times it) are flagged
 * too, as the kernel sends them to the reference order. Test infrastructure
 * (tests/test_idct_fast_path.py); the basis is the oracle's table.
 *
 * usage: idct_fast_check <liboracle.so> <blocks> <seed> [mpeg2|mjpeg]
 *        (prints: blocks flagged_blocks mismatches)
 * stdin (optional, after the count line): extra blocks, 64 integers each. */
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/tables.h"

#define M ORACLE_IDCT_MATRIX
static const double kWindow = 0x1p-30;
static const uint32_t kRowAbsLimit = 391168 / 8;
static int mjpeg;

static uint64_t rng_state;
static uint64_t next(void) {   /* splitmix64 */
    uint64_t z = (rng_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

/* The device's fast path: row pass lane v, column pass lane x over all 8 rows, fused steps, then
 * the near-half test. Returns 1 when any sample is near. */
static int fast_idct(const int32_t c[64], double out[64]) {
    double tmp[64];
    for (int v = 0; v < 8; ++v)
        for (int x = 0; x < 8; ++x) {
            double s = M[x][0] * (double)c[v * 8];
            for (int u = 1; u < 8; ++u) s = fma(M[x][u], (double)c[v * 8 + u], s);
            tmp[v * 8 + x] = s;
        }
    int near = 0;
    for (int x = 0; x < 8; ++x)
        for (int y = 0; y < 8; ++y) {
            double s = M[y][0] * tmp[x];
            for (int v = 1; v < 8; ++v) s = fma(M[y][v], tmp[v * 8 + x], s);
            if (mjpeg) s = s + 128.0;
            out[y * 8 + x] = s;
            double a = fabs(s), f = a - floor(a);
            near |= fabs(f - 0.5) <= kWindow;
        }
    if (mjpeg)
        for (int v = 0; v < 8; ++v) {
            uint64_t sum = 0;
            for (int u = 0; u < 8; ++u) sum += (uint64_t)llabs((long long)c[v * 8 + u]);
            near |= sum > kRowAbsLimit;
        }
    return near;
}

static int16_t round_clamp(double v) {
    double r = round(v);
    if (mjpeg) return (int16_t)(r < 0 ? 0 : r > 255 ? 255 : r);
    return (int16_t)(r < -256 ? -256 : r > 255 ? 255 : r);
}

static void mjpeg_ref(void (*f)(const int32_t*, uint8_t*), const int32_t* c, int16_t* out) {
    uint8_t o[64];
    f(c, o);
    for (int i = 0; i < 64; ++i) out[i] = o[i];
}

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    void* lib = dlopen(argv[1], RTLD_NOW);
    if (!lib) { fprintf(stderr, "%s\n", dlerror()); return 2; }
    mjpeg = argc > 4 && argv[4][1] == 'j';
    void (*ref)(const int32_t*, int16_t*) = (void (*)(const int32_t*, int16_t*))dlsym(lib, "oracle_mpeg2_inverse_dct");
    void (*ref_mjpeg)(const int32_t*, uint8_t*) = (void (*)(const int32_t*, uint8_t*))dlsym(lib, "oracle_mjpeg_inverse_dct");
    if (!ref || !ref_mjpeg) return 2;
    long n = atol(argv[2]);
    rng_state = strtoull(argv[3], 0, 10);
    long flagged = 0, bad = 0, count = 0;
    int32_t c[64];
    double fast[64];
    int16_t want[64];
    for (long b = 0;; ++b) {
        if (b < n) {
            const uint64_t kind = next() % 8;
            for (int i = 0; i < 64; ++i) c[i] = 0;
            if (mjpeg && kind < 2) {    /* dense: level x quantiser, up to and past the row limit */
                const int32_t lim = kind == 0 ? 6112 : 1 << 20;
                for (int i = 0; i < 64; ++i) c[i] = (int32_t)(next() % (2 * (uint64_t)lim + 1)) - lim;
            } else if (kind == 0) {     /* dense, up to the saturation bound */
                for (int i = 0; i < 64; ++i) c[i] = (int32_t)(next() % 4095) - 2047;
            } else if (kind == 1) {     /* the largest sum|c|: every AC saturated, DC int16 << 3 */
                for (int i = 1; i < 64; ++i) c[i] = (next() & 1) ? 2047 : -2048;
                c[0] = ((int32_t)(next() % 65536) - 32768) * 8;
            } else {                    /* sparse, small: where exact ties live */
                const int k = 1 + (int)(next() % 6);
                for (int j = 0; j < k; ++j) c[next() % 64] = (int32_t)(next() % 25) - 12;
                if (kind == 2) c[0] = ((int32_t)(next() % 512) - 256) * 8;
            }
        } else {
            int ok = 1;
            for (int i = 0; i < 64 && ok; ++i) ok = scanf("%d", &c[i]) == 1;
            if (!ok) break;
        }
        ++count;
        const int near = fast_idct(c, fast);
        if (mjpeg) mjpeg_ref(ref_mjpeg, c, want);
        else ref(c, want);
        if (near) { ++flagged; continue; }
        for (int i = 0; i < 64; ++i)
            if (round_clamp(fast[i]) != want[i]) ++bad;
    }
    printf("%ld %ld %ld\n", count, flagged, bad);
    return 0;
}
