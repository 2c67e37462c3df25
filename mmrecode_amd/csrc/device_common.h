// Device-side helpers shared by the reconstruction kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mmr_recon.h"

#define MMR_DEV __device__ __forceinline__

// Non-fused f64 arithmetic. The reference evaluates every product and sum with its own rounding
// (Rust never contracts a*b+c into an FMA); these builtins keep the compiler from fusing too.
MMR_DEV double dmul(double a, double b) { return __dmul_rn(a, b); }
MMR_DEV double dadd(double a, double b) { return __dadd_rn(a, b); }

// 8x8 IDCT basis: IDCT_MATRIX of crates/codecs/mjpeg/src/transform.rs:2-84, which equals MATRIX of
// crates/codecs/mpeg2/src/transform.rs:6-87. The literal values are needed for bit-exactness (half
// of them differ in the last bit from a recomputed cosine basis).
__constant__ static const double kIdctBasis[64] = {
    0.3535533905932738, 0.4903926402016152, 0.4619397662556434, 0.4157348061512726, 0.3535533905932738, 0.2777851165098011, 0.1913417161825449, 0.09754516100806417,
    0.3535533905932738, 0.4157348061512726, 0.1913417161825449, -0.0975451610080641, -0.3535533905932737, -0.4903926402016152, -0.4619397662556434, -0.2777851165098011,
    0.3535533905932738, 0.2777851165098011, -0.1913417161825448, -0.4903926402016152, -0.3535533905932738, 0.09754516100806415, 0.4619397662556433, 0.4157348061512727,
    0.3535533905932738, 0.09754516100806417, -0.4619397662556434, -0.2777851165098011, 0.3535533905932737, 0.4157348061512727, -0.191341716182545, -0.4903926402016152,
    0.3535533905932738, -0.0975451610080641, -0.4619397662556434, 0.2777851165098009, 0.3535533905932738, -0.4157348061512726, -0.1913417161825449, 0.4903926402016152,
    0.3535533905932738, -0.277785116509801, -0.1913417161825452, 0.4903926402016152, -0.3535533905932733, -0.097545161008064, 0.4619397662556437, -0.415734806151272,
    0.3535533905932738, -0.4157348061512727, 0.191341716182545, 0.097545161008064, -0.3535533905932736, 0.4903926402016153, -0.4619397662556435, 0.2777851165098008,
    0.3535533905932738, -0.4903926402016152, 0.4619397662556433, -0.415734806151272, 0.3535533905932732, -0.2777851165098008, 0.1913417161825431, -0.0975451610080625,
};

// Zig-zag scan position -> natural index (JPEG / MPEG-2 / DV 8-8 share it).
__constant__ static const uint8_t kZigzag[64] = {
    0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20,
    13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59,
    52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// MPEG-2 alternate scan (crates/codecs/mpeg2/src/tables.rs:16-20).
__constant__ static const uint8_t kAlternateScan[64] = {
    0, 8, 16, 24, 1, 9, 2, 10, 17, 25, 32, 40, 48, 56, 57, 49, 41, 33, 26, 18, 3, 11, 4, 12, 19,
    27, 34, 42, 50, 58, 35, 43, 51, 59, 20, 28, 5, 13, 6, 14, 21, 29, 36, 44, 52, 60, 37, 45, 53,
    61, 22, 30, 7, 15, 23, 31, 38, 46, 54, 62, 39, 47, 55, 63};

// Blocks handled by one 256-thread workgroup in the 8-lanes-per-block kernels.
constexpr int kBlocksPerGroup = 32;
constexpr int kGroupThreads = 256;

// Separable f64 IDCT of one 8x8 block, evaluated by the 8 lanes of a block group: lane x returns
// the column-x sums (before level shift / rounding) for rows y = 0..7.
//   row pass:    tmp[v][x] = sum_u M[x][u] * c[v][u]   (u ascending, from +0.0)
//   column pass: out[y][x] = sum_v M[y][v] * tmp[v][x] (v ascending, from +0.0)
// Rows of zero coefficients are skipped: their products are +-0.0 and adding a signed zero to an
// accumulator that started at +0.0 leaves it unchanged (the accumulator is never -0.0), so the
// result equals the dense evaluation bit for bit.
MMR_DEV void idct8x8_column(const int32_t* __restrict__ coef, int x, uint32_t row_mask, double out[8]) {
    double mx[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) mx[u] = kIdctBasis[x * 8 + u];
    double tmp[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        tmp[v] = 0.0;
        if (row_mask & (1u << v)) {
            const int4 lo = *reinterpret_cast<const int4*>(coef + v * 8);
            const int4 hi = *reinterpret_cast<const int4*>(coef + v * 8 + 4);
            const int32_t c[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
            double t = 0.0;
#pragma unroll
            for (int u = 0; u < 8; ++u) t = dadd(t, dmul(mx[u], (double)c[u]));
            tmp[v] = t;
        }
    }
#pragma unroll
    for (int y = 0; y < 8; ++y) {
        double s = 0.0;
#pragma unroll
        for (int v = 0; v < 8; ++v)
            if (row_mask & (1u << v)) s = dadd(s, dmul(kIdctBasis[y * 8 + v], tmp[v]));
        out[y] = s;
    }
}

MMR_DEV double round_half_away(double v) { return round(v); }

MMR_DEV uint8_t clamp_u8_d(double r) {
    r = r < 0.0 ? 0.0 : r;
    r = r > 255.0 ? 255.0 : r;
    return (uint8_t)r;
}

MMR_DEV int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// XOR / OR reductions across the 8 lanes of a block group (lanes 8g..8g+7 of the wave).
MMR_DEV uint32_t group8_xor(uint32_t v) {
    v ^= __shfl_xor(v, 1, 8);
    v ^= __shfl_xor(v, 2, 8);
    v ^= __shfl_xor(v, 4, 8);
    return v;
}
MMR_DEV uint32_t group8_or(uint32_t v) {
    v |= __shfl_xor(v, 1, 8);
    v |= __shfl_xor(v, 2, 8);
    v |= __shfl_xor(v, 4, 8);
    return v;
}

// Agent-scope acquire / release on work-queue counters (coherent across the XCDs' L2 caches).
MMR_DEV uint32_t load_acquire(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
MMR_DEV void add_release(uint32_t* p, uint32_t v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
MMR_DEV uint32_t fetch_add_relaxed(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
