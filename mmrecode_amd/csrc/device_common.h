// Device-side helpers shared by the reconstruction kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mmr_recon.h"

#define MMR_DEV __device__ __forceinline__

// Pointers read from job records are generic to the compiler, which then emits flat accesses
// (they wait on the LDS counter too). All buffers the kernels touch through them are device
// memory, so they are cast to the global address space at the point of use.
#define MMR_GLOBAL __attribute__((address_space(1)))
template <typename T>
MMR_DEV const MMR_GLOBAL T* as_global(const T* p) { return (const MMR_GLOBAL T*)p; }
template <typename T>
MMR_DEV MMR_GLOBAL T* as_global(T* p) { return (MMR_GLOBAL T*)p; }

// 16- and 8-byte global loads (clang vector types: HIP's int4 / uint2 cannot be copied out of an
// address-space-qualified lvalue).
typedef int32_t mmr_v4i32 __attribute__((ext_vector_type(4)));
typedef uint32_t mmr_v2u32 __attribute__((ext_vector_type(2)));
// 12 bytes at a 4-byte aligned address (the vector type's own alignment would be 16).
typedef uint32_t mmr_v3u32 __attribute__((ext_vector_type(3), aligned(4)));
typedef int16_t mmr_v2i16 __attribute__((ext_vector_type(2)));
MMR_DEV int4 load_global_int4(const void* p) {
    const mmr_v4i32 v = *reinterpret_cast<const MMR_GLOBAL mmr_v4i32*>(as_global(p));
    return make_int4(v.x, v.y, v.z, v.w);
}
MMR_DEV uint2 load_global_uint2(const void* p) {
    const mmr_v2u32 v = *reinterpret_cast<const MMR_GLOBAL mmr_v2u32*>(as_global(p));
    return make_uint2(v.x, v.y);
}

// Element i (0..2, per lane) of a three-pointer array that is uniform across the workgroup, as a
// select between scalar values (a plain p[i] is a per-lane load from the job record).
template <typename T>
MMR_DEV T* pick3(T* const* p, int i) {
    uint64_t v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint64_t x = reinterpret_cast<uint64_t>(p[k]);
        v[k] = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(x >> 32)) << 32) |
               (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)x);
    }
    return reinterpret_cast<T*>(i == 0 ? v[0] : (i == 1 ? v[1] : v[2]));
}

// A pointer that is the same in every lane, in scalar registers.
template <typename T>
MMR_DEV T* uniform_ptr(T* p) {
    const uint64_t x = reinterpret_cast<uint64_t>(p);
    return reinterpret_cast<T*>(((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(x >> 32)) << 32) |
                                (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)x));
}

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

// The 8-lanes-per-block kernels (MJPEG, MPEG-2, DV) take BPG = 16, 32 or 64 blocks per workgroup
// (a template parameter; the launch shape is a tuning choice). Work is split into tiles of BPG
// blocks that never straddle two frames; the host uploads a tile -> job table with the jobs, so a
// workgroup finds its (uniform) job with one scalar load.

// Workgroups are dispatched to the eight XCDs of an MI355X round robin (workgroup i runs on XCD
// i % 8), and every XCD has its own L2. xcd_tile remaps the workgroup index so that each XCD walks a
// contiguous range of `per` tiles: neighbouring tiles share reference and output cache lines, which
// then meet in one L2 instead of being fetched and partially written back by several. A launch of
// kXcds * per workgroups covers `tiles` tiles when per = ceil(tiles / kXcds); indices past the end
// are idle.
constexpr uint32_t kXcds = 8;
MMR_DEV uint32_t xcd_tile(uint32_t block, uint32_t per) { return (block % kXcds) * per + block / kXcds; }
inline uint32_t xcd_per(uint32_t tiles) { return (tiles + kXcds - 1) / kXcds; }

// Store 8 bytes (little-endian in v) at p: one 8-byte store when aligned, else two 4-byte stores
// (p must be 4-byte aligned), else bytes.
MMR_DEV void store_row8(uint8_t* p, uint64_t v) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    MMR_GLOBAL uint8_t* q = as_global(p);
    if ((a & 7) == 0) {
        *reinterpret_cast<MMR_GLOBAL uint64_t*>(q) = v;
    } else if ((a & 3) == 0) {
        reinterpret_cast<MMR_GLOBAL uint32_t*>(q)[0] = (uint32_t)v;
        reinterpret_cast<MMR_GLOBAL uint32_t*>(q)[1] = (uint32_t)(v >> 32);
    } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) q[i] = (uint8_t)(v >> (8 * i));
    }
}

// Orders LDS accesses between the lanes of one wave (the 8 lanes of a block group always share a
// wave): a compiler-level fence around a wave barrier; the hardware executes a wave's LDS
// operations in order, so no wait instruction is needed.
MMR_DEV void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Separable f64 IDCT of one 8x8 block by the 8 lanes of its group, in two passes with an LDS
// transpose between them (kIdctTmpStride doubles of LDS per group; the 64-byte pad after the 8x8
// doubles puts consecutive groups 16 banks apart):
//   row pass    (lane v, natural row v):  tmp[v][x] = sum_u M[x][u] * c[v][u]   (u ascending)
//   column pass (lane x):                 out[y][x] = sum_v M[y][v] * tmp[v][x] (v ascending)
// Every product and every sum is rounded on its own (dmul / dadd), in the reference's order. M is
// uniform in both passes (scalar loads). The reference starts each sum from +0.0; starting from
// the first product instead can only turn a +0.0 result into -0.0 when every term is zero, and
// such a value leaves the later sums unchanged and rounds to the integer 0, so the integer
// outputs are identical. Natural rows whose coefficients are all zero in every block of the wave
// are skipped in the column pass (their tmp values are zero, so their terms are signed zeros); the
// skip mask is wave-uniform, so the skip is a scalar branch.
constexpr int kIdctTmpStride = 72;

MMR_DEV void idct8x8_row_pass(const int32_t c[8], double* __restrict__ tmp_row) {
    double cd[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) cd[u] = (double)c[u];
    double t[8];
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        double s = dmul(kIdctBasis[x * 8], cd[0]);
#pragma unroll
        for (int u = 1; u < 8; ++u) s = dadd(s, dmul(kIdctBasis[x * 8 + u], cd[u]));
        t[x] = s;
    }
#pragma unroll
    for (int x = 0; x < 8; x += 2) *reinterpret_cast<double2*>(tmp_row + x) = make_double2(t[x], t[x + 1]);
}

// Bit v set when natural row v is nonzero in any block of the wave; `row_nonzero` is the calling
// lane's own row (v = lane & 7). Row 0 is always included (it starts every column sum).
MMR_DEV uint32_t wave_nonzero_rows(bool row_nonzero) {
    uint64_t b = __ballot(row_nonzero);
    b |= b >> 32;
    b |= b >> 16;
    b |= b >> 8;
    return ((uint32_t)b & 0xffu) | 1u;
}

MMR_DEV void idct8x8_column_pass(const double* __restrict__ tmp, int x, uint32_t rows, double out[8]) {
    const double c0 = tmp[x];
#pragma unroll
    for (int y = 0; y < 8; ++y) out[y] = dmul(kIdctBasis[y * 8], c0);
    // Remaining rows in ascending order; `rows` is wave-uniform, so this is a scalar loop.
    uint32_t rest = rows & ~1u;
    while (rest) {
        const int v = __builtin_ctz(rest);
        rest &= rest - 1;
        const double cv = tmp[v * 8 + x];
#pragma unroll
        for (int y = 0; y < 8; ++y) out[y] = dadd(out[y], dmul(kIdctBasis[y * 8 + v], cv));
    }
}

// Fast form of the same two passes: the same sums over the same basis, each term fused into its
// sum (one rounding per step instead of two), in half the instructions. It is not the reference's
// arithmetic, so its results are used only where they provably round like the reference's:
// both evaluations are within (2 g8 + g8^2) A of the exact sum S (g8 = 8u / (1 - 8u), u = 2^-53;
// A = sum_v |M[y][v]| sum_u |M[x][u]| |c[v][u]| <= max|M|^2 sum|c|, max|M| = 0.49039264020161522),
// so they differ by at most 32.0001 u A. For an MPEG-2 block (63 coefficients saturated to
// [-2048, 2047], the intra DC an int16 shifted left by at most 3) sum|c| <= 391168, so the
// difference is below 3.4e-10, and a fast value farther than kIdctTieWindow = 2^-30 (9.3e-10) from
// every half-integer rounds (half away from zero) exactly as the reference's does. A fast value
// inside the window is reported (idct_near_half) and the caller recomputes in reference order.
constexpr double kIdctTieWindow = 0x1p-30;

// Blocks whose coefficients are not saturated (MJPEG: level x quantiser) keep within that sum|c|
// when every row's sum of |c| is at most kIdctFastRowAbsSum; a row past it sends the block to the
// reference order. (MJPEG's + 128 rounds both evaluations once more, by at most 2^-45 each.)
constexpr uint32_t kIdctFastRowAbsSum = 391168 / 8;

MMR_DEV bool idct_row_within_fast_bound(const int32_t c[8]) {
    uint32_t s = 0;
    bool ok = true;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const uint32_t a = (uint32_t)(c[u] < 0 ? -(int64_t)c[u] : c[u]);
        ok &= a <= kIdctFastRowAbsSum;
        s += ok ? a : 0u;
    }
    return ok && s <= kIdctFastRowAbsSum;
}

MMR_DEV void idct8x8_row_pass_fast(const int32_t c[8], double* __restrict__ tmp_row) {
    double cd[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) cd[u] = (double)c[u];
    double t[8];
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        double s = dmul(kIdctBasis[x * 8], cd[0]);
#pragma unroll
        for (int u = 1; u < 8; ++u) s = __builtin_fma(kIdctBasis[x * 8 + u], cd[u], s);
        t[x] = s;
    }
#pragma unroll
    for (int x = 0; x < 8; x += 2) *reinterpret_cast<double2*>(tmp_row + x) = make_double2(t[x], t[x + 1]);
}

// Every row, zero or not (a zero row adds exact zeros): fully unrolled, the basis addresses are
// static, so its scalar loads are batched and issued ahead instead of eight waits per row.
MMR_DEV void idct8x8_column_pass_fast(const double* __restrict__ tmp, int x, double out[8]) {
    double t[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) t[v] = tmp[v * 8 + x];
#pragma unroll
    for (int y = 0; y < 8; ++y) {
        double s = dmul(kIdctBasis[y * 8], t[0]);
#pragma unroll
        for (int v = 1; v < 8; ++v) s = __builtin_fma(kIdctBasis[y * 8 + v], t[v], s);
        out[y] = s;
    }
}

// True when one of the 8 values lies within kIdctTieWindow of a half-integer: |fract(|v|) - 1/2|
// (fract and the subtraction are exact near 1/2).
MMR_DEV bool idct_near_half(const double v[8]) {
    bool near = false;
#pragma unroll
    for (int i = 0; i < 8; ++i) near |= __builtin_fabs(__builtin_amdgcn_fract(__builtin_fabs(v[i])) - 0.5) <= kIdctTieWindow;
    return near;
}

// round() (half away from zero) of a value with |v| < 2^31, as an int: v + copysign(h, v) with
// h the largest double below 0.5, truncated. For |v| < 2^52 the sum rounds to at least n + 1 exactly
// when |v| >= n + 0.5 and stays below n + 1 otherwise (h is 2^-54 short of 0.5, less than half a
// unit in the last place of any sum that could cross n + 1), so the truncation is the rounded value.
MMR_DEV int round_half_away_i32(double v) {
    const double h = __builtin_copysign(0.49999999999999994, v);
    return (int)dadd(v, h);
}

MMR_DEV double round_half_away(double v) { return round(v); }

MMR_DEV uint8_t clamp_u8_d(double r) {
    r = r < 0.0 ? 0.0 : r;
    r = r > 255.0 ? 255.0 : r;
    return (uint8_t)r;
}

MMR_DEV int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// XOR / OR reductions across the 8 lanes of a block group (lanes 8g..8g+7 of the wave), with DPP
// lane moves: swap neighbours, swap pairs (now every lane of a quad holds the quad's value), then
// the half-row mirror pairs each lane with one of the other quad. All 8 lanes of a group must be
// active (the callers' conditions are uniform per block).
MMR_DEV uint32_t dpp_quad_swap1(uint32_t v) { return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xf, 0xf, false); }
MMR_DEV uint32_t dpp_quad_swap2(uint32_t v) { return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xf, 0xf, false); }
MMR_DEV uint32_t dpp_half_mirror(uint32_t v) { return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xf, 0xf, false); }
MMR_DEV uint32_t group8_xor(uint32_t v) {
    v ^= dpp_quad_swap1(v);
    v ^= dpp_quad_swap2(v);
    return v ^ dpp_half_mirror(v);
}
MMR_DEV uint32_t group8_or(uint32_t v) {
    v |= dpp_quad_swap1(v);
    v |= dpp_quad_swap2(v);
    return v | dpp_half_mirror(v);
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

// Data handed between workgroups of one kernel (possibly on different XCDs) without cache
// maintenance. Agent-scope relaxed atomic stores and loads are coherent across the XCDs' L2
// caches (the sc1 forms of the memory instructions): a producer that writes the data with
// agent_store and then publishes a counter with publish_agent (its earlier memory operations
// complete before the counter store is issued), and a consumer that polls the counter with
// agent-scope loads and only then reads the data with agent_load (behind a compiler barrier), see
// the data without the write-back of the XCD's whole L2 that an agent-scope release costs, or the
// invalidate of an agent-scope acquire.
template <typename T>
MMR_DEV void agent_store(MMR_GLOBAL T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
MMR_DEV T agent_load(const MMR_GLOBAL T* p) {
    return __hip_atomic_load(const_cast<MMR_GLOBAL T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MMR_DEV void publish_agent(uint32_t* counter, uint32_t v) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __hip_atomic_store(counter, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A counter polled by a whole wave, read once per lane and made wave-uniform: the polling loop is
// then a scalar loop (a per-lane compare would make it a divergent loop, with exec-mask
// bookkeeping on every iteration that takes issue slots from the other waves of the SIMD).
MMR_DEV uint32_t poll_workgroup(uint32_t* counter) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
}
MMR_DEV uint32_t poll_agent(uint32_t* counter) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
