// DV25 frame reconstruction: dequantisation, 8-8 / 2-4-8 direct-form f64 IDCTs, shuffled placement.
//
// Reference: consume_ac / inverse_factor_for_encoding (crates/codecs/dv/src/decoder.rs:262-310),
// inverse_88 / inverse_248 (:520-602), place_macroblock / put_block / put_split_chroma (:432-505),
// macroblock_coordinates (:395-422), tables (crates/codecs/dv/src/tables.rs:6-50).
//
// The cosine tables are computed on the host with the C library cos on exactly the reference's
// argument expression ((2x+1) as f64 * u as f64 * PI / 16.0) and uploaded: the device never
// evaluates cos itself, so the basis is the same double values the reference multiplies by.
#include <math.h>

#include "device_common.h"
#include "recon_internal.h"

namespace mmr {

__constant__ double kDvCos16[64];   // [x*8+u] = cos((2x+1) u pi / 16)
__constant__ double kDvCos8[32];    // [y*4+v] = cos((2y+1) v pi / 8)
__constant__ double kDvCuCv[64];    // [v*8+u] = cu * cv

__constant__ static const uint8_t kZigzag248[64] = {
    0, 8, 1, 9, 16, 24, 2, 10, 17, 25, 32, 40, 48, 56, 33, 41, 18, 26, 3, 11, 4, 12, 19, 27, 34,
    42, 49, 57, 50, 58, 35, 43, 20, 28, 5, 13, 6, 14, 21, 29, 36, 44, 51, 59, 52, 60, 37, 45, 22,
    30, 7, 15, 23, 31, 38, 46, 53, 61, 54, 62, 39, 47, 55, 63};
__constant__ static const uint8_t kQuantShifts[22][4] = {
    {3, 3, 4, 4}, {3, 3, 4, 4}, {2, 3, 3, 4}, {2, 3, 3, 4}, {2, 2, 3, 3}, {2, 2, 3, 3},
    {1, 2, 2, 3}, {1, 2, 2, 3}, {1, 1, 2, 2}, {1, 1, 2, 2}, {0, 1, 1, 2}, {0, 1, 1, 2},
    {0, 0, 1, 1}, {0, 0, 1, 1}, {0, 0, 0, 1}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0},
    {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
__constant__ static const uint8_t kQuantOffset[4] = {6, 3, 0, 1};
__constant__ static const int32_t kIWeight88[64] = {
    32768, 16705, 16705, 17734, 17032, 17734, 18205, 18081, 18081, 18205, 18725, 18562, 19195,
    18562, 18725, 19266, 19091, 19705, 19705, 19091, 19266, 21407, 19643, 20267, 20228, 20267,
    19643, 21407, 22725, 21826, 20853, 20806, 20806, 20853, 21826, 22725, 23170, 23170, 21407,
    21400, 21407, 23170, 23170, 24598, 23786, 22018, 22018, 23786, 24598, 25251, 24465, 22654,
    24465, 25251, 25972, 25172, 25172, 25972, 26722, 27969, 26722, 29692, 29692, 31521};
__constant__ static const int32_t kIWeight248[64] = {
    32768, 16384, 16705, 16705, 17734, 17734, 17734, 17734, 18081, 18081, 18725, 18725, 21407,
    21407, 19091, 19091, 19195, 19195, 18205, 18205, 18725, 18725, 19705, 19705, 20267, 20267,
    21826, 21826, 23170, 23170, 20806, 20806, 20267, 20267, 19266, 19266, 21407, 21407, 20853,
    20853, 21400, 21400, 23786, 23786, 24465, 24465, 22018, 22018, 23170, 23170, 22725, 22725,
    24598, 24598, 24465, 24465, 25172, 25172, 27969, 27969, 25972, 25972, 29692, 29692};

constexpr double kFrac1Sqrt2 = 0.70710678118654752440;

// macroblock_coordinates (decoder.rs:395-422): top-left of macroblock `mb` of segment (seq, slot),
// in units of 8 samples.
MMR_DEV void dv_mb_coordinates(uint32_t system, int seq, int slot, int mb, int* ox, int* oy) {
    const int OFF[5] = {2, 6, 8, 0, 4};
    const int SHUF3[5] = {18, 9, 27, 0, 36};
    const int LSTART[5] = {9, 4, 13, 0, 18};
    if (system == MMR_DV_SYSTEM_625_50) {
        const int serpent1 = (slot % 6 < 3) ? (slot % 6) : (5 - slot % 6);
        const int x = SHUF3[mb] + slot / 3;
        const int y = serpent1 + ((seq + OFF[mb]) % 12) * 3;
        *ox = x * 2;
        *oy = y * 2;
    } else {
        const int row = (seq + OFF[mb]) % 10;
        const int k = slot + ((mb == 1 || mb == 2) ? 3 : 0);
        const int x = LSTART[mb] + k / 6;
        const int serpent2 = (k % 12 < 6) ? (k % 12) : (11 - k % 12);
        int y = serpent2 + row * 6;
        if (x > 21) y = y * 2 - row * 6;
        *ox = x * 4;
        *oy = y;
    }
}

// Fused separable form of both transforms, used where it provably rounds like the reference.
// inverse_88 is sum_v sum_u cu cv c[v][u] C16[x][u] C16[y][v] / 4; inverse_248, with its row-pair
// butterfly expanded, is sum_r V[2y+f][r] (sum_u cu c[r][u] C16[x][u]) / 4 with V[2y+f][2v] =
// cv C8[y][v] and V[2y+f][2v+1] = (f ? -1 : 1) cv C8[y][v]. So both are one row pass with
// kDvRowA[x][u] = cu C16[x][u] (lane v: tmp[v][x]) and one column pass (lane x) with the mode's
// vertical table, every step fused (tables rounded once on the host). Error: the reference's
// evaluation is within gamma_66 A of the exact sum S (64 sequential sums of three-factor
// products; inverse_248 less), the fused one within (2 gamma_8 + gamma_8^2 + 3u) A of S (tables
// and two 8-term fused chains), A <= sum |c| <= 64 * 132092 (|c| < 2^17: the dequantised values
// are int32 >> 14, the DC an int16 * 4 + 1024); divided by 4 both lie within 2.1e-8 of S / 4. A
// fused value farther than kDvTieWindow = 2^-25 (3.0e-8) from every half-integer therefore
// rounds (half away from zero) like the reference's; a block with a value inside the window is
// recomputed in the reference's order.
__constant__ double kDvRowA[64];     // [x*8+u] = cu * C16[x][u]
__constant__ double kDvCol88[64];    // [y*8+v] = cv * C16[y][v]
__constant__ double kDvCol248[64];   // [(2y+f)*8 + 2v+s] = (s && f ? -1 : 1) * cv * C8[y][v]
constexpr double kDvTieWindow = 0x1p-25;

MMR_DEV void dv_row_pass(const int32_t c[8], double* __restrict__ tmp_row) {
    double cd[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) cd[u] = (double)c[u];
    double t[8];
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        double s = dmul(kDvRowA[x * 8], cd[0]);
#pragma unroll
        for (int u = 1; u < 8; ++u) s = __builtin_fma(kDvRowA[x * 8 + u], cd[u], s);
        t[x] = s;
    }
#pragma unroll
    for (int x = 0; x < 8; x += 2) *reinterpret_cast<double2*>(tmp_row + x) = make_double2(t[x], t[x + 1]);
}

// Column pass of lane x with vertical table B, over every row (a zero row adds exact zeros), fully
// unrolled so that the table's scalar loads have static addresses and are issued ahead.
MMR_DEV void dv_column_pass(const double* __restrict__ tmp, int x, const double* __restrict__ B, double out[8]) {
    double t[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) t[v] = tmp[v * 8 + x];
#pragma unroll
    for (int y = 0; y < 8; ++y) {
        double s = dmul(B[y * 8], t[0]);
#pragma unroll
        for (int v = 1; v < 8; ++v) s = __builtin_fma(B[y * 8 + v], t[v], s);
        out[y] = s;
    }
}

// Index of the k-th set bit (k from 0) of m; m has more than k set bits.
MMR_DEV uint32_t kth_set_bit(uint64_t m, uint32_t k) {
    uint32_t base = 0;
#pragma unroll
    for (uint32_t w = 32; w >= 1; w >>= 1) {
        const uint64_t lo = m & ((1ull << w) - 1);
        const uint32_t c = (uint32_t)__popcll(lo);
        if (k >= c) {
            k -= c;
            m >>= w;
            base += w;
        } else {
            m = lo;
        }
    }
    return base;
}

// inverse_88 in the reference's order (decoder.rs:520-548) for lane x of a group: sum over
// (v, u) in ascending natural order of cu*cv*c*cos((2x+1)u pi/16)*cos((2y+1)v pi/16), then
// (sum / 4).round().clamp(0, 255). Zero coefficients contribute signed zeros and are skipped.
MMR_DEV void dv_inverse_88_reference(const int32_t* __restrict__ cf, const double* __restrict__ cos16,
                                     const double* __restrict__ cucv, int x, int lane, uint8_t out[8]) {
    uint32_t rowbits = 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) rowbits |= (uint32_t)(cf[lane * 8 + u] != 0) << u;
    const uint32_t lo = group8_or(lane < 4 ? rowbits << (lane * 8) : 0u);
    const uint32_t hi = group8_or(lane >= 4 ? rowbits << ((lane - 4) * 8) : 0u);
    double sum[8];
#pragma unroll
    for (int y = 0; y < 8; ++y) sum[y] = 0.0;
    uint64_t mask = ((uint64_t)hi << 32) | lo;
    while (mask) {
        const int p = __builtin_ctzll(mask);
        mask &= mask - 1;
        const int v = p >> 3, u = p & 7;
        double t = dmul(cucv[p], (double)cf[p]);
        t = dmul(t, cos16[x * 8 + u]);
#pragma unroll
        for (int y = 0; y < 8; ++y) sum[y] = dadd(sum[y], dmul(t, cos16[y * 8 + v]));
    }
#pragma unroll
    for (int y = 0; y < 8; ++y) out[y] = clamp_u8_d(round_half_away(sum[y] / 4.0));
}

// inverse_248 in the reference's order (decoder.rs:555-602) for lane x: butterfly of row pairs,
// horizontal 8-point pass (/2), per-field vertical 4-point pass (/2), round, clamp.
MMR_DEV void dv_inverse_248_reference(const int32_t* __restrict__ cf, const double* __restrict__ cos16, int x, uint8_t out[8]) {
    double h[8];
#pragma unroll
    for (int row = 0; row < 8; ++row) {
        const int pair = row >> 1, second = row & 1;
        double s = 0.0;
        for (int u = 0; u < 8; ++u) {
            const double first = (double)cf[pair * 16 + u];
            const double other = (double)cf[pair * 16 + 8 + u];
            const double bfly = second ? first - other : first + other;
            if (bfly != 0.0) {
                const double cu = u == 0 ? kFrac1Sqrt2 : 1.0;
                s = dadd(s, dmul(dmul(cu, bfly), cos16[x * 8 + u]));
            }
        }
        h[row] = s / 2.0;
    }
#pragma unroll
    for (int field = 0; field < 2; ++field)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            double s = 0.0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double cv = v == 0 ? kFrac1Sqrt2 : 1.0;
                s = dadd(s, dmul(dmul(cv, h[v * 2 + field]), kDvCos8[y * 4 + v]));
            }
            out[y * 2 + field] = clamp_u8_d(round_half_away(s / 2.0));
        }
}

// BPG 8x8 blocks (8 lanes each) per workgroup, BPG <= 64. The tile's blocks are taken in
// mode-sorted order (8-8 blocks first, then 2-4-8 blocks), so a wave's blocks share their
// vertical table and only the wave on the boundary between the two runs both column passes.
template <int BPG>
__global__ __launch_bounds__(BPG * 8) void dv_kernel(const DvJob* __restrict__ jobs,
                                                     const uint32_t* __restrict__ tile_job, uint32_t tile_base,
                                                     uint32_t tiles, uint32_t per_xcd) {
    static_assert(BPG <= 64, "one ballot covers the tile's blocks");
    // Per block group: natural-order coefficients, then the row-pass output, then the samples.
    __shared__ __align__(16) double s_area[BPG * kIdctTmpStride];
    __shared__ double s_cos16[64];
    __shared__ double s_cucv[64];
    const int g = threadIdx.x >> 3, lane = threadIdx.x & 7;
    const uint32_t local = xcd_tile(blockIdx.x, per_xcd);
    if (local >= tiles) return;   // uniform over the workgroup
    if (threadIdx.x < 64) {
        s_cos16[threadIdx.x] = kDvCos16[threadIdx.x];
        s_cucv[threadIdx.x] = kDvCuCv[threadIdx.x];
    }
    const uint32_t tile = tile_base + local;
    const DvJob& job = jobs[tile_job[tile]];
    const uint32_t first = (tile - job.tile_begin) * BPG;

    // Mode-sorted order of the tile's blocks: every wave reads the tile's BPG mode bits (lane i:
    // block first + i) and group g takes the g-th block of the order.
    uint32_t k = (uint32_t)g;
    uint64_t m88, m248;
    {
        const uint32_t i = threadIdx.x & 63;
        const bool in = i < (uint32_t)BPG && first + i < job.block_count;
        const uint8_t inf = in ? as_global(job.info)[first + i] : (uint8_t)0;
        m248 = __ballot(in && (inf & 1));
        m88 = __ballot(in && !(inf & 1));
    }
    const uint32_t n88 = (uint32_t)__popcll(m88), n248 = (uint32_t)__popcll(m248);
    const bool active = k < n88 + n248;
    const uint32_t blk = first + (k < n88 ? kth_set_bit(m88, k) : (active ? kth_set_bit(m248, k - n88) : 0u));
    const uint8_t info = active ? as_global(job.info)[blk] : 0;
    const int mode248 = info & 1, cls = (info >> 1) & 3, qno = (info >> 3) & 15;
    int32_t* coef = reinterpret_cast<int32_t*>(s_area + g * kIdctTmpStride);
    double* tmp = s_area + g * kIdctTmpStride;

    if (active) {
        const int4 raw = load_global_int4(job.levels + (size_t)blk * 64 + lane * 8);
        const int16_t* l8 = reinterpret_cast<const int16_t*>(&raw);
        const uint8_t* scan = mode248 ? kZigzag248 : kZigzag;
        const int32_t* weight = mode248 ? kIWeight248 : kIWeight88;
        const int q = qno + kQuantOffset[cls];
        const int class_shift = cls == 3;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int p = lane * 8 + i;
            const int32_t level = l8[i];
            int32_t v;
            if (p == 0) {
                v = level * 4 + 1024;  // dc * 4 + 1024 (decoder.rs:214)
            } else if (level == 0) {
                v = 0;
            } else {
                const int area = p <= 5 ? 0 : (p <= 20 ? 1 : (p <= 42 ? 2 : 3));
                const int32_t factor = weight[p] << (kQuantShifts[q][area] + 1 + class_shift);
                v = (level * factor + 8192) >> 14;
            }
            coef[scan[p]] = v;
        }
    }
    __syncthreads();   // s_cos16 / s_cucv for the reference-order fallback; every wave passes

    // Row v = lane of the block, in registers.
    int32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (active) {
        const int4 lo = *reinterpret_cast<const int4*>(coef + lane * 8);
        const int4 hi = *reinterpret_cast<const int4*>(coef + lane * 8 + 4);
        c[0] = lo.x; c[1] = lo.y; c[2] = lo.z; c[3] = lo.w; c[4] = hi.x; c[5] = hi.y; c[6] = hi.z; c[7] = hi.w;
    }
    wave_lds_sync();   // every row is in registers before the row pass overwrites the area
    if (active) dv_row_pass(c, tmp + lane * 8);
    wave_lds_sync();
    double col[8];
#pragma unroll
    for (int y = 0; y < 8; ++y) col[y] = 0.0;
    const int x = lane;
    if (__ballot(active && !mode248) != 0) {   // wave-uniform
        if (active && !mode248) dv_column_pass(tmp, x, kDvCol88, col);
    }
    if (__ballot(active && mode248) != 0) {
        if (active && mode248) dv_column_pass(tmp, x, kDvCol248, col);
    }
    uint8_t out[8];
    bool near = false;
#pragma unroll
    for (int y = 0; y < 8; ++y) {
        const double v = col[y] * 0.25;   // exact
        near |= __builtin_fabs(__builtin_amdgcn_fract(__builtin_fabs(v)) - 0.5) <= kDvTieWindow;
        out[y] = (uint8_t)clampi(round_half_away_i32(v), 0, 255);
    }
    const bool redo = active && group8_or(near ? 1u : 0u) != 0;
    if (__ballot(redo) != 0) {   // wave-uniform: some block falls back to the reference's order
        wave_lds_sync();   // the column reads are done before the coefficients come back
        if (redo) {
            *reinterpret_cast<int4*>(coef + lane * 8) = make_int4(c[0], c[1], c[2], c[3]);
            *reinterpret_cast<int4*>(coef + lane * 8 + 4) = make_int4(c[4], c[5], c[6], c[7]);
        }
        wave_lds_sync();
        if (redo) {
            if (mode248) dv_inverse_248_reference(coef, s_cos16, x, out);
            else dv_inverse_88_reference(coef, s_cos16, s_cucv, x, lane, out);
        }
    }

    // Column x of the block -> LDS, then lane r stores row r (place_macroblock / put_split_chroma).
    wave_lds_sync();   // the area's last reads are done
    uint8_t* pix = reinterpret_cast<uint8_t*>(tmp);
#pragma unroll
    for (int r = 0; r < 8; ++r) pix[r * 8 + x] = out[r];
    wave_lds_sync();
    if (!active) return;
    const uint64_t row = *reinterpret_cast<const uint64_t*>(pix + lane * 8);
    const uint32_t mb_index = blk / 6;
    const int b = (int)(blk % 6);
    const int seq = (int)(mb_index / 135), slot = (int)((mb_index % 135) / 5), mbs = (int)(mb_index % 5);
    int mbx, mby;
    dv_mb_coordinates(job.system, seq, slot, mbs, &mbx, &mby);
    const int X = mbx * 8, Y = mby * 8;
    const int width = 720;
    const int r = lane;
    if (job.system == MMR_DV_SYSTEM_625_50) {
        if (b < 4) {
            store_row8(job.planes[0] + (size_t)(Y + (b >> 1) * 8 + r) * width + X + (b & 1) * 8, row);
        } else {
            store_row8(job.planes[b == 4 ? 2 : 1] + (size_t)(Y / 2 + r) * 360 + X / 2, row);
        }
    } else if (mbx < 88) {
        if (b < 4) {
            store_row8(job.planes[0] + (size_t)(Y + r) * width + X + b * 8, row);
        } else {
            store_row8(job.planes[b == 4 ? 2 : 1] + (size_t)(Y + r) * 180 + X / 4, row);
        }
    } else {
        if (b < 4) {
            store_row8(job.planes[0] + (size_t)(Y + (b >> 1) * 8 + r) * width + X + (b & 1) * 8, row);
        } else {
            // Split chroma: columns 0..3 go to rows y..y+7, columns 4..7 to rows y+8..y+15.
            MMR_GLOBAL uint8_t* p = as_global(job.planes[b == 4 ? 2 : 1]) + (size_t)(Y + r) * 180 + X / 4;
            *reinterpret_cast<MMR_GLOBAL uint32_t*>(p) = (uint32_t)row;
            *reinterpret_cast<MMR_GLOBAL uint32_t*>(p + 8 * 180) = (uint32_t)(row >> 32);
        }
    }
}

hipError_t dv_upload_tables(hipStream_t s) {
    double cos16[64], cos8[32], cucv[64], rowa[64], col88[64], col248[64];
    for (int x = 0; x < 8; ++x)
        for (int u = 0; u < 8; ++u) cos16[x * 8 + u] = cos((double)(2 * x + 1) * (double)u * M_PI / 16.0);
    for (int y = 0; y < 8; ++y)
        for (int v = 0; v < 4; ++v) cos8[y * 4 + v] = cos((double)(2 * y + 1) * (double)v * M_PI / 8.0);
    for (int v = 0; v < 8; ++v)
        for (int u = 0; u < 8; ++u) {
            const volatile double cu = u == 0 ? M_SQRT1_2 : 1.0, cv = v == 0 ? M_SQRT1_2 : 1.0;
            cucv[v * 8 + u] = cu * cv;
        }
    // Fused-form tables: each entry one rounded product of the reference's factors.
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) {
            const volatile double cj = j == 0 ? M_SQRT1_2 : 1.0;
            rowa[i * 8 + j] = cj * cos16[i * 8 + j];
            col88[i * 8 + j] = cj * cos16[i * 8 + j];
        }
    for (int y = 0; y < 4; ++y)
        for (int f = 0; f < 2; ++f)
            for (int v = 0; v < 4; ++v) {
                const volatile double cv = v == 0 ? M_SQRT1_2 : 1.0;
                const double e = cv * cos8[y * 4 + v];
                col248[(2 * y + f) * 8 + 2 * v] = e;
                col248[(2 * y + f) * 8 + 2 * v + 1] = f ? -e : e;
            }
    hipError_t e = hipMemcpyToSymbolAsync(HIP_SYMBOL(kDvCos16), cos16, sizeof cos16, 0, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyToSymbolAsync(HIP_SYMBOL(kDvCos8), cos8, sizeof cos8, 0, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyToSymbolAsync(HIP_SYMBOL(kDvCuCv), cucv, sizeof cucv, 0, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyToSymbolAsync(HIP_SYMBOL(kDvRowA), rowa, sizeof rowa, 0, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyToSymbolAsync(HIP_SYMBOL(kDvCol88), col88, sizeof col88, 0, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyToSymbolAsync(HIP_SYMBOL(kDvCol248), col248, sizeof col248, 0, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

hipError_t launch_dv(const DvJob* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, uint32_t group_blocks,
                     hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    const uint32_t per = xcd_per(tiles);
    switch (group_blocks) {
    case 16: hipLaunchKernelGGL(dv_kernel<16>, dim3(per * kXcds), dim3(16 * 8), 0, s, d_jobs, d_tile_job, tile_base, tiles, per); break;
    case 32: hipLaunchKernelGGL(dv_kernel<32>, dim3(per * kXcds), dim3(32 * 8), 0, s, d_jobs, d_tile_job, tile_base, tiles, per); break;
    case 64: hipLaunchKernelGGL(dv_kernel<64>, dim3(per * kXcds), dim3(64 * 8), 0, s, d_jobs, d_tile_job, tile_base, tiles, per); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mmr
